"""Pins of the multi-determinant oracle (oracle/multidet.py), CPU only.

Mirrors the reference's multi-det tests (test_afqmc_multidet.py): overlap and
energy against an exact evaluation (Cauchy-Binet CI vectors and the
brute-force second-quantised Hamiltonian instead of PySCF FCI), zero
variance for the exact trial, and reduction to the single-determinant path.
"""

import itertools

import numpy as np
import pytest

from afq_cases import walkers_near_trial
from fci_brute import det_vector, hamiltonian
from oracle import afqmc as O
from oracle import multidet as MD
from omegaqmc_amd.synth import make_system

M, NA, NB, G = 6, 2, 2, 10


def _sys(seed=3):
    return make_system(M, NA, NB, G, seed=seed)


def _dets(ndet, rng):
    """The aufbau det first, then random distinct occupation pairs."""
    ups = list(itertools.combinations(range(M), NA))
    dns = list(itertools.combinations(range(M), NB))
    pairs = [(tuple(range(NA)), tuple(range(NB)))]
    while len(pairs) < ndet:
        p = (ups[rng.integers(len(ups))], dns[rng.integers(len(dns))])
        if p not in pairs:
            pairs.append(p)
    occ_up = np.array([p[0] for p in pairs])
    occ_dn = np.array([p[1] for p in pairs])
    return occ_up, occ_dn


def _walkers(W, rng, scale=0.5):
    pa = walkers_near_trial(np.eye(M, NA).astype(complex), W, rng, scale)
    pb = walkers_near_trial(np.eye(M, NB).astype(complex), W, rng, scale)
    return pa, pb


def _trial_vector(occ_up, occ_dn, ci):
    """sum_I c_I |D_I> in the kron(alpha, beta) string basis of fci_brute."""
    tu = MD.trials_from_occ(M, occ_up)
    td = MD.trials_from_occ(M, occ_dn)
    return sum(c * np.kron(det_vector(a), det_vector(b)) for c, a, b in zip(ci, tu, td))


@pytest.mark.parametrize("seed", [0, 1])
def test_overlap_equals_cauchy_binet(seed):
    rng = np.random.default_rng(seed)
    occ_up, occ_dn = _dets(7, rng)
    ci = rng.normal(size=7) + 1j * rng.normal(size=7)
    tu, td = MD.trials_from_occ(M, occ_up), MD.trials_from_occ(M, occ_dn)
    pa, pb = _walkers(5, rng)
    ov = MD.greens_function_multidet_overlap_only(pa, pb, tu, td, ci, det_chunk_size=3)
    _, _, ov2 = MD.greens_function_multidet(pa, pb, tu, td, ci, det_chunk_size=4)
    t = _trial_vector(occ_up, occ_dn, ci)
    for w in range(5):
        ref = np.vdot(t, np.kron(det_vector(pa[w]), det_vector(pb[w])))
        assert abs(ov[w] - ref) < 1e-12 * abs(ref)
        assert abs(ov2[w] - ref) < 1e-12 * abs(ref)


@pytest.mark.parametrize("seed", [0, 1])
def test_mixed_energy_equals_brute_force(seed):
    s = _sys(seed)
    rng = np.random.default_rng(10 + seed)
    occ_up, occ_dn = _dets(6, rng)
    ci = rng.normal(size=6) + 0.3j * rng.normal(size=6)
    tu, td = MD.trials_from_occ(M, occ_up), MD.trials_from_occ(M, occ_dn)
    pa, pb = _walkers(4, rng)
    e, e1, e2 = MD.local_energy_multidet(s.h1e, s.chol, pa, pb, tu, td, ci, s.enuc)
    H = hamiltonian(s.h1e_mod, s.chol, s.enuc, M, NA, NB)
    t = _trial_vector(occ_up, occ_dn, ci)
    for w in range(4):
        v = np.kron(det_vector(pa[w]), det_vector(pb[w]))
        ref = np.vdot(t, H @ v) / np.vdot(t, v)
        assert abs(e[w] - ref) < 1e-9 * abs(ref)


def test_exact_trial_has_zero_variance():
    """All 15 x 15 dets with FCI ground-state coefficients: E_loc = E_0."""
    s = _sys(5)
    H = hamiltonian(s.h1e_mod, s.chol, s.enuc, M, NA, NB)
    evals, evecs = np.linalg.eigh(H)
    ups = list(itertools.combinations(range(M), NA))
    dns = list(itertools.combinations(range(M), NB))
    occ_up = np.array([u for u in ups for _ in dns])
    occ_dn = np.array([d for _ in ups for d in dns])
    ci = evecs[:, 0].astype(complex)
    order = np.argsort(-np.abs(ci), kind='stable')  # dominant det first
    occ_up, occ_dn, ci = occ_up[order], occ_dn[order], ci[order]
    tu, td = MD.trials_from_occ(M, occ_up), MD.trials_from_occ(M, occ_dn)
    rng = np.random.default_rng(2)
    pa, pb = _walkers(6, rng, 0.7)
    e, _, _ = MD.local_energy_multidet(s.h1e, s.chol, pa, pb, tu, td, ci, s.enuc)
    np.testing.assert_allclose(e, evals[0], rtol=0, atol=1e-9)


def test_one_det_reduces_to_single_det():
    s = _sys(7)
    rng = np.random.default_rng(4)
    occ_up, occ_dn = np.array([range(NA)]), np.array([range(NB)])
    tu, td = MD.trials_from_occ(M, occ_up), MD.trials_from_occ(M, occ_dn)
    ci = np.array([0.7 - 0.2j])
    pa, pb = _walkers(5, rng)
    Ga, Gb, ov = MD.greens_function_multidet(pa, pb, tu, td, ci)
    rGa, rGb, _, _, rov = O.greens_function(pa, pb, tu[0].astype(complex), td[0].astype(complex))
    np.testing.assert_allclose(Ga, rGa, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ov, np.conj(ci[0]) * rov, rtol=1e-13)
    ra, rb = O.half_rotate_cholesky(s.chol, tu[0], td[0])
    e, _, _ = MD.local_energy_multidet(s.h1e, s.chol, pa, pb, tu, td, ci, s.enuc)
    _, _, Gha, Ghb, _ = O.greens_function(pa, pb, tu[0], td[0])
    re, _, _ = O.local_energy(s.h1e, rGa, rGb, Gha, Ghb, ra, rb, s.enuc)
    np.testing.assert_allclose(e, re, rtol=1e-12)
    prop = MD.build_propagator_multidet(s.h1e_mod, s.chol, tu, td, ci, 0.01)
    xi = rng.normal(size=(5, G))
    w = rng.uniform(0.5, 1.5, 5)
    eh = rng.normal(size=5) + 0j
    out = MD.propagate_walkers_multidet(pa, pb, w, None, eh, prop, s.chol, tu, td, ci, -2.0, xi)
    ref = O.propagate_walkers(pa, pb, w, None, eh, prop, s.chol, ra, rb, tu[0].astype(complex),
                              td[0].astype(complex), -2.0, xi)
    for x, y in zip(out, (ref[0], ref[1], ref[2], np.conj(ci[0]) * ref[3], ref[4])):
        np.testing.assert_allclose(x, y, rtol=1e-11, atol=1e-12)


def test_orthogonal_dets_are_masked():
    """Walkers that are copies of det 0 are orthogonal to every det that
    differs in occupation: those dets are masked and G equals det 0's."""
    s = _sys(8)
    rng = np.random.default_rng(6)
    occ_up, occ_dn = _dets(5, rng)
    tu, td = MD.trials_from_occ(M, occ_up), MD.trials_from_occ(M, occ_dn)
    ci = rng.normal(size=5) + 0j
    pa = np.repeat(tu[:1].astype(complex), 3, axis=0)
    pb = np.repeat(td[:1].astype(complex), 3, axis=0)
    Ga, Gb, ov = MD.greens_function_multidet(pa, pb, tu, td, ci)
    assert np.all(np.isfinite(Ga)) and np.all(np.isfinite(Gb))
    np.testing.assert_allclose(Ga, np.repeat((tu[0] @ tu[0].T)[None], 3, axis=0), atol=1e-14)
    np.testing.assert_allclose(ov, ci[0], rtol=1e-14)
    e, _, _ = MD.local_energy_multidet(s.h1e, s.chol, pa, pb, tu, td, ci, s.enuc)
    ehf = O.determinant_energy_from_integrals(s.h1e, s.chol, s.enuc, NA, NB)
    np.testing.assert_allclose(e.real, ehf, rtol=1e-12)


@pytest.mark.parametrize("seed,e_chunk_g", [(0, 3), (1, 16)])
def test_streamed_energy_equals_brute_force(seed, e_chunk_g):
    """energy.py:318-395 with greens.py:314-369: the per-det one-body trace
    against h1e @ T_I^*, the per-det slabbed two-body, weighted by
    c_I^* o_Ia o_Ib, equals <Psi_T|H|phi>/<Psi_T|phi> from the brute-force
    Hamiltonian.  A walker exactly orthogonal to one det (singular S:
    non-finite Ghalf, zeroed, and a zero weight) drops that det's
    <D_I|H|phi>, as the masked full-G estimator does: there the two agree."""
    s = _sys(seed)
    rng = np.random.default_rng(20 + seed)
    occ_up, occ_dn = _dets(6, rng)
    ci = rng.normal(size=6) + 0.3j * rng.normal(size=6)
    tu, td = MD.trials_from_occ(M, occ_up), MD.trials_from_occ(M, occ_dn)
    pa, pb = _walkers(4, rng)
    # walker 3: alpha orbitals without any weight on det 1's first orbital row
    pa[3, occ_up[1][0], :] = 0.0
    pa[3, occ_up[1][1], :] = 0.0
    Gha, Ghb, ov, oa, ob = MD.greens_function_multidet_force_bias(pa, pb, tu, td, ci)
    assert oa[1, 3] == 0.0 and np.all(np.isfinite(Gha))
    ra, rb = MD.half_rotate_cholesky_multidet(s.chol, tu, td)
    hta = np.einsum('pq,dqi->dpi', s.h1e, tu.conj())
    htb = np.einsum('pq,dqi->dpi', s.h1e, td.conj())
    e, e1, e2 = MD.local_energy_multidet_streamed(hta, htb, Gha, Ghb, ra, rb, ci, oa, ob, s.enuc,
                                                  e_chunk_g)
    H = hamiltonian(s.h1e_mod, s.chol, s.enuc, M, NA, NB)
    t = _trial_vector(occ_up, occ_dn, ci)
    for w in range(4):
        v = np.kron(det_vector(pa[w]), det_vector(pb[w]))
        assert abs(ov[w] - np.vdot(t, v)) < 1e-12 * abs(np.vdot(t, v))
        if w < 3:
            ref = np.vdot(t, H @ v) / np.vdot(t, v)
            assert abs(e[w] - ref) < 1e-9 * abs(ref), w
    em, e1m, e2m = MD.local_energy_multidet(s.h1e, s.chol, pa, pb, tu, td, ci, s.enuc)
    np.testing.assert_allclose(e, em, rtol=1e-12)
    np.testing.assert_allclose(e1, e1m, rtol=1e-12)
