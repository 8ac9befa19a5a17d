"""Pins of the NumPy oracle (CPU): each check is independent of the formula it
tests -- brute-force second quantisation, closed forms, invariants, textbook
routines (scipy expm, LAPACK QR) -- so that a dropped term, a wrong sign or
a transposed operand in the oracle fails one of them."""

import numpy as np
import pytest
import scipy.linalg

from afq_cases import make_harness, walkers_near_trial
from fci_brute import hamiltonian, mixed_energy
from oracle import afqmc as O
from oracle import philox as P
from omegaqmc_amd.synth import make_system


def _tiny(M=6, na=2, nb=2, G=8, seed=3, sigma=0.3):
    return make_system(M, na, nb, G, seed=seed, chol_sigma=sigma)


# --------------------------------------------------------------------------- greens

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_greens_invariants(seed):
    tu, td, pa, pb, *_ = make_harness(seed)
    Ga, Gb, Gha, Ghb, ov = O.greens_function(pa, pb, tu, td)
    # Ghalf Psi* = S^-1 S = I ; Tr G = N ; G^2 = G (single determinant)
    I3 = np.eye(3)
    assert np.allclose(np.einsum('wiq,qj->wij', Gha, tu.conj()), I3, atol=1e-10)
    assert np.allclose(np.trace(Ga, axis1=1, axis2=2), 3, atol=1e-10)
    assert np.allclose(Ga @ Ga, Ga, atol=1e-9)
    assert np.allclose(Gb @ Gb, Gb, atol=1e-9)


def test_overlap_permutation_expansion():
    """det by the Leibniz formula (N <= 3), independent of LAPACK."""
    import itertools
    tu, td, pa, pb, *_ = make_harness(4)

    def leibniz(A):
        n = A.shape[0]
        tot = 0j
        for perm in itertools.permutations(range(n)):
            sgn = np.linalg.det(np.eye(n)[list(perm)])
            tot += sgn * np.prod([A[i, perm[i]] for i in range(n)])
        return tot

    ov = O.greens_function_overlap(pa, pb, tu, td)
    for w in range(pa.shape[0]):
        ref = leibniz(pa[w].T @ tu.conj()) * leibniz(pb[w].T @ td.conj())
        assert abs(ov[w] - ref) < 1e-10 * abs(ref)


def test_specialised_variants_bit_equal():
    """test_afqmc_kernels.py:97-134."""
    tu, td, pa, pb, *_ = make_harness(1)
    _, _, Gha, Ghb, ov = O.greens_function(pa, pb, tu, td)
    Gha2, Ghb2, ov2 = O.greens_function_force_bias(pa, pb, tu, td)
    assert np.array_equal(Gha, Gha2) and np.array_equal(Ghb, Ghb2) and np.array_equal(ov, ov2)
    assert np.array_equal(ov, O.greens_function_overlap(pa, pb, tu, td))


# --------------------------------------------------------------------------- energy

@pytest.mark.parametrize("seed", [0, 1])
def test_mixed_energy_equals_brute_force_ci(seed):
    """E_loc(phi) = <Psi_T|H|phi>/<Psi_T|phi> in the full determinant space."""
    s = _tiny(seed=seed)
    H = hamiltonian(s.h1e_mod, s.chol, s.enuc, 6, 2, 2)
    rng = np.random.default_rng(seed)
    tu, td = s.trial_up.astype(complex), s.trial_dn.astype(complex)
    pa = walkers_near_trial(tu, 4, rng, 0.4)
    pb = walkers_near_trial(td, 4, rng, 0.4)
    ra, rb = O.half_rotate_cholesky(s.chol, tu, td)
    Ga, Gb, Gha, Ghb, ov = O.greens_function(pa, pb, tu, td)
    e_tot, e1, e2 = O.local_energy(s.h1e, Ga, Gb, Gha, Ghb, ra, rb, s.enuc)
    for w in range(4):
        e_ref, ov_ref = mixed_energy(H, tu, td, pa[w], pb[w])
        assert abs(ov[w] - ov_ref) < 1e-10 * abs(ov_ref)
        assert abs(e_tot[w] - e_ref) < 1e-9 * abs(e_ref)


def test_energy_at_trial_is_determinant_energy():
    s = make_system(12, 3, 2, 30, seed=9)
    tu, td = s.trial_up.astype(complex), s.trial_dn.astype(complex)
    pa, pb = tu[None], td[None]
    ra, rb = O.half_rotate_cholesky(s.chol, tu, td)
    Ga, Gb, Gha, Ghb, _ = O.greens_function(pa, pb, tu, td)
    e, _, _ = O.local_energy(s.h1e, Ga, Gb, Gha, Ghb, ra, rb, s.enuc)
    ehf = O.determinant_energy_from_integrals(s.h1e, s.chol, s.enuc, 3, 2)
    assert abs(e[0] - ehf) < 1e-12 * abs(ehf)
    # the closed form itself is <HF|H|HF> of the brute-force Hamiltonian
    s6 = _tiny()
    H6 = hamiltonian(s6.h1e_mod, s6.chol, s6.enuc, 6, 2, 2)
    assert abs(H6[0, 0] - O.determinant_energy_from_integrals(s6.h1e, s6.chol, s6.enuc, 2, 2)) < 1e-12


def test_no_two_body_gives_orbital_energies():
    """L = 0: every walker has E_loc = 2 sum_{i<N} eps_i + E_nuc (zero variance)."""
    s = make_system(10, 3, 3, 4, seed=2, chol_sigma=0.0)
    eps = np.diag(s.h1e)
    tu, td = s.trial_up.astype(complex), s.trial_dn.astype(complex)
    rng = np.random.default_rng(0)
    pa, pb = walkers_near_trial(tu, 5, rng, 0.5), walkers_near_trial(td, 5, rng, 0.5)
    ra, rb = O.half_rotate_cholesky(s.chol, tu, td)
    Ga, Gb, Gha, Ghb, _ = O.greens_function(pa, pb, tu, td)
    e, _, _ = O.local_energy(s.h1e, Ga, Gb, Gha, Ghb, ra, rb, s.enuc)
    assert np.allclose(e, 2 * np.sum(np.sort(eps)[:3]) + s.enuc, atol=1e-10)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("e_chunk_g", [4, 7, 24])
def test_streamed_energy_matches(seed, e_chunk_g):
    """test_afqmc_kernels.py:180-238 on the reference's own harness inputs."""
    tu, td, pa, pb, ra, rb, h1e, enuc = make_harness(seed)
    Ga, Gb, Gha, Ghb, _ = O.greens_function(pa, pb, tu, td)
    ec, ex = O.local_energy_2body(Gha, Ghb, ra, rb)
    ec2, ex2 = O.local_energy_2body_streamed(Gha, Ghb, ra, rb, e_chunk_g)
    np.testing.assert_allclose(ec2, ec, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(ex2, ex, rtol=1e-13, atol=1e-12)
    et, e1, e2 = O.local_energy(h1e, Ga, Gb, Gha, Ghb, ra, rb, enuc)
    et2, e12, e22 = O.local_energy_streamed(h1e @ tu.conj(), h1e @ td.conj(), Gha, Ghb, ra, rb,
                                            enuc, 4)
    np.testing.assert_allclose(e12, e1, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(et2, et, rtol=1e-13, atol=1e-12)


# --------------------------------------------------------------------------- propagation

def test_vhs_linear_and_symmetric():
    s = _tiny(M=7, G=9)
    rng = np.random.default_rng(1)
    x1 = rng.normal(size=(3, 9)) + 1j * rng.normal(size=(3, 9))
    x2 = rng.normal(size=(3, 9)) + 1j * rng.normal(size=(3, 9))
    V1, V2 = O.build_vhs(x1, s.chol, 0.01), O.build_vhs(x2, s.chol, 0.01)
    V12 = O.build_vhs(2.0 * x1 - x2, s.chol, 0.01)
    assert np.allclose(V12, 2 * V1 - V2, atol=1e-13)
    assert np.allclose(V1, V1.transpose(0, 2, 1), atol=1e-14)
    # slabbing over g does not change the result
    assert np.allclose(O.build_vhs(x1, s.chol, 0.01, chol_chunk_g=4), V1, atol=1e-14)
    # single field: VHS = i sqrt(dt) x L_g
    e = np.zeros((1, 9), complex)
    e[0, 4] = 1.0
    assert np.allclose(O.build_vhs(e, s.chol, 0.01)[0], 0.1j * s.chol[4], atol=1e-15)


def test_taylor_vs_expm_within_remainder():
    rng = np.random.default_rng(2)
    M, N = 9, 3
    for scale in (0.05, 0.3):
        V = scale * (rng.normal(size=(2, M, M)) + 1j * rng.normal(size=(2, M, M)))
        phi = rng.normal(size=(2, M, N)) + 1j * rng.normal(size=(2, M, N))
        got = O.apply_exp_vhs(V, phi, 6)
        for w in range(2):
            ref = scipy.linalg.expm(V[w]) @ phi[w]
            nv = np.linalg.norm(V[w], 2)
            bound = nv ** 7 / 5040 * np.exp(nv) * np.linalg.norm(phi[w], 2)
            assert np.linalg.norm(got[w] - ref, 2) <= bound * 1.0001 + 1e-13
    assert np.array_equal(O.apply_exp_vhs(V, phi, 0), phi)


def test_build_propagator_closed_forms():
    s = _tiny(M=8, G=10)
    prop = O.build_propagator(s.h1e_mod, s.chol, s.trial_up, s.trial_dn, 0.005)
    mf = prop['mf_shift']
    assert np.all(mf.real == 0)
    # mf_shift_g = i Tr(L_g G_c), G_c = diag(2,2,0,...) -> 2 i sum_{i<2} L_gii
    assert np.allclose(mf, 2j * np.einsum('gii->g', s.chol[:, :2, :2]), atol=1e-14)
    Hs = s.h1e_mod + np.einsum('g,gpq->pq', mf.imag, s.chol)
    ev, U = np.linalg.eigh(Hs)
    ref = (U * np.exp(-0.0025 * ev)) @ U.T
    assert np.allclose(prop['expH1'], ref, atol=1e-13)
    assert np.max(np.abs(prop['expH1'].imag)) == 0.0


def test_force_bias_full_g_equals_half_rotated():
    """afqmc_gto.py:794-795 (full-G force bias) equals the half-rotated form."""
    s = _tiny(M=8, na=2, nb=2, G=10)
    rng = np.random.default_rng(5)
    tu, td = s.trial_up.astype(complex), s.trial_dn.astype(complex)
    pa, pb = walkers_near_trial(tu, 4, rng, 0.3), walkers_near_trial(td, 4, rng, 0.3)
    Ga, Gb, Gha, Ghb, _ = O.greens_function(pa, pb, tu, td)
    ra, rb = O.half_rotate_cholesky(s.chol, tu, td)
    prop = O.build_propagator(s.h1e_mod, s.chol, tu, td, 0.005)
    xi = rng.normal(size=(4, 10))
    _, _, _, _, vb = O.force_bias_shift(Gha, Ghb, ra, rb, prop['mf_shift'], xi, 0.005)
    vfull = np.einsum('gpq,wqp->gw', s.chol, Ga) + np.einsum('gpq,wqp->gw', s.chol, Gb)
    assert np.allclose(vb, vfull, atol=1e-12)


def test_force_bias_cap_and_constants_by_hand():
    dt = 0.01
    vb = np.array([[0.3 + 2.0j]])        # (G=1, W=1)
    mf = np.array([0.5j])
    Gh = np.ones((1, 1, 1))
    # choose rchol so vbias = 0.3 + 2j exactly: rchol_a = 1, Ghalf_a = vb, spin b empty
    xbar, xs, cfb, cmf, v = O.force_bias_shift(vb.reshape(1, 1, 1), np.zeros((1, 0, 1)),
                                               Gh, np.zeros((1, 0, 1)), mf, np.array([[0.7]]), dt,
                                               fbbound=0.1)
    raw = -np.sqrt(dt) * (1j * (0.3 + 2.0j) - 0.5j)
    capped = raw * 0.1 / abs(raw)
    assert abs(raw) > 0.1 and np.isclose(xbar[0, 0], capped)
    assert np.isclose(xs[0, 0], 0.7 - capped)
    assert np.isclose(cmf[0], -np.sqrt(dt) * (0.7 - capped) * 0.5j)
    assert np.isclose(cfb[0], 0.7 * capped - 0.5 * capped ** 2)


def test_update_weights_by_hand():
    dt, eshift = 0.01, -1.0
    w = np.array([2.0, 1.0, 1.0])
    oo = np.array([1.0 + 0j, 1.0, 1.0])
    on = np.array([np.exp(0.02), np.exp(-5.0), np.exp(0.0 + 1.0j)])
    z = np.zeros(3, complex)
    wn, en = O.update_weights_phaseless(w, oo, on, z, z, z, eshift, dt)
    # walker 0: E = -0.02/0.01 = -2 (inside eshift +- sqrt(200))
    assert np.isclose(en[0], -2.0)
    assert np.isclose(wn[0], 2.0 * np.exp(-dt * (0.5 * -2.0 - eshift)))
    # walker 1: E = 500 clipped to eshift + sqrt(2/dt)
    assert np.isclose(en[1].real, eshift + np.sqrt(2 / dt))
    # walker 2: E = -100i -> dtheta = -dt*Im(E) = 1 -> cos(1) projection
    assert np.isclose(en[2], -100j)
    assert np.isclose(wn[2], np.exp(-dt * (0 - eshift)) * np.cos(1.0))
    # cos < 0 kills the walker
    wn2, _ = O.update_weights_phaseless(w[:1], oo[:1], np.exp([2.0j]), z[:1], z[:1], z[:1], 0, dt)
    assert wn2[0] == 0.0
    # dt -> 0 with unit ratio leaves the weight unchanged
    wn3, _ = O.update_weights_phaseless(w, oo, oo, z, z, z, 0.0, 1e-12)
    assert np.allclose(wn3, w)


# --------------------------------------------------------------------------- QR, comb

def test_qr_unique_positive_factor():
    rng = np.random.default_rng(3)
    pa = rng.normal(size=(4, 12, 5)) + 1j * rng.normal(size=(4, 12, 5))
    pb = rng.normal(size=(4, 12, 3)) + 1j * rng.normal(size=(4, 12, 3))
    qa, qb, ld, _ = O.orthogonalize_walkers(pa, pb)
    for q, p in ((qa, pa), (qb, pb)):
        R = np.einsum('wpi,wpj->wij', q.conj(), p)
        assert np.allclose(np.einsum('wpi,wpj->wij', q.conj(), q), np.eye(q.shape[2]), atol=1e-12)
        assert np.allclose(q @ R, p, atol=1e-12)
        assert np.allclose(np.tril(R, -1), 0, atol=1e-12)
        d = np.diagonal(R, axis1=1, axis2=2)
        assert np.all(d.real > 0) and np.allclose(d.imag, 0, atol=1e-12)
    ref = (np.sum(np.log(np.abs(np.diagonal(np.linalg.qr(pa)[1], axis1=1, axis2=2))), axis=1)
           + np.sum(np.log(np.abs(np.diagonal(np.linalg.qr(pb)[1], axis1=1, axis2=2))), axis=1))
    assert np.allclose(ld, ref)
    # det relation with an eye trial: ovlp(phi) = ovlp(Q) exp(log_detR)
    tu, td = np.eye(12, 5), np.eye(12, 3)
    assert np.allclose(O.greens_function_overlap(pa, pb, tu, td),
                       O.greens_function_overlap(qa, qb, tu, td) * np.exp(ld)
                       * np.prod(np.diagonal(np.einsum('wpi,wpj->wij', qa.conj(), pa), axis1=1, axis2=2) /
                                 np.abs(np.diagonal(np.einsum('wpi,wpj->wij', qa.conj(), pa), axis1=1, axis2=2)), axis=1)
                       * np.prod(np.diagonal(np.einsum('wpi,wpj->wij', qb.conj(), pb), axis1=1, axis2=2) /
                                 np.abs(np.diagonal(np.einsum('wpi,wpj->wij', qb.conj(), pb), axis1=1, axis2=2)), axis=1))


@pytest.mark.parametrize("W", [1, 5, 64, 1000])
def test_comb_properties_and_brute_force(W):
    rng = np.random.default_rng(W)
    w = rng.exponential(size=W)
    w[rng.uniform(size=W) < 0.2] = 0.0
    w[-1] = max(w[-1], 0.1)
    r = 0.618
    idx = O.comb_indices(w, r)
    assert np.all(np.diff(idx) >= 0)
    # brute-force tooth search on a sequential running sum
    tot = 0.0
    for x in w:
        tot += x
    c, acc = [], 0.0
    for x in w:
        acc += x / (tot / W)
        c.append(acc)
    for j in range(W):
        t = j + r
        k = next((i for i, ci in enumerate(c) if ci >= t), W - 1)
        assert idx[j] == min(k, W - 1)
    counts = np.bincount(idx, minlength=W)
    expect = w * W / np.sum(w)
    assert np.all(counts >= np.floor(expect) - 1e-9) and np.all(counts <= np.ceil(expect) + 1e-9)
    assert np.all(counts[w == 0] == 0)
    pa = rng.normal(size=(W, 2, 1)) + 0j
    wn, pan, pbn, eh, idx2 = O.population_control_comb(w, pa, pa, r, np.arange(W) + 0j)
    assert np.sum(wn) == W and np.array_equal(idx2, idx) and np.array_equal(eh.real, idx)


def test_pair_branch_known_answer():
    """Worked by hand: <w> = 14.01/6, bounds 0.2335 and 9.34; the sorted
    order is 1 (0.01), 0, 3, 4, 5 (1 each, index order), 2 (10): only the
    pair (1, 2) qualifies, both at 5.005, walker 1 a copy of walker 2."""
    w = np.array([1.0, 0.01, 10.0, 1.0, 1.0, 1.0])
    pa = np.arange(6.0)[:, None, None] + 0j
    pb = -np.arange(6.0)[:, None, None] + 0j
    wn, pan, pbn, src = O.population_control_pair_branch(w, pa, pb)
    np.testing.assert_array_equal(wn, [1.0, 5.005, 5.005, 1.0, 1.0, 1.0])
    np.testing.assert_array_equal(src, [0, 2, 2, 3, 4, 5])
    assert pan[1, 0, 0] == 2.0 and pbn[1, 0, 0] == -2.0 and pan[2, 0, 0] == 2.0
    # nothing beyond the bounds: unchanged
    wn, _, _, src = O.population_control_pair_branch(np.ones(5), pa[:5], pb[:5])
    assert np.array_equal(wn, np.ones(5)) and np.array_equal(src, np.arange(5))


@pytest.mark.parametrize("W", [9, 200])
def test_pair_branch_invariants(W):
    """Total weight kept (each clone pair keeps its sum), every copy made from
    a walker above 4<w> onto one below 0.1<w>, sources never overwritten, and
    the k-th lightest paired with the k-th heaviest (ties by index)."""
    rng = np.random.default_rng(W)
    w = rng.lognormal(0.0, 1.5, size=W)
    w[rng.uniform(size=W) < 0.3] = 0.0
    pa = rng.normal(size=(W, 3, 2)) + 1j * rng.normal(size=(W, 3, 2))
    wn, pan, _, src = O.population_control_pair_branch(w, pa, pa)
    avg = np.sum(w) / W
    np.testing.assert_allclose(np.sum(wn), np.sum(w), rtol=1e-14)
    moved = np.nonzero(src != np.arange(W))[0]
    assert len(moved) > 0
    order = sorted(range(W), key=lambda i: (w[i], i))
    partner = {order[k]: order[W - 1 - k] for k in range(W // 2)}
    for i in moved:
        j = src[i]
        assert partner[i] == j and w[j] > 4 * avg and w[i] < 0.1 * avg
        assert src[j] == j and np.array_equal(pan[i], pa[j]) and wn[i] == wn[j]
    for i in set(range(W)) - set(moved) - set(src[moved]):
        assert wn[i] == w[i] and np.array_equal(pan[i], pa[i])


# --------------------------------------------------------------------------- statistics

def test_binning_closed_forms():
    a = np.array([1.0, 2.0, 3.0, 4.0])
    mean, err, std, kappa = O.do_binning_analysis(a)
    # x = [-1.5,-.5,.5,1.5]; ac(lag1)/ac(0) = (0.75-0.25+0.75)/5 = 0.25 ; lag2 = (-0.75-0.75)/5 < 0
    assert mean == 2.5 and np.isclose(std, np.sqrt(1.25))
    assert np.isclose(kappa, 1.5) and np.isclose(err, np.sqrt(1.25) * np.sqrt(1.5 / 4))
    rng = np.random.default_rng(0)
    _, _, _, k = O.do_binning_analysis(rng.normal(size=4000))
    assert 0.8 < k < 1.4


# --------------------------------------------------------------------------- philox

def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32-10."""
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff, 0xffffffff),
              (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for c, k, ref in cases:
        out = P.philox4x32_10(*c, *k)
        assert tuple(int(v) for v in out) == ref


def test_aux_field_stream_statistics():
    xi = P.aux_fields(42, 3, 64, 1001)
    assert xi.shape == (64, 1001)
    assert abs(xi.mean()) < 0.01 and abs(xi.std() - 1.0) < 0.01
    assert not np.array_equal(P.aux_fields(42, 3, 2, 8), P.aux_fields(42, 4, 2, 8))
    # odd naux drops the last sine of each walker row only
    assert np.array_equal(P.aux_fields(7, 1, 3, 5), P.aux_fields(7, 1, 3, 6)[:, :5])
    r = [P.comb_offset(42, s) for s in range(1, 200)]
    assert min(r) >= 0.0 and max(r) < 1.0


# --------------------------------------------------------------------------- driver

def test_driver_recovers_fci_weakly_correlated():
    """Phaseless AFQMC (oracle driver) agrees with brute-force FCI for a weakly
    correlated synthetic system (the reference's statistical test pattern,
    tests/test_afqmc_LiH.py:86-87: |E - E_FCI| < 5 sigma + 2 mHa)."""
    s = make_system(6, 2, 2, 12, seed=1, chol_sigma=0.1)
    H = hamiltonian(s.h1e_mod, s.chol, s.enuc, 6, 2, 2)
    e_fci = np.linalg.eigvalsh(H)[0]
    d = O.Driver(s.h1e, s.h1e_mod, s.chol, s.enuc, 2, 2, dt=0.01)
    res = d.run(lambda st, W, G: P.aux_fields(5, st, W, G), lambda st: P.comb_offset(5, st),
                num_walkers=100, num_blocks=20, num_steps_per_block=20, stabilize_freq=5,
                pop_control_freq=5, num_blocks_equil=4)
    assert e_fci < d.ehf
    assert abs(res['energy_mean'] - e_fci) < 5 * res['energy_err'] + 2e-3
