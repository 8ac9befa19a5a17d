"""Electron gas (plane waves): the system builder against the reference's
own checks (tests/test_afqmc_3deg.py of the reference, restated), and the
pins of oracle.heg (closed forms and reductions, see its header).  CPU only."""

import numpy as np
import pytest

from omegaqmc_amd import heg as H
from oracle import afqmc as A
from oracle import heg as OH


# -- system and basis (reference tests/test_afqmc_3deg.py:25-111) ----------

@pytest.mark.parametrize("n,expect", [(1, 1), (7, 7), (19, 19), (27, 27), (33, 33), (20, 19)])
def test_shell_counts(n, expect):
    assert H.kgrid_shells(n)[1] == expect


def test_grid_shells_sorted_symmetric():
    grid, n = H.kgrid_shells(27)
    assert grid.shape == (n, 3)
    assert any(np.all(g == 0) for g in grid)
    n2 = np.sum(grid * grid, axis=1)
    assert np.all(n2[:-1] <= n2[1:])
    pts = {tuple(g) for g in grid}
    assert all(tuple(-g) in pts for g in grid)


def test_system_parameters():
    s = H.build_3deg_system(1.0, 14, 19)
    assert (s['nup'], s['ndown'], s['N_elec'], s['N_pw']) == (7, 7, 14, 19)
    p = H.build_3deg_system(1.0, 7, 19, polarization='polarized')
    assert (p['nup'], p['ndown']) == (7, 0)
    for rs in (0.5, 1.0, 4.0):
        t = H.build_3deg_system(rs, 14, 19)
        np.testing.assert_allclose(t['volume'] / 14, 4.0 * np.pi / 3.0 * rs ** 3, rtol=1e-12)
        np.testing.assert_allclose(t['L'] ** 3, t['volume'], rtol=1e-12)
    assert s['kvecs'].shape == (19, 3)
    kap = np.array([0.1, -0.2, 0.3])
    tw = H.build_3deg_system(1.0, 14, 19, kappa=kap)
    np.testing.assert_allclose(tw['kvecs'] - s['kvecs'], np.tile(kap * s['dk'], (19, 1)),
                               rtol=1e-12, atol=1e-15)
    with pytest.raises(ValueError):
        H.build_3deg_system(1.0, 7, 19)


# -- integrals (reference tests/test_afqmc_3deg.py:113-256) -----------------

@pytest.fixture(scope="module")
def gas():
    s = H.build_3deg_system(2.0, 14, 19)
    return s, H.prepare_3deg_integrals(s)


def test_h1e_kinetic(gas):
    s, ints = gas
    h = ints['h1e']
    assert np.max(np.abs(h - np.diag(np.diag(h)))) == 0.0
    np.testing.assert_allclose(np.diag(h), 0.5 * np.sum(s['kvecs'] ** 2, axis=1), rtol=1e-14)


def test_cholesky_structure(gas):
    s, ints = gas
    L, sign = ints['chol'], ints['chol_sign']
    n = s['N_pw']
    assert L.shape[1:] == (n, n) and ints['q_vecs'].shape == (L.shape[0], 3)
    assert set(np.unique(sign)) <= {-1.0, 1.0}
    assert np.sum(sign > 0) >= np.sum(sign < 0)
    for g in range(L.shape[0]):
        assert np.count_nonzero(L[g]) <= 2 * n
        if sign[g] > 0:
            np.testing.assert_array_equal(L[g], L[g].T)
        else:
            np.testing.assert_array_equal(L[g], -L[g].T)


def test_cholesky_reproduces_coulomb(gas):
    """sum_g sign_g L_pq L_rs = 4 pi / (V |k_p - k_q|^2) when k_p - k_q =
    k_s - k_r != 0 (momentum conservation), else 0 -- by brute force over all
    index quadruples."""
    s, ints = gas
    L, sign = ints['chol'], ints['chol_sign']
    eri = np.einsum('g,gpq,grs->pqrs', sign, L, L)
    grid, k, V = s['grid_int'], s['kvecs'], s['volume']
    n = s['N_pw']
    ref = np.zeros((n, n, n, n))
    for p in range(n):
        for q in range(n):
            d = grid[p] - grid[q]
            if not np.any(d):
                continue
            qq = np.sum((k[p] - k[q]) ** 2)
            for r in range(n):
                for t in range(n):
                    if np.array_equal(grid[t] - grid[r], d):
                        ref[p, q, r, t] = 4.0 * np.pi / (V * qq)
    np.testing.assert_allclose(eri, ref, atol=1e-12)


def test_madelung(gas):
    """Negative, of the simple-cubic magnitude, and independent of the Ewald
    splitting parameter (a wrong term would make it depend on eta)."""
    s, ints = gas
    e = ints['e_madelung']
    assert -3.0 < e < 0.0
    for f in (0.7, 1.4):
        np.testing.assert_allclose(H.compute_madelung_3d(s, eta=f * np.sqrt(np.pi) / s['L']), e,
                                   rtol=1e-10)


def test_trial_fermi_sea(gas):
    s, ints = gas
    up, dn = H.build_trial_3deg(ints['h1e'], 7, 7)
    assert up.shape == (19, 7) and dn.shape == (19, 7)
    np.testing.assert_allclose(up.T @ up, np.eye(7), atol=1e-12)
    eps = np.diag(ints['h1e'])
    occ = np.sum(up ** 2, axis=1) > 0.5
    assert np.max(eps[occ]) <= np.min(eps[~occ]) + 1e-10
    up1, dn0 = H.build_trial_3deg(ints['h1e'], 7, 0)
    assert dn0.shape == (19, 0)


# -- oracle.heg pins --------------------------------------------------------

def _trial_blocks(ints, nup, ndown):
    up, dn = H.build_trial_3deg(ints['h1e'], nup, ndown)
    ra = np.einsum('pi,gpq->giq', up, ints['chol'])
    rb = np.einsum('pi,gpq->giq', dn, ints['chol'])
    return up, dn, ra, rb


def test_fermi_sea_energy_closed_form(gas):
    """Coulomb 0 and exchange 1/2 sum_s sum_{k != k'} 4 pi / (V |k - k'|^2)
    over occupied pairs, summed directly."""
    s, ints = gas
    up, dn, ra, rb = _trial_blocks(ints, 7, 7)
    Ga, Gb, Gha, Ghb, _ = A.greens_function(up[None].astype(complex), dn[None].astype(complex),
                                            up.astype(complex), dn.astype(complex))
    ec, ex = OH.local_energy_2body_pw(Gha, Ghb, ra, rb, ints['chol_sign'])
    np.testing.assert_allclose(ec, 0.0, atol=1e-13)
    k, V = s['kvecs'], s['volume']
    ref = 0.0
    for t in (up, dn):
        occ = np.nonzero(np.sum(t ** 2, axis=1) > 0.5)[0]
        for a in occ:
            for b in occ:
                if a != b:
                    ref += 0.5 * 4.0 * np.pi / (V * np.sum((k[a] - k[b]) ** 2))
    np.testing.assert_allclose(ex.real, ref, rtol=1e-12)
    e_hf = H.hf_energy_3deg(ints['h1e'], ints['chol'], ints['chol_sign'], up, dn, 0.0)
    np.testing.assert_allclose(e_hf, np.trace(ints['h1e'] @ (up @ up.T + dn @ dn.T)) - ref,
                               rtol=1e-12)


def test_energy_all_positive_signs_is_gto():
    rng = np.random.default_rng(3)
    G, N, M, W = 5, 3, 8, 4
    ra = rng.normal(size=(G, N, M))
    rb = rng.normal(size=(G, N, M))
    Gha = rng.normal(size=(W, N, M)) + 1j * rng.normal(size=(W, N, M))
    Ghb = rng.normal(size=(W, N, M)) + 1j * rng.normal(size=(W, N, M))
    a = OH.local_energy_2body_pw(Gha, Ghb, ra, rb, np.ones(G))
    b = A.local_energy_2body(Gha, Ghb, ra, rb)
    np.testing.assert_allclose(a[0], b[0], rtol=1e-13)
    np.testing.assert_allclose(a[1], b[1], rtol=1e-13)


def _random_sym_system(rng, M, Na, Nb, G):
    L = rng.normal(size=(G, M, M)) * 0.3
    L = 0.5 * (L + L.transpose(0, 2, 1))
    h = rng.normal(size=(M, M))
    h = 0.5 * (h + h.T)
    up = np.eye(M)[:, :Na]
    dn = np.eye(M)[:, :Nb]
    prop = A.build_propagator(h, L, up, dn, 0.01)
    ra, rb = A.half_rotate_cholesky(L, up, dn)
    return L, prop, up, dn, ra, rb


def test_step_all_phase_i_is_gto():
    rng = np.random.default_rng(5)
    M, Na, Nb, G, W = 9, 3, 2, 6, 5
    L, prop, up, dn, ra, rb = _random_sym_system(rng, M, Na, Nb, G)
    pa = up + 0.1 * (rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na)))
    pb = dn + 0.1 * (rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb)))
    xi = rng.normal(size=(W, G))
    args = (np.ones(W), np.ones(W, complex), np.zeros(W, complex), prop, L, ra, rb,
            up.astype(complex), dn.astype(complex), -0.3, xi)
    a = OH.propagate_3deg_walkers(pa, pb, *args, True, chol_phase=1j * np.ones(G))
    b = A.propagate_walkers(pa, pb, *args)
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-14)


def test_step_antisymmetric_vectors_are_orthogonal():
    """Only sign -1 (phase 1) vectors, real walkers and fields, no one-body
    term: VHS is real antisymmetric and exp(VHS) is orthogonal, so phi^H phi
    is kept (up to the Taylor remainder of a small VHS).  With phase i the
    VHS i sqrt(dt) sum x B is Hermitian and phi^H phi changes."""
    rng = np.random.default_rng(7)
    M, Na, G, W = 8, 3, 5, 4
    L = rng.normal(size=(G, M, M)) * 0.2
    L = 0.5 * (L - L.transpose(0, 2, 1))
    up = np.eye(M)[:, :Na]
    dn = np.zeros((M, 1))
    dn[0, 0] = 1.0
    prop = {'expH1': np.eye(M), 'mf_shift': np.zeros(G, complex), 'dt': 0.01}
    ra = np.einsum('pi,gpq->giq', up, L)
    rb = np.zeros((G, 1, M))
    pa = (up + 0.2 * rng.normal(size=(W, M, Na))).astype(complex)
    pb = np.tile(dn[None], (W, 1, 1)).astype(complex)
    xi = rng.normal(size=(W, G))
    args = (np.ones(W), np.ones(W, complex), np.zeros(W, complex), prop, L, ra, rb,
            up.astype(complex), dn.astype(complex), 0.0, xi, True)
    out = OH.propagate_3deg_walkers(pa, pb, *args, ndown=0, chol_phase=np.ones(G),
                                    fbbound=1e9)
    gram0 = np.einsum('wpi,wpj->wij', pa.conj(), pa)
    gram1 = np.einsum('wpi,wpj->wij', out[0].conj(), out[0])
    np.testing.assert_allclose(gram1, gram0, atol=1e-10)
    np.testing.assert_array_equal(out[1], pb)  # beta untouched when ndown == 0
    wrong = OH.propagate_3deg_walkers(pa, pb, *args, ndown=0, chol_phase=1j * np.ones(G),
                                      fbbound=1e9)
    assert np.max(np.abs(np.einsum('wpi,wpj->wij', wrong[0].conj(), wrong[0]) - gram0)) > 1e-4


def test_step_without_coulomb_is_one_body(gas):
    s, ints = gas
    up, dn = H.build_trial_3deg(ints['h1e'], 7, 7)
    dt = 0.01
    eps = np.diag(ints['h1e'])
    prop = {'expH1': np.diag(np.exp(-0.5 * dt * eps)), 'mf_shift': np.zeros(0), 'dt': dt}
    W = 3
    pa = np.tile(up[None], (W, 1, 1)).astype(complex)
    pb = np.tile(dn[None], (W, 1, 1)).astype(complex)
    e0 = np.zeros((0, 19, 19))
    out = OH.propagate_3deg_walkers(pa, pb, np.ones(W), np.ones(W, complex),
                                    np.zeros(W, complex), prop, e0, np.zeros((0, 7, 19)),
                                    np.zeros((0, 7, 19)), up.astype(complex), dn.astype(complex),
                                    0.0, np.zeros((W, 0)), False)
    np.testing.assert_allclose(out[0], np.exp(-dt * eps)[None, :, None] * pa, rtol=1e-14)
    Ga, Gb, Gha, Ghb, _ = A.greens_function(out[0], out[1], up.astype(complex),
                                            dn.astype(complex))
    et, e1, e2 = OH.local_energy_pw(ints['h1e'], Ga, Gb, Gha, Ghb, np.zeros((0, 7, 19)),
                                    np.zeros((0, 7, 19)), 0.0, np.zeros(0))
    occ = np.sum(up ** 2, axis=1) > 0.5
    np.testing.assert_allclose(et.real, 2.0 * np.sum(eps[occ]), rtol=1e-12)


# -- twists and infinite-gas references (afqmc_pw_heg.py:970-1169) --------

def test_halton_twists():
    from scipy.stats.qmc import Halton
    t = H.generate_halton_twists_3d(12)
    np.testing.assert_allclose(t, Halton(d=3, scramble=False).random(n=12) - 0.5, atol=1e-15)
    assert np.all((t >= -0.5) & (t < 0.5))
    s = H.build_3deg_system(1.0, 14, 19)
    tw = H.twisted_system_3d(s, t[3])
    ref = H.build_3deg_system(1.0, 14, 19, kappa=t[3])
    np.testing.assert_allclose(tw['kvecs'], ref['kvecs'], rtol=1e-15)


def test_infinite_gas_references():
    """3/5 E_F kinetic and -3 kF / (4 pi) exchange (kF from the density
    n = 3 / (4 pi rs^3), per spin channel for the unpolarized gas), and the
    Perdew-Zunger fit continuous at rs = 1 for both branches."""
    for rs in (0.8, 2.0, 5.0):
        n = 3.0 / (4.0 * np.pi * rs ** 3)
        kf_u = (3.0 * np.pi ** 2 * n) ** (1.0 / 3.0)
        kf_p = (6.0 * np.pi ** 2 * n) ** (1.0 / 3.0)
        np.testing.assert_allclose(H.hf_kinetic_energy_per_elec_ry(rs), 2 * 0.3 * kf_u ** 2,
                                   rtol=1e-13)
        np.testing.assert_allclose(H.hf_kinetic_energy_per_elec_ry(rs, 'polarized'),
                                   2 * 0.3 * kf_p ** 2, rtol=1e-13)
        np.testing.assert_allclose(H.hf_exchange_energy_per_elec_ry(rs),
                                   -2 * 3.0 * kf_u / (4 * np.pi), rtol=1e-13)
    for pol in ('unpolarized', 'polarized'):
        lo, hi = H.pz_correlation_energy(1.0 - 1e-9, pol), H.pz_correlation_energy(1.0, pol)
        assert abs(lo - hi) < 2e-4 and hi < 0
        assert H.pz_correlation_energy(2.0, pol) > H.pz_correlation_energy(0.5, pol)
