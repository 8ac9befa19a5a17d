"""Pins of the cavity-QED oracle (oracle/qed.py) by closed forms."""

import numpy as np
import pytest
import scipy.linalg

from afq_cases import rel_err
from oracle import afqmc as O
from oracle import philox as P
from oracle import qed as Q


@pytest.mark.parametrize("s,q0,omega,dt", [(1.0, 0.0, 0.7, 0.05), (1.3, 0.4, 0.2, 0.005),
                                            (0.6, -0.8, 2.0, 0.01)])
def test_photon_step_is_importance_sampled_mehler_kernel(s, q0, omega, dt):
    """The proposal density N(q'; q/cosh d + tanh d drift, tanh d) times the
    weight factor equals f(q') K(q, q') / f(q), with K Mehler's kernel of
    Omega/2 (p^2 + q^2) - Omega/2 over dt/2 and f the Gaussian trial."""
    delta = omega * dt / 2.0
    rng = np.random.default_rng(1)
    for q in rng.normal(size=4):
        mean = q / np.cosh(delta) + np.tanh(delta) * (-s * (q - q0))
        qp = mean + np.sqrt(np.tanh(delta)) * np.linspace(-6.0, 6.0, 41)
        eta = (qp - mean) / np.sqrt(np.tanh(delta))
        qn, lw = Q.propagate_photon(np.full_like(qp, q), omega, dt, s, q0, eta)
        np.testing.assert_allclose(qn, qp, rtol=0, atol=1e-12)
        dens = np.exp(-0.5 * eta ** 2) / np.sqrt(2 * np.pi * np.tanh(delta))

        def f(x):
            return (s / np.pi) ** 0.25 * np.exp(-s * (x - q0) ** 2 / 2.0)
        sh = np.sinh(delta)
        K = (np.exp(-((q ** 2 + qp ** 2) * np.cosh(delta) - 2 * q * qp) / (2 * sh))
             / np.sqrt(2 * np.pi * sh) * np.exp(delta / 2.0))
        np.testing.assert_allclose(np.exp(lw) * dens, f(qp) * K / f(q), rtol=1e-11)


def test_photon_local_energy_is_mixed_estimator():
    """E_ph = <f|H_ph|q>/<f|q> with H_ph = Omega/2 (-d^2/dq^2 + q^2 - 1):
    the second derivative of the trial by central differences."""
    omega, s, q0 = 0.9, 1.4, 0.3
    q = np.linspace(-2, 2, 9)
    h = 1e-4

    def f(x):
        return np.exp(-s * (x - q0) ** 2 / 2.0)
    d2 = (f(q + h) - 2 * f(q) + f(q - h)) / h ** 2
    want = omega / 2.0 * (-d2 / f(q) + q ** 2 - 1.0)
    W, M = q.size, 4
    z = np.zeros((W, M, M), complex)
    gh = np.zeros((W, 1, M), complex)
    L = np.zeros((1, 1, M))
    _, _, _, eph = Q.qed_local_energy(np.zeros((M, M)), z, z, gh, gh, L, L, 0.0, q,
                                      np.zeros((M, M)), omega, s, q0)
    np.testing.assert_allclose(eph, want, rtol=0, atol=1e-6)


def test_zero_coupling_reduces_to_electronic():
    """Zero dipole: the QED step's electrons follow the plain step (the
    augmented vector is zero), the weights get only the photon factor, and
    the local energy is the electronic one plus the photon term."""
    from omegaqmc_amd.synth import make_system
    s = make_system(10, 2, 2, 20, seed=5)
    tu, td = s.trial_up, s.trial_dn
    dip = np.zeros((10, 10))
    qi = Q.prepare_qed_integrals_from_arrays(s.h1e, s.chol, dip)
    ra, rb = O.half_rotate_cholesky(qi['chol_qed'], tu, td)
    prop = Q.build_qed_propagator(qi['h1e_mod_0'], qi['chol_qed'], tu, td, 0.01)
    plain = O.build_propagator(s.h1e_mod, s.chol, tu, td, 0.01)
    assert rel_err(prop['expH1_0'], plain['expH1']) < 1e-14
    W = 6
    rng = np.random.default_rng(2)
    pa = tu + 0.1 * (rng.normal(size=(W, 10, 2)) + 1j * rng.normal(size=(W, 10, 2)))
    pb = td + 0.1 * (rng.normal(size=(W, 10, 2)) + 1j * rng.normal(size=(W, 10, 2)))
    w = np.ones(W)
    eh = np.zeros(W, complex)
    q = rng.normal(size=W)
    xi = P.aux_fields(3, 1, W, 21)
    e1, e2 = rng.normal(size=W), rng.normal(size=W)
    out = Q.propagate_qed_walkers(pa, pb, q, w, None, eh, prop, qi['chol_qed'], ra.real, rb.real,
                                  tu, td, -1.0, xi, e1, e2, dip, 0.5, 1.0, 0.0)
    ref = O.propagate_walkers(pa, pb, w, None, eh, plain, s.chol, ra.real[:20], rb.real[:20], tu,
                              td, -1.0, xi[:, :20])
    assert rel_err(out[0], ref[0]) < 1e-13 and rel_err(out[1], ref[1]) < 1e-13
    _, l1 = Q.propagate_photon(q, 0.5, 0.01, 1.0, 0.0, e1)
    qm, _ = Q.propagate_photon(q, 0.5, 0.01, 1.0, 0.0, e1)
    _, l2 = Q.propagate_photon(qm, 0.5, 0.01, 1.0, 0.0, e2)
    assert rel_err(out[3], ref[2] * np.exp(np.clip(l1 + l2, -10, 10))) < 1e-13
    Ga, Gb, Gha, Ghb, _ = O.greens_function(pa, pb, tu, td)
    et, e1b, e2b, eph = Q.qed_local_energy(s.h1e, Ga, Gb, Gha, Ghb, ra.real, rb.real, s.enuc, q,
                                           dip, 0.5, 1.0, 0.0)
    rt, r1, r2 = O.local_energy(s.h1e, Ga, Gb, Gha, Ghb, ra.real[:20], rb.real[:20], s.enuc)
    assert rel_err(e1b, r1) < 1e-13 and rel_err(e2b, r2) < 1e-12
    assert rel_err(et, rt + eph) < 1e-12


def test_photon_taylor_matches_expm():
    """exp(c d) phi by four Taylor terms vs scipy expm, within the remainder
    |c d|^5 / 5! e^|c d| of the truncated series."""
    rng = np.random.default_rng(4)
    M, N, W = 8, 3, 5
    d = rng.normal(size=(M, M))
    d = 0.5 * (d + d.T)
    c = 0.01 * rng.normal(size=W)
    phi = rng.normal(size=(W, M, N)) + 1j * rng.normal(size=(W, M, N))
    got = Q.apply_exp_vhs_photon(c, d, phi, 4)
    for i in range(W):
        a = abs(c[i]) * np.linalg.norm(d, 2)
        bound = a ** 5 / 120.0 * np.exp(a) * np.linalg.norm(phi[i])
        err = np.linalg.norm(got[i] - scipy.linalg.expm(c[i] * d) @ phi[i])
        assert err <= 1.01 * bound + 1e-14 * np.linalg.norm(phi[i])  # + rounding


def test_energy_at_coherent_trial_closed_form():
    """At phi = trial and q = q0 = -<D>/sqrt(Omega): E = E_det(h1e, [chol; d])
    - <D>^2 + Omega/2 (s - 1 + q0^2), E_det the closed-form determinant
    energy (sum of occupied h1e plus Coulomb minus exchange of the vectors)."""
    from omegaqmc_amd.synth import make_system
    M, Na, Nb, G = 9, 3, 2, 14
    s = make_system(M, Na, Nb, G, seed=11)
    rng = np.random.default_rng(12)
    d = rng.normal(size=(M, M))
    dip = 0.1 * (d + d.T)
    qi = Q.prepare_qed_integrals_from_arrays(s.h1e, s.chol, dip, dip @ dip)
    tu, td = s.trial_up, s.trial_dn
    ra, rb = O.half_rotate_cholesky(qi['chol_qed'], tu, td)
    omega, sq = 0.7, 1.3
    D = np.sum(np.diag(dip)[:Na]) + np.sum(np.diag(dip)[:Nb])
    q0 = -D / np.sqrt(omega)
    W = 2
    pa = np.tile(tu[None].astype(complex), (W, 1, 1))
    pb = np.tile(td[None].astype(complex), (W, 1, 1))
    Ga, Gb, Gha, Ghb, _ = O.greens_function(pa, pb, tu, td)
    et, _, _, eph = Q.qed_local_energy(qi['h1e'], Ga, Gb, Gha, Ghb, ra.real, rb.real, s.enuc,
                                       np.full(W, q0), dip, omega, sq, q0)
    # E_det by explicit sums over occupied orbitals (no Green's function)
    L = qi['chol_qed']
    e1 = s.enuc + np.trace(qi['h1e'][:Na, :Na]) + np.trace(qi['h1e'][:Nb, :Nb])
    J = 0.5 * np.sum((np.einsum('gii->g', L[:, :Na, :Na]) + np.einsum('gii->g', L[:, :Nb, :Nb])) ** 2)
    K = 0.5 * (np.sum(L[:, :Na, :Na] ** 2) + np.sum(L[:, :Nb, :Nb] ** 2))
    want = e1 + J - K - D * D + omega / 2.0 * (sq - 1.0 + q0 ** 2)
    np.testing.assert_allclose(et.real, want, rtol=1e-13)
    np.testing.assert_allclose(eph, omega / 2.0 * (sq - 1.0 + q0 ** 2), rtol=1e-14)
