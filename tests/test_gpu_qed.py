"""Cavity-QED step on the GPU against the QED oracle (oracle/qed.py).

Pieces (photon half step, scaled Taylor exponential, photon weight factor,
comb gather) and the QED local energy at 1e-12; one full step with given
draws and a 20-step driver trajectory (QR every 5 steps, comb every 10, the
fields and photon normals from the Philox stream) at 1e-10 relative, the
comb survivor indices exactly.  At a larger size the walkers at the trial
are checked against the closed form of the energy of a determinant with a
coherent-state photon.
"""

import numpy as np
import pytest

from afq_cases import rel_err
from oracle import afqmc as O
from oracle import philox as P
from oracle import qed as Q

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qed(afq_lib):
    import torch
    from omegaqmc_amd import qed as QD
    torch.cuda.set_device(0)
    return QD


def _t(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _n(t):
    return t.detach().cpu().numpy()


def _dipole(M, scale, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(M, M))
    return scale * 0.5 * (d + d.T)


def _qed_pair(qed, M, Na, Nb, G, dip_scale=0.05, omega=0.5, dt=0.01, seed=7):
    from omegaqmc_amd.synth import make_system
    s = make_system(M, Na, Nb, G, seed=seed)
    dip = _dipole(M, dip_scale, seed + 1)
    quad = dip @ dip  # a symmetric positive quadrupole stand-in
    oi = Q.prepare_qed_integrals_from_arrays(s.h1e, s.chol, dip, quad)
    tu, td = s.trial_up, s.trial_dn
    ora, orb = O.half_rotate_cholesky(oi['chol_qed'], tu, td)
    oprop = Q.build_qed_propagator(oi['h1e_mod_0'], oi['chol_qed'], tu, td, dt)
    gi = qed.prepare_qed_integrals_from_arrays(_t(s.h1e), _t(s.chol), _t(dip), s.enuc, (Na, Nb),
                                               quad_mo=_t(quad))
    drv = qed.QEDDriver(gi, omega, dt=dt, verbose=False)
    return s, oi, ora.real, orb.real, oprop, drv, dip


def test_pieces_match_oracle(qed):
    rng = np.random.default_rng(3)
    W, M, N = 37, 14, 3
    q = rng.normal(size=W)
    eta = rng.normal(size=W)
    for omega, dt, s, q0 in ((0.5, 0.01, 1.0, -0.3), (2.0, 0.005, 1.4, 0.8)):
        qn, lw, _ = qed.propagate_photon(_t(q), omega, dt, s, q0, eta=_t(eta))
        rq, rl = Q.propagate_photon(q, omega, dt, s, q0, eta)
        assert rel_err(_n(qn), rq) < 1e-14 and rel_err(_n(lw), rl) < 1e-12
    d = _dipole(M, 0.3, 4)
    c = 0.05 * rng.normal(size=W)
    phi = rng.normal(size=(W, M, N)) + 1j * rng.normal(size=(W, M, N))
    for nmax in (0, 1, 4):
        got = qed.apply_exp_scaled(_t(c), _t(d), _t(phi), nmax)
        assert rel_err(_n(got), Q.apply_exp_vhs_photon(c, d, phi, nmax)) < 1e-13, nmax
    w = rng.uniform(0.5, 2.0, size=W)
    l1, l2 = 8.0 * rng.normal(size=W), 8.0 * rng.normal(size=W)
    got = qed.photon_weights(_t(w), _t(l1), _t(l2))
    assert rel_err(_n(got), w * np.exp(np.clip(l1 + l2, -10.0, 10.0))) < 1e-14
    idx = rng.integers(0, W, size=W).astype(np.int32)
    got = qed.gather_walker_values(_t(idx), _t(q))
    assert np.array_equal(_n(got), q[idx])


def test_local_energy_matches_oracle(qed):
    s, oi, ra, rb, oprop, drv, dip = _qed_pair(qed, 16, 3, 2, 40)
    rng = np.random.default_rng(5)
    W = 11
    pa = s.trial_up + 0.2 * (rng.normal(size=(W, 16, 3)) + 1j * rng.normal(size=(W, 16, 3)))
    pb = s.trial_dn + 0.2 * (rng.normal(size=(W, 16, 2)) + 1j * rng.normal(size=(W, 16, 2)))
    q = drv.q0 + rng.normal(size=W)
    Ga, Gb, Gha, Ghb, _ = O.greens_function(pa, pb, s.trial_up, s.trial_dn)
    ref = Q.qed_local_energy(oi['h1e'], Ga, Gb, Gha, Ghb, ra, rb, s.enuc, q, dip, drv.omega,
                             drv.s, drv.q0)
    got = qed.qed_local_energy(drv.h1e, _t(Ga), _t(Gb), _t(Gha), _t(Ghb), drv.rchol_a, drv.rchol_b,
                               drv.enuc, _t(q), drv.dip_mo, drv.omega, drv.s, drv.q0)
    for x, y in zip(got, ref):
        np.testing.assert_allclose(_n(x), y, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("M,Na,Nb,G,W", [(12, 3, 2, 30, 16), (24, 5, 5, 60, 40)])
def test_step_matches_oracle(qed, M, Na, Nb, G, W):
    s, oi, ra, rb, oprop, drv, dip = _qed_pair(qed, M, Na, Nb, G)
    assert rel_err(_n(drv.propagator['expH1_0']), oprop['expH1_0'].real) < 1e-12
    assert rel_err(_n(drv.propagator['mf_shift']), oprop['mf_shift']) < 1e-12
    rng = np.random.default_rng(6)
    pa = s.trial_up + 0.1 * (rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na)))
    pb = s.trial_dn + 0.1 * (rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb)))
    q = drv.q0 + 0.5 * rng.normal(size=W)
    w = rng.uniform(0.5, 1.5, size=W)
    eh = rng.normal(size=W) + 0j
    xi = rng.normal(size=(W, G + 1))
    e1, e2 = rng.normal(size=W), rng.normal(size=W)
    ref = Q.propagate_qed_walkers(pa, pb, q, w, None, eh, oprop, oi['chol_qed'], ra, rb,
                                  s.trial_up, s.trial_dn, -1.2, xi, e1, e2, dip, drv.omega, drv.s,
                                  drv.q0)
    got = qed.propagate_qed_walkers(_t(pa), _t(pb), _t(q), _t(w), None, _t(eh), drv.propagator,
                                    drv.chol_qed, drv.rchol_a, drv.rchol_b, drv.trial_up,
                                    drv.trial_dn, -1.2, None, drv.dip_mo, drv.omega, drv.s,
                                    drv.q0, xi=_t(xi), eta=(_t(e1), _t(e2)))
    for k, (g, r) in enumerate(zip(got[:6], ref)):
        assert rel_err(_n(g), r) < 1e-10, k


def test_driver_trajectory_matches_oracle(qed):
    M, Na, Nb, G, W = 12, 3, 2, 30, 48
    s, oi, ra, rb, oprop, drv, dip = _qed_pair(qed, M, Na, Nb, G)
    seed, eshift = 42, -0.9
    nq = G + 1
    st = drv.init_walkers(W)
    pa = np.tile(s.trial_up[None].astype(complex), (W, 1, 1))
    pb = np.tile(s.trial_dn[None].astype(complex), (W, 1, 1))
    q = np.full(W, drv.q0)
    w = np.ones(W)
    eh = np.zeros(W, complex)
    for step in range(1, 21):
        if step % 5 == 0:
            pa, pb, _, _ = O.orthogonalize_walkers(pa, pb)
        z = P.aux_fields(seed, step, W, nq + 2)
        pa, pb, q, w, ov, eh = Q.propagate_qed_walkers(
            pa, pb, q, w, None, eh, oprop, oi['chol_qed'], ra, rb, s.trial_up, s.trial_dn, eshift,
            z[:, :nq], z[:, nq], z[:, nq + 1], dip, drv.omega, drv.s, drv.q0)
        if step > 1:
            w = np.clip(w, 0.0, np.sum(np.abs(w)) * O.WEIGHT_CLIP_FRACTION)
        if step % 10 == 0:
            w, pa, pb, q, eh, _ = Q.population_control_comb_qed(w, pa, pb, q,
                                                                P.comb_offset(seed, step), eh)
        drv.step(st, step, seed, eshift, 5, 10)
        for name, ref in (('phia', pa), ('phib', pb), ('q', q), ('weights', w),
                          ('e_hybrid', eh)):
            assert rel_err(_n(st[name]), ref) < 1e-10, (step, name)
    Ga, Gb, Gha, Ghb, _ = O.greens_function(pa, pb, s.trial_up, s.trial_dn)
    ref = Q.qed_local_energy(oi['h1e'], Ga, Gb, Gha, Ghb, ra, rb, s.enuc, q, dip, drv.omega,
                             drv.s, drv.q0)
    got = drv.block_energy(st)
    for x, y in zip(got, ref):
        np.testing.assert_allclose(_n(x), y, rtol=1e-10, atol=1e-9)


def test_driver_run_blocks(qed):
    """Two blocks through the driver's own loop: finite, and the photon
    observables are the |w|-weighted averages of the final state."""
    import torch
    s, oi, ra, rb, oprop, drv, dip = _qed_pair(qed, 12, 3, 2, 30)
    res = drv(num_walkers=64, num_blocks=2, num_steps_per_block=10, stabilize_freq=5,
              pop_control_freq=5, num_blocks_equil=1)
    assert np.all(np.isfinite(res['energy_blocks'])) and len(res['energy_blocks']) == 3
    st = res['state']
    w = torch.abs(st['weights'])
    qm = float(torch.sum(w * st['q']) / torch.sum(w))
    assert abs(qm - res['q_mean_blocks'][-1]) < 1e-12 * max(1.0, abs(qm))


def test_energy_at_trial_closed_form(qed):
    """M=100, N=10+10, G=300, W=256 at the trial with q = q0 (the coherent
    state): E = E_det(h1e + quad/2, [chol; d]) - <D>^2 + Omega/2 (s - 1 + q0^2),
    <D> the occupied diagonal dipole sum, E_det the closed-form determinant
    energy."""
    import torch
    M, N, G, W = 100, 10, 300, 256
    s, oi, ra, rb, oprop, drv, dip = _qed_pair(qed, M, N, N, G, dip_scale=0.02, omega=0.4)
    st = drv.init_walkers(W)
    e_tot, e_1b, e_2b, e_ph = drv.block_energy(st)
    torch.cuda.synchronize()
    D = 2.0 * np.sum(np.diag(dip)[:N])
    edet = O.determinant_energy_from_integrals(oi['h1e'], oi['chol_qed'], s.enuc, N, N)
    want = edet - D * D + drv.omega / 2.0 * (drv.s - 1.0 + drv.q0 ** 2)
    np.testing.assert_allclose(_n(e_tot).real, want, rtol=1e-11)
    np.testing.assert_allclose(_n(e_ph).real, drv.omega / 2.0 * (drv.s - 1.0 + drv.q0 ** 2),
                               rtol=1e-13)
    assert abs(drv.q0 + D / np.sqrt(drv.omega)) < 1e-12 * max(1.0, abs(drv.q0))


def test_qed_walkers_container(qed):
    """qed_afqmc_gto.py:80-102: trial copies with unit weights and q = q0."""
    import torch
    tu = torch.eye(5, 2, dtype=torch.float64)
    wk = qed.QEDWalkers(tu, tu, 3, 5, 2, 2, q0=-0.25)
    assert wk.phia.shape == (3, 5, 2) and wk.phia.is_cuda
    assert torch.all(wk.q == -0.25) and torch.all(wk.weights == 1)
