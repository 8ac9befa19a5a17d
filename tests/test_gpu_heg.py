"""Electron gas on the GPU step: the FP64 step with signed (symmetric and
antisymmetric) Cholesky vectors and the sign-weighted local energy against
oracle.heg on identical inputs, the fully polarized gas, and the
non-interacting gas whose AFQMC energy is exactly its kinetic energy."""

import numpy as np
import pytest

from afq_cases import rel_err
from oracle import afqmc as A
from oracle import heg as OH

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def afq(afq_lib):
    import torch
    import omegaqmc_amd as M
    torch.cuda.set_device(0)
    return M


def _t(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _n(t):
    return t.detach().cpu().numpy()


def _driver(rs, n_elec, n_pw, polarization='unpolarized', coulomb=True, dt=0.01,
            precision="fp64"):
    from omegaqmc_amd import heg as H
    s = H.build_3deg_system(rs, n_elec, n_pw, polarization=polarization)
    return s, H.HEGDriver(s, dt=dt, include_coulomb=coulomb, verbose=False, precision=precision)


def _oracle_inputs(d):
    prop = {'expH1': _n(d.propagator['expH1']), 'mf_shift': _n(d.propagator['mf_shift']),
            'dt': d.dt}
    sign = _n(d.chol_sign)
    phase = np.where(sign > 0, 1j, 1.0)
    return prop, _n(d.chol), _n(d.rchol_a), _n(d.rchol_b), _n(d.trial_up), _n(d.trial_dn), sign, phase


@pytest.mark.parametrize("rs,n_elec,n_pw", [(2.0, 14, 19), (1.0, 2, 7)])
def test_step_matches_oracle(afq, rs, n_elec, n_pw):
    s, d = _driver(rs, n_elec, n_pw)
    prop, L, ra, rb, tu, td, sign, phase = _oracle_inputs(d)
    assert np.any(sign < 0)
    W, M, Na, Nb, G = 13, s['N_pw'], s['nup'], s['ndown'], L.shape[0]
    rng = np.random.default_rng(n_pw)
    pa = tu + 0.15 * (rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na)))
    pb = td + 0.15 * (rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb)))
    w = rng.uniform(0.5, 1.5, size=W)
    eh = rng.normal(size=W) + 0.1j * rng.normal(size=W)
    for step in range(3):
        xi = rng.normal(size=(W, G))
        ref = OH.propagate_3deg_walkers(pa, pb, w, None, eh, prop, L, ra, rb, tu.astype(complex),
                                        td.astype(complex), -0.4, xi, True, ndown=Nb,
                                        chol_phase=phase)
        ga, gb, gw, ge = _t(pa), _t(pb), _t(w), _t(eh)
        ov = d.system.step(ga, gb, gw, ge, -0.4, xi=_t(xi))
        for x, y in ((ga, ref[0]), (gb, ref[1]), (gw, ref[2]), (ov, ref[3]), (ge, ref[4])):
            assert rel_err(_n(x), y) < 1e-10, step
        pa, pb, w, _, eh = ref


def test_polarized_step_matches_oracle(afq):
    """ndown = 0: no beta determinant here; the oracle carries the
    reference's one-column dummy beta walker, untouched, zero Cholesky rows."""
    s, d = _driver(2.0, 7, 19, polarization='polarized')
    prop, L, ra, _, tu, _, sign, phase = _oracle_inputs(d)
    W, M, Na, G = 9, s['N_pw'], s['nup'], L.shape[0]
    dummy = np.zeros((M, 1))
    dummy[0, 0] = 1.0
    rng = np.random.default_rng(11)
    pa = tu + 0.15 * (rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na)))
    pb_dummy = np.tile(dummy[None], (W, 1, 1)).astype(complex)
    w = np.ones(W)
    eh = np.zeros(W, complex)
    xi = rng.normal(size=(W, G))
    ref = OH.propagate_3deg_walkers(pa, pb_dummy, w, None, eh, prop, L, ra, np.zeros((G, 1, M)),
                                    tu.astype(complex), dummy.astype(complex), 0.1, xi, True,
                                    ndown=0, chol_phase=phase)
    import torch
    ga, gw, ge = _t(pa), _t(w), _t(eh)
    gb = torch.zeros((W, M, 0), dtype=torch.complex128, device='cuda')
    ov = d.system.step(ga, gb, gw, ge, 0.1, xi=_t(xi))
    for x, y in ((ga, ref[0]), (gw, ref[2]), (ov, ref[3]), (ge, ref[4])):
        assert rel_err(_n(x), y) < 1e-10


def test_local_energy_matches_oracle(afq):
    s, d = _driver(2.0, 14, 19)
    prop, L, ra, rb, tu, td, sign, phase = _oracle_inputs(d)
    W, M = 7, s['N_pw']
    rng = np.random.default_rng(2)
    pa = tu + 0.2 * (rng.normal(size=(W, M, 7)) + 1j * rng.normal(size=(W, M, 7)))
    pb = td + 0.2 * (rng.normal(size=(W, M, 7)) + 1j * rng.normal(size=(W, M, 7)))
    Ga, Gb, Gha, Ghb, _ = A.greens_function(pa, pb, tu.astype(complex), td.astype(complex))
    ref = OH.local_energy_pw(_n(d.h1e), Ga, Gb, Gha, Ghb, ra, rb, d.enuc, sign)
    out = d.system.local_energy(_t(pa), _t(pb), d.h1e, d.enuc)
    for x, y in zip(out, ref):
        assert rel_err(_n(x), y) < 1e-11
    # the Fermi sea itself: the trial (HF) energy
    e_tr = d.system.local_energy(_t(tu[None].astype(complex)), _t(td[None].astype(complex)),
                                 d.h1e, d.enuc)[0]
    np.testing.assert_allclose(_n(e_tr).real, d.e_trial, rtol=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("n_elec,pol", [(14, 'unpolarized'), (7, 'polarized')])
def test_free_electron_energy_is_kinetic(afq, n_elec, pol, precision):
    """Without the Coulomb term every walker stays a scaled Fermi sea: the
    AFQMC energy is the sum of the occupied kinetic energies (reference
    tests/test_afqmc_3deg.py:296-373, here exact; no Cholesky vectors, so
    the reduced precisions take the same path with zero fields)."""
    s, d = _driver(2.0, n_elec, 19, polarization=pol, coulomb=False, precision=precision)
    eps = np.sort(np.diag(_n(d.h1e)))
    exact = (2.0 if pol == 'unpolarized' else 1.0) * np.sum(eps[:s['nup']])
    out = d(num_walkers=16, num_blocks=3, num_steps_per_block=5, num_blocks_equil=1)
    np.testing.assert_allclose(out['energy_blocks'], exact, rtol=1e-10)
    np.testing.assert_allclose(out['e_per_elec_ha'] * n_elec, out['energy_mean'], rtol=1e-14)


def test_coulomb_lowers_energy(afq):
    """reference tests/test_afqmc_3deg.py:325-351 (with exchange the energy
    drops below the kinetic-only run)."""
    _, free = _driver(2.0, 2, 7, coulomb=False)
    _, coul = _driver(2.0, 2, 7, coulomb=True)
    kw = dict(num_walkers=32, num_blocks=4, num_steps_per_block=5, num_blocks_equil=1)
    ef, ec = free(**kw)['energy_mean'], coul(**kw)['energy_mean']
    assert np.isfinite(ec) and ec < ef + 0.1


def test_bf16x3_refuses_antisymmetric_vectors(afq):
    from omegaqmc_amd import heg as H
    s = H.build_3deg_system(2.0, 2, 7)
    with pytest.raises(afq.AFQError):
        H.HEGDriver(s, dt=0.01, verbose=False, precision="bf16x3")


def test_twist_averaged_run(afq):
    from omegaqmc_amd import heg as H
    base = H.build_3deg_system(2.0, 2, 7)
    out = H.run_twist_averaged_afqmc_3deg(base, n_twists=2, dt=0.01, num_walkers=16,
                                          num_blocks=2, num_steps_per_block=3,
                                          num_blocks_equil=1, verbose=False)
    for i, kap in enumerate(out['twists']):
        ints = H.prepare_3deg_integrals(H.twisted_system_3d(base, kap))
        tu, td = H.build_trial_3deg(ints['h1e'], 1, 1)
        e_hf = H.hf_energy_3deg(ints['h1e'], ints['chol'], ints['chol_sign'], tu, td,
                                ints['e_madelung'] * 2)
        np.testing.assert_allclose(out['e_trial_twists'][i], e_hf, rtol=1e-12)
    assert np.all(np.isfinite(out['e_twists']))
    np.testing.assert_allclose(out['energy_mean'], np.mean(out['e_twists']), rtol=1e-14)


def test_fp32_step_and_energy_with_signed_vectors(afq):
    """FP32-MFMA takes the signed (and antisymmetric) vectors: one step
    against the oracle at the FP32 step bound, and the block energy of the
    stepped walkers equal to the FP64 system's (the same FP64 estimator)."""
    s, d = _driver(2.0, 14, 19, precision="fp32")
    _, d64 = _driver(2.0, 14, 19)
    prop, L, ra, rb, tu, td, sign, phase = _oracle_inputs(d)
    assert np.any(sign < 0)
    W, M, Na, Nb, G = 11, s['N_pw'], s['nup'], s['ndown'], L.shape[0]
    rng = np.random.default_rng(5)
    pa = tu + 0.15 * (rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na)))
    pb = td + 0.15 * (rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb)))
    w = np.ones(W)
    eh = np.zeros(W, complex)
    xi = rng.normal(size=(W, G))
    ref = OH.propagate_3deg_walkers(pa, pb, w, None, eh, prop, L, ra, rb, tu.astype(complex),
                                    td.astype(complex), -0.4, xi, True, ndown=Nb,
                                    chol_phase=phase)
    ga, gb, gw, ge = _t(pa), _t(pb), _t(w), _t(eh)
    ov = d.system.step(ga, gb, gw, ge, -0.4, xi=_t(xi))
    for x, y in ((ga, ref[0]), (gb, ref[1]), (ov, ref[3])):
        assert rel_err(_n(x), y) < 1e-5
    e32 = d.system.local_energy(ga, gb, d.h1e, d.enuc)
    e64 = d64.system.local_energy(ga, gb, d64.h1e, d64.enuc)
    for x, y in zip(e32, e64):
        assert np.array_equal(_n(x), _n(y))


def test_bf16x3_refuses_nonsymmetric_vectors(afq):
    """The split-BF16 VHS mirrors the upper triangle, so a BF16x3 system
    with any L_g not exactly symmetric is refused at creation."""
    import torch
    from omegaqmc_amd.afqmc_gto import System
    from omegaqmc_amd.synth import make_system
    sy = make_system(16, 3, 3, 20, seed=1)
    chol = sy.chol.copy()
    chol[3, 2, 5] += 1e-3
    tu, td = torch.eye(16, 3, dtype=torch.complex128), torch.eye(16, 3, dtype=torch.complex128)
    prop = {'expH1': _t(np.eye(16)), 'mf_shift': _t(np.zeros(20, complex)), 'dt': 0.01}
    ra = _t(np.ascontiguousarray(chol[:, :3, :]))
    with pytest.raises(afq.AFQError, match="symmetric"):
        System(prop, _t(chol), ra, ra, tu.cuda(), td.cuda(), precision="bf16x3")
    System(prop, _t(chol), ra, ra, tu.cuda(), td.cuda(), precision="fp32")


def test_local_energy_pw_functions_match_oracle(afq):
    """afqmc_pw_heg.py:414-471 mirrors (signed Coulomb and exchange) against
    oracle.heg on the same Green's functions."""
    from omegaqmc_amd import heg as H
    s, d = _driver(2.0, 14, 19)
    prop, L, ra, rb, tu, td, sign, phase = _oracle_inputs(d)
    W, M = 6, s['N_pw']
    rng = np.random.default_rng(31)
    pa = tu + 0.2 * (rng.normal(size=(W, M, 7)) + 1j * rng.normal(size=(W, M, 7)))
    pb = td + 0.2 * (rng.normal(size=(W, M, 7)) + 1j * rng.normal(size=(W, M, 7)))
    Ga, Gb, Gha, Ghb, _ = A.greens_function(pa, pb, tu.astype(complex), td.astype(complex))
    ref = OH.local_energy_pw(_n(d.h1e), Ga, Gb, Gha, Ghb, ra, rb, d.enuc, sign)
    got = H.local_energy_pw(d.h1e, d.chol, _t(Ga), _t(Gb), _t(Gha), _t(Ghb), _t(ra), _t(rb),
                            d.enuc, d.chol_sign)
    for x, y in zip(got, ref):
        assert rel_err(_n(x), y) < 1e-11
    ec, ex = H.local_energy_2body_pw(_t(Gha), _t(Ghb), _t(ra), _t(rb), d.chol_sign)
    np.testing.assert_allclose(_n(ec - ex), ref[2], rtol=1e-11, atol=1e-12)
    # all signs +1 is the plain two-body energy
    ec1, ex1 = H.local_energy_2body_pw(_t(Gha), _t(Ghb), _t(ra), _t(rb), _t(np.ones(len(sign))))
    ec0, ex0 = afq.local_energy_2body(_t(Gha), _t(Ghb), _t(ra), _t(rb))
    assert np.array_equal(_n(ec1), _n(ec0)) and np.array_equal(_n(ex1), _n(ex0))
