"""Split-BF16 (BF16x3, FP32 accumulate) step against the FP64 step.

Both precisions are fed identical auxiliary fields and comb offsets.  The
per-product error of BF16x3 is ~2^-16 relative, so one step must agree with
the FP64 oracle to ~1e-5 relative in the walkers and the block energy of a
short trajectory must stay within the 0.5 mHa budget of the FP64 run.
"""

import numpy as np
import pytest

from afq_cases import bf16x3_two_body, rel_err
from oracle import afqmc as O
from oracle import philox as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def afq(afq_lib):
    import torch
    import omegaqmc_amd as A
    torch.cuda.set_device(0)
    return A


def _t(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _n(t):
    return t.detach().cpu().numpy()


def _drivers(M, Na, Nb, G, normalise=True):
    from omegaqmc_amd.synth import make_system
    from omegaqmc_amd.driver import AFQMCDriver
    s = make_system(M, Na, Nb, G, seed=42, normalise=normalise)
    integ = {'h1e': _t(s.h1e), 'h1e_mod': _t(s.h1e_mod), 'chol': _t(s.chol), 'enuc': s.enuc,
             'nbasis': M, 'nup': Na, 'ndown': Nb}
    return (s, AFQMCDriver(integ, verbose=False, precision="fp64"),
            AFQMCDriver(integ, verbose=False, precision="bf16x3"))


@pytest.mark.parametrize("M,Na,Nb,G,W", [(40, 5, 4, 120, 64), (64, 8, 7, 200, 130),
                                         (100, 13, 13, 300, 32)])
def test_one_step_close_to_oracle(afq, M, Na, Nb, G, W):
    s, d64, dbf = _drivers(M, Na, Nb, G)
    od = O.Driver(s.h1e, s.h1e_mod, s.chol, s.enuc, Na, Nb)
    rng = np.random.default_rng(3)
    pa = od.trial_up + 0.1 * (rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na)))
    pb = od.trial_dn + 0.1 * (rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb)))
    w = np.ones(W)
    eh = np.zeros(W, complex)
    xi = P.aux_fields(5, 1, W, G)
    eshift = od.ehf - od.e_const
    ref = O.propagate_walkers(pa, pb, w, None, eh, od.propagator, od.chol, od.rchol_a, od.rchol_b,
                              od.trial_up, od.trial_dn, eshift, xi)
    ta, tb, tw, te = _t(pa), _t(pb), _t(w), _t(eh)
    ov = dbf.system.step(ta, tb, tw, te, eshift, xi=_t(xi))
    for x, y, tol in ((ta, ref[0], 1e-4), (tb, ref[1], 1e-4), (ov, ref[3], 1e-3),
                      (tw, ref[2], 1e-4)):
        assert rel_err(_n(x), y) < tol


def test_trajectory_energy_within_budget(afq):
    """20 steps (QR every 5, comb every 10) with identical fields; at each comb
    the BF16x3 run replays the FP64 run's survivor indices (same trajectory).
    The mixed estimator of the BF16x3 walkers stays within 0.5 mHa of FP64."""
    import torch
    s, d64, dbf = _drivers(60, 8, 8, 240)
    W = 128
    s64, sbf = d64.init_walkers(W), dbf.init_walkers(W)
    eshift = d64.ehf - d64.e_const
    for step in range(1, 21):
        d64.step(s64, step, 42, eshift, 5, 10)
        dbf.step(sbf, step, 42, eshift, 5, 10, comb_idx=s64['idx'] if step % 10 == 0 else None)
    d64.block_energy(s64)
    dbf.block_energy(sbf)
    torch.cuda.synchronize()
    e = {"fp64": float(s64['stats'][0].item()), "bf16x3": float(sbf['stats'][0].item())}
    assert np.isfinite(e["fp64"])
    assert abs(e["bf16x3"] - e["fp64"]) < 0.5e-3, e


@pytest.mark.parametrize("M,Na,Nb,G,W", [(40, 5, 4, 120, 64), (100, 40, 38, 130, 37),
                                         (60, 16, 1, 90, 21), (64, 40, 40, 200, 130),
                                         (256, 36, 32, 97, 23)])
def test_block_energy_close_to_oracle(afq, M, Na, Nb, G, W):
    """Mixed-estimator local energy with the split-BF16 Coulomb and exchange
    contractions (ragged row/walker tiles, N = 1 .. 40; M = 256 with N = 36, 32
    runs the exchange's wide shape, the others its narrow one) against the
    FP64 oracle.  E_2b = E_coul - E_exch cancels two large terms, so the bound is
    per walker relative to |E_coul| + |E_exch|: 3e-5 (BF16x3 products carry
    ~2^-17 relative error, FP32 partial sums over K = N M add ~1e-6; these
    walkers are far from the trial, |Ghalf| up to 35).  The
    error itself must be the split's own: a float64 emulation of the same
    hi/lo products lands within a third of the emulated Coulomb and exchange
    errors (summed: they can cancel in E_2b), plus 2e-7 of the scale for the
    FP32 accumulation."""
    s, d64, dbf = _drivers(M, Na, Nb, G)
    od = O.Driver(s.h1e, s.h1e_mod, s.chol, s.enuc, Na, Nb)
    rng = np.random.default_rng(M + W)
    pa = od.trial_up + 0.2 * (rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na)))
    pb = od.trial_dn + 0.2 * (rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb)))
    Ga, Gb, Gha, Ghb, _ = O.greens_function(pa, pb, od.trial_up.astype(complex),
                                            od.trial_dn.astype(complex))
    rt, r1, r2 = O.local_energy(s.h1e, Ga, Gb, Gha, Ghb, od.rchol_a, od.rchol_b, s.enuc)
    et, e1, e2 = dbf.system.local_energy(_t(pa), _t(pb), _t(s.h1e), s.enuc)
    ec, ex = O.local_energy_2body(Gha, Ghb, od.rchol_a, od.rchol_b)
    scale = np.abs(ec) + np.abs(ex)
    assert rel_err(_n(e1), r1) < 1e-12
    assert np.all(np.abs(_n(e2) - r2) < 3e-5 * scale)
    assert np.all(np.abs(_n(et) - rt) < 3e-5 * scale + 1e-12 * np.abs(r1))
    mc, mx = bf16x3_two_body(Gha, Ghb, od.rchol_a, od.rchol_b)
    emu = mc - mx
    split_err = np.abs(mc - ec) + np.abs(mx - ex)  # the components' errors; they can cancel in E_2b
    assert np.all(np.abs(_n(e2) - emu) < 0.34 * split_err + 2e-7 * scale)
    ft, _, f2 = d64.system.local_energy(_t(pa), _t(pb), _t(s.h1e), s.enuc)
    assert rel_err(_n(f2), r2) < 1e-11


def _full_size_drivers(M, Na, Nb, G, trial=None):
    from omegaqmc_amd.synth import make_system
    from omegaqmc_amd.driver import AFQMCDriver
    s = make_system(M, Na, Nb, G, seed=42, rotate_device="cuda", normalise=True)
    integ = {'h1e': _t(s.h1e), 'h1e_mod': _t(s.h1e_mod), 'chol': _t(s.chol), 'enuc': s.enuc,
             'nbasis': M, 'nup': Na, 'ndown': Nb}
    return [AFQMCDriver(integ, verbose=False, precision=p, trial=trial)
            for p in ("fp64", "bf16x3")]


@pytest.mark.parametrize("workload", ["m300", "m150_32det"])
def test_lockstep_energy_full_length(afq, workload):
    """The measured workloads over their full 100 steps (1024 walkers, the
    reference driver's 25-step blocks and eshift updates, QR every 5, comb
    every 10 with the FP64 run's survivors replayed).  Through step 50 every
    checkpoint's block energy is within 0.1 mHa of the FP64 run of the same
    trajectory (the 0.5 mHa bar with margin).  Later the FP64 population
    develops walkers far in the tails of the local-energy distribution and
    the estimator amplifies any perturbation (an FP64 run perturbed by 1e-9
    moves by 1e-3 mHa there); the BF16x3 error at the last two block ends
    must stay below a tenth of the FP64 block energy's statistical error."""
    from omegaqmc_amd.driver import lockstep_energy_error
    from omegaqmc_amd.synth import multidet_trial
    if workload == "m300":
        ref, low = _full_size_drivers(300, 40, 40, 1800)
    else:
        ref, low = _full_size_drivers(150, 20, 18, 800, trial=multidet_trial(150, 20, 18, 32))
    r = lockstep_energy_error(ref, low, 1024, 100, every=10)
    print(f"[lockstep] {workload} block ends (step, dE mHa, FP64 sigma mHa, beyond cap): "
          f"{[(k, round(d * 1e3, 4), round(sg * 1e3, 2), n) for k, d, _, n, sg in r['block_ends']]}",
          flush=True)
    assert len(r['block_ends']) == 4
    assert all(d < 0.1e-3 for k, d, *_ in r['trace'] if k <= 50), r['trace']
    assert all(d < 0.1 * sg for k, d, _, _, sg in r['block_ends'] if k > 50), r['block_ends']


@pytest.mark.parametrize("M,Na,Nb,G,W", [(40, 5, 4, 120, 9), (64, 20, 19, 200, 5)])
def test_propagator_block_knob_same_result(afq, M, Na, Nb, G, W):
    """The propagator's column blocking (16, 32 or 48 columns per workgroup)
    changes which workgroup multiplies a column, not the products: one step
    gives bit-identical walkers and overlaps for every block width, and stays
    within the BF16x3 bound of the oracle."""
    s, d64, dbf = _drivers(M, Na, Nb, G)
    od = O.Driver(s.h1e, s.h1e_mod, s.chol, s.enuc, Na, Nb)
    rng = np.random.default_rng(8)
    pa = od.trial_up + 0.1 * (rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na)))
    pb = od.trial_dn + 0.1 * (rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb)))
    xi = P.aux_fields(6, 1, W, G)
    eshift = od.ehf - od.e_const
    ref = O.propagate_walkers(pa, pb, np.ones(W), None, np.zeros(W, complex), od.propagator,
                              od.chol, od.rchol_a, od.rchol_b, od.trial_up, od.trial_dn, eshift,
                              xi)
    outs = []
    for cols in (48, 32, 16, 48):
        dbf.system.set_propagator_block(cols)
        ta, tb, tw, te = _t(pa), _t(pb), _t(np.ones(W)), _t(np.zeros(W, complex))
        ov = dbf.system.step(ta, tb, tw, te, eshift, xi=_t(xi))
        outs.append((_n(ta), _n(tb), _n(ov), _n(tw)))
        assert rel_err(outs[-1][0], ref[0]) < 1e-4 and rel_err(outs[-1][1], ref[1]) < 1e-4
    for o in outs[1:]:
        for x, y in zip(o, outs[0]):
            assert np.array_equal(x, y)
    with pytest.raises(afq.AFQError):
        dbf.system.set_propagator_block(24)
