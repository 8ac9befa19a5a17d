"""Oracle parity at the shapes the measured workloads run.

The kernels are instantiated by shape (row tiles of M, register rows of N,
split-K widths of G), so parity at small sizes does not reach the
instantiations behind the measured numbers.  These tests run them at their
real shapes, with few walkers so the NumPy oracle finishes in seconds:

* M=300, N=40+40, Nchol=1800 (k_prop_f64<3>, k_prop_bf3<3>, the split-BF16
  GEMMs at full K, k_greens_gj<40>, k_exch_bf3): one FP64 step at 1e-10, one
  BF16x3 step at 1e-4, the BF16x3 block energy within 3e-5 of the component
  scale and matching a float64 emulation of the split products.
* M=150, N=20+18, Nchol=800, 32 determinants (k_prop_bf3<2>, k_md_gj): one
  step in FP64 (1e-10) and BF16x3 (1e-4).
* M=120, N=16+16, Nchol=600: a 10-step FP64 driver trajectory at 1e-10.
* M=11, N=2+2, Nchol=50, 256 walkers, 400 steps (QR every 5, comb every
  10): every comb's survivor indices bit-exact, walkers at 1e-10.

The size-normalised generator is used at M >= 120 (with the literal one the
overlap overflows FP64 in the first step, reference semantics).
"""

import numpy as np
import pytest

from afq_cases import bf16x3_two_body, rel_err
from oracle import afqmc as O
from oracle import multidet as MD
from oracle import philox as P

pytestmark = pytest.mark.gpu

DT = 0.005


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


class _Oracle:
    """The oracle driver's inputs for an identity trial, formed without the
    O(G N M^2) half-rotation einsum (for trial = eye(M, N) it is the row
    slice L[:, :N, :] exactly)."""

    def __init__(self, s):
        M, Na, Nb = s.nbasis, s.nup, s.ndown
        self.trial_up = np.eye(M, Na)
        self.trial_dn = np.eye(M, Nb)
        self.chol = s.chol
        self.rchol_a = np.ascontiguousarray(s.chol[:, :Na, :])
        self.rchol_b = np.ascontiguousarray(s.chol[:, :Nb, :])
        self.propagator = O.build_propagator(s.h1e_mod, s.chol, self.trial_up, self.trial_dn, DT)
        self.e_const = O.e_const(s.enuc, self.propagator['mf_shift'])
        self.ehf = O.determinant_energy_from_integrals(s.h1e, s.chol, s.enuc, Na, Nb)


def _gpu_driver(s, precision):
    from omegaqmc_amd.driver import AFQMCDriver
    integ = {'h1e': _t(s.h1e), 'h1e_mod': _t(s.h1e_mod), 'chol': _t(s.chol), 'enuc': s.enuc,
             'nbasis': s.nbasis, 'nup': s.nup, 'ndown': s.ndown}
    return AFQMCDriver(integ, verbose=False, precision=precision, dt=DT)


def _walkers(od, W, seed, scale):
    rng = np.random.default_rng(seed)
    M, Na = od.trial_up.shape
    Nb = od.trial_dn.shape[1]
    pa = od.trial_up + scale * (rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na)))
    pb = od.trial_dn + scale * (rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb)))
    w = rng.uniform(0.5, 1.5, W)
    eh = rng.normal(size=W) + 0.1j * rng.normal(size=W)
    return pa, pb, w, eh


@pytest.fixture(scope="module")
def m300(afq):
    from omegaqmc_amd.synth import make_system
    print("[case] building the M=300 N=40+40 G=1800 system", flush=True)
    s = make_system(300, 40, 40, 1800, seed=42, rotate_device="cuda", normalise=True)
    return s, _Oracle(s)


def _one_step(afq, s, od, precision, W=8, seed=3):
    gd = _gpu_driver(s, precision)
    pa, pb, w, eh = _walkers(od, W, seed, 0.05)
    xi = P.aux_fields(5, 1, W, s.naux)
    eshift = od.ehf - od.e_const
    ref = O.propagate_walkers(pa, pb, w, None, eh, od.propagator, od.chol, od.rchol_a,
                              od.rchol_b, od.trial_up, od.trial_dn, eshift, xi)
    ta, tb, tw, te = _t(pa), _t(pb), _t(w), _t(eh)
    ov = gd.system.step(ta, tb, tw, te, eshift, xi=_t(xi))
    return {"phia": (_n(ta), ref[0]), "phib": (_n(tb), ref[1]), "weights": (_n(tw), ref[2]),
            "overlap": (_n(ov), ref[3]), "e_hybrid": (_n(te), ref[4])}


def test_fp64_step_m300(afq, m300):
    """k_prop_f64<3>, k_vhs_f64 and k_fb_f64 at G=1800, k_greens_gj<40>."""
    s, od = m300
    out = _one_step(afq, s, od, "fp64")
    for name, (x, y) in out.items():
        assert rel_err(x, y) < 1e-10, (name, rel_err(x, y))


def test_bf16x3_step_m300(afq, m300):
    """k_prop_bf3<3> and k_bf3_gemm<0/1> at G=1800: split-BF16 products carry
    ~2^-17 relative error per product; the walkers after one step stay within
    1e-4 of the FP64 oracle, the overlap (a product of two 40x40
    determinants) within 1e-3, e_hybrid (the log overlap ratio over dt)
    within 1e-2 relative."""
    s, od = m300
    out = _one_step(afq, s, od, "bf16x3")
    tol = {"phia": 1e-4, "phib": 1e-4, "weights": 1e-4, "overlap": 1e-3, "e_hybrid": 1e-2}
    for name, (x, y) in out.items():
        assert rel_err(x, y) < tol[name], (name, rel_err(x, y))


def test_bf16x3_block_energy_m300(afq, m300):
    """k_exch_bf3 at N=40, K=M=300 over G=1800 rows: per-walker two-body
    energy within 3e-5 of |E_coul| + |E_exch|, the error the split's own (a
    float64 emulation of the same hi/lo products lands within a third of it,
    plus 2e-7 of the scale for the FP32 accumulation); the FP64 system at
    1e-10."""
    s, od = m300
    W = 8
    pa, pb, _, _ = _walkers(od, W, 11, 0.1)
    Ga, Gb, Gha, Ghb, _ = O.greens_function(pa, pb, od.trial_up.astype(complex),
                                            od.trial_dn.astype(complex))
    e1r = O.local_energy_1body(s.h1e, Ga, Gb, s.enuc)
    ec, ex = O.local_energy_2body_streamed(Gha, Ghb, od.rchol_a, od.rchol_b, e_chunk_g=300)
    r2 = ec - ex
    scale = np.abs(ec) + np.abs(ex)
    bf = _gpu_driver(s, "bf16x3")
    et, e1, e2 = (_n(x) for x in bf.system.local_energy(_t(pa), _t(pb), _t(s.h1e), s.enuc))
    assert rel_err(e1, e1r) < 1e-12
    assert np.all(np.abs(e2 - r2) < 3e-5 * scale), np.max(np.abs(e2 - r2) / scale)
    # E_coul = 1/2 sum_g X_g^2 and E_exch are sums over g: emulated per g slab
    emu = 0.0
    for g0 in range(0, s.naux, 300):
        c, x = bf16x3_two_body(Gha, Ghb, od.rchol_a[g0:g0 + 300], od.rchol_b[g0:g0 + 300])
        emu = emu + (c - x)
    assert np.all(np.abs(e2 - emu) < 0.34 * np.abs(emu - r2) + 2e-7 * scale)
    fp = _gpu_driver(s, "fp64")
    ft, f1, f2 = (_n(x) for x in fp.system.local_energy(_t(pa), _t(pb), _t(s.h1e), s.enuc))
    assert rel_err(f2, r2) < 1e-10 and rel_err(ft, e1r + r2) < 1e-10


@pytest.fixture(scope="module")
def m150_md(afq):
    from omegaqmc_amd.synth import make_system, multidet_trial
    M, Na, Nb, G = 150, 20, 18, 800
    print("[case] building the M=150 N=20+18 G=800 32-det system", flush=True)
    s = make_system(M, Na, Nb, G, seed=42, rotate_device="cuda", normalise=True)
    tr = multidet_trial(M, Na, Nb, 32)
    tu, td = MD.trials_from_occ(M, tr['occ_up']), MD.trials_from_occ(M, tr['occ_dn'])
    return s, tr, tu, td


@pytest.mark.parametrize("method", ["ratio", "explicit"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("bf16x3", 1e-4)])
def test_multidet_step_m150(afq, m150_md, precision, tol, method):
    """One multi-determinant step (k_md_gj<20>, k_md_gsym, the force bias over
    the packed triangle, k_prop_f64<2> / k_prop_bf3<2>) at the 32-det
    workload's shape against oracle/multidet.py."""
    s, tr, tu, td = m150_md
    ci = tr['ci_coeffs']
    prop = MD.build_propagator_multidet(s.h1e_mod, s.chol, tu, td, ci, DT)
    propd = {'expH1': _t(prop['expH1'].real.copy()), 'mf_shift': _t(prop['mf_shift']), 'dt': DT}
    W = 8
    rng = np.random.default_rng(21)
    pa = tu[0] + 0.05 * (rng.normal(size=(W, 150, 20)) + 1j * rng.normal(size=(W, 150, 20)))
    pb = td[0] + 0.05 * (rng.normal(size=(W, 150, 18)) + 1j * rng.normal(size=(W, 150, 18)))
    w = rng.uniform(0.5, 1.5, W)
    eh = rng.normal(size=W) + 0.1j * rng.normal(size=W)
    xi = P.aux_fields(9, 1, W, s.naux)
    ref = MD.propagate_walkers_multidet(pa, pb, w, None, eh, prop, s.chol, tu, td, ci, -200.0, xi)
    out = afq.propagate_walkers_multidet(_t(pa), _t(pb), _t(w), None, _t(eh), propd, _t(s.chol),
                                         _t(tr['occ_up']), _t(tr['occ_dn']), _t(ci), -200.0, None,
                                         xi=_t(xi), precision=precision, method=method)
    for k, name in enumerate(("phia", "phib", "weights", "overlap", "e_hybrid")):
        lim = tol if k < 3 else (10 * tol if precision == "fp64" else 100 * tol)
        assert rel_err(_n(out[k]), ref[k]) < lim, (name, rel_err(_n(out[k]), ref[k]))


def _trajectory(afq, s, W, nsteps, seed=42, tol=1e-10, check_every=1, energy_every=25):
    """Driver trajectory on the GPU and in the oracle with the same fields and
    comb offsets.  After every ``check_every`` steps the walkers, weights and
    overlaps agree within ``tol``; every ``energy_every`` steps the block
    (mixed-estimator) energy agrees within ``tol``; comb indices bit-exact.

    e_hybrid = -(log(ovlp'/ovlp) + cfb + cmf)/dt turns the rounding of the
    log overlap ratio (~1e-12 absolute after hundreds of steps) into 1/dt =
    200 times that in absolute terms, so it is held to 100 tol relative."""
    od = _Oracle(s)
    gd = _gpu_driver(s, "fp64")
    st = gd.init_walkers(W)
    pa = np.tile(od.trial_up[None].astype(complex), (W, 1, 1))
    pb = np.tile(od.trial_dn[None].astype(complex), (W, 1, 1))
    w = np.ones(W)
    ov = np.ones(W, complex)
    eh = np.zeros(W, complex)
    eshift = od.ehf - od.e_const
    lim = {"phia": tol, "phib": tol, "weights": tol, "overlap": tol, "e_hybrid": 100 * tol,
           "energy": tol}
    worst = dict.fromkeys(lim, 0.0)
    ncomb = 0
    for step in range(1, nsteps + 1):
        if step % 5 == 0:
            pa, pb, _, _ = O.orthogonalize_walkers(pa, pb)
        xi = P.aux_fields(seed, step, W, s.naux)
        pa, pb, w, ov, eh = O.propagate_walkers(pa, pb, w, ov, eh, od.propagator, od.chol,
                                                od.rchol_a, od.rchol_b, od.trial_up, od.trial_dn,
                                                eshift, xi)
        if step > 1:
            w = np.clip(w, 0.0, np.sum(np.abs(w)) * O.WEIGHT_CLIP_FRACTION)
        ridx = None
        if step % 10 == 0:
            w, pa, pb, eh, ridx = O.population_control_comb(w, pa, pb, P.comb_offset(seed, step),
                                                            eh)
        gd.step(st, step, seed, eshift, 5, 10)
        if ridx is not None:
            assert np.array_equal(_n(st['idx']), ridx), step
            ncomb += 1
        errs = {}
        if step % check_every == 0 or step == nsteps:
            for name, x, y in (("phia", st['phia'], pa), ("phib", st['phib'], pb),
                               ("weights", st['weights'], w), ("e_hybrid", st['e_hybrid'], eh)):
                errs[name] = rel_err(_n(x), y)
            if ridx is None:  # the overlap is that of the walkers before a comb
                errs["overlap"] = rel_err(_n(st['overlap']), ov)
        if step % energy_every == 0:
            gd.block_energy(st)
            Ga, Gb, Gha, Ghb, _ = O.greens_function(pa, pb, od.trial_up, od.trial_dn)
            et, _, _ = O.local_energy(s.h1e, Ga, Gb, Gha, Ghb, od.rchol_a, od.rchol_b, s.enuc)
            wa = np.abs(w)
            eb = np.sum(wa * et.real) / np.sum(wa)
            errs["energy"] = abs(float(st['stats'][0].item()) - eb) / abs(eb)
        for name, e in errs.items():
            worst[name] = max(worst[name], e)
            assert e < lim[name], (step, name, e)
    print(f"[trajectory] M={s.nbasis} W={W} steps={nsteps} combs={ncomb} max rel err "
          + " ".join(f"{k}={v:.2e}" for k, v in worst.items()), flush=True)
    return worst


def test_trajectory_m120(afq):
    """M=120, N=16+16, Nchol=600 (the FP64/FP32 comparison workload's shape),
    16 walkers, 10 steps: QR at 5 and 10, comb at 10."""
    from omegaqmc_amd.synth import make_system
    s = make_system(120, 16, 16, 600, seed=42, normalise=True)
    _trajectory(afq, s, 16, 10, energy_every=5)


def test_trajectory_lih_like_full(afq):
    """The LiH-like workload in full: M=11, N=2+2, Nchol=50, 256 walkers, 400
    steps, QR every 5, comb every 10 (40 combs, indices bit-exact)."""
    from omegaqmc_amd.synth import make_system
    s = make_system(11, 2, 2, 50, seed=42)
    _trajectory(afq, s, 256, 400, check_every=7)
