"""Uneven batch of independent systems (afq_batch): every mode -- one stream,
eager streams, hipGraph cycles -- reproduces the per-system device driver
bit for bit (same kernels, same fields, comb offsets and energy shifts), and
a batch member follows the FP64 oracle's driver trajectory at 1e-10."""

import numpy as np
import pytest

from afq_cases import rel_err
from oracle import afqmc as O
from oracle import philox as P

pytestmark = pytest.mark.gpu

SPECS = [(40, 5, 5, 200, 128, "fp64"), (56, 7, 7, 280, 96, "bf16x3"), (24, 3, 2, 60, 200, "fp64"),
         (72, 9, 9, 360, 64, "fp32")]


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


def _drivers(specs):
    from omegaqmc_amd.driver import AFQMCDriver
    from omegaqmc_amd.synth import make_system
    out = []
    for k, (M, Na, Nb, G, W, prec) in enumerate(specs):
        s = make_system(M, Na, Nb, G, seed=100 + k, normalise=True)
        integ = {'h1e': _t(s.h1e), 'h1e_mod': _t(s.h1e_mod), 'chol': _t(s.chol), 'enuc': s.enuc,
                 'nbasis': M, 'nup': Na, 'ndown': Nb}
        out.append((s, AFQMCDriver(integ, verbose=False, precision=prec)))
    return out


@pytest.mark.parametrize("mode", ["sequential", "streams", "graph"])
def test_batch_matches_per_system_driver(afq, mode):
    import torch
    from omegaqmc_amd.batch import Batch
    drv = _drivers(SPECS)
    walkers = [sp[4] for sp in SPECS]
    seeds = [7 + 3 * k for k in range(len(SPECS))]
    eshift = [d.ehf - d.e_const for _, d in drv]
    b = Batch([d.system for _, d in drv], walkers, seeds=seeds, nstreams=3,
              stream_of=[0, 1, 2, 1])
    b.set_eshift(eshift)
    nsteps = 47  # 4 whole driver cycles plus a partial one (QR at 5, comb at 10)
    b.run(nsteps, mode)
    torch.cuda.synchronize()
    for i, (_, d) in enumerate(drv):
        st = d.init_walkers(walkers[i])
        for step in range(1, nsteps + 1):
            d.step(st, step, seeds[i], eshift[i], 5, 10)
        got = b.state(i)
        torch.cuda.synchronize()
        for k in ('phia', 'phib', 'weights', 'e_hybrid', 'acc'):
            assert torch.equal(got[k], st[k]), (i, mode, k)


def test_batch_graph_restarts_and_eshift(afq):
    """Several run() calls in graph mode (the captured cycle replayed with the
    step counter and energy shift read from device memory) equal one long
    eager run with the same eshift changes between calls."""
    import torch
    from omegaqmc_amd.batch import Batch
    drv = _drivers(SPECS[:3])
    walkers = [sp[4] for sp in SPECS[:3]]
    seeds = [11, 12, 13]
    sysl = [d.system for _, d in drv]
    bg = Batch(sysl, walkers, seeds=seeds, nstreams=2)
    be = Batch(sysl, walkers, seeds=seeds, nstreams=1)
    for chunk, es in ((25, -1.0), (30, -2.5), (15, -0.3)):
        e = [es + 0.1 * i for i in range(3)]
        for b, mode in ((bg, "graph"), (be, "sequential")):
            b.set_eshift(e)
            b.run(chunk, mode)
    torch.cuda.synchronize()
    for i in range(3):
        a, c = bg.state(i), be.state(i)
        torch.cuda.synchronize()
        for k in ('phia', 'phib', 'weights', 'e_hybrid', 'acc'):
            assert torch.equal(a[k], c[k]), (i, k)


def test_batch_member_follows_oracle(afq):
    """System 0 of a graph-mode batch (M=40, N=5+5, Nchol=200, 128 walkers)
    against the oracle driver: 30 steps, walkers and weights at 1e-10."""
    import torch
    from omegaqmc_amd.batch import Batch
    drv = _drivers(SPECS)
    s, d = drv[0]
    walkers = [sp[4] for sp in SPECS]
    seeds = [42, 1, 2, 3]
    eshift = [dd.ehf - dd.e_const for _, dd in drv]
    b = Batch([dd.system for _, dd in drv], walkers, seeds=seeds, nstreams=2)
    b.set_eshift(eshift)
    b.run(30, "graph")
    torch.cuda.synchronize()
    od = O.Driver(s.h1e, s.h1e_mod, s.chol, s.enuc, s.nup, s.ndown)
    W, G = walkers[0], s.naux
    pa = np.tile(od.trial_up[None].astype(complex), (W, 1, 1))
    pb = np.tile(od.trial_dn[None].astype(complex), (W, 1, 1))
    w = np.ones(W)
    eh = np.zeros(W, complex)
    for step in range(1, 31):
        if step % 5 == 0:
            pa, pb, _, _ = O.orthogonalize_walkers(pa, pb)
        xi = P.aux_fields(42, step, W, G)
        pa, pb, w, _, eh = O.propagate_walkers(pa, pb, w, None, eh, od.propagator, od.chol,
                                               od.rchol_a, od.rchol_b, od.trial_up, od.trial_dn,
                                               eshift[0], xi)
        if step > 1:
            w = np.clip(w, 0.0, np.sum(np.abs(w)) * O.WEIGHT_CLIP_FRACTION)
        if step % 10 == 0:
            w, pa, pb, eh, _ = O.population_control_comb(w, pa, pb, P.comb_offset(42, step), eh)
    got = b.state(0)
    for k, ref in (('phia', pa), ('phib', pb), ('weights', w), ('e_hybrid', eh)):
        assert rel_err(_n(got[k]), ref) < 1e-10, k


def test_driver_graph_mode_bit_identical(afq):
    """The LiH-like workload (M=11, N=2+2, Nchol=50, 256 walkers, 400 steps,
    QR every 5, comb every 10) through the driver with graph=True -- each
    block's steps as hipGraph replays of a one-system batch -- gives the
    eager driver's block energies and final walkers bit for bit."""
    import torch
    from omegaqmc_amd.driver import AFQMCDriver
    from omegaqmc_amd.synth import make_system
    s = make_system(11, 2, 2, 50, seed=42)
    integ = {'h1e': _t(s.h1e), 'h1e_mod': _t(s.h1e_mod), 'chol': _t(s.chol), 'enuc': s.enuc,
             'nbasis': 11, 'nup': 2, 'ndown': 2}
    args = dict(rng_key=afq.PhiloxKey(42), num_walkers=256, num_blocks=12,
                num_steps_per_block=25, stabilize_freq=5, pop_control_freq=10,
                num_blocks_equil=4)
    eager = AFQMCDriver(integ, verbose=False)(**args)
    graph = AFQMCDriver(integ, verbose=False)(graph=True, **args)
    torch.cuda.synchronize()
    assert np.array_equal(graph['energy_blocks'], eager['energy_blocks'])
    for k in ('phia', 'phib', 'weights', 'e_hybrid'):
        assert torch.equal(graph['state'][k], eager['state'][k]), k


def test_driver_graph_mode_refuses_multidet(afq):
    from omegaqmc_amd.driver import AFQMCDriver
    from omegaqmc_amd.synth import make_system
    s = make_system(12, 3, 2, 30, seed=4)
    integ = {'h1e': _t(s.h1e), 'h1e_mod': _t(s.h1e_mod), 'chol': _t(s.chol), 'enuc': s.enuc,
             'nbasis': 12, 'nup': 3, 'ndown': 2}
    trial = {'occ_up': np.array([[0, 1, 2], [0, 1, 3]]), 'occ_dn': np.array([[0, 1], [0, 2]]),
             'ci_coeffs': np.array([0.9, 0.3], dtype=complex)}
    d = AFQMCDriver(integ, verbose=False, trial=trial, multidet_method='explicit')
    with pytest.raises(ValueError):
        d(num_walkers=8, num_blocks=1, num_steps_per_block=2, num_blocks_equil=0, graph=True)
