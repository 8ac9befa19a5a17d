"""Orbital counts beyond the fused FP64 propagator (M = 528 > 512, whose
column block does not fit the kernel's register tiling): the batched-GEMM
form of expH1 . exp(VHS) . expH1 in the single-det and multi-det steps, and
of the photon's scaled exponential, against the oracle at 1e-10."""

import itertools

import numpy as np
import pytest

from afq_cases import rel_err, walkers_near_trial
from oracle import afqmc as O
from oracle import multidet as MD
from oracle import philox as P
from oracle import qed as Q

pytestmark = pytest.mark.gpu

M, NA, NB, G, W = 528, 6, 5, 40, 4


@pytest.fixture(scope="module")
def afq(afq_lib):
    import torch
    import omegaqmc_amd as A
    torch.cuda.set_device(0)
    return A


@pytest.fixture(scope="module")
def system():
    from omegaqmc_amd.synth import make_system
    return make_system(M, NA, NB, G, seed=3, normalise=True)


def _t(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _n(t):
    return t.detach().cpu().numpy()


def test_step_large_m(afq, system):
    s = system
    tu, td = s.trial_up, s.trial_dn
    prop = O.build_propagator(s.h1e_mod, s.chol, tu, td, 0.01)
    ra, rb = O.half_rotate_cholesky(s.chol, tu, td)
    rng = np.random.default_rng(5)
    pa = walkers_near_trial(tu.astype(complex), W, rng, 0.05)
    pb = walkers_near_trial(td.astype(complex), W, rng, 0.05)
    w = np.ones(W)
    eh = np.zeros(W, complex)
    xi = P.aux_fields(1, 1, W, G)
    ref = O.propagate_walkers(pa, pb, w, None, eh, prop, s.chol, ra.real, rb.real, tu, td, -1.0,
                              xi)
    propd = afq.build_propagator(_t(s.h1e_mod), _t(s.chol), _t(tu), _t(td), 0.01)
    rga, rgb = afq.half_rotate_cholesky(_t(s.chol), _t(tu), _t(td))
    out = afq.propagate_walkers(_t(pa), _t(pb), _t(w), None, _t(eh), propd, _t(s.chol), rga, rgb,
                                _t(tu), _t(td), -1.0, None, xi=_t(xi))
    for k in range(5):
        assert rel_err(_n(out[k]), ref[k]) < 1e-10, k


def test_multidet_step_large_m(afq, system):
    s = system
    rng = np.random.default_rng(6)
    ups = list(itertools.combinations(range(NA + 3), NA))
    dns = list(itertools.combinations(range(NB + 3), NB))
    pairs = [(tuple(range(NA)), tuple(range(NB)))]
    while len(pairs) < 5:
        p = (ups[rng.integers(len(ups))], dns[rng.integers(len(dns))])
        if p not in pairs:
            pairs.append(p)
    ou = np.array([p[0] for p in pairs], dtype=np.int32)
    od = np.array([p[1] for p in pairs], dtype=np.int32)
    ci = np.exp(-np.arange(5) / 4.0) * (rng.normal(size=5) + 0.2j * rng.normal(size=5))
    tu, td = MD.trials_from_occ(M, ou), MD.trials_from_occ(M, od)
    prop = MD.build_propagator_multidet(s.h1e_mod, s.chol, tu, td, ci, 0.01)
    propd = {'expH1': _t(prop['expH1'].real.copy()), 'mf_shift': _t(prop['mf_shift']), 'dt': 0.01}
    pa = walkers_near_trial(np.eye(M, NA).astype(complex), W, rng, 0.05)
    pb = walkers_near_trial(np.eye(M, NB).astype(complex), W, rng, 0.05)
    w = np.ones(W)
    eh = np.zeros(W, complex)
    xi = P.aux_fields(2, 1, W, G)
    ref = MD.propagate_walkers_multidet(pa, pb, w, None, eh, prop, s.chol, tu, td, ci, -1.0, xi)
    out = afq.propagate_walkers_multidet(_t(pa), _t(pb), _t(w), None, _t(eh), propd, _t(s.chol),
                                         _t(ou), _t(od), _t(ci), -1.0, None, xi=_t(xi),
                                         method="explicit")
    for k in range(5):
        assert rel_err(_n(out[k]), ref[k]) < 1e-10, k


def test_scaled_exponential_large_m(afq, system):
    from omegaqmc_amd import qed
    rng = np.random.default_rng(7)
    d = rng.normal(size=(M, M))
    d = 0.02 * (d + d.T)
    c = 0.05 * rng.normal(size=W)
    phi = rng.normal(size=(W, M, 3)) + 1j * rng.normal(size=(W, M, 3))
    got = qed.apply_exp_scaled(_t(c), _t(d), _t(phi), 4)
    assert rel_err(_n(got), Q.apply_exp_vhs_photon(c, d, phi, 4)) < 1e-12
