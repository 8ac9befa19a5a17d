"""Entry points of the Python mirror not reached by the step tests: the
one-body half step on its own, and the prepared-system cache of
propagate_walkers when the caller passes fresh arrays on every call."""

import numpy as np
import pytest

from afq_cases import rel_err
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


@pytest.mark.parametrize("W,M,N", [(5, 11, 2), (7, 64, 13), (3, 300, 40), (4, 150, 18)])
def test_apply_one_body_matches_oracle(afq, W, M, N):
    """afq_apply_one_body (afqmc_gto.py:678-679): expH1 @ phi per walker, the
    oracle's half step of propagate_walkers, at 1e-13."""
    from omegaqmc_amd.synth import make_system
    s = make_system(M, N, N, 2 * M, seed=M, normalise=True)
    prop = O.build_propagator(s.h1e_mod, s.chol, s.trial_up, s.trial_dn, 0.005)
    rng = np.random.default_rng(W + M)
    phi = rng.normal(size=(W, M, N)) + 1j * rng.normal(size=(W, M, N))
    out = afq.apply_one_body(_t(prop['expH1'].real.copy()), _t(phi))
    ref = np.matmul(prop['expH1'], phi)
    assert rel_err(_n(out), ref) < 1e-13


def test_propagate_walkers_with_fresh_arrays(afq):
    """Two calls with different temporary trials (NumPy arrays converted on
    every call): each step follows its own trial, never a cached system
    built from the other call's arrays."""
    from omegaqmc_amd.synth import make_system
    M, Na, Nb, G, W = 12, 3, 3, 30, 6
    s = make_system(M, Na, Nb, G, seed=8)
    rng = np.random.default_rng(1)
    prop_np = O.build_propagator(s.h1e_mod, s.chol, s.trial_up, s.trial_dn, 0.005)
    propd = {'expH1': _t(prop_np['expH1'].real.copy()), 'mf_shift': _t(prop_np['mf_shift']),
             'dt': 0.005}
    xi = P.aux_fields(3, 1, W, G)
    for k in range(3):
        # a different (orthonormal) trial per call, passed as fresh NumPy arrays
        q, _ = np.linalg.qr(np.eye(M, Na) + 0.2 * k * rng.normal(size=(M, Na)))
        tu = q.astype(complex)
        td = q.astype(complex)
        ra, rb = O.half_rotate_cholesky(s.chol, tu, td)
        pa = tu[None] + 0.05 * (rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na)))
        pb = td[None] + 0.05 * (rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb)))
        w = np.ones(W)
        eh = np.zeros(W, complex)
        ref = O.propagate_walkers(pa, pb, w, None, eh, prop_np, s.chol, ra.real, rb.real, tu, td,
                                  -1.0, xi)
        out = afq.propagate_walkers(pa, pb, w, None, eh, propd, s.chol, ra.real.copy(),
                                    rb.real.copy(), tu.copy(), td.copy(), -1.0, None, xi=_t(xi))
        for x, y in zip(out[:5], ref):
            assert rel_err(_n(x), y) < 1e-10, k
    assert len(propd['_systems']) <= afq.afqmc_gto.SYSTEM_CACHE_SIZE
