"""Edge shapes in every precision: an empty beta spin, single orbitals,
orbital counts off the 16-row tiles, very few Cholesky vectors and single
walkers.  One step (walkers, weights, overlaps, hybrid energies) and the
local energy of the stepped walkers against the FP64 oracle, at the bounds
the per-precision tests use: FP64 1e-10 throughout; FP32-MFMA 1e-5 (walkers,
weights), 1e-4 (overlap), 1e-3 (hybrid energy, which divides the overlap
ratio's log by dt); BF16x3 1e-4, 1e-3, 1e-3.  Energy: the FP32 system's is
the FP64 estimator (1e-11); BF16x3 splits the two-body contractions (1e-4 of
the energy scale)."""

import numpy as np
import pytest

from afq_cases import rel_err
from oracle import afqmc as O
from oracle import philox as P

pytestmark = pytest.mark.gpu

SHAPES = [(37, 5, 0, 60, 3), (33, 1, 1, 40, 1), (17, 3, 2, 7, 5), (50, 2, 6, 3, 9)]
# (walkers, weights), overlap, hybrid energy
STEP_TOL = {"fp64": (1e-10, 1e-10, 1e-10), "fp32": (1e-5, 1e-4, 1e-3),
            "bf16x3": (1e-4, 1e-3, 1e-3)}
ENERGY_TOL = {"fp64": 1e-11, "fp32": 1e-11, "bf16x3": 1e-4}


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


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("M,Na,Nb,G,W", SHAPES)
def test_edge_shape_step_and_energy(afq, M, Na, Nb, G, W, precision):
    from omegaqmc_amd.driver import AFQMCDriver
    from omegaqmc_amd.synth import make_system
    s = make_system(M, Na, Nb, G, seed=M + G, normalise=True)
    integ = {'h1e': _t(s.h1e), 'h1e_mod': _t(s.h1e_mod), 'chol': _t(s.chol), 'enuc': s.enuc,
             'nbasis': M, 'nup': Na, 'ndown': Nb}
    d = AFQMCDriver(integ, verbose=False, precision=precision)
    od = O.Driver(s.h1e, s.h1e_mod, s.chol, s.enuc, Na, Nb)
    rng = np.random.default_rng(M * 7 + W)
    pa = od.trial_up + 0.1 * (rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na)))
    pb = od.trial_dn + 0.1 * (rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb)))
    w = rng.uniform(0.5, 1.5, size=W)
    eh = rng.normal(size=W) + 0.1j * rng.normal(size=W)
    xi = P.aux_fields(9, 1, W, G)
    eshift = od.ehf - od.e_const
    ref = O.propagate_walkers(pa, pb, w, None, eh, od.propagator, od.chol, od.rchol_a, od.rchol_b,
                              od.trial_up, od.trial_dn, eshift, xi)
    ta, tb, tw, te = _t(pa), _t(pb), _t(w), _t(eh)
    ov = d.system.step(ta, tb, tw, te, eshift, xi=_t(xi))
    tw_, to_, te_ = STEP_TOL[precision]
    for k, (x, y, tol) in enumerate(((ta, ref[0], tw_), (tb, ref[1], tw_), (tw, ref[2], tw_),
                                     (ov, ref[3], to_), (te, ref[4], te_))):
        if np.asarray(y).size:
            assert rel_err(_n(x), y) < tol, (k, rel_err(_n(x), y))
    Ga, Gb, Gha, Ghb, _ = O.greens_function(ref[0], ref[1], od.trial_up, od.trial_dn)
    eref = O.local_energy(s.h1e, Ga, Gb, Gha, Ghb, od.rchol_a, od.rchol_b, s.enuc)
    e_tot, e_1b, e_2b = d.system.local_energy(_t(ref[0]), _t(ref[1]), d.h1e, d.enuc)
    scale = max(np.max(np.abs(eref[0])), 1.0)
    assert np.max(np.abs(_n(e_tot) - eref[0])) / scale < ENERGY_TOL[precision]
