"""FP32-MFMA step (v_mfma_f32_16x16x4f32 for the force bias, the VHS build and
the fused propagator; FP64 Green's function, determinants, weights and
walkers) against the FP64 oracle.

FP32 products carry ~2^-24 relative rounding with FP32 accumulation over K
terms, so one step stays within 1e-5 of the oracle in the walkers and the
weights (the FP64 path is held to 1e-10).  Shapes cover every row-tile
instantiation (k_prop<float, 1..3>), ragged walker and column tiles and an
open shell.
"""

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


def _driver(s, precision):
    from omegaqmc_amd.driver import AFQMCDriver
    integ = {'h1e': _t(s.h1e), 'h1e_mod': _t(s.h1e_mod), 'chol': _t(s.chol), 'enuc': s.enuc,
             'nbasis': s.nbasis, 'nup': s.nup, 'ndown': s.ndown}
    return AFQMCDriver(integ, verbose=False, precision=precision)


@pytest.mark.parametrize("M,Na,Nb,G,W", [(40, 5, 4, 120, 37), (120, 16, 16, 600, 16),
                                         (150, 20, 18, 800, 9), (300, 40, 40, 1800, 8),
                                         (64, 17, 0, 90, 21)])
def test_fp32_step_close_to_oracle(afq, M, Na, Nb, G, W):
    from omegaqmc_amd.synth import make_system
    s = make_system(M, Na, Nb, G, seed=42, normalise=True,
                    rotate_device="cuda" if M >= 150 else None)
    gd = _driver(s, "fp32")
    tu, td = np.eye(M, Na), np.eye(M, Nb)
    ra, rb = s.chol[:, :Na, :], s.chol[:, :Nb, :]
    prop = O.build_propagator(s.h1e_mod, s.chol, tu, td, 0.005)
    rng = np.random.default_rng(M + W)
    pa = tu + 0.05 * (rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na)))
    pb = td + 0.05 * (rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb)))
    w = rng.uniform(0.5, 1.5, W)
    eh = rng.normal(size=W) + 0.1j * rng.normal(size=W)
    xi = P.aux_fields(5, 1, W, G)
    eshift = -1.0
    ref = O.propagate_walkers(pa, pb, w, None, eh, prop, s.chol, ra, rb, tu, td, eshift, xi)
    ta, tb, tw, te = _t(pa), _t(pb), _t(w), _t(eh)
    ov = gd.system.step(ta, tb, tw, te, eshift, xi=_t(xi))
    tol = {"phia": 1e-5, "phib": 1e-5, "weights": 1e-5, "overlap": 1e-4, "e_hybrid": 1e-3}
    for name, x, y in (("phia", ta, ref[0]), ("phib", tb, ref[1]), ("weights", tw, ref[2]),
                       ("overlap", ov, ref[3]), ("e_hybrid", te, ref[4])):
        if y.size == 0:
            continue
        assert rel_err(_n(x), y) < tol[name], (name, rel_err(_n(x), y))


def test_fp32_trajectory_follows_fp64(afq):
    """20 driver steps (QR every 5, comb every 10) at the M=120, N=16+16,
    Nchol=600 shape with identical fields: the FP32 run replays the FP64
    run's survivors at each comb and its block energy stays within 0.05 mHa
    of the FP64 run of the same trajectory."""
    import torch
    from omegaqmc_amd.synth import make_system
    from omegaqmc_amd.driver import lockstep_energy_error
    s = make_system(120, 16, 16, 600, seed=42, normalise=True)
    ref, low = _driver(s, "fp64"), _driver(s, "fp32")
    r = lockstep_energy_error(ref, low, 256, 20, every=10, num_steps_per_block=10)
    torch.cuda.synchronize()
    assert r['max_raw'] < 0.05e-3, r['trace']
