"""Per-walker numerical status flags (afq_walker_status) on the GPU.

Planted NaN/inf/zero values in host-made arrays are checked against a
NumPy restatement of the flag rules, bit for bit, and the summary word.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NONFINITE_PHI, ZERO_OV, NONFINITE_OV, BAD_W, NONFINITE_EH, ZERO_W = 1, 2, 4, 8, 16, 32


@pytest.fixture(scope="module")
def afq(afq_lib):
    import torch
    import omegaqmc_amd as A
    torch.cuda.set_device(0)
    return A


def _t(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _expected(pa, pb, ov, w, eh, tol):
    f = np.zeros(pa.shape[0], dtype=np.int64)
    f |= np.where(~np.isfinite(pa).all(axis=(1, 2)) | ~np.isfinite(pb).all(axis=(1, 2)),
                  NONFINITE_PHI, 0)
    fin = np.isfinite(ov)
    f |= np.where(~fin, NONFINITE_OV, 0)
    f |= np.where(fin & (np.abs(np.where(fin, ov, 1.0)) <= tol), ZERO_OV, 0)
    f |= np.where(~np.isfinite(w) | (w < 0), BAD_W, 0)
    f |= np.where(w == 0, ZERO_W, 0)
    f |= np.where(~np.isfinite(eh), NONFINITE_EH, 0)
    return f


@pytest.mark.parametrize("Nb", [2, 0])
def test_planted_flags(afq, Nb):
    rng = np.random.default_rng(0)
    W, M, Na = 12, 10, 3
    pa = rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na))
    pb = rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb))
    ov = rng.normal(size=W) + 1j * rng.normal(size=W)
    w = rng.uniform(0.5, 1.5, size=W)
    eh = rng.normal(size=W) + 0j
    pa[2, 3, 1] = np.nan
    if Nb:
        pb[4, 9, 1] = complex(0.0, np.inf)
    ov[5] = 0.0
    ov[6] = complex(np.nan, 0.0)
    ov[11] = 1e-310  # below the tolerance
    w[7] = -1.0
    w[8] = 0.0
    w[9] = np.inf
    eh[10] = complex(1.0, np.nan)
    flags, summ = afq.walker_status(_t(pa), _t(pb), _t(ov), _t(w), _t(eh), zero_overlap_tol=1e-300)
    got = flags.cpu().numpy().astype(np.int64)
    want = _expected(pa, pb, ov, w, eh, 1e-300)
    assert np.array_equal(got, want), (got, want)
    s = summ.cpu().numpy()
    assert int(s[0]) == int(np.bitwise_or.reduce(want))
    assert int(s[1]) == int(np.sum((want & ~ZERO_W) != 0))
    assert afq.decode_walker_flags(got[9]) == ["bad_weight"]


def test_optional_arrays_and_empty(afq):
    import torch
    rng = np.random.default_rng(1)
    W, M = 5, 6
    pa = rng.normal(size=(W, M, 2)) + 0j
    pb = rng.normal(size=(W, M, 2)) + 0j
    pb[1, 0, 0] = np.nan
    flags, summ = afq.walker_status(_t(pa), _t(pb))
    assert flags.cpu().numpy().tolist() == [0, NONFINITE_PHI, 0, 0, 0]
    assert summ.cpu().numpy().tolist() == [NONFINITE_PHI, 1]
    e = torch.zeros((0, M, 2), dtype=torch.complex128, device="cuda")
    flags, summ = afq.walker_status(e, e)
    assert flags.numel() == 0 and summ.cpu().numpy().tolist() == [0, 0]


def test_driver_flags_orthogonal_walker(afq):
    """A walker orthogonal to the trial has zero overlap: the driver's check
    flags exactly that walker."""
    import torch
    from omegaqmc_amd.synth import make_system
    from omegaqmc_amd.driver import AFQMCDriver
    s = make_system(10, 2, 2, 20, seed=3)
    integ = {'h1e': _t(s.h1e), 'h1e_mod': _t(s.h1e_mod), 'chol': _t(s.chol), 'enuc': s.enuc,
             'nbasis': 10, 'nup': 2, 'ndown': 2}
    gd = AFQMCDriver(integ, verbose=False)
    st = gd.init_walkers(8)
    pa = st['phia'].clone()
    pa[3] = 0.0
    pa[3, 5, 0] = 1.0  # occupies only orbitals outside the trial's
    pa[3, 6, 1] = 1.0
    st['phia'] = pa
    flags, summ = gd.walker_status(st)
    torch.cuda.synchronize()
    f = flags.cpu().numpy()
    assert f[3] == ZERO_OV and np.all(np.delete(f, 3) == 0)
    assert summ.cpu().numpy().tolist() == [ZERO_OV, 1]
    res = gd(num_walkers=16, num_blocks=1, num_steps_per_block=5, num_blocks_equil=0,
             check_status=True)
    assert len(res['status_blocks']) == 1 and res['status_blocks'][0][1] == 0
