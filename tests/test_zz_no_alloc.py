"""afq_population_control_comb_noalloc against afq_population_control_comb.

Both walker counts exceed what the running sum can keep in LDS, so the
static device array is what the allocation-free entry point uses; the second
call is smaller and must not see the first call's contents.  Indices must be
bit-exact, the gathered walkers and e_hybrid identical, and an oversized
walker count is rejected before any launch."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def afq(afq_lib):
    import torch
    import omegaqmc_amd as A
    torch.cuda.set_device(0)
    return A


def _run(lib, ctx, fn, w, pa, pb, eh, r):
    import torch
    from omegaqmc_amd._lib import check, ptr
    W, M, Na = pa.shape
    Nb = pb.shape[2]
    wn = torch.empty_like(w)
    pan, pbn, ehn = torch.empty_like(pa), torch.empty_like(pb), torch.empty_like(eh)
    idx = torch.empty(W, dtype=torch.int32, device=pa.device)
    check(getattr(lib, fn)(ctx, W, M, Na, Nb, ptr(w), ptr(pa), ptr(pb), r, ptr(eh), ptr(wn),
                           ptr(pan), ptr(pbn), ptr(ehn), ptr(idx)), fn)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (idx, wn, pan, pbn, ehn)]


def test_comb_noalloc_matches_comb_twice(afq):
    import torch
    from omegaqmc_amd import _lib
    from oracle import afqmc as O
    lib = _lib.load()
    ctx = _lib.context().handle
    rng = np.random.default_rng(12)
    for W, r in ((60000, 0.37), (25000, 0.81)):  # both beyond the LDS running sum
        print(f"[case] comb W={W} r={r}", flush=True)
        M, Na, Nb = 3, 1, 2
        w_np = rng.uniform(0.1, 2.0, W)
        w_np[rng.integers(0, W, 50)] = 0.0
        w = torch.from_numpy(w_np).cuda()
        pa = torch.from_numpy(rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na))).cuda()
        pb = torch.from_numpy(rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb))).cuda()
        eh = torch.from_numpy(rng.normal(size=W) + 1j * rng.normal(size=W)).cuda()
        ref = _run(lib, ctx, "afq_population_control_comb", w, pa, pb, eh, r)
        got = _run(lib, ctx, "afq_population_control_comb_noalloc", w, pa, pb, eh, r)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
        assert np.array_equal(got[0], O.comb_indices(w_np, r))


def test_comb_noalloc_rejects_oversize(afq):
    from omegaqmc_amd import _lib
    lib = _lib.load()
    ctx = _lib.context().handle
    W = (1 << 20) + 1
    rc = lib.afq_population_control_comb_noalloc(ctx, W, 3, 1, 1, None, None, None, 0.5, None,
                                                 None, None, None, None, None)
    assert rc == _lib.AFQ_OK - 1  # AFQ_EINVAL, before any launch
    assert b"static limit" in lib.afq_last_error()
