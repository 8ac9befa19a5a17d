"""Parity of each HIP entry point with the NumPy FP64 oracle (same inputs).

Tolerances: 1e-10 relative (max-abs error over max-abs value) for the FP64
path, as the plain FP64 GPU step must match the oracle; comb indices are
compared bit for bit.
"""

import numpy as np
import pytest

from afq_cases import make_harness, rel_err, walkers_near_trial
from oracle import afqmc as O
from oracle import philox as P

pytestmark = pytest.mark.gpu

TOL = 1e-10


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


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_greens_harness(afq, seed):
    tu, td, pa, pb, *_ = make_harness(seed)
    Ga, Gb, Gha, Ghb, ov = afq.greens_function(_t(pa), _t(pb), _t(tu), _t(td))
    rGa, rGb, rGha, rGhb, rov = O.greens_function(pa, pb, tu, td)
    for x, y in ((Ga, rGa), (Gb, rGb), (Gha, rGha), (Ghb, rGhb), (ov, rov)):
        assert rel_err(_n(x), y) < TOL


def test_force_bias_and_overlap_variants_match_full(afq):
    """test_afqmc_kernels.py:97-134: specialised variants equal the full one bit for bit."""
    rng = np.random.default_rng(7)
    M, Na, Nb, W = 12, 3, 2, 8
    tu = np.linalg.qr(rng.normal(size=(M, Na)))[0].astype(np.complex128)
    td = np.linalg.qr(rng.normal(size=(M, Nb)))[0].astype(np.complex128)
    pa, pb = walkers_near_trial(tu, W, rng), walkers_near_trial(td, W, rng)
    _, _, Gha, Ghb, ov = afq.greens_function(_t(pa), _t(pb), _t(tu), _t(td))
    Gha2, Ghb2, ov2 = afq.greens_function_force_bias(_t(pa), _t(pb), _t(tu), _t(td))
    ov3 = afq.greens_function_overlap(_t(pa), _t(pb), _t(tu), _t(td))
    assert np.array_equal(_n(Gha), _n(Gha2)) and np.array_equal(_n(Ghb), _n(Ghb2))
    assert np.array_equal(_n(ov), _n(ov2)) and np.array_equal(_n(ov), _n(ov3))
    assert rel_err(_n(ov), O.greens_function_overlap(pa, pb, tu, td)) < TOL


@pytest.mark.parametrize("M,N", [(16, 1), (40, 7), (64, 33), (70, 64)])
def test_greens_shapes(afq, M, N):
    rng = np.random.default_rng(M + N)
    tr = np.eye(M, N)
    W = 5
    pa = walkers_near_trial(tr, W, rng, 0.3)
    pb = walkers_near_trial(np.eye(M, max(N - 1, 0)), W, rng, 0.3)
    Gha, Ghb, ov = afq.greens_function_force_bias(_t(pa), _t(pb), _t(tr), _t(np.eye(M, max(N - 1, 0))))
    rGha, rGhb, rov = O.greens_function_force_bias(pa, pb, tr.astype(complex),
                                                   np.eye(M, max(N - 1, 0)).astype(complex))
    assert rel_err(_n(Gha), rGha) < TOL
    if N > 1:
        assert rel_err(_n(Ghb), rGhb) < TOL
    assert rel_err(_n(ov), rov) < TOL


def test_singular_overlap_gives_zero(afq):
    """A walker orthogonal to the trial: slogdet sign 0 -> overlap exactly 0."""
    M, N = 6, 2
    tr = np.eye(M, N).astype(complex)
    pa = np.zeros((1, M, N), complex)
    pa[0, 2, 0] = 1.0
    pa[0, 3, 1] = 1.0
    ov = afq.greens_function_overlap(_t(pa), _t(pa), _t(tr), _t(tr))
    assert _n(ov)[0] == 0


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("e_chunk_g", [4, 7, 24])
def test_local_energy_2body_streamed_harness(afq, seed, e_chunk_g):
    """test_afqmc_kernels.py:180-209 shape: complex rchol, streamed vs reference."""
    tu, td, pa, pb, ra, rb, h1e, enuc = make_harness(seed)
    _, _, Gha, Ghb, _ = O.greens_function(pa, pb, tu, td)
    ec_ref, ex_ref = O.local_energy_2body(Gha, Ghb, ra, rb)
    hta, htb = h1e @ tu.conj(), h1e @ td.conj()
    et, e1, e2 = afq.local_energy_streamed(_t(hta), _t(htb), _t(Gha), _t(Ghb), _t(ra), _t(rb), enuc,
                                           e_chunk_g)
    rt, r1, r2 = O.local_energy(h1e, *O.greens_function(pa, pb, tu, td)[:4], ra, rb, enuc)
    np.testing.assert_allclose(_n(e2), ec_ref - ex_ref, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(_n(e1), r1, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(_n(et), rt, rtol=1e-13, atol=1e-12)
    ec, ex = afq.local_energy_2body(_t(Gha), _t(Ghb), _t(ra), _t(rb))
    np.testing.assert_allclose(_n(ec), ec_ref, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(_n(ex), ex_ref, rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("seed", [0, 1])
def test_local_energy_full_g(afq, seed):
    tu, td, pa, pb, ra, rb, h1e, enuc = make_harness(seed)
    Ga, Gb, Gha, Ghb, _ = O.greens_function(pa, pb, tu, td)
    ref = O.local_energy(h1e, Ga, Gb, Gha, Ghb, ra, rb, enuc)
    out = afq.local_energy(_t(h1e), _t(Ga), _t(Gb), _t(Gha), _t(Ghb), _t(ra), _t(rb), enuc)
    for x, y in zip(out, ref):
        np.testing.assert_allclose(_n(x), y, rtol=1e-13, atol=1e-12)


def _small_system(M=14, Na=3, Nb=2, G=40, seed=5):
    from omegaqmc_amd.synth import make_system
    s = make_system(M, Na, Nb, G, seed=seed)
    tu, td = s.trial_up, s.trial_dn
    ra, rb = O.half_rotate_cholesky(s.chol, tu, td)
    prop = O.build_propagator(s.h1e_mod, s.chol, tu, td, 0.005)
    return s, tu, td, ra.real.copy(), rb.real.copy(), prop


def test_build_vhs_and_taylor(afq):
    s, tu, td, ra, rb, prop = _small_system()
    rng = np.random.default_rng(3)
    W = 6
    xs = rng.normal(size=(W, s.naux)) + 0.3j * rng.normal(size=(W, s.naux))
    V = afq._build_vhs(_t(xs), _t(s.chol), 0.005)
    Vr = O.build_vhs(xs, s.chol, 0.005)
    assert rel_err(_n(V), Vr) < TOL
    pa = walkers_near_trial(tu, W, rng, 0.2)
    out = afq._apply_exp_vhs(_t(Vr), _t(pa), 6)
    assert rel_err(_n(out), O.apply_exp_vhs(Vr, pa, 6)) < TOL
    out0 = afq._apply_exp_vhs(_t(Vr), _t(pa), 0)
    assert np.array_equal(_n(out0), pa)


def test_force_bias_shift(afq):
    s, tu, td, ra, rb, prop = _small_system()
    rng = np.random.default_rng(4)
    W = 7
    pa, pb = walkers_near_trial(tu, W, rng, 0.3), walkers_near_trial(td, W, rng, 0.3)
    Gha, Ghb, _ = O.greens_function_force_bias(pa, pb, tu.astype(complex), td.astype(complex))
    xi = rng.normal(size=(W, s.naux))
    for fb in (1.0, 0.05):  # 0.05 exercises the cap
        xbar, xs, cfb, cmf = afq.force_bias(_t(Gha), _t(Ghb), _t(ra), _t(rb), _t(prop['mf_shift']),
                                            _t(xi), 0.005, fb)
        rxbar, rxs, rcfb, rcmf, _ = O.force_bias_shift(Gha, Ghb, ra, rb, prop['mf_shift'], xi, 0.005, fb)
        for x, y in ((xbar, rxbar), (xs, rxs), (cfb, rcfb), (cmf, rcmf)):
            assert rel_err(_n(x), y) < TOL


def test_update_weights_edge_cases(afq):
    rng = np.random.default_rng(11)
    W = 64
    w = rng.uniform(0.1, 2.0, W)
    oo = rng.normal(size=W) + 1j * rng.normal(size=W)
    on = oo * np.exp(rng.normal(scale=0.05, size=W) + 1j * rng.normal(scale=1.5, size=W))
    on[:4] = oo[:4] * 1e6            # huge ratio: energy clipped low
    on[4:8] = oo[4:8] * 1e-6         # tiny ratio: energy clipped high
    cfb = rng.normal(scale=0.2, size=W) + 1j * rng.normal(scale=2.0, size=W)  # cos < 0 for some
    cmf = rng.normal(scale=0.1, size=W) + 1j * rng.normal(scale=0.1, size=W)
    eh = rng.normal(size=W) + 1j * rng.normal(size=W)
    wn, en = afq._update_weights_phaseless(_t(w), _t(oo), _t(on), _t(cfb), _t(cmf), _t(eh), -1.3, 0.005)
    rwn, ren = O.update_weights_phaseless(w, oo, on, cfb, cmf, eh, -1.3, 0.005)
    assert np.any(rwn == 0.0)
    np.testing.assert_allclose(_n(wn), rwn, rtol=1e-12, atol=0)
    np.testing.assert_allclose(_n(en), ren, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("W,M,Na,Nb", [(10, 37, 6, 5), (9, 50, 1, 0), (6, 64, 17, 16),
                                        (5, 90, 33, 24), (4, 120, 40, 40), (3, 120, 45, 41),
                                        (2, 100, 64, 7)])
def test_orthogonalize(afq, W, M, Na, Nb):
    """Every register-elimination size (N <= 40), the LDS factorisation
    (N > 40) and an empty spin, against the Householder oracle."""
    rng = np.random.default_rng(9 + M + Na)
    pa = rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na))
    pb = rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb))
    qa, qb, ld, _ = afq.orthogonalize_walkers(_t(pa), _t(pb))
    rqa, rqb, rld, _ = O.orthogonalize_walkers(pa, pb)
    assert rel_err(_n(qa), rqa) < TOL
    if Nb:
        assert rel_err(_n(qb), rqb) < TOL
    assert rel_err(_n(ld), rld) < TOL


def test_orthogonalize_walkers_are_independent(afq):
    """A walker whose columns are nearly dependent (condition ~1e12: its Gram
    matrix is singular to working precision) takes the shifted or the failed
    branch on its own; the other walkers of the batch match the oracle."""
    rng = np.random.default_rng(21)
    W, M, Na, Nb = 6, 80, 12, 9
    pa = rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na))
    pb = rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb))
    pa[2, :, 5] = pa[2, :, 4] * (1 + 1e-12) + 1e-12 * rng.normal(size=M)
    qa, qb, ld, _ = afq.orthogonalize_walkers(_t(pa), _t(pb))
    rqa, rqb, rld, _ = O.orthogonalize_walkers(pa, pb)
    qa, qb, ld = _n(qa), _n(qb), _n(ld)
    ok = [w for w in range(W) if w != 2]
    assert rel_err(qa[ok], rqa[ok]) < TOL
    assert rel_err(qb, rqb) < TOL
    assert rel_err(ld[ok], rld[ok]) < TOL
    if np.isfinite(ld[2]):
        # the shifted passes: orthonormal columns spanning phi
        q = qa[2]
        assert np.abs(q.conj().T @ q - np.eye(Na)).max() < 1e-10
        r = q.conj().T @ pa[2]
        assert np.linalg.norm(q @ r - pa[2]) < 1e-10 * np.linalg.norm(pa[2])


@pytest.mark.parametrize("W", [1, 7, 256, 1024, 20000])
def test_comb_indices_bit_exact(afq, W):
    rng = np.random.default_rng(W)
    w = rng.exponential(size=W)
    w[rng.uniform(size=W) < 0.1] = 0.0
    if W > 1:
        w[0] = 0.0
    M, Na, Nb = 3, 1, 1
    pa = rng.normal(size=(W, M, Na)) + 0j
    pb = rng.normal(size=(W, M, Nb)) + 0j
    eh = rng.normal(size=W) + 1j * rng.normal(size=W)
    for r in (0.0, 0.37, P.comb_offset(42, 10)):
        wn, pan, pbn, ehn, idx = afq.population_control_comb(_t(w), _t(pa), _t(pb), r, _t(eh),
                                                             return_indices=True)
        ridx = O.comb_indices(w, r)
        assert np.array_equal(_n(idx), ridx)
        assert np.array_equal(_n(pan), pa[ridx]) and np.array_equal(_n(ehn), eh[ridx])
        assert np.all(_n(wn) == 1.0)


def test_population_control_apply_replays_indices(afq):
    """afq_population_control_apply with given survivor indices gathers the
    same walkers as the comb that chose them (weights reset to 1)."""
    import torch
    from omegaqmc_amd import _lib
    rng = np.random.default_rng(5)
    W, M, Na, Nb = 37, 6, 2, 1
    pa = rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na))
    pb = rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb))
    eh = rng.normal(size=W) + 1j * rng.normal(size=W)
    idx = rng.integers(0, W, size=W).astype(np.int32)
    idx[:3] = W - 1
    out = [torch.empty_like(_t(x)) for x in (pa, pb, eh)]
    wn = torch.empty(W, dtype=torch.float64, device="cuda")
    ti, tpa, tpb, teh = _t(idx), _t(pa), _t(pb), _t(eh)
    L = _lib.load()
    _lib.check(L.afq_population_control_apply(
        _lib.context().handle, W, M, Na, Nb, _lib.ptr(ti), _lib.ptr(tpa), _lib.ptr(tpb),
        _lib.ptr(teh), _lib.ptr(wn), _lib.ptr(out[0]), _lib.ptr(out[1]), _lib.ptr(out[2])),
        "apply")
    torch.cuda.synchronize()
    assert np.array_equal(_n(out[0]), pa[idx]) and np.array_equal(_n(out[1]), pb[idx])
    assert np.array_equal(_n(out[2]), eh[idx]) and np.all(_n(wn) == 1.0)
    # an out-of-range index is never read: its output row is poisoned with NaN
    bad_idx = idx.copy()
    bad_idx[4], bad_idx[7] = W, -1
    _lib.check(L.afq_population_control_apply(
        _lib.context().handle, W, M, Na, Nb, _lib.ptr(_t(bad_idx)), _lib.ptr(tpa), _lib.ptr(tpb),
        _lib.ptr(teh), _lib.ptr(wn), _lib.ptr(out[0]), _lib.ptr(out[1]), _lib.ptr(out[2])),
        "apply")
    torch.cuda.synchronize()
    got = _n(out[0])
    assert np.all(np.isnan(got[[4, 7]])) and np.all(np.isnan(_n(out[2])[[4, 7]]))
    keep = [i for i in range(W) if i not in (4, 7)]
    assert np.array_equal(got[keep], pa[idx[keep]])


def test_aux_fields_match_oracle_stream(afq):
    import omegaqmc_amd as A
    for (W, G, step) in ((4, 7, 1), (33, 50, 12345)):
        xi = A.aux_fields(A.PhiloxKey(42, step), W, G)
        ref = P.aux_fields(42, step, W, G)
        np.testing.assert_allclose(_n(xi), ref, rtol=1e-13, atol=1e-14)
    assert A.comb_offset(A.PhiloxKey(7, 3)) == P.comb_offset(7, 3)


@pytest.mark.parametrize("M,Na,Nb,G,W", [(11, 2, 2, 50, 16), (24, 5, 3, 60, 9)])
def test_propagate_one_step(afq, M, Na, Nb, G, W):
    s, tu, td, ra, rb, prop = _small_system(M, Na, Nb, G)
    rng = np.random.default_rng(1)
    pa = walkers_near_trial(tu, W, rng, 0.1)
    pb = walkers_near_trial(td, W, rng, 0.1)
    w = rng.uniform(0.5, 1.5, W)
    eh = rng.normal(size=W) + 0.1j * rng.normal(size=W)
    xi = P.aux_fields(3, 1, W, G)
    ref = O.propagate_walkers(pa, pb, w, None, eh, prop, s.chol, ra, rb, tu.astype(complex),
                              td.astype(complex), -1.0, xi)
    propd = {'expH1': _t(prop['expH1'].real.copy()), 'mf_shift': _t(prop['mf_shift']), 'dt': 0.005}
    out = afq.propagate_walkers(_t(pa), _t(pb), _t(w), None, _t(eh), propd, _t(s.chol), _t(ra),
                                _t(rb), _t(tu), _t(td), -1.0, None, xi=_t(xi))
    for x, y in zip(out[:5], ref):
        assert rel_err(_n(x), y) < TOL



@pytest.mark.parametrize("W", [7, 1000, 8191, 8192])
def test_pair_branch_matches_oracle(afq, W):
    """afqmc_gto.py:241-284 on the GPU: weights, walkers and sources equal
    the oracle's bit for bit (heavy-tailed weights with many exact zeros, so
    ties are ordered by index)."""
    rng = np.random.default_rng(W)
    w = rng.lognormal(0.0, 1.5, size=W)
    w[rng.uniform(size=W) < 0.25] = 0.0
    M, Na, Nb = 5, 2, 1
    pa = rng.normal(size=(W, M, Na)) + 1j * rng.normal(size=(W, M, Na))
    pb = rng.normal(size=(W, M, Nb)) + 1j * rng.normal(size=(W, M, Nb))
    rw, rpa, rpb, rsrc = O.population_control_pair_branch(w, pa, pb)
    assert np.any(rsrc != np.arange(W)) or W == 7
    gw, gpa, gpb, gsrc = afq.population_control_pair_branch(_t(w), _t(pa), _t(pb),
                                                            return_sources=True)
    assert np.array_equal(_n(gsrc), rsrc)
    assert np.array_equal(_n(gw), rw)
    assert np.array_equal(_n(gpa), rpa) and np.array_equal(_n(gpb), rpb)


def test_pair_branch_limits(afq):
    import torch
    W = 8193
    z = torch.zeros((W, 2, 1), dtype=torch.complex128, device="cuda")
    with pytest.raises(afq.AFQError):
        afq.population_control_pair_branch(torch.ones(W, dtype=torch.float64, device="cuda"), z, z)
    # the reference's container: trial copies, unit weights and overlaps
    tu = torch.eye(6, 2, dtype=torch.float64)
    wk = afq.Walkers(tu, tu[:, :1], 4, 6, 2, 1)
    assert wk.phia.shape == (4, 6, 2) and wk.phib.shape == (4, 6, 1)
    assert torch.all(wk.weights == 1) and torch.all(wk.overlap == 1)
    assert torch.all(wk.e_hybrid == 0) and torch.all(wk.log_detR == 0)
