"""Shared inputs of the parity tests (NumPy, host)."""

from __future__ import annotations

import numpy as np


def make_harness(seed=0, nwalkers=8, nbasis=10, nup=3, ndown=3, naux=24):
    """The reference's random harness (tests/test_afqmc_kernels.py:155-177), same
    draw order, so the arrays equal the ones the reference's own tests use."""
    rng = np.random.default_rng(seed)

    def _crand(shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    trial_up = _crand((nbasis, nup))
    trial_dn = _crand((nbasis, ndown))
    phia = _crand((nwalkers, nbasis, nup))
    phib = _crand((nwalkers, nbasis, ndown))
    rchol_a = _crand((naux, nup, nbasis))
    rchol_b = _crand((naux, ndown, nbasis))
    h1e = rng.standard_normal((nbasis, nbasis)) + 1j * rng.standard_normal((nbasis, nbasis))
    h1e = 0.5 * (h1e + h1e.conj().T)
    enuc = float(rng.standard_normal())
    return trial_up, trial_dn, phia, phib, rchol_a, rchol_b, h1e, enuc


def walkers_near_trial(trial, nwalkers, rng, scale=0.05):
    """Random complex walkers around a trial (test_afqmc_kernels.py:75-94 recipe)."""
    M, N = trial.shape
    return (np.repeat(trial[None].astype(np.complex128), nwalkers, axis=0)
            + scale * (rng.normal(size=(nwalkers, M, N)) + 1j * rng.normal(size=(nwalkers, M, N))))


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)


def bf16_round(x):
    """Round float64 -> float32 -> bfloat16 (nearest even), as the GPU's
    split does (v_cvt to f32, then to bf16); returned as float64."""
    x32 = np.asarray(x, dtype=np.float32)
    b = x32.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16x3_split(x):
    """hi = bf16(x), lo = bf16(x - hi)."""
    hi = bf16_round(x)
    return hi, bf16_round(np.asarray(x) - hi)


def bf16x3_two_body(Ghalfa, Ghalfb, rchol_a, rchol_b):
    """(e_coul, e_exch) of energy.py:51-91 with every product a*b replaced by
    ah*bh + ah*bl + al*bh of the split operands, summed in float64: the error
    the split-BF16 kernels are designed to have (without FP32 accumulation)."""
    def prod(axes, L, G):
        Lh, Ll = bf16x3_split(L)
        Gh, Gl = bf16x3_split(G)
        f = lambda a, b: np.tensordot(a, b, axes=axes)  # noqa: E731
        return f(Lh, Gh) + f(Lh, Gl) + f(Ll, Gh)
    e_exch = 0.0
    X = 0.0
    tq, tiq = ([2], [2]), ([1, 2], [1, 2])
    for L, Gh in ((rchol_a.real, Ghalfa), (rchol_b.real, Ghalfb)):
        # T[g, w, i, j] = sum_q L[g, i, q] Gh[w, j, q]
        T = (prod(tq, L, Gh.real) + 1j * prod(tq, L, Gh.imag)).transpose(0, 2, 1, 3)
        e_exch = e_exch + 0.5 * np.sum(T * T.transpose(0, 1, 3, 2), axis=(0, 2, 3))
        X = X + prod(tiq, L, Gh.real) + 1j * prod(tiq, L, Gh.imag)
    return 0.5 * np.sum(X ** 2, axis=0), e_exch
