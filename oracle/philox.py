"""Philox4x32-10 and the auxiliary-field normal stream (oracle, NumPy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Philox4x32-10 is the counter-based generator of Salmon, Moraes, Dror and
Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11): ten rounds of

    (hi0, lo0) = mulhilo(0xD2511F53, c0);  (hi1, lo1) = mulhilo(0xCD9E8D57, c2)
    c = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0)
    k = (k0 + 0x9E3779B9, k1 + 0xBB67AE85)          (key bumped between rounds)

Pinned by the published known-answer vectors of the Random123 distribution
(``tests/test_oracle_philox.py``).

Stream layout (identical in ``omegaqmc_amd/csrc/afqmc_rng.hip``):

* auxiliary field ``xi[w, g]`` at 1-based step ``s``: counter
  ``(g // 2, w, s, 0)``, key ``(seed & 0xffffffff, seed >> 32)``.  The two
  64-bit words ``a = (x1 << 32) | x0`` and ``b = (x3 << 32) | x2`` give
  ``u1 = ((a >> 11) + 1) * 2**-53`` in (0, 1] and ``u2 = (b >> 11) * 2**-53``
  in [0, 1); Box-Muller gives ``xi[w, 2k] = r cos(2 pi u2)`` and
  ``xi[w, 2k+1] = r sin(2 pi u2)`` with ``r = sqrt(-2 log u1)``.
* comb offset at step ``s``: counter ``(0, 0, s, 1)``, ``r = (a >> 11) * 2**-53``.

The reference draws these numbers with ``jax.random`` (threefry); that stream
cannot be reproduced without JAX, so the step functions take ``xi`` and ``r``
as inputs and this stream is the package's own (documented) choice.
"""

from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)

STREAM_XI = 0
STREAM_COMB = 1


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10; every argument is broadcast uint32."""
    c0, c1, c2, c3 = (np.asarray(v, dtype=np.uint32) for v in (c0, c1, c2, c3))
    c0, c1, c2, c3 = np.broadcast_arrays(c0, c1, c2, c3)
    c0, c1, c2, c3 = (v.copy() for v in (c0, c1, c2, c3))
    k0 = np.asarray(k0, dtype=np.uint32).copy()
    k1 = np.asarray(k1, dtype=np.uint32).copy()
    with np.errstate(over="ignore"):
        for _ in range(10):
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & MASK32).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & MASK32).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            k0 = (k0 + W0).astype(np.uint32)
            k1 = (k1 + W1).astype(np.uint32)
    return c0, c1, c2, c3


def _key(seed):
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return np.uint32(seed & 0xFFFFFFFF), np.uint32(seed >> 32)


def _u53(hi, lo):
    a = (hi.astype(np.uint64) << np.uint64(32)) | lo.astype(np.uint64)
    return a >> np.uint64(11)


def aux_fields(seed: int, step: int, nwalkers: int, naux: int) -> np.ndarray:
    """xi[w, g] ~ N(0, 1) of the stream described in the module docstring."""
    k0, k1 = _key(seed)
    npair = (naux + 1) // 2
    kk, ww = np.meshgrid(np.arange(npair, dtype=np.uint32),
                         np.arange(nwalkers, dtype=np.uint32))
    x0, x1, x2, x3 = philox4x32_10(kk, ww, np.uint32(step), np.uint32(STREAM_XI), k0, k1)
    u1 = (_u53(x1, x0) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = _u53(x3, x2).astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    th = 2.0 * np.pi * u2
    out = np.empty((nwalkers, 2 * npair))
    out[:, 0::2] = r * np.cos(th)
    out[:, 1::2] = r * np.sin(th)
    return np.ascontiguousarray(out[:, :naux])


def comb_offset(seed: int, step: int) -> float:
    """Uniform comb offset r in [0, 1) for 1-based step ``step``."""
    k0, k1 = _key(seed)
    x0, x1, _, _ = philox4x32_10(0, 0, np.uint32(step), np.uint32(STREAM_COMB), k0, k1)
    return float(_u53(x1, x0).astype(np.float64) * 2.0 ** -53)
