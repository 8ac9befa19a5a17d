"""Uneven batches of independent AFQMC systems, stepped together.

The reference runs one system per driver (``_AFQMCDriverGTO``,
OmegaQMC/afqmc_gto.py:1257-1546), one eager dispatch after another.  A
:class:`Batch` holds many single-determinant systems of different sizes,
walker counts and precisions; each driver step (QR every
``stabilize_freq`` steps, the fused phaseless step, weight clip from step 2
on, comb every ``pop_control_freq`` steps, hybrid-energy accumulators --
afqmc_gto.py:1383-1451) of all of them runs on a few HIP streams, and one
driver cycle of the whole batch is a single hipGraph launch
(``afq_batch_run`` mode 2).

:func:`autotune` picks, per (M, N, Nchol, precision) bucket, the K-split
count of the force-bias GEMM and then the BF16x3 propagator's column block
(wide blocks reuse a walker's VHS over more columns, narrow ones give small
systems more workgroups) by timing the candidates on the bucket's first
system, then the number of streams for the whole batch, assigning systems to
streams longest-first by their measured step time.  Choices are cached in a
JSON file keyed by bucket and device name, so a second run reuses them.
"""

from __future__ import annotations

import ctypes
import json
import os
import pathlib
import time

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr

MODES = {"sequential": 0, "streams": 1, "graph": 2}
FB_SPLITS = {"fp64": (2, 4, 8), "fp32": (2, 4, 8), "bf16x3": (2, 4)}
# widest propagator column block (afq_system_set_propagator_block; BF16x3 only)
PROP_BLOCKS = {"fp64": (48,), "fp32": (48,), "bf16x3": (48, 32, 16)}
STREAM_CHOICES = (1, 2, 4)
DEFAULT_CACHE = pathlib.Path(os.environ.get(
    "OMEGAQMC_AMD_TUNE_CACHE", pathlib.Path.home() / ".cache" / "omegaqmc_amd" / "tune.json"))


class Batch:
    """An ``afq_batch``: ``systems`` (prepared :class:`~omegaqmc_amd.afqmc_gto.System`
    objects) with ``walkers[i]`` walkers each, initialised to their trials."""

    def __init__(self, systems, walkers, seeds=None, stabilize_freq=5, pop_control_freq=10,
                 nstreams=4, stream_of=None):
        self.systems = list(systems)
        self.walkers = [int(w) for w in walkers]
        n = len(self.systems)
        if len(self.walkers) != n:
            raise ValueError("one walker count per system")
        seeds = list(range(42, 42 + n)) if seeds is None else [int(s) for s in seeds]
        self.seeds = seeds
        self.nstreams = int(nstreams)
        L = _lib.load()
        ctx = _lib.context().handle
        handles = (ctypes.c_void_p * n)(*[s.handle.value for s in self.systems])
        W = (ctypes.c_int * n)(*self.walkers)
        S = (ctypes.c_uint64 * n)(*[s & (2 ** 64 - 1) for s in seeds])
        so = None if stream_of is None else (ctypes.c_int * n)(*[int(k) for k in stream_of])
        h = _lib._P()
        check(L.afq_batch_create(ctx, ctypes.byref(h), n, handles, W, S, int(stabilize_freq),
                                 int(pop_control_freq), self.nstreams, so), "afq_batch_create")
        self.handle = h
        self.stream_of = list(stream_of) if stream_of is not None else [i % self.nstreams
                                                                      for i in range(n)]

    def set_eshift(self, eshift):
        e = (ctypes.c_double * len(self.systems))(*[float(x) for x in eshift])
        check(_lib.load().afq_batch_set_eshift(self.handle, e), "afq_batch_set_eshift")

    def run(self, nsteps, mode="graph"):
        _lib.context()
        check(_lib.load().afq_batch_run(self.handle, int(nsteps), MODES[mode]), "afq_batch_run")

    def zero_accumulators(self):
        """Reset every system's hybrid-energy accumulators (the per-block
        reset, afqmc_gto.py:1384-1385)."""
        _lib.context()
        check(_lib.load().afq_batch_zero_accumulators(self.handle), "afq_batch_zero_accumulators")

    def copy_state_into(self, i, st):
        """Copy system i's walkers, weights, e_hybrid, overlaps and
        accumulators into the matching tensors of the driver state dict
        ``st`` (as made by ``AFQMCDriver.init_walkers``)."""
        _lib.context()
        check(_lib.load().afq_batch_copy_state(self.handle, int(i), ptr(st['phia']),
                                               ptr(st['phib']), ptr(st['weights']),
                                               ptr(st['e_hybrid']), ptr(st['overlap']),
                                               ptr(st['acc'])), "afq_batch_copy_state")

    def state(self, i):
        """Copies of system i's walkers, weights, e_hybrid, last overlaps and
        accumulators (torch tensors on the device)."""
        sysobj = self.systems[i]
        W, M = self.walkers[i], sysobj.M
        dev = torch.device("cuda", torch.cuda.current_device())
        out = {
            'phia': torch.empty((W, M, sysobj.Na), dtype=torch.complex128, device=dev),
            'phib': torch.empty((W, M, sysobj.Nb), dtype=torch.complex128, device=dev),
            'weights': torch.empty(W, dtype=torch.float64, device=dev),
            'e_hybrid': torch.empty(W, dtype=torch.complex128, device=dev),
            'overlap': torch.empty(W, dtype=torch.complex128, device=dev),
            'acc': torch.empty(2, dtype=torch.float64, device=dev),
        }
        _lib.context()
        check(_lib.load().afq_batch_copy_state(self.handle, int(i), ptr(out['phia']),
                                               ptr(out['phib']), ptr(out['weights']),
                                               ptr(out['e_hybrid']), ptr(out['overlap']),
                                               ptr(out['acc'])), "afq_batch_copy_state")
        return out

    def __del__(self):
        try:
            if getattr(self, "handle", None) is not None and _lib._lib is not None:
                _lib._lib.afq_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


# ---------------------------------------------------------------------------
# autotuner
# ---------------------------------------------------------------------------

def bucket_key(sysobj, precision):
    """Systems of one bucket share their tile choice: M to 16 rows, Nchol to 64."""
    return f"{precision}:M{(sysobj.M + 15) // 16 * 16}:N{sysobj.Na}+{sysobj.Nb}:G{(sysobj.G + 63) // 64 * 64}"


def _device_name():
    try:
        return torch.cuda.get_device_name(torch.cuda.current_device())
    except Exception:
        return "unknown"


def _load_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _time_steps(systems, walkers, nsteps, mode, nstreams=1, stream_of=None, seeds=None):
    b = Batch(systems, walkers, seeds=seeds, nstreams=nstreams, stream_of=stream_of)
    b.run(20, mode)  # an eager cycle (sizes the workspaces), the capture, one launch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(nsteps, mode)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    del b
    return dt / nsteps


def autotune(systems, walkers, precisions, cache=DEFAULT_CACHE, nsteps=10, verbose=False):
    """Choose the force-bias K splits and the propagator column block per
    bucket and the batch's stream count and stream assignment.  Returns
    {'fb_splits': {bucket: n}, 'prop_block': {bucket: cols}, 'nstreams': k,
    'stream_of': [...], 'step_ms': [...]} and applies the choices to the
    systems."""
    cache = pathlib.Path(cache)
    dev = _device_name()
    db = _load_cache(cache)
    entries = db.setdefault(dev, {})
    L = _lib.load()
    buckets = {}
    for i, (s, p) in enumerate(zip(systems, precisions)):
        buckets.setdefault(bucket_key(s, p), []).append(i)
    choice = {}
    for key, idx in buckets.items():
        if key in entries and "prop_block" in entries[key]:
            choice[key] = (int(entries[key]["fb_splits"]), int(entries[key]["prop_block"]))
            continue
        i0 = idx[0]
        s0, w0, p0 = systems[i0], walkers[i0], precisions[i0]
        # the two knobs in turn: force-bias K splits, then the propagator's
        # column block with the chosen splits
        best, best_t = None, None
        for n in FB_SPLITS[p0]:
            check(L.afq_system_set_force_bias_splits(s0.handle, n), "set splits")
            check(L.afq_system_set_propagator_block(s0.handle, PROP_BLOCKS[p0][0]), "set block")
            t = _time_steps([s0], [w0], nsteps, "graph")
            if verbose:
                print(f"[autotune] {key} W={w0} fb_splits={n}: {t * 1e3:.3f} ms/step")
            if best_t is None or t < best_t:
                best, best_t = (n, PROP_BLOCKS[p0][0]), t
        check(L.afq_system_set_force_bias_splits(s0.handle, best[0]), "set splits")
        for c in PROP_BLOCKS[p0][1:]:
            check(L.afq_system_set_propagator_block(s0.handle, c), "set block")
            t = _time_steps([s0], [w0], nsteps, "graph")
            if verbose:
                print(f"[autotune] {key} W={w0} prop_block={c}: {t * 1e3:.3f} ms/step")
            if t < best_t:
                best, best_t = (best[0], c), t
        choice[key] = best
        entries[key] = {"fb_splits": best[0], "prop_block": best[1],
                        "step_ms": round(best_t * 1e3, 4), "walkers": w0}
    for key, idx in buckets.items():
        for i in idx:
            check(L.afq_system_set_force_bias_splits(systems[i].handle, choice[key][0]),
                  "set splits")
            check(L.afq_system_set_propagator_block(systems[i].handle, choice[key][1]),
                  "set block")
    # per-system step time with the chosen tiles, for the stream assignment
    step_ms = [_time_steps([s], [w], nsteps, "graph") * 1e3 for s, w in zip(systems, walkers)]
    best_ns, best_t, best_assign = None, None, None
    for ns in STREAM_CHOICES:
        loads = [0.0] * ns
        assign = [0] * len(systems)
        for i in sorted(range(len(systems)), key=lambda k: -step_ms[k]):
            k = int(np.argmin(loads))
            assign[i] = k
            loads[k] += step_ms[i]
        t = _time_steps(systems, walkers, nsteps, "graph", nstreams=ns, stream_of=assign)
        if verbose:
            print(f"[autotune] batch of {len(systems)}: {ns} streams {t * 1e3:.3f} ms/step")
        if best_t is None or t < best_t:
            best_ns, best_t, best_assign = ns, t, assign
    db[dev] = entries
    try:
        cache.parent.mkdir(parents=True, exist_ok=True)
        with open(cache, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
    except OSError:
        pass
    return {"fb_splits": {k: v[0] for k, v in choice.items()},
            "prop_block": {k: v[1] for k, v in choice.items()}, "nstreams": best_ns,
            "stream_of": best_assign,
            "step_ms": step_ms, "batch_ms_per_step": best_t * 1e3}
