"""ctypes binding of ``libomega_afqmc.so`` (the C ABI of include/omega_afqmc.h).

Device memory comes from torch tensors (``data_ptr()``); complex128 tensors
are interleaved (re, im) pairs, exactly the layout the ABI expects.  There is
no fallback: if the library is missing or fails to load, every call raises.
"""

from __future__ import annotations

import ctypes
import os
import pathlib
import threading

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libomega_afqmc.so"

AFQ_OK = 0
AFQ_FP64, AFQ_FP32, AFQ_BF16X3 = 0, 1, 2
AFQ_REAL, AFQ_COMPLEX = 0, 1

# per-walker status bits (enum afq_walker_flag)
WALKER_FLAGS = {"nonfinite_phi": 1, "zero_overlap": 2, "nonfinite_overlap": 4, "bad_weight": 8,
                "nonfinite_e_hybrid": 16, "zero_weight": 32}

PRECISIONS = {"fp64": AFQ_FP64, "fp32": AFQ_FP32, "bf16x3": AFQ_BF16X3}

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_U64 = ctypes.c_uint64
_I64 = ctypes.c_int64
_SZ = ctypes.c_size_t

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "afq_version": [],
    "afq_last_error": [],
    "afq_ctx_create": [ctypes.POINTER(_P), _I, _P],
    "afq_ctx_destroy": [_P],
    "afq_ctx_set_stream": [_P, _P],
    "afq_ctx_synchronize": [_P],
    "afq_ctx_reserve": [_P, _SZ],
    "afq_greens_function_force_bias": [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P],
    "afq_greens_function_overlap": [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P],
    "afq_greens_function": [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "afq_local_energy_2body": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P],
    "afq_local_energy": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _D, _P, _P, _P],
    "afq_local_energy_2body_pw": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P, _P],
    "afq_local_energy_pw": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _D, _P, _P, _P,
                            _P],
    "afq_local_energy_streamed": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _D, _I,
                                  _P, _P, _P],
    "afq_build_vhs": [_P, _I, _I, _I, _P, _P, _D, _P],
    "afq_apply_exp_vhs": [_P, _I, _I, _I, _P, _P, _I, _P],
    "afq_apply_one_body": [_P, _I, _I, _I, _P, _P, _P],
    "afq_force_bias": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _D, _D, _P, _P, _P, _P],
    "afq_update_weights_phaseless": [_P, _I, _P, _P, _P, _P, _P, _P, _D, _D, _P, _P],
    "afq_orthogonalize_walkers": [_P, _I, _I, _I, _I, _P, _P, _P, _P],
    "afq_population_control_comb": [_P, _I, _I, _I, _I, _P, _P, _P, _D, _P, _P, _P, _P, _P, _P],
    "afq_population_control_comb_noalloc": [_P, _I, _I, _I, _I, _P, _P, _P, _D, _P, _P, _P, _P,
                                            _P, _P],
    "afq_population_control_apply": [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    "afq_population_control_pair_branch": [_P, _I, _I, _I, _I, _P, _P, _P, _P],
    "afq_clip_weights": [_P, _I, _P, _D],
    "afq_accumulate_hybrid": [_P, _I, _P, _P, _P],
    "afq_block_stats": [_P, _I, _P, _P, _P],
    "afq_greens_function_multidet_force_bias": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P,
                                                _P, _P, _P, _P],
    "afq_local_energy_multidet_streamed": [_P, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _I,
                                           _P, _P, _P, _D, _I, _P, _P, _P],
    "afq_system_set_propagator_block": [_P, _I],
    "afq_walker_status": [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _D, _P, _P],
    "afq_aux_fields": [_P, _U64, _I64, _I, _I, _P],
    "afq_comb_offset": [_U64, _I64],
    "afq_system_create": [_P, ctypes.POINTER(_P), _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P,
                          _D, _D, _I, _I],
    "afq_system_destroy": [_P],
    "afq_system_set_chol_sign": [_P, _P],
    "afq_greens_function_multidet": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P],
    "afq_greens_function_multidet_overlap_only": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P,
                                                  _I, _P],
    "afq_local_energy_multidet": [_P, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _D,
                                  _I, _P, _P, _P],
    "afq_system_create_multidet": [_P, ctypes.POINTER(_P), _I, _I, _I, _I, _P, _P, _P, _I, _P,
                                   _P, _P, _D, _D, _I, _I, _I],
    "afq_propagate_walkers": [_P, _I, _P, _P, _P, _P, _D, _P, _U64, _I64, _P],
    "afq_system_local_energy": [_P, _I, _P, _P, _P, _D, _P, _P, _P],
    "afq_qed_propagate_photon": [_P, _I, _P, _P, _D, _D, _D, _D, _P, _P],
    "afq_apply_exp_scaled": [_P, _I, _I, _I, _P, _P, _P, _I, _P],
    "afq_qed_photon_weights": [_P, _I, _P, _P, _P],
    "afq_gather_walker_values": [_P, _I, _P, _P, _P],
    "afq_qed_local_energy": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _D,
                             _D, _D, _P, _P, _P, _P],
    "afq_batch_create": [_P, ctypes.POINTER(_P), _I, _P, _P, _P, _I, _I, _I, _P],
    "afq_batch_set_eshift": [_P, _P],
    "afq_batch_run": [_P, _I, _I],
    "afq_batch_state": [_P, _I, _P, _P, _P, _P, _P, _P, _P],
    "afq_batch_copy_state": [_P, _I, _P, _P, _P, _P, _P, _P],
    "afq_batch_zero_accumulators": [_P],
    "afq_batch_destroy": [_P],
    "afq_system_set_force_bias_splits": [_P, _I],
    "afq_system_set_multidet_method": [_P, _I],
    "afq_system_get_multidet_method": [_P],
    "afq_system_timing_enable": [_P, _I],
    "afq_system_timing_read": [_P, ctypes.POINTER(_D), ctypes.POINTER(_I64), _I, _I],
}

PHASES = ("greens", "one_body", "force_bias", "shift", "vhs", "taylor", "overlap", "weights", "rng")
_RESTYPE = {"afq_last_error": ctypes.c_char_p, "afq_comb_offset": _D}

_lib = None
_lock = threading.Lock()


class AFQError(RuntimeError):
    pass


def load():
    """Load the shared library (raises if it was never built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise AFQError(
                f"{LIB_PATH} is missing: build it with `python -m omegaqmc_amd.build_lib` "
                "(there is no fallback implementation)")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, _I)
        _lib = lib
        return lib


def check(rc: int, what: str = ""):
    if rc != AFQ_OK:
        msg = load().afq_last_error().decode(errors="replace")
        raise AFQError(f"{what}: status {rc}: {msg}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


class Context:
    """Owns an ``afq_ctx`` bound to a device and its current torch stream."""

    def __init__(self, device: int = 0, stream=None):
        import torch
        lib = load()
        self.device = int(device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        self._stream = stream
        h = _P()
        check(lib.afq_ctx_create(ctypes.byref(h), self.device, _P(stream.cuda_stream)),
              "afq_ctx_create")
        self.handle = h

    def sync_stream(self):
        """Re-bind the context to torch's current stream on its device."""
        import torch
        s = torch.cuda.current_stream(self.device)
        if s.cuda_stream != self._stream.cuda_stream:
            check(load().afq_ctx_set_stream(self.handle, _P(s.cuda_stream)), "afq_ctx_set_stream")
            self._stream = s

    def synchronize(self):
        check(load().afq_ctx_synchronize(self.handle), "afq_ctx_synchronize")

    def __del__(self):
        try:
            if getattr(self, "handle", None) is not None and _lib is not None:
                _lib.afq_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


_contexts: dict[int, Context] = {}


def context(device: int | None = None) -> Context:
    """Per-device default context, bound to torch's current stream."""
    import torch
    if device is None:
        device = torch.cuda.current_device()
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        _contexts[device] = ctx
    else:
        ctx.sync_stream()
    return ctx
