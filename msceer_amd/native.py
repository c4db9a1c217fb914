"""ctypes binding of libmsceer_hip.so (the C ABI in include/msceer.h).

The library is loaded from this package directory.  There is no fallback:
if the library is missing or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmsceer_hip.so")

DEVICE_PTRS = 0x1
SKIP_LDIR = 0x2
TILED_LAYOUT = 0x40
TIMING = 0x4
STOP_REGIONS = 0x8
STOP_INTEGRATED = 0x10
FP32_FIRST = 0x20
# accuracy branches of msceer_ondemand
NEED_ASC1 = 0x1
NEED_DSC1 = 0x2
NEED_ASC2 = 0x4
NEED_DSC2 = 0x8
BRICKS = ("64x2x2", "32x4x2", "16x4x4", "8x8x4")  # lower-star brick shapes by index
# error codes
OK, EINVAL, EHIP, ENOMEM, ETOOBIG, ENODEV = 0, -1, -2, -3, -4, -5

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_U32 = ctypes.c_uint32
_I32 = ctypes.c_int32

# Every exported symbol with (restype, argtypes); include/msceer.h is the source of truth
# and tests check that the two agree.
SIGNATURES = {
    "msceer_version": (ctypes.c_int, []),
    "msceer_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "msceer_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int]),
    "msceer_destroy": (ctypes.c_int, [_P]),
    "msceer_set_stream": (ctypes.c_int, [_P, _P]),
    "msceer_last_timing": (ctypes.c_int, [_P, _P]),
    "msceer_set_autotune": (ctypes.c_int, [_P, _I32]),
    "msceer_lstar_brick": (ctypes.c_int, [_P, _I64]),
    "msceer_lstar_grad": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, _U32]),
    "msceer_lstar_grad_noalloc": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, _U32]),
    "msceer_ondemand_asc": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, ctypes.c_float, ctypes.c_float, _I32,
                                           ctypes.c_float, _P, _P, _P, _P, _U32]),
    "msceer_ondemand": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, ctypes.c_float, ctypes.c_float, _I32,
                                       ctypes.c_float, _U32, _P, _P, _P, _P, _P, _U32]),
    "msceer_ondemand2": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, ctypes.c_float, ctypes.c_float, _I32,
                                        ctypes.c_float, _U32, _P, _P, _P, _P, _P, _P, _U32]),
    "msceer_integrate_labels": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, _P, ctypes.c_float, ctypes.c_float, _I32,
                                               _I32, _P, _P, _U32]),
    "msceer_clean_labels": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, _U32]),
    "msceer_restriction_labels": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, _U32]),
    "msceer_region_boundary_kind": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, _U32]),
    "msceer_restriction_add_max": (ctypes.c_int, [_P, _P, _P, _I64, _I64, _I64, _P, _U32]),
    "msceer_constrained_lstar": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, _P, _U32]),
    "msceer_batch": (ctypes.c_int, [_P, _P, _I32, _I32, ctypes.c_float, ctypes.c_float, _I32, ctypes.c_float, _U32]),
    "msceer_batch_grad": (ctypes.c_int, [_P, _P, _I32, _I64, _U32]),
    "msceer_msc_base": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, _P, _P, _U32]),
    "msceer_expanding_region_stop": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, ctypes.c_float, ctypes.c_float, _I32,
                                                    _P, _P, _U32]),
    "msceer_last_batch_stats": (ctypes.c_int, [_P, _P]),
    "msceer_cnn_num_params": (ctypes.c_int, []),
    "msceer_cnn_create": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(_P)]),
    "msceer_cnn_destroy": (ctypes.c_int, [_P]),
    "msceer_cnn_get_params": (ctypes.c_int, [_P, _P, _P]),
    "msceer_cnn_set_params": (ctypes.c_int, [_P, _P, _P]),
    "msceer_cnn_predict": (ctypes.c_int, [_P, _P, _P, _I64, _I64, _I64, _P, _U32]),
    "msceer_cnn_gradient": (ctypes.c_int, [_P, _P, _P, _I64, _I64, _I64, _P, ctypes.c_float, _P, _P, _U32]),
    "msceer_cnn_adam_step": (ctypes.c_int, [_P, _P, ctypes.c_float]),
    "msceer_cnn_mask": (ctypes.c_int, [_P, _P, _P, _I64, _I64, _I64, ctypes.c_float, _I32, _P, _P, _U32]),
}


class Timing(ctypes.Structure):
    _fields_ = [
        ("lstar_ms", ctypes.c_float),
        ("ldir_ms", ctypes.c_float),
        ("total_ms", ctypes.c_float),
        ("ldir_rounds", ctypes.c_int32),
        ("critical", ctypes.c_int64 * 4),
    ]


class LabelStats(ctypes.Structure):
    _fields_ = [
        ("minima", ctypes.c_int64),
        ("saddles", ctypes.c_int64),
        ("regions", ctypes.c_int64),
        ("traced_fp32", ctypes.c_int64),
        ("over_budget_fp32", ctypes.c_int64),
        ("reintegrated_fp64", ctypes.c_int64),
        ("over_budget_fp64", ctypes.c_int64),
        ("integrated_vertices", ctypes.c_int64),
        ("certain_sweeps", ctypes.c_int32),
        ("pass3_rounds", ctypes.c_int32),
        ("ms_lstar", ctypes.c_float),
        ("ms_regions", ctypes.c_float),
        ("ms_integrate", ctypes.c_float),
        ("ms_cleanup", ctypes.c_float),
        ("ms_grad", ctypes.c_float),
        ("ms_total", ctypes.c_float),
        ("steps_fp32", ctypes.c_int64),
        ("steps_fp64", ctypes.c_int64),
        ("ahead_fp64", ctypes.c_int64),
        ("ms_pass1", ctypes.c_float),
        ("ms_pass3", ctypes.c_float),
        ("reintegrated_pass3", ctypes.c_int64),
        ("escalated_fp64", ctypes.c_int64),
        ("lines_asc", ctypes.c_int64),
        ("lines_dsc", ctypes.c_int64),
        ("lines_replayed", ctypes.c_int64),
        ("line_windows", ctypes.c_int32),
        ("line_cuts", ctypes.c_int32),
        ("ms_lines", ctypes.c_float),
        ("line_alone", ctypes.c_int32),
        ("line_overflows", ctypes.c_int32),
        ("maxima", ctypes.c_int64),
        ("saddles2", ctypes.c_int64),
        ("regions_dsc", ctypes.c_int64),
        ("certain_sweeps_dsc", ctypes.c_int32),
        ("ms_dsc", ctypes.c_float),
        ("ms_dsc_wait", ctypes.c_float),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Volume(ctypes.Structure):
    _fields_ = [
        ("values", ctypes.c_void_p),
        ("nx", ctypes.c_int64),
        ("ny", ctypes.c_int64),
        ("nz", ctypes.c_int64),
        ("grad_out", ctypes.c_void_p),
        ("labels_out", ctypes.c_void_p),
        ("ms", ctypes.c_float),
        ("status", ctypes.c_int32),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("groups", ctypes.c_int32),
        ("replays", ctypes.c_int32),
        ("captures", ctypes.c_int32),
        ("eager_rounds", ctypes.c_int32),
        ("shape", ctypes.c_int32),
        ("blocks", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
        ("rounds", ctypes.c_int32),
        ("critical", ctypes.c_int64 * 4),
        ("ms_gpu", ctypes.c_float),
    ]


class MsceerError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MsceerError(f"{LIB_PATH} is missing: run `python -m msceer_amd.build` first")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().msceer_strerror(rc).decode()
        raise MsceerError(f"{what}: {msg} ({rc})")


class Context:
    """A msceer_ctx on one device (owns a HIP stream and device scratch)."""

    def __init__(self, device: int = 0):
        self._ptr = _P()
        check(lib().msceer_create(ctypes.byref(self._ptr), device), "msceer_create")

    @property
    def handle(self):
        return self._ptr

    def close(self):
        if self._ptr:
            lib().msceer_destroy(self._ptr)
            self._ptr = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle: int | None):
        check(lib().msceer_set_stream(self._ptr, _P(stream_handle) if stream_handle else None),
              "msceer_set_stream")

    def last_timing(self) -> Timing:
        t = Timing()
        check(lib().msceer_last_timing(self._ptr, ctypes.byref(t)), "msceer_last_timing")
        return t
