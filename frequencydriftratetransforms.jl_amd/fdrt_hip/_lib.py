"""ctypes binding of libzdt_hip.so (include/zdt.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared object is missing (build it with ``python build.py`` in the
package directory, or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("ZDT_LIB", _HERE.parent / "lib" / "libzdt_hip.so"))

ZDT_OK = 0
ZDT_EINVAL = 1
ZDT_ENOMEM = 2
ZDT_EHIP = 3
ZDT_ENOTSUP = 4
ZDT_ETRUNC = 5
ZDT_ESTATE = 6

ZDT_WINDOW_RECT = 0
ZDT_WINDOW_HAMMING = 1
ZDT_WINDOW_CUSTOM = 2

ZDT_FLAG_UNNORMALIZED = 0x1
ZDT_FLAG_GENERIC = 0x2
ZDT_FLAG_PROFILE = 0x4
ZDT_FLAG_N2_512 = 0x8
ZDT_FLAG_RADIX8 = 0x10

ZDT_PLAN_AUTOTUNE = 0x100

ZDT_PRECISION_FP32 = 0
ZDT_PRECISION_SPLIT_BF16 = 1
ZDT_PRECISION_SPLIT_FP16 = 2
PRECISIONS = {"fp32": ZDT_PRECISION_FP32, "split-bf16": ZDT_PRECISION_SPLIT_BF16,
              "split-fp16": ZDT_PRECISION_SPLIT_FP16}

ZDT_PATH_GENERIC = 0
ZDT_PATH_FUSED = 1
ZDT_PATH_FUSED_OUTPUT = 2
PATHS = {ZDT_PATH_GENERIC: "generic", ZDT_PATH_FUSED: "fused", ZDT_PATH_FUSED_OUTPUT: "fused-output"}

ZDT_BUF_F = 0
ZDT_BUF_Y = 1
ZDT_BUF_V = 2

_i64 = C.c_int64
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p
_fp = C.c_void_p  # device float*

# name -> (restype, argtypes); every symbol include/zdt.h declares
SIGNATURES = {
    "zdt_last_error": (C.c_char_p, []),
    "zdt_version": (C.c_int, []),
    "zdt_calc_nl": (_i64, [_i64, _i64, C.POINTER(C.c_int), C.c_int]),
    "zdt_grow_nr": (_i64, [_i64, _i64, C.POINTER(C.c_int), C.c_int]),
    "zdt_estimate_memory": (_i64, [_i64, _i64, _i64, _i64, _i64, C.POINTER(C.c_int), C.c_int]),
    "zdt_workspace_bytes": (_i64, [_i64, _i64, _i64, C.POINTER(C.c_int), C.c_int, C.c_int]),
    "zdt_create": (C.c_int, [_fp, _i64, _i64, C.c_float, C.c_float, _i64, C.POINTER(C.c_int), C.c_int,
                             C.c_int, C.c_int, _vp, C.POINTER(_vp)]),
    "zdt_destroy": (C.c_int, [_vp]),
    "zdt_info": (C.c_int, [_vp, _i64p, _i64p, C.POINTER(C.c_float), C.POINTER(C.c_float), _i64p, _i64p]),
    "zdt_path": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "zdt_sizeof": (C.c_int, [_vp, _i64p]),
    "zdt_set_stream": (C.c_int, [_vp, _vp]),
    "zdt_sync": (C.c_int, [_vp]),
    "zdt_input": (C.c_int, [_vp, _fp]),
    "zdt_run": (C.c_int, [_vp, C.c_float, C.c_int, _fp, _fp]),
    "zdt_run_batches": (C.c_int, [_vp, C.POINTER(C.c_float), C.c_int, C.c_int, _fp, C.POINTER(_vp)]),
    "zdt_output": (C.c_int, [_vp, _fp]),
    "zdt_preprocess": (C.c_int, [_vp, C.c_float]),
    "zdt_convolve": (C.c_int, [_vp]),
    "zdt_postprocess": (C.c_int, [_vp, C.c_int, _fp]),
    "zdt_buffer": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), _i64p, _i64p]),
    "zdt_fftfdr": (C.c_int, [_vp, C.POINTER(C.c_float), _i64, _fp]),
    "zdt_fdshift": (C.c_int, [_vp, C.c_float, _fp]),
    "zdt_profile_reset": (C.c_int, [_vp]),
    "zdt_profile_read": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.c_char_p, C.c_int, C.POINTER(C.c_double),
                                   _i64p, C.POINTER(C.c_double)]),
    "zdt_fdrstats": (C.c_int, [C.POINTER(_vp), C.c_int, _i64, _i64, _i64, _vp, C.POINTER(C.c_double),
                               C.POINTER(C.c_double)]),
    "zdt_fdrnormalize": (C.c_int, [C.POINTER(_vp), C.c_int, _i64, _i64, _i64, C.c_double, C.c_double, _vp]),
    "zdt_findprotohits": (C.c_int, [C.POINTER(_vp), C.c_int, _i64, _i64, _i64, C.c_double, _vp, _i64, _i64p,
                                    _vp]),
    "zdt_calib_valu": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "zdt_calib_lds": (C.c_int, [C.c_int, C.POINTER(C.c_double)]),
    "zdt_plan_tuning": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "zdt_tune_table": (C.c_int, [C.c_char_p]),
    "zdt_bshuf_decompress": (C.c_int, [_vp, C.c_size_t, _vp, C.c_size_t, C.c_size_t, C.c_size_t, C.c_int]),
    "zdt_clusterprotohits": (C.c_int, [_vp, _i64, C.POINTER(_vp), C.c_int, _i64, _i64, _i64, C.c_int, C.c_int,
                                       _vp, _vp, _vp, _vp, _i64, _i64p, _vp, _vp]),
    "zdt_findprotohits_noalloc": (C.c_int, [C.POINTER(_vp), C.c_int, _i64, _i64, _i64, C.c_double, _vp, _i64,
                                            _i64p, _vp]),
    "zdt_plan_ragged": (C.c_int, [C.c_int, _i64p, _i64p, C.POINTER(C.c_float), C.POINTER(C.c_float), _i64p,
                                  C.c_int, _vp, C.POINTER(_vp)]),
    "zdt_plan_destroy": (C.c_int, [_vp]),
    "zdt_plan_info": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_int)]),
    "zdt_plan_bind": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(C.c_double), C.c_int, _vp, _i64,
                                _vp, _vp]),
    "zdt_plan_launch": (C.c_int, [_vp]),
    "zdt_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(_vp)]),
    "zdt_host_free": (C.c_int, [_vp]),
    "zdt_stream_create": (C.c_int, [_i64, _i64, C.c_float, C.c_float, _i64, _i64, _i64, C.c_int, C.c_int,
                                    C.POINTER(_vp)]),
    "zdt_stream_workspace": (C.c_int, [_vp, C.POINTER(_vp)]),
    "zdt_stream_destroy": (C.c_int, [_vp]),
    "zdt_stream_info": (C.c_int, [_vp, _i64p, _i64p, C.POINTER(C.c_int), _i64p]),
    "zdt_stream_search": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.c_double, _vp, _i64, _vp, _vp]),
    "zdt_stream_findhits": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.c_double, C.c_int, C.c_int, _i64, _vp, _vp,
                                      _vp, _vp, _i64, _vp, _vp, _vp]),
}


class ZDTError(RuntimeError):
    """Raised for a non-zero status; carries zdt_last_error()."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[zdt {code}] {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        # torch first: its HIP runtime (soname libamdhip64.so.7) must be the
        # one the library binds to, so device pointers and streams are shared
        import torch  # noqa: F401
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} not found: the HIP library is required (no CPU fallback). "
                "Build it with `python build.py` in frequencydriftratetransforms.jl_amd/.")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != ZDT_OK:
        msg = lib().zdt_last_error().decode(errors="replace")
        raise ZDTError(rc, msg)


def int_array(vals):
    vals = [int(v) for v in vals]
    return (C.c_int * len(vals))(*vals), len(vals)
