"""ctypes binding of ``libwavefd.so`` (the C ABI declared in ``include/wavefd.h``).

The library is loaded from this package directory.  There is no fallback: if it is missing,
``load()`` raises, so nothing can silently run on another path.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libwavefd.so"

WFD_OK, WFD_EINVAL, WFD_EHIP, WFD_ENOMEM, WFD_ENONFINITE = 0, 1, 2, 3, 4
WFD_F32, WFD_F64 = 0, 1
WFD_AC_SCRATCH_BYTES = 64 << 20  # include/wavefd.h


class WaveError(RuntimeError):
    """A wavefd call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[wavefd status {code}] {msg}")
        self.code = code


class Grid(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("nshots", ctypes.c_int32),
        ("n", ctypes.c_int64 * 3),
        ("h", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("ld", ctypes.c_int64),
        ("plane", ctypes.c_int64),
        ("offset", ctypes.c_int64),
        ("elems", ctypes.c_int64),
    ]


class Medium(ctypes.Structure):
    _fields_ = [
        ("coef", ctypes.c_void_p),
        ("kconst", ctypes.c_double),
        ("coef_shot_stride", ctypes.c_int64),
        ("beta", ctypes.c_void_p),
    ]


class Points(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_int32),
        ("nrows", ctypes.c_int32),
        ("row_ptr", ctypes.c_void_p),
        ("col", ctypes.c_void_p),
        ("id", ctypes.c_void_p),
        ("row", ctypes.c_void_p),
        ("data", ctypes.c_void_p),
        ("width", ctypes.c_int64),
        ("nsteps", ctypes.c_int64),
        ("shot_stride", ctypes.c_int64),
    ]


class BatchShot(ctypes.Structure):
    _fields_ = [
        ("grid", Grid),
        ("nsteps", ctypes.c_int64),
        ("seg", ctypes.c_int64),
        ("med", Medium),
        ("src", Points),
        ("rec", Points),
        ("observed", ctypes.c_void_p),
        ("adj", ctypes.c_void_p),
        ("mu_src", ctypes.c_void_p),
        ("G", ctypes.c_void_p),
    ]


class BatchOpts(ctypes.Structure):
    _fields_ = [
        ("max_active_cells", ctypes.c_int64),
        ("blocks_per_cu", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("timeout_ms", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class AcMedium(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_void_p),
        ("cn", ctypes.c_void_p),
        ("cw", ctypes.c_void_p),
        ("beta", ctypes.c_void_p),
    ]


# name -> (restype, argtypes); every symbol declared in include/wavefd.h
_P = ctypes.c_void_p
_GP = ctypes.POINTER(Grid)
_MP = ctypes.POINTER(Medium)
_PP = ctypes.POINTER(Points)
_I64 = ctypes.c_int64
_I64P = ctypes.POINTER(ctypes.c_int64)
_I32P = ctypes.POINTER(ctypes.c_int32)
_DP = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "wfd_last_error": (ctypes.c_char_p, []),
    "wfd_version": (ctypes.c_char_p, []),
    "wfd_grid_init": (ctypes.c_int, [_GP, ctypes.c_int, _I64P, _DP, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int]),
    "wfd_stable_dt": (ctypes.c_double, [ctypes.c_int, _DP, ctypes.c_double, ctypes.c_int]),
    "wfd_stencil_weights": (ctypes.c_int, [_GP, _DP, _DP]),
    "wfd_points_csr": (ctypes.c_int, [_GP, ctypes.c_int32, _I64P, _I32P, _I32P, _I32P, _I32P]),
    "wfd_points_nrows": (ctypes.c_int32, [_GP]),
    "wfd_step": (ctypes.c_int, [_GP, _P, _P, _P, _MP, _PP, _PP, _I64, _P]),
    "wfd_forward": (ctypes.c_int, [_GP, _P, _P, _MP, _PP, _PP, _I64, _I64, _P]),
    "wfd_forward_graph": (ctypes.c_int, [_GP, _P, _P, _MP, _PP, _PP, _I64, _I64, ctypes.c_int32, _P]),
    "wfd_fgraph_create": (ctypes.c_int, [ctypes.POINTER(_P), _GP, _P, _P, _MP, _PP, _PP, ctypes.c_int32]),
    "wfd_fgraph_run": (ctypes.c_int, [_P, _I64, _I64, _P]),
    "wfd_fgraph_destroy": (ctypes.c_int, [_P]),
    "wfd_forward2": (ctypes.c_int, [_GP, _P, _P, _P, _P, _MP, _PP, _PP, _I64, _I64, ctypes.c_int32, _P]),
    "wfd_adjoint_step": (ctypes.c_int, [_GP, _P, _P, _P, _MP, _PP, _PP, _I64, _P, _P, _P]),
    "wfd_grad_flush": (ctypes.c_int, [_GP, _P, _P, _P]),
    "wfd_check_medium_async": (ctypes.c_int, [_GP, _MP, _P, _P]),
    "wfd_check_medium": (ctypes.c_int, [_GP, _MP, _P]),
    "wfd_tune": (ctypes.c_int, [_GP, _P, _P, _MP, ctypes.c_int32, _I32P, _I32P, _DP, _P]),
    "wfd_set_config": (ctypes.c_int, [_GP, ctypes.c_int32, ctypes.c_int32]),
    "wfd_get_config": (ctypes.c_int, [_GP, _I32P, _I32P]),
    "wfd_plan_create": (ctypes.c_int, [ctypes.POINTER(_P), _GP, _I64, _I64, ctypes.c_size_t]),
    "wfd_plan_destroy": (ctypes.c_int, [_P]),
    "wfd_plan_info": (ctypes.c_int, [_P, _I64P, _I64P, ctypes.POINTER(ctypes.c_size_t)]),
    "wfd_plan_forward": (ctypes.c_int, [_P, _MP, _PP, _PP, _P, _P]),
    "wfd_plan_backward": (ctypes.c_int, [_P, _MP, _PP, _PP, _PP, _P, _P]),
    "wfd_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int32, ctypes.POINTER(BatchShot),
                                        ctypes.POINTER(BatchOpts)]),
    "wfd_batch_destroy": (ctypes.c_int, [_P]),
    "wfd_batch_run": (ctypes.c_int, [_P, ctypes.c_int32, _P]),
    "wfd_batch_status": (ctypes.c_int, [_P, _P]),
    "wfd_batch_tune": (ctypes.c_int, [_P, ctypes.c_int32, _P]),
    "wfd_batch_set_config": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "wfd_batch_info": (ctypes.c_int, [_P, _I64P, _I64P, ctypes.POINTER(ctypes.c_size_t), _I64P]),
    "wfd_batch_shot_info": (ctypes.c_int, [_P, ctypes.c_int32, _I32P, _I32P, _I32P, _I64P, _I64P, _I64P]),
    "wfd_batch_shot_ops": (ctypes.c_int, [_I64, _I64, ctypes.c_int32, ctypes.c_int32, _I32P, _I64, _I64P, _I32P]),
    "wfd_fma_peak": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, _P]),
    "wfd_stream_bench": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int64, ctypes.c_int32, _P]),
    "wfd_tfm": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, _P, _P, _I64, _P, ctypes.c_int32, _P]),
    "wfd_ac_forward": (ctypes.c_int, [_GP, _P, _P, ctypes.POINTER(AcMedium), _PP, _PP, _PP, _I64, _I64, _P]),
    "wfd_ac_backward": (ctypes.c_int, [_GP, _P, _P, _P, _P, ctypes.POINTER(AcMedium), _PP, _PP, _PP, _PP, _I64,
                                       _P, _P, _P]),
    "wfd_ac_backward_noalloc": (ctypes.c_int, [_GP, _P, _P, ctypes.POINTER(AcMedium), _PP, _PP, _PP, _PP, _I64,
                                               _P, _P]),
    "wfd_ac_set_rows": (ctypes.c_int, [ctypes.c_int32]),
    "wfd_ref_step2d": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int32, _P]),
    "wfd_ref_step1d": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int32, _P]),
    "wfd_ref_temporal": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_int32, _P]),
    "wfd_ref_mixed_step": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P]),
    "wfd_ref_scaled_step": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_int32, _P]),
}

_lib = None


def load():
    """Load ``libwavefd.so`` once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python neuralmech_amd/_build.py`")
    lib = ctypes.CDLL(str(LIB_PATH), mode=getattr(os, "RTLD_LOCAL", 0))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().wfd_last_error().decode(errors="replace")


def check(rc: int) -> None:
    if rc != WFD_OK:
        raise WaveError(rc, last_error())
