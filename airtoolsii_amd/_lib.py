"""ctypes binding of libairhip.so (include/airhip.h).

The library is built in-tree by ``airtoolsii_amd.build``; there is no CPU
fallback: every product entry point fails loudly when it is missing or when no
HIP device is visible.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libairhip.so")

AIR_OK = 0
AIR_E_ARG, AIR_E_NOMEM, AIR_E_HIP, AIR_E_UNSUPPORTED, AIR_E_STATE, AIR_E_GEOMETRY = -1, -2, -3, -4, -5, -6
DTYPES = {"f32": 0, "float32": 0, "bf16": 1, "bfloat16": 1, "f16": 2, "float16": 2, "f64": 3, "float64": 3}
METHODS = {"landweber": 0, "cimmino": 1, "cav": 2, "drop": 3, "sart": 4, "custom": 5}
RELAX = {"default": 0, "scalar": 1, "line": 2, "psi1": 3, "psi1mod": 4, "psi2": 5, "psi2mod": 6, "vector": 7,
         "learned": 8, "greedy": 9}
STOP = {"NONE": 0, "NCP": 1, "DP": 2, "ME": 3}
STOP_LEARNED = 4
BOUND_NONE, BOUND_SCALAR, BOUND_VECTOR = 0, 1, 2

# every symbol include/airhip.h declares
EXPORTS = ["air_last_error", "air_create", "air_destroy", "air_synchronize", "air_set_matrix_csr",
           "air_set_matrix_csc", "air_build_paralleltomo", "air_build_fanlineartomo", "air_build_fancurvedtomo", "air_build_seismictomo",
           "air_build_sphericaltomo", "air_build_seismicwavetomo", "air_fbp", "air_art", "air_cart",
           "air_set_graph", "air_graph_launches",
           "air_size", "air_get_csr",
           "air_apply", "air_apply_device", "air_weights", "air_spectral_radius", "air_calczeta", "air_sirt", "air_sirt_prepare",
           "air_sirt_iterate", "air_sirt_fetch", "air_set_kernel_timing", "air_get_kernel_timing",
           "air_layout_info", "air_mlp_forward", "air_mlp_grad", "air_mlp_train", "air_sirt_records",
           "air_build_paralleltomo_batch", "air_batch_info", "air_sirt_multi", "air_set_afun"]


# int fn(void* ctx, int transp, const double* in, double* out): air_set_afun's operator
AFUN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double))


class AirError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"airhip error {code}: {msg}")
        self.code = code


class SirtOptions(C.Structure):
    _fields_ = [("method", C.c_int), ("M", C.c_void_p), ("D", C.c_void_p),
                ("relax_kind", C.c_int), ("relaxpar", C.c_double), ("relax_vec", C.c_void_p),
                ("rho", C.c_double),
                ("lb_kind", C.c_int), ("lb", C.c_double), ("lb_vec", C.c_void_p),
                ("ub_kind", C.c_int), ("ub", C.c_double), ("ub_vec", C.c_void_p),
                ("stoprule", C.c_int), ("taudelta", C.c_double), ("res_dims", C.c_int * 2),
                ("ncp_smooth", C.c_int), ("verbose", C.c_int),
                ("mlp_params", C.c_void_p), ("noise_norm", C.c_double), ("x_true", C.c_void_p)]


class Layout(C.Structure):
    _fields_ = [("tiled", C.c_int), ("value_bytes", C.c_int), ("m", C.c_int64), ("n", C.c_int64),
                ("nnz", C.c_int64), ("n_internal", C.c_int64), ("pieces", C.c_int64), ("runs", C.c_int64),
                ("tiles_fwd", C.c_int64), ("tiles_bwd", C.c_int64), ("staged_rows", C.c_int64),
                ("lds_bytes", C.c_int64), ("fwd_stream", C.c_int), ("fwd_lanes", C.c_int), ("fwd_loads", C.c_int),
                ("bwd_lanes", C.c_int), ("bwd_loads", C.c_int), ("fwd_pieces_per_block", C.c_int),
                ("fwd_entries", C.c_int64), ("bwd_entries", C.c_int64),
                ("tune_fwd_us", C.c_float * 64), ("tune_bwd_us", C.c_float * 16),
                ("fwd_step", C.c_int), ("fwd_step_ok", C.c_int), ("bwd_step", C.c_int), ("bwd_step_ok", C.c_int),
                ("bwd_step_entries", C.c_int64), ("fwd_waves_per_block", C.c_int),
                ("bwd_step_base_bits", C.c_int), ("bwd_step_lds_bytes", C.c_int64)]


class RowOptions(C.Structure):
    _fields_ = [("relax_kind", C.c_int), ("relaxpar", C.c_double), ("relax_vec", C.c_void_p),
                ("lb_kind", C.c_int), ("lb", C.c_double), ("lb_vec", C.c_void_p),
                ("ub_kind", C.c_int), ("ub", C.c_double), ("ub_vec", C.c_void_p),
                ("stoprule", C.c_int), ("taudelta", C.c_double), ("damp", C.c_double), ("THR", C.c_double),
                ("Kbegin", C.c_int), ("Nunflag", C.c_int), ("seed", C.c_uint64), ("res_dims", C.c_int * 2),
                ("ncp_smooth", C.c_int)]


class SirtInfo(C.Structure):
    _fields_ = [("stoprule", C.c_int), ("finaliter", C.c_int), ("relaxpar", C.c_double), ("rho", C.c_double),
                ("nsaved", C.c_int), ("timetaken", C.c_double), ("setup_time", C.c_double),
                ("iter_time", C.c_double)]


_lib = None
_P = C.c_void_p
_i64 = C.c_int64
_dp = C.POINTER(C.c_double)


def load():
    """Load libairhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m airtoolsii_amd.build` "
                           "(the engine has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    L.air_last_error.restype = C.c_char_p
    L.air_create.argtypes = [C.c_int, C.POINTER(_P)]
    L.air_destroy.argtypes = [_P]
    L.air_destroy.restype = None
    L.air_synchronize.argtypes = [_P]
    L.air_set_matrix_csr.argtypes = [_P, _i64, _i64, _P, _P, _P, C.c_int]
    L.air_set_matrix_csc.argtypes = [_P, _i64, _i64, _P, _P, _P, C.c_int]
    L.air_set_afun.argtypes = [_P, _i64, _i64, AFUN, C.c_void_p]
    L.air_build_paralleltomo.argtypes = [_P, C.c_int32, _P, C.c_int32, C.c_int32, C.c_double, C.c_int, C.c_int]
    L.air_build_fanlineartomo.argtypes = [_P, C.c_int32, _P, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                          C.c_double, C.c_int, C.c_int]
    L.air_build_fancurvedtomo.argtypes = [_P, C.c_int32, _P, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                          C.c_int, C.c_int]
    L.air_build_seismictomo.argtypes = [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int, C.c_int]
    L.air_build_sphericaltomo.argtypes = [_P, C.c_int32, _P, C.c_int32, C.c_int32, C.c_int, C.c_int]
    L.air_build_seismicwavetomo.argtypes = [_P, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int, C.c_int]
    L.air_fbp.argtypes = [_P, _P, C.c_int32, C.c_int, _P]
    L.air_art.argtypes = [_P, C.c_int, _P, _i64, C.POINTER(RowOptions), _P, _P, _P, C.c_int32, _P, _P,
                          C.POINTER(C.c_int32), _P, C.POINTER(SirtInfo)]
    L.air_cart.argtypes = [_P, _P, _i64, C.POINTER(RowOptions), _P, _P, _P, C.c_int32, _P, _P,
                           C.POINTER(C.c_int32), _P, C.POINTER(SirtInfo)]
    L.air_size.argtypes = [_P, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)]
    L.air_get_csr.argtypes = [_P, _P, _P, _P]
    L.air_apply.argtypes = [_P, C.c_int, _P, _P]
    L.air_apply_device.argtypes = [_P, C.c_int, _P, _P]
    L.air_weights.argtypes = [_P, C.c_int, _P, _P]
    L.air_spectral_radius.argtypes = [_P, C.c_int, _P, _P, _dp]
    L.air_calczeta.argtypes = [C.c_int32, _P]
    L.air_sirt.argtypes = [_P, C.POINTER(SirtOptions), _P, _P, _P, C.c_int32, _P, _P, C.POINTER(C.c_int32), _P,
                           C.POINTER(SirtInfo)]
    L.air_sirt_prepare.argtypes = [_P, C.POINTER(SirtOptions), _P, _P, C.c_int32]
    L.air_sirt_iterate.argtypes = [_P, C.c_int32, _dp]
    L.air_sirt_fetch.argtypes = [_P, _P, _P, C.POINTER(SirtInfo)]
    L.air_set_kernel_timing.argtypes = [_P, C.c_int]
    L.air_set_graph.argtypes = [_P, C.c_int32]
    L.air_graph_launches.argtypes = [_P, C.POINTER(_i64)]
    L.air_get_kernel_timing.argtypes = [_P, _dp, _dp, _dp, C.POINTER(_i64)]
    L.air_layout_info.argtypes = [_P, C.POINTER(Layout)]
    L.air_mlp_forward.argtypes = [_P, _P, _P, _i64, _P]
    L.air_sirt_records.argtypes = [_P, _P, _P, _P, _P]
    L.air_build_paralleltomo_batch.argtypes = [_P, C.c_int32, _P, _P, _P, _P, C.c_int, C.c_int]
    L.air_batch_info.argtypes = [_P, _P, _P, _P, _P, _P]
    L.air_sirt_multi.argtypes = [_P, C.POINTER(SirtOptions), C.c_int32, _P, _P, C.c_int32, _P, _P, _dp]
    L.air_mlp_grad.argtypes = [_P, _P, _P, _P, _i64, C.c_double, _P, _dp]
    L.air_mlp_train.argtypes = [_P, _P, _P, _P, _i64, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64, _P]
    for name in EXPORTS:
        if name not in ("air_last_error", "air_destroy"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc):
    if rc != AIR_OK:
        msg = load().air_last_error().decode(errors="replace")
        raise AirError(rc, msg)
    return rc


class _ArrayPtr(C.c_void_p):
    """A data pointer that keeps its array alive for as long as the pointer
    object exists, so that ``ptr(np.ascontiguousarray(...))`` may be passed
    straight into a call even when the conversion made a temporary copy."""


def ptr(a):
    """Data pointer of a C-contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    p = _ArrayPtr(a.ctypes.data)
    p._array = a
    return p


def f64(a):
    return np.ascontiguousarray(np.ravel(np.asarray(a, dtype=np.float64)))


class Engine:
    """Owns one air_engine handle (one device, one stream, one operator)."""

    def __init__(self, device: int = 0):
        L = load()
        h = C.c_void_p()
        check(L.air_create(int(device), C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            load().air_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self):
        m, n, nnz = _i64(), _i64(), _i64()
        check(load().air_size(self.h, C.byref(m), C.byref(n), C.byref(nnz)))
        return m.value, n.value, nnz.value


def nan():
    return math.nan
