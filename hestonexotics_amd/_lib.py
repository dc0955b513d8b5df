"""ctypes binding of the C ABI in include/hexo.h (libhexo.so, built in hestonexotics_amd/csrc).

There is no fallback: if the HIP library is missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "csrc", "libhexo.so")

HEXO_OK = 0
ERRORS = {
    -1: "HEXO_ERR_INVALID", -2: "HEXO_ERR_NO_DEVICE", -3: "HEXO_ERR_HIP", -4: "HEXO_ERR_NOMEM",
    -5: "HEXO_ERR_ADMISSION", -6: "HEXO_ERR_NOT_FOUND", -7: "HEXO_ERR_CANCELLED", -8: "HEXO_ERR_TIMEOUT",
    -9: "HEXO_ERR_STATE", -10: "HEXO_ERR_INJECTED",
}
ASIAN, EUROPEAN, ASIAN_GEO_CV, GEO_ASIAN = 1, 2, 4, 8
DRIFT, DISCOUNT, TRAPEZOID, MARTINGALE, CV_REGRESSION = 1, 2, 4, 8, 16
FP64, FP32, FP32_COMPENSATED = 0, 1, 2


class hexo_request(C.Structure):
    _fields_ = [
        ("v0", C.c_double), ("theta", C.c_double), ("rho", C.c_double), ("kappa", C.c_double),
        ("sigma", C.c_double), ("S0", C.c_double), ("r", C.c_double),
        ("n_paths", C.c_uint64), ("seed", C.c_uint64), ("path_offset", C.c_uint64),
        ("steps", C.c_uint32), ("n_expiries", C.c_uint32),
        ("expiries", C.POINTER(C.c_double)), ("n_strikes", C.POINTER(C.c_uint32)),
        ("strikes", C.POINTER(C.c_double)),
        ("payoffs", C.c_uint32), ("flags", C.c_uint32), ("precision", C.c_uint32), ("reserved", C.c_uint32),
    ]


class hexo_tuning(C.Structure):
    _fields_ = [("paths_per_thread", C.c_uint32), ("block_threads", C.c_uint32),
                ("waves_per_simd", C.c_uint32), ("reserved", C.c_uint32)]


class hexo_batch_stats(C.Structure):
    _fields_ = [("n_requests", C.c_uint64), ("path_steps", C.c_double), ("nominal_path_steps", C.c_double),
                ("last_run_ms", C.c_double), ("path_kernel_ms", C.c_double), ("reduce_kernel_ms", C.c_double),
                ("path_kernel_launches", C.c_uint64)]


class hexo_server_config(C.Structure):
    _fields_ = [("memory_budget_bytes", C.c_uint64), ("streams", C.c_uint32), ("max_batch", C.c_uint32),
                ("batch_window_s", C.c_double)]


class hexo_ticket_info(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("admit_s", C.c_double), ("launch_s", C.c_double),
                ("done_s", C.c_double), ("batch_id", C.c_uint32), ("batch_size", C.c_uint32),
                ("status", C.c_int32)]


class hexo_server_stats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("cancelled", C.c_uint64), ("batches", C.c_uint64), ("max_queue_depth", C.c_uint64),
                ("bytes_in_use", C.c_uint64), ("bytes_peak", C.c_uint64), ("budget_bytes", C.c_uint64),
                ("path_steps_done", C.c_double)]


class hexo_swift_grid(C.Structure):
    _fields_ = [("m", C.c_uint32), ("exp2_m", C.c_uint32), ("sqrt_exp2_m", C.c_double), ("lower", C.c_double),
                ("upper", C.c_double), ("k1", C.c_int32), ("k2", C.c_int32), ("J", C.c_uint32),
                ("reserved", C.c_uint32)]


class hexo_swift_request(C.Structure):
    _fields_ = [("v0", C.c_double), ("theta", C.c_double), ("rho", C.c_double), ("kappa", C.c_double),
                ("sigma", C.c_double), ("S0", C.c_double), ("r", C.c_double), ("T", C.c_double),
                ("n_strikes", C.c_uint32), ("flags", C.c_uint32), ("strikes", C.POINTER(C.c_double)),
                ("trunc", C.c_double), ("grid", hexo_swift_grid)]


class hexo_calib_chain(C.Structure):
    _fields_ = [("T", C.c_double), ("n_options", C.c_uint32), ("reserved", C.c_uint32),
                ("strikes", C.POINTER(C.c_double)), ("prices", C.POINTER(C.c_double)),
                ("volumes", C.POINTER(C.c_double))]


class hexo_calib_config(C.Structure):
    _fields_ = [("max_iter", C.c_uint32), ("flags", C.c_uint32), ("initial", C.c_double * 5),
                ("mu_scale", C.c_double), ("eps_grad", C.c_double), ("eps_step", C.c_double),
                ("eps_err", C.c_double), ("r", C.c_double), ("trunc", C.c_double)]


class hexo_calib_result(C.Structure):
    _fields_ = [("params", C.c_double * 5), ("initial", C.c_double * 5), ("err0", C.c_double), ("err", C.c_double),
                ("iterations", C.c_uint32), ("reason", C.c_uint32), ("n_price_evals", C.c_uint32),
                ("n_jacobian_evals", C.c_uint32)]


CALIB_INITIAL_GIVEN = 1
SWIFT_FIXED_GRID = 1
SWIFT_MAX_J = 16384

_P = C.c_void_p
_d = C.POINTER(C.c_double)
_u32 = C.POINTER(C.c_uint32)
_u64 = C.POINTER(C.c_uint64)
_req = C.POINTER(hexo_request)

SIGNATURES = {
    "hexo_strerror": (C.c_char_p, [C.c_int]),
    "hexo_version": (C.c_char_p, []),
    "hexo_validate_request": (C.c_int, [_req]),
    "hexo_request_n_outputs": (C.c_uint64, [_req]),
    "hexo_time_grid": (C.c_int, [_d, C.c_uint32, _u32, C.c_uint32, _u64, _d, _u64]),
    "hexo_asian_weights": (C.c_int64, [_req, C.c_uint32, _d, _d, C.c_uint64]),
    "hexo_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "hexo_ctx_destroy": (C.c_int, [_P]),
    "hexo_ctx_last_error": (C.c_char_p, [_P]),
    "hexo_device_info": (C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "hexo_price": (C.c_int, [_P, _req, _d, _d]),
    "hexo_price_noalloc": (C.c_int, [_P, _req, _d, _d]),
    "hexo_price_batch": (C.c_int, [_P, _req, C.c_uint32, _d, _d]),
    "hexo_batch_create": (C.c_int, [_P, _req, C.c_uint32, C.POINTER(_P)]),
    "hexo_batch_run": (C.c_int, [_P]),
    "hexo_batch_run_eager": (C.c_int, [_P]),
    "hexo_batch_sync": (C.c_int, [_P]),
    "hexo_batch_results": (C.c_int, [_P, _d, _d]),
    "hexo_batch_stats_get": (C.c_int, [_P, C.POINTER(hexo_batch_stats)]),
    "hexo_batch_branch_counts": (C.c_int, [_P, _u64]),
    "hexo_batch_destroy": (C.c_int, [_P]),
    "hexo_paths": (C.c_int, [_P, _req, C.c_uint64, C.c_uint64, _d]),
    "hexo_geo_asian_price": (C.c_int, [_P, _req, _d]),
    "hexo_rng_fill": (C.c_int, [_P, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, _u32]),
    "hexo_philox4x32_10": (C.c_int, [_P, _u32, _u32, C.c_uint64, _u32]),
    "hexo_normals": (C.c_int, [_P, _u32, C.c_uint64, C.c_uint32, _d]),
    "hexo_get_tuning": (C.c_int, [_P, C.POINTER(hexo_tuning)]),
    "hexo_set_tuning": (C.c_int, [_P, C.POINTER(hexo_tuning)]),
    "hexo_autotune": (C.c_int, [_P, _req, C.c_uint32, C.c_uint32, C.POINTER(hexo_tuning)]),
    "hexo_copy_probe": (C.c_int, [_P, C.c_uint64, C.c_uint32, _d]),
    "hexo_server_start": (C.c_int, [_P, C.POINTER(hexo_server_config)]),
    "hexo_submit": (C.c_int, [_P, _req, _u64]),
    "hexo_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_int)]),
    "hexo_wait": (C.c_int, [_P, C.c_uint64, C.c_double, _d, _d, C.POINTER(hexo_ticket_info)]),
    "hexo_cancel": (C.c_int, [_P, C.c_uint64]),
    "hexo_server_stats_get": (C.c_int, [_P, C.POINTER(hexo_server_stats)]),
    "hexo_server_stop": (C.c_int, [_P]),
    "hexo_server_run_trace": (C.c_int, [_P, _req, _d, C.c_uint32, _d, _d, C.POINTER(hexo_ticket_info)]),
    "hexo_inject_fault": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "hexo_swift_grid_for": (C.c_int, [C.POINTER(hexo_swift_request), C.POINTER(hexo_swift_grid)]),
    "hexo_calib_defaults": (C.c_int, [C.POINTER(hexo_calib_config)]),
    "hexo_calibrate": (C.c_int, [_P, C.c_double, C.POINTER(hexo_calib_chain), C.c_uint32,
                                 C.POINTER(hexo_calib_config), C.POINTER(hexo_calib_result)]),
    "hexo_swift_price": (C.c_int, [_P, C.POINTER(hexo_swift_request), C.c_uint32, _d, _d,
                                   C.POINTER(hexo_swift_grid)]),
}

_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Loads libhexo.so and declares every entry point. Raises if the library is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(the HIP path has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class HexoError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
