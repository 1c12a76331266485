"""ctypes binding of libsptrsv_amd.so (declarations in include/sptrsv_amd.h).

The library is the only compute path: if it is missing or fails to load,
``lib()`` raises ``SptrsvUnavailable`` -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import (POINTER, Structure, c_char_p, c_double, c_float, c_int, c_int64, c_size_t,
                    c_uint64, c_void_p)
from pathlib import Path

LIB_PATH = Path(__file__).resolve().with_name("libsptrsv_amd.so")

# status codes (sptrsv_amd.h)
OK = 0
E_NO_ANALYSIS_NEEDED = -1
E_INVALID_ARG = -2
E_INVALID_CSR = -3
E_ALLOC = -4
E_HIP = -5
E_WATCHDOG = -6
E_NOT_ANALYSED = -7
E_UNAVAILABLE = -8
E_IO = -9
E_CAPACITY = -10

# MODE (dfr_syncfree.h:7-12)
SIMPLE, ORDER, MULTIROW, ALL = 0, 1, 2, 3

KERNEL_AUTO, KERNEL_SIMPLE, KERNEL_ORDER, KERNEL_MULTIROW, KERNEL_JACOBI = 0, 1, 2, 3, 4
E_NOT_CONVERGED = -11
PREC_F64, PREC_F32, PREC_MIXED = 0, 1, 2
NOALLOC_MAX_WORDS = 1 << 22  # SPTRSV_NOALLOC_MAX_WORDS


class SptrsvUnavailable(ImportError):
    """The HIP library is not built or cannot be loaded."""


class SptrsvError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{message} (status {status})")
        self.status = status


class SpMat(Structure):  # sp_mat_t, dfr_syncfree.h:14-23
    _fields_ = [("nr", c_int), ("nc", c_int), ("nnz", c_int), ("ia", c_void_p), ("ja", c_void_p),
                ("a", c_void_p), ("a32", c_void_p), ("priv", c_void_p)]


class AnalysisInfo(Structure):  # dfr_analysis_info_t, dfr_syncfree.h:25-41
    _fields_ = [("mode", c_int), ("lev_ctr", c_int), ("nlevs", c_int), ("lev_size", c_void_p),
                ("warp_lev", c_void_p), ("n_warps", c_int), ("inv_iorder", c_void_p),
                ("iorder", c_void_p), ("ibase_row", c_void_p), ("ivect_size_warp", c_void_p),
                ("row_ctr", c_void_p), ("n", c_int), ("level", c_void_p), ("bucket_off", c_void_p),
                ("n_waves", c_int), ("ibase_row64", c_void_p), ("ivect_size_wave", c_void_p),
                ("priv", c_void_p)]


class Config(Structure):  # sptrsv_config_t
    _fields_ = [("kernel", c_int), ("lanes", c_int), ("wg", c_int), ("fused", c_int), ("variant", c_int),
                ("cu_waves", c_int), ("sweeps", c_int), ("outer", c_int), ("window", c_int),
                ("sweeps_next", c_int), ("inner", c_int), ("claim_rounds", c_int),
                ("split_rows", c_int), ("reorder", c_int)]


class BenchResult(Structure):  # sptrsv_bench_t
    _fields_ = [("mean_ms", c_float), ("stdev_ms", c_float), ("min_ms", c_float),
                ("power_mw", c_float), ("errors", c_int), ("reps", c_int)]


_P_INT = POINTER(c_int)
_P_PINT = POINTER(POINTER(c_int))
_P_PDBL = POINTER(POINTER(c_double))

# name -> (restype, argtypes)
_PROTOS = {
    "sptrsv_status_string": (c_char_p, [c_int]),
    "sptrsv_last_error": (c_char_p, []),
    "allocate_memory": (c_int, [c_void_p, c_void_p, c_void_p, c_int, POINTER(SpMat), c_int]),
    "allocate_memory_f32": (c_int, [c_void_p, c_void_p, c_void_p, c_int, POINTER(SpMat), c_int]),
    "allocate_memory_upper": (c_int, [c_void_p, c_void_p, c_void_p, c_int, POINTER(SpMat), c_int]),
    "sptrsv_transpose": (c_int, [POINTER(SpMat), POINTER(SpMat)]),
    "sptrsv_update_values": (c_int, [POINTER(SpMat), c_void_p, c_int64, c_int]),
    "sptrsv_spsv_create": (c_int, [POINTER(c_void_p), c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                   c_int]),
    "sptrsv_spsv_analysis": (c_int, [c_void_p, c_int, c_int, c_int]),
    "sptrsv_spsv_solve": (c_int, [c_void_p, c_int, ctypes.c_double, c_void_p, c_void_p, c_int, c_void_p]),
    "sptrsv_spsv_matrix": (POINTER(SpMat), [c_void_p]),
    "sptrsv_spsv_destroy": (None, [c_void_p]),
    "test_solve_L_analysis_multirow": (c_int, [c_char_p, c_void_p, c_void_p, c_void_p, c_int]),
    "test_solve_L_analysis_multirow_f32": (c_int, [c_char_p, c_void_p, c_void_p, c_void_p, c_int]),
    "run_analysis": (c_int, [POINTER(AnalysisInfo), POINTER(SpMat)]),
    "run_solver": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), c_void_p, c_void_p, c_int]),
    "free_device_analysis": (None, [POINTER(SpMat), POINTER(AnalysisInfo)]),
    "multirow_analysis_base_GPU": (c_int, [POINTER(POINTER(AnalysisInfo)), POINTER(SpMat), c_int]),
    "sptrsv_save_analysis": (c_int, [c_char_p, POINTER(SpMat), POINTER(AnalysisInfo)]),
    "sptrsv_load_analysis": (c_int, [c_char_p, POINTER(SpMat), POINTER(AnalysisInfo)]),
    "sptrsv_copy_analysis": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), POINTER(SpMat), POINTER(AnalysisInfo)]),
    "sptrsv_analysis_segments": (c_int, [POINTER(AnalysisInfo), POINTER(c_int), POINTER(c_int)]),
    "csr_L_solve_simple": (c_int, [POINTER(SpMat), c_void_p, c_void_p, c_int, c_void_p]),
    "csr_L_solve_order": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), c_void_p, c_void_p, c_int, c_void_p]),
    "csr_L_solve_multirow": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), c_void_p, c_void_p, c_int, c_void_p]),
    "csr_L_solve_simple_f32": (c_int, [POINTER(SpMat), c_void_p, c_void_p, c_int, c_void_p]),
    "csr_L_solve_order_f32": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), c_void_p, c_void_p, c_int, c_void_p]),
    "csr_L_solve_multirow_f32": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), c_void_p, c_void_p, c_int, c_void_p]),
    "sptrsv_solve": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), c_int, POINTER(Config), c_void_p,
                             c_void_p, c_int, c_int, c_int, c_void_p]),
    "sptrsv_solve_graph": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), c_int, POINTER(Config), c_void_p,
                                   c_void_p, c_int, c_int, c_int, c_void_p]),
    "csr_L_solve_mixed": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), c_void_p, c_void_p, c_int, c_int,
                                  c_int, c_void_p]),
    "sptrsv_solve_mixed_noalloc": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), POINTER(Config), c_void_p,
                                           c_void_p, c_int, c_int, c_void_p]),
    "sptrsv_reserve": (c_int, [POINTER(SpMat), c_int]),
    "sptrsv_check": (c_int, [POINTER(SpMat), c_void_p]),
    "sptrsv_set_watchdog_ms": (c_int, [POINTER(SpMat), c_double]),
    "sptrsv_autotune": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), c_int, c_int, c_int, POINTER(Config),
                                POINTER(c_float)]),
    "sptrsv_default_config": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), c_int, c_int, POINTER(Config)]),
    "sptrsv_batch_create": (c_int, [c_int, _P_INT, POINTER(c_void_p), POINTER(c_void_p), POINTER(c_void_p),
                                    POINTER(c_void_p)]),
    "sptrsv_batch_rows": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_int64)]),
    "sptrsv_batch_tune": (c_int, [c_void_p, c_int, c_int]),
    "sptrsv_batch_solve": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "sptrsv_batch_solve_graph": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "sptrsv_batch_row_offsets": (c_int, [c_void_p, POINTER(c_int64)]),
    "sptrsv_batch_choice": (c_int, [c_void_p, c_int, POINTER(Config)]),
    "sptrsv_batch_check": (c_int, [c_void_p, c_void_p]),
    "sptrsv_batch_schedule": (c_int, [c_void_p, _P_INT, _P_INT]),
    "sptrsv_batch_patterns": (c_int, [c_void_p, _P_INT, c_void_p]),
    "sptrsv_batch_set_schedule": (c_int, [c_void_p, c_int, c_int]),
    "sptrsv_batch_destroy": (None, [c_void_p]),
    "csr_L_rocsparse_create": (c_int, [POINTER(SpMat), POINTER(c_void_p), POINTER(c_float)]),
    "csr_L_solve_rocsparse": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "csr_L_rocsparse_itsv_create": (c_int, [POINTER(SpMat), POINTER(c_void_p), POINTER(c_float)]),
    "csr_L_solve_rocsparse_itsv": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_int, c_int, _P_INT,
                                           c_void_p]),
    "csr_L_rocsparse_destroy": (None, [c_void_p]),
    "amdsmiAPIRun": (c_int, [c_char_p, c_int, c_int]),
    "amdsmiAPIEnd": (c_float, []),
    "amdsmiAPISamples": (c_int, []),
    "sptrsv_bench_solve": (c_int, [POINTER(SpMat), POINTER(AnalysisInfo), c_int, POINTER(Config), c_void_p,
                                   c_void_p, c_int, c_int, c_int, c_int, c_int, POINTER(BenchResult)]),
    "mm_read_banner": (c_int, [c_void_p, c_void_p]),
    "mm_read_mtx_crd_size": (c_int, [c_void_p, _P_INT, _P_INT, _P_INT]),
    "mm_is_valid": (c_int, [c_char_p]),
    "mm_typecode_to_str_r": (c_char_p, [c_char_p, c_char_p]),
    "sptrsv_read_mtx_lower": (c_int, [c_char_p, c_int, _P_INT, _P_INT, _P_PINT, _P_PINT, _P_PDBL]),
    "sptrsv_read_mtx_coo": (c_int, [c_char_p, c_int, _P_INT, _P_INT, _P_INT, _P_PINT, _P_PINT, _P_PDBL]),
    "sptrsv_coo_to_lower": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, POINTER(SpMat)]),
    "sptrsv_save_csr": (c_int, [c_char_p, c_int, c_void_p, c_void_p, c_void_p]),
    "sptrsv_load_csr": (c_int, [c_char_p, _P_INT, _P_INT, _P_PINT, _P_PINT, _P_PDBL]),
    "sptrsv_free_host": (None, [c_void_p]),
    "sptrsv_generate": (c_int, [c_int, c_int, c_uint64, _P_INT, _P_INT, _P_PINT, _P_PINT, _P_PDBL, _P_INT,
                                _P_PDBL]),
    "sptrsv_generate_batch_member": (c_int, [c_int, c_int, c_int, c_uint64, _P_INT, _P_INT, _P_PINT, _P_PINT,
                                             _P_PDBL, _P_PDBL]),
    "sptrsv_device_count": (c_int, [_P_INT]),
    "sptrsv_device_name": (c_int, [c_int, c_char_p, c_int]),
    "sptrsv_malloc": (c_int, [POINTER(c_void_p), c_size_t]),
    "sptrsv_free": (c_int, [c_void_p]),
    "sptrsv_memcpy_htod": (c_int, [c_void_p, c_void_p, c_size_t]),
    "sptrsv_memcpy_dtoh": (c_int, [c_void_p, c_void_p, c_size_t]),
    "sptrsv_memset": (c_int, [c_void_p, c_int, c_size_t]),
    "sptrsv_synchronize": (c_int, []),
    "sptrsv_copy_device": (c_int, [c_void_p, c_void_p, c_size_t]),
    "sptrsv_timer_start": (c_int, [c_void_p]),
    "sptrsv_timer_stop": (c_int, [c_void_p, POINTER(c_float)]),
    "sptrsv_flush_cache": (c_int, [c_void_p]),
    "sptrsv_kernel_timer": (c_int, [c_int]),
    "sptrsv_kernel_timer_read": (c_int, [POINTER(c_float), _P_INT]),
    "sptrsv_measure_dep_latency": (c_int, [c_int, POINTER(c_float)]),
    "sptrsv_set_backoff": (c_int, [POINTER(SpMat), c_int]),
    "sptrsv_set_lds_pad": (c_int, [POINTER(SpMat), c_int]),
    "sptrsv_set_tolerance": (c_int, [POINTER(SpMat), c_double]),
    "sptrsv_dominance_margin": (c_int, [POINTER(SpMat), POINTER(c_double)]),
    "sptrsv_trace": (c_int, [POINTER(SpMat), c_int, c_int]),
    "sptrsv_trace_read": (c_int, [POINTER(SpMat), POINTER(c_int64), c_size_t, POINTER(c_double)]),
}

EXPORTED = tuple(_PROTOS)

_lib = None


def lib() -> ctypes.CDLL:
    """Loads the library once; raises SptrsvUnavailable when it is absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise SptrsvUnavailable(
                f"{LIB_PATH} is not built; run `python -m sptrsv_amd.build` (or "
                "__graft_entry__.build()) -- there is no CPU fallback")
        try:
            handle = ctypes.CDLL(str(LIB_PATH))
        except OSError as e:
            raise SptrsvUnavailable(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in _PROTOS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().sptrsv_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str = "") -> int:
    if status != OK:
        detail = last_error()
        name = lib().sptrsv_status_string(status).decode()
        raise SptrsvError(status, f"{what}: {name}" + (f" -- {detail}" if detail else ""))
    return status
