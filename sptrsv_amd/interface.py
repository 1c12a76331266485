"""Python mirror of the reference host interface (header/interface.h:6-9,
header/analysis.h:20, header/solver.h:5-8) over the C-ABI library.

Same names and argument meaning as the reference; errors raise
``SptrsvError`` (the reference prints and continues, common.h:163-167).
Host arrays are numpy; device buffers are ``DeviceArray`` objects owned by
the library's allocator (no torch involved).
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_float, c_int, c_int64, c_void_p
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import (ALL, E_NOT_CONVERGED, KERNEL_AUTO, KERNEL_JACOBI, KERNEL_MULTIROW, KERNEL_ORDER,
                   KERNEL_SIMPLE, MULTIROW, ORDER, PREC_F32, PREC_F64, PREC_MIXED, SIMPLE, AnalysisInfo, BenchResult, Config, SpMat,
                   SptrsvError, check)

__all__ = [
    "SIMPLE", "ORDER", "MULTIROW", "ALL", "KERNEL_AUTO", "KERNEL_SIMPLE", "KERNEL_ORDER",
    "KERNEL_MULTIROW", "KERNEL_JACOBI", "E_NOT_CONVERGED", "PREC_F64", "PREC_F32", "PREC_MIXED", "SpMat", "AnalysisInfo", "Config",
    "SptrsvError", "DeviceArray", "allocate_memory", "allocate_memory_f32", "run_analysis",
    "run_solver", "free_device_analysis", "multirow_analysis_base_GPU", "csr_L_solve_simple",
    "csr_L_solve_order", "csr_L_solve_multirow", "csr_L_solve_simple_f32", "csr_L_solve_order_f32",
    "csr_L_solve_multirow_f32", "csr_L_solve_mixed", "solve", "analysis_arrays", "autotune",
    "default_config", "set_watchdog_ms", "check_watchdog", "reserve", "generate",
    "generate_batch_member", "read_mtx_lower", "Batch", "RocsparseBaseline", "power_start",
    "power_end", "bench_solve", "device_name", "synchronize", "flush_cache", "Timer",
    "kernel_timer", "kernel_timer_read", "dep_latency_ns", "set_tolerance", "dominance_margin", "solve_graph",
    "copy_device", "allocate_memory_upper", "transpose", "matrix_arrays", "read_mtx_coo", "coo_to_lower",
    "save_csr", "load_csr", "update_values", "Spsv", "FILL_LOWER", "FILL_UPPER", "DIAG_NON_UNIT", "DIAG_UNIT", "OP_NONE",
    "OP_TRANSPOSE",
]


def _ptr(a: np.ndarray) -> c_void_p:
    return c_void_p(a.ctypes.data)


class DeviceArray:
    """A device buffer allocated through the library."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = c_void_p()
        check(L.lib().sptrsv_malloc(byref(p), max(self.nbytes, 1)), "sptrsv_malloc")
        self.ptr = p

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.nbytes:
            check(L.lib().sptrsv_memcpy_htod(d.ptr, _ptr(a), a.nbytes), "memcpy htod")
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        if self.nbytes:
            check(L.lib().sptrsv_memcpy_dtoh(_ptr(out), self.ptr, self.nbytes), "memcpy dtoh")
        return out

    def fill_bytes(self, value: int) -> None:
        check(L.lib().sptrsv_memset(self.ptr, value, self.nbytes), "memset")

    def free(self) -> None:
        if self.ptr is not None and self.ptr.value:
            L.lib().sptrsv_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _dev_int_array(ptr: int | None, count: int) -> np.ndarray:
    out = np.empty(max(count, 0), np.int32)
    if ptr and count > 0:
        check(L.lib().sptrsv_memcpy_dtoh(_ptr(out), c_void_p(ptr), out.nbytes), "memcpy dtoh")
    return out


# ---- interface.h --------------------------------------------------------------

def allocate_memory(ia, ja, a, n: int | None = None, mode: int = ALL) -> SpMat:
    """interface.h:6: upload a host CSR lower triangle (diagonal last)."""
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja, np.int32)
    a = np.ascontiguousarray(a, np.float64)
    n = len(ia) - 1 if n is None else n
    m = SpMat()
    check(L.lib().allocate_memory(_ptr(ia), _ptr(ja), _ptr(a), n, byref(m), mode), "allocate_memory")
    return m


def allocate_memory_f32(ia, ja, a, n: int | None = None, mode: int = ALL) -> SpMat:
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja, np.int32)
    a = np.ascontiguousarray(a, np.float32)
    n = len(ia) - 1 if n is None else n
    m = SpMat()
    check(L.lib().allocate_memory_f32(_ptr(ia), _ptr(ja), _ptr(a), n, byref(m), mode),
          "allocate_memory_f32")
    return m


def allocate_memory_upper(ia, ja, a, n: int | None = None, mode: int = ALL) -> SpMat:
    """Upload a host CSR upper triangle (diagonal first in every row).

    The matrix is kept as the lower triangle of the reversed index order, so
    every analysis and solve call on it solves U x = b with b and x in the
    caller's order.  No counterpart in the reference, which solves L only.
    """
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja, np.int32)
    a = np.ascontiguousarray(a, np.float64)
    n = len(ia) - 1 if n is None else n
    m = SpMat()
    check(L.lib().allocate_memory_upper(_ptr(ia), _ptr(ja), _ptr(a), n, byref(m), mode),
          "allocate_memory_upper")
    return m


def transpose(mat: SpMat) -> SpMat:
    """A new handle that solves A^T x = b for the A that `mat` solves, built on the device."""
    t = SpMat()
    check(L.lib().sptrsv_transpose(byref(mat), byref(t)), "sptrsv_transpose")
    return t


def update_values(mat: SpMat, vals) -> None:
    """New values (host numpy or DeviceArray) in the layout the handle was created from."""
    if isinstance(vals, DeviceArray):
        check(L.lib().sptrsv_update_values(byref(mat), vals.ptr, vals.count, 1), "sptrsv_update_values")
    else:
        v = np.ascontiguousarray(vals, np.float64)
        check(L.lib().sptrsv_update_values(byref(mat), _ptr(v), len(v), 0), "sptrsv_update_values")


def matrix_arrays(mat: SpMat) -> dict:
    """Copies a handle's stored CSR (the lower triangle the kernels read) to numpy."""
    n, nnz = mat.nr, mat.nnz
    out = {"ia": _dev_int_array(mat.ia, n + 1), "ja": _dev_int_array(mat.ja, nnz)}
    for name, dt in (("a", np.float64), ("a32", np.float32)):
        arr = np.empty(nnz, dt)
        ptr = getattr(mat, name)
        if ptr and nnz:
            check(L.lib().sptrsv_memcpy_dtoh(_ptr(arr), c_void_p(ptr), arr.nbytes), "memcpy dtoh")
        out[name] = arr if ptr else None
    return out


def run_analysis(info: AnalysisInfo, mat: SpMat) -> int:
    """interface.h:7: returns -1 for SIMPLE (no analysis), like the reference."""
    st = L.lib().run_analysis(byref(info), byref(mat))
    if st == L.E_NO_ANALYSIS_NEEDED:
        return st
    return check(st, "run_analysis")


def multirow_analysis_base_GPU(mat: SpMat, mode: int = ALL, info: AnalysisInfo | None = None) -> AnalysisInfo:
    """analysis.h:20."""
    info = info if info is not None else AnalysisInfo()
    p = ctypes.pointer(info)
    check(L.lib().multirow_analysis_base_GPU(byref(p), byref(mat), mode), "multirow_analysis_base_GPU")
    return info


def save_analysis(path: str, mat: SpMat, info: AnalysisInfo) -> None:
    """Writes the analysis of mat to a checksummed file (sptrsv_save_analysis)."""
    check(L.lib().sptrsv_save_analysis(str(path).encode(), byref(mat), byref(info)), "sptrsv_save_analysis")


def load_analysis(path: str, mat: SpMat, info: AnalysisInfo | None = None) -> AnalysisInfo:
    """Reads an analysis file for a matrix of the same structure (sptrsv_load_analysis)."""
    info = info if info is not None else AnalysisInfo()
    check(L.lib().sptrsv_load_analysis(str(path).encode(), byref(mat), byref(info)), "sptrsv_load_analysis")
    return info


def analysis_segments(info: AnalysisInfo) -> tuple[int, int]:
    """(rows split, segments) of a MULTIROW / ALL analysis (sptrsv_analysis_segments)."""
    rows, segs = c_int(), c_int()
    check(L.lib().sptrsv_analysis_segments(byref(info), byref(rows), byref(segs)), "sptrsv_analysis_segments")
    return rows.value, segs.value


def copy_analysis(src_mat: SpMat, src: AnalysisInfo, dst_mat: SpMat, dst: AnalysisInfo | None = None) -> AnalysisInfo:
    """dst_mat (same structure as src_mat) gets its own copy of src (sptrsv_copy_analysis)."""
    dst = dst if dst is not None else AnalysisInfo()
    check(L.lib().sptrsv_copy_analysis(byref(src_mat), byref(src), byref(dst_mat), byref(dst)),
          "sptrsv_copy_analysis")
    return dst


def run_solver(mat: SpMat, info: AnalysisInfo | None, b: np.ndarray) -> np.ndarray:
    """interface.h:8: host b in, host x out."""
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty_like(b)
    check(L.lib().run_solver(byref(mat), byref(info) if info is not None else None, _ptr(b), _ptr(x),
                             mat.nr), "run_solver")
    return x


def free_device_analysis(mat: SpMat | None, info: AnalysisInfo | None = None) -> None:
    """interface.h:9."""
    L.lib().free_device_analysis(byref(mat) if mat is not None else None,
                                 byref(info) if info is not None else None)


def analysis_arrays(info: AnalysisInfo) -> dict:
    """Copies the device arrays of an analysis to numpy."""
    n, nl = info.n, info.nlevs
    out = {
        "nlevs": nl, "n_warps": info.n_warps, "n_waves": info.n_waves, "lev_ctr": info.lev_ctr,
        "level": _dev_int_array(info.level, n),
        "iorder": _dev_int_array(info.iorder, n),
        "inv_iorder": _dev_int_array(info.inv_iorder, n),
        "lev_size": _dev_int_array(info.lev_size, nl),
        "bucket_off": _dev_int_array(info.bucket_off, 7 * nl + 1 if nl else 0),
    }
    if info.ibase_row:
        out["ibase_row"] = _dev_int_array(info.ibase_row, info.n_warps + 1)
        out["ivect_size_warp"] = _dev_int_array(info.ivect_size_warp, info.n_warps)
        out["warp_lev"] = _dev_int_array(info.warp_lev, nl + 1)
    if info.ibase_row64:
        out["ibase_row64"] = _dev_int_array(info.ibase_row64, info.n_waves + 1)
        out["ivect_size_wave"] = _dev_int_array(info.ivect_size_wave, info.n_waves)
    return out


# ---- solver.h -------------------------------------------------------------------

def csr_L_solve_simple(mat: SpMat, b: DeviceArray, x: DeviceArray, n: int | None = None, is_solved=None) -> None:
    check(L.lib().csr_L_solve_simple(byref(mat), b.ptr, x.ptr, mat.nr if n is None else n, None),
          "csr_L_solve_simple")


def csr_L_solve_order(mat: SpMat, info: AnalysisInfo, b: DeviceArray, x: DeviceArray, n=None, stream=None) -> None:
    check(L.lib().csr_L_solve_order(byref(mat), byref(info), b.ptr, x.ptr, mat.nr if n is None else n, stream),
          "csr_L_solve_order")


def csr_L_solve_multirow(mat: SpMat, info: AnalysisInfo, b: DeviceArray, x: DeviceArray, n=None, stream=None) -> None:
    check(L.lib().csr_L_solve_multirow(byref(mat), byref(info), b.ptr, x.ptr, mat.nr if n is None else n,
                                       stream), "csr_L_solve_multirow")


def csr_L_solve_simple_f32(mat: SpMat, b: DeviceArray, x: DeviceArray, n=None, is_solved=None) -> None:
    check(L.lib().csr_L_solve_simple_f32(byref(mat), b.ptr, x.ptr, mat.nr if n is None else n, None),
          "csr_L_solve_simple_f32")


def csr_L_solve_order_f32(mat, info, b, x, n=None, stream=None) -> None:
    check(L.lib().csr_L_solve_order_f32(byref(mat), byref(info), b.ptr, x.ptr, mat.nr if n is None else n,
                                        stream), "csr_L_solve_order_f32")


def csr_L_solve_multirow_f32(mat, info, b, x, n=None, stream=None) -> None:
    check(L.lib().csr_L_solve_multirow_f32(byref(mat), byref(info), b.ptr, x.ptr, mat.nr if n is None else n,
                                           stream), "csr_L_solve_multirow_f32")


def csr_L_solve_mixed(mat, info, b: DeviceArray, x: DeviceArray, k: int = 1, refine_steps: int = 1,
                      stream=None) -> None:
    check(L.lib().csr_L_solve_mixed(byref(mat), byref(info) if info is not None else None, b.ptr, x.ptr,
                                    mat.nr, k, refine_steps, stream), "csr_L_solve_mixed")


def _cfg(kernel=KERNEL_AUTO, lanes=0, wg=0, fused=1, variant=0, cu_waves=0, sweeps=0, outer=0,
         window=0, sweeps_next=0, inner=0, claim_rounds=0, split_rows=0, reorder=0) -> Config:
    return Config(kernel, lanes, wg, fused, variant, cu_waves, sweeps, outer, window, sweeps_next, inner,
                  claim_rounds, split_rows, reorder)


def solve(mat, info, precision: int, b: DeviceArray, x: DeviceArray, k: int = 1, refine_steps: int = 1,
          kernel: int = KERNEL_AUTO, lanes: int = 0, wg: int = 0, fused: int = 1, variant: int = 0,
          cu_waves: int = 0, sweeps: int = 0, outer: int = 0, window: int = 0, sweeps_next: int = 0,
          inner: int = 0, stream=None, cfg: Config | None = None) -> None:
    """cfg, when given, is used as is (e.g. the result of autotune)."""
    cfg = cfg if cfg is not None else _cfg(kernel, lanes, wg, fused, variant, cu_waves, sweeps, outer, window,
                                           sweeps_next, inner)
    check(L.lib().sptrsv_solve(byref(mat), byref(info) if info is not None else None, precision, byref(cfg),
                               b.ptr, x.ptr, mat.nr, k, refine_steps, stream), "sptrsv_solve")


def solve_graph(mat, info, precision: int, b: DeviceArray, x: DeviceArray, cfg: Config, k: int = 1,
                refine_steps: int = 1, stream=None) -> None:
    """The same solve captured once into a hipGraph (per b, x and configuration) and replayed."""
    check(L.lib().sptrsv_solve_graph(byref(mat), byref(info) if info is not None else None, precision,
                                     byref(cfg), b.ptr, x.ptr, mat.nr, k, refine_steps, stream),
          "sptrsv_solve_graph")


def solve_mixed_noalloc(mat, info, b: DeviceArray, x: DeviceArray, k: int = 1, cfg: Config | None = None,
                        stream=None) -> None:
    """The fused mixed solve with its published words in the library's fixed
    per-device scratch (sptrsv_solve_mixed_noalloc): allocates nothing, n*k <=
    NOALLOC_MAX_WORDS, sync-free kernels only, overlapping calls on one GPU race."""
    check(L.lib().sptrsv_solve_mixed_noalloc(byref(mat), byref(info) if info is not None else None,
                                             byref(cfg) if cfg is not None else None, b.ptr, x.ptr, mat.nr, k,
                                             stream), "sptrsv_solve_mixed_noalloc")


def reserve(mat: SpMat, k: int) -> None:
    check(L.lib().sptrsv_reserve(byref(mat), k), "sptrsv_reserve")


def set_watchdog_ms(mat: SpMat, ms: float) -> None:
    check(L.lib().sptrsv_set_watchdog_ms(byref(mat), ms), "sptrsv_set_watchdog_ms")


def check_watchdog(mat: SpMat) -> int:
    """Returns the status of the last solves (0, E_WATCHDOG or E_NOT_CONVERGED) without raising."""
    return L.lib().sptrsv_check(byref(mat), None)


def autotune(mat, info, precision: int, k: int = 1, reps: int = 3) -> tuple[Config, float]:
    best = Config()
    ms = c_float()
    check(L.lib().sptrsv_autotune(byref(mat), byref(info) if info is not None else None, precision, k, reps,
                                  byref(best), byref(ms)), "sptrsv_autotune")
    return best, ms.value


def default_config(mat, info, precision: int, k: int = 1) -> Config:
    c = Config()
    check(L.lib().sptrsv_default_config(byref(mat), byref(info) if info is not None else None, precision, k,
                                        byref(c)), "sptrsv_default_config")
    return c


def bench_solve(mat, info, precision: int, b: DeviceArray, x: DeviceArray, k: int = 1, refine_steps: int = 1,
                reps: int = 5, check_ones: bool = False, kernel=KERNEL_AUTO, lanes=0, wg=0, fused=1,
                variant=0) -> BenchResult:
    cfg = _cfg(kernel, lanes, wg, fused, variant)
    out = BenchResult()
    check(L.lib().sptrsv_bench_solve(byref(mat), byref(info) if info is not None else None, precision,
                                     byref(cfg), b.ptr, x.ptr, mat.nr, k, refine_steps, reps,
                                     1 if check_ones else 0, byref(out)), "sptrsv_bench_solve")
    return out


# ---- inputs -----------------------------------------------------------------------

def _take_int(p, count) -> np.ndarray:
    a = np.ctypeslib.as_array(p, shape=(count,)).copy() if count > 0 else np.empty(0, np.int32)
    L.lib().sptrsv_free_host(ctypes.cast(p, c_void_p))
    return a


def _take_dbl(p, count) -> np.ndarray:
    a = np.ctypeslib.as_array(p, shape=(count,)).copy() if count > 0 else np.empty(0, np.float64)
    L.lib().sptrsv_free_host(ctypes.cast(p, c_void_p))
    return a


@dataclass
class System:
    n: int
    ia: np.ndarray
    ja: np.ndarray
    a: np.ndarray
    b: np.ndarray
    k: int = 1

    @property
    def nnz(self) -> int:
        return int(self.ia[-1])


def generate(cfg_id: int, param: int = 0, seed: int | None = None) -> System:
    n, nnz, k = c_int(), c_int(), c_int()
    ia, ja = ctypes.POINTER(c_int)(), ctypes.POINTER(c_int)()
    a, b = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_double)()
    seed = cfg_id if seed is None else seed
    check(L.lib().sptrsv_generate(cfg_id, param, seed, byref(n), byref(nnz), byref(ia), byref(ja), byref(a),
                                  byref(k), byref(b)), "sptrsv_generate")
    return System(n.value, _take_int(ia, n.value + 1), _take_int(ja, nnz.value), _take_dbl(a, nnz.value),
                  _take_dbl(b, n.value * k.value), k.value)


def generate_batch_member(m: int, nmin: int = 2000, nmax: int = 200000, seed: int = 5) -> System:
    n, nnz = c_int(), c_int()
    ia, ja = ctypes.POINTER(c_int)(), ctypes.POINTER(c_int)()
    a, b = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_double)()
    check(L.lib().sptrsv_generate_batch_member(m, nmin, nmax, seed, byref(n), byref(nnz), byref(ia), byref(ja),
                                               byref(a), byref(b)), "sptrsv_generate_batch_member")
    return System(n.value, _take_int(ia, n.value + 1), _take_int(ja, nnz.value), _take_dbl(a, nnz.value),
                  _take_dbl(b, n.value))


def read_mtx_lower(path: str, unit: bool = True) -> tuple[int, np.ndarray, np.ndarray, np.ndarray]:
    """main.cu:66-246: .mtx -> CSR -> lower triangle with explicit diagonal."""
    n, nnz = c_int(), c_int()
    ia, ja = ctypes.POINTER(c_int)(), ctypes.POINTER(c_int)()
    a = ctypes.POINTER(ctypes.c_double)()
    check(L.lib().sptrsv_read_mtx_lower(str(path).encode(), 1 if unit else 0, byref(n), byref(nnz), byref(ia),
                                        byref(ja), byref(a)), "sptrsv_read_mtx_lower")
    return n.value, _take_int(ia, n.value + 1), _take_int(ja, nnz.value), _take_dbl(a, nnz.value)


def read_mtx_coo(path: str, nthreads: int = 0) -> tuple[int, np.ndarray, np.ndarray, np.ndarray, bool]:
    """main.cu:131-155 read over line-aligned chunks: (n, row, col, val, symmetric), 0-based, file order."""
    n, nz, sym = c_int(), c_int(), c_int()
    r, c = ctypes.POINTER(c_int)(), ctypes.POINTER(c_int)()
    v = ctypes.POINTER(ctypes.c_double)()
    check(L.lib().sptrsv_read_mtx_coo(str(path).encode(), nthreads, byref(n), byref(nz), byref(sym), byref(r),
                                      byref(c), byref(v)), "sptrsv_read_mtx_coo")
    return n.value, _take_int(r, nz.value), _take_int(c, nz.value), _take_dbl(v, nz.value), bool(sym.value)


def coo_to_lower(row, col, val, n: int, symmetric: bool, unit: bool = True) -> SpMat:
    """main.cu:124-246 on the device: COO -> A (mirrored if symmetric) -> lower triangle, as a handle."""
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    m = SpMat()
    check(L.lib().sptrsv_coo_to_lower(_ptr(row), _ptr(col), _ptr(val), len(row), n, int(symmetric), int(unit),
                                      byref(m)), "sptrsv_coo_to_lower")
    return m


def save_csr(path: str, ia, ja, a) -> None:
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja, np.int32)
    a = np.ascontiguousarray(a, np.float64)
    check(L.lib().sptrsv_save_csr(str(path).encode(), len(ia) - 1, _ptr(ia), _ptr(ja), _ptr(a)), "sptrsv_save_csr")


def load_csr(path: str) -> tuple[int, np.ndarray, np.ndarray, np.ndarray]:
    n, nnz = c_int(), c_int()
    ia, ja = ctypes.POINTER(c_int)(), ctypes.POINTER(c_int)()
    a = ctypes.POINTER(ctypes.c_double)()
    check(L.lib().sptrsv_load_csr(str(path).encode(), byref(n), byref(nnz), byref(ia), byref(ja), byref(a)),
          "sptrsv_load_csr")
    return n.value, _take_int(ia, n.value + 1), _take_int(ja, nnz.value), _take_dbl(a, nnz.value)


FILL_LOWER, FILL_UPPER = 0, 1
DIAG_NON_UNIT, DIAG_UNIT = 0, 1
OP_NONE, OP_TRANSPOSE = 0, 1


class Spsv:
    """The spsv-style handle (sptrsv_spsv_*) over a device CSR: op(A) x = alpha b."""

    def __init__(self, n: int, ia: DeviceArray, ja: DeviceArray, a: DeviceArray, fill: int = FILL_LOWER,
                 diag: int = DIAG_NON_UNIT, op: int = OP_NONE):
        self.n = n
        h = c_void_p()
        check(L.lib().sptrsv_spsv_create(byref(h), n, ja.count, ia.ptr, ja.ptr, a.ptr, fill, diag, op),
              "sptrsv_spsv_create")
        self.h = h

    def analysis(self, precision: int = PREC_F64, k: int = 1, tune: bool = False) -> None:
        check(L.lib().sptrsv_spsv_analysis(self.h, precision, k, int(tune)), "sptrsv_spsv_analysis")

    def solve(self, b: DeviceArray, x: DeviceArray, precision: int = PREC_F64, alpha: float = 1.0, k: int = 1,
              stream=None) -> None:
        check(L.lib().sptrsv_spsv_solve(self.h, precision, alpha, b.ptr, x.ptr, k, stream), "sptrsv_spsv_solve")

    def matrix(self) -> SpMat:
        return L.lib().sptrsv_spsv_matrix(self.h).contents

    def close(self) -> None:
        if self.h is not None and self.h.value:
            L.lib().sptrsv_spsv_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- batch, vendor baseline, power, timing ------------------------------------------

class Batch:
    """Many independent systems solved in one launch (sptrsv_batch_*)."""

    def __init__(self, systems: list):
        self._keep = [(np.ascontiguousarray(s.ia, np.int32), np.ascontiguousarray(s.ja, np.int32),
                       np.ascontiguousarray(s.a, np.float64)) for s in systems]
        nm = len(systems)
        ns = (c_int * nm)(*[s.n for s in systems])
        ia = (c_void_p * nm)(*[k[0].ctypes.data for k in self._keep])
        ja = (c_void_p * nm)(*[k[1].ctypes.data for k in self._keep])
        a = (c_void_p * nm)(*[k[2].ctypes.data for k in self._keep])
        h = c_void_p()
        check(L.lib().sptrsv_batch_create(nm, ns, ia, ja, a, byref(h)), "sptrsv_batch_create")
        self._keep = None
        self.h = h
        self.nmat = nm
        rows, nnz = c_int64(), c_int64()
        check(L.lib().sptrsv_batch_rows(h, byref(rows), byref(nnz)), "sptrsv_batch_rows")
        self.rows, self.nnz = rows.value, nnz.value
        off = (c_int64 * (nm + 1))()
        check(L.lib().sptrsv_batch_row_offsets(h, off), "row offsets")
        self.row_off = np.array(off[:], np.int64)

    def tune(self, precision: int = PREC_MIXED, tune: bool = True) -> None:
        check(L.lib().sptrsv_batch_tune(self.h, precision, 1 if tune else 0), "sptrsv_batch_tune")

    def choice(self, m: int) -> Config:
        c = Config()
        check(L.lib().sptrsv_batch_choice(self.h, m, byref(c)), "sptrsv_batch_choice")
        return c

    def solve(self, precision: int, b: DeviceArray, x: DeviceArray, graph: bool = True, stream=None) -> None:
        fn = L.lib().sptrsv_batch_solve_graph if graph else L.lib().sptrsv_batch_solve
        check(fn(self.h, precision, b.ptr, x.ptr, stream), "sptrsv_batch_solve")

    def check(self, stream=None) -> int:
        """Status of the last solves (0, E_WATCHDOG or E_NOT_CONVERGED) without raising."""
        return L.lib().sptrsv_batch_check(self.h, stream)

    def patterns(self) -> tuple[int, np.ndarray]:
        """(distinct structures, first member with each member's structure)."""
        cnt = c_int()
        of = np.empty(self.nmat, np.int32)
        check(L.lib().sptrsv_batch_patterns(self.h, byref(cnt), of.ctypes.data), "sptrsv_batch_patterns")
        return cnt.value, of

    def schedule(self) -> tuple[int, int]:
        """(serial, separate) of a MIXED batched solve: halves one after the
        other; sync-free members solved as one analysed matrix of their own."""
        a, b = c_int(), c_int()
        check(L.lib().sptrsv_batch_schedule(self.h, byref(a), byref(b)), "sptrsv_batch_schedule")
        return a.value, b.value

    def set_schedule(self, serial: int, separate: int) -> None:
        check(L.lib().sptrsv_batch_set_schedule(self.h, serial, separate), "sptrsv_batch_set_schedule")

    def close(self) -> None:
        if self.h is not None:
            L.lib().sptrsv_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RocsparseBaseline:
    """rocSPARSE csrsv (or, with itsv=True, csritsv: FP64 Jacobi with a residual
    stopping test) on the same device matrix (solver.h:8 counterpart)."""

    def __init__(self, mat: SpMat, itsv: bool = False):
        h = c_void_p()
        ms = c_float()
        fn = L.lib().csr_L_rocsparse_itsv_create if itsv else L.lib().csr_L_rocsparse_create
        check(fn(byref(mat), byref(h), byref(ms)), "rocSPARSE create")
        self.h, self.analysis_ms, self.itsv = h, ms.value, itsv

    def solve(self, b: DeviceArray, x: DeviceArray, stream=None) -> None:
        check(L.lib().csr_L_solve_rocsparse(self.h, b.ptr, x.ptr, stream), "csr_L_solve_rocsparse")

    def solve_itsv(self, b: DeviceArray, x: DeviceArray, tol: float, max_iter: int = 1000, free_iter: int = 20,
                   stream=None) -> int:
        """Jacobi iterations performed: (k + 1) * free_iter + k, where rocSPARSE's k is the
        index of the stopping test that met |b - L x|_inf <= tol (max_iter tests at most)."""
        it = c_int()
        check(L.lib().csr_L_solve_rocsparse_itsv(self.h, b.ptr, x.ptr, tol, max_iter, free_iter, byref(it), stream),
              "csr_L_solve_rocsparse_itsv")
        return it.value

    def close(self) -> None:
        if self.h is not None:
            L.lib().csr_L_rocsparse_destroy(self.h)
            self.h = None


def power_start(label: str = "", device: int = 0, period_us: int = 1000) -> bool:
    return L.lib().amdsmiAPIRun(label.encode(), device, period_us) == L.OK


def power_end() -> float:
    """Mean power in mW (< 0 when unavailable)."""
    return float(L.lib().amdsmiAPIEnd())


def device_name(device: int = 0) -> str:
    buf = ctypes.create_string_buffer(256)
    check(L.lib().sptrsv_device_name(device, buf, 256), "device_name")
    return buf.value.decode()


def copy_device(dst: DeviceArray, src: DeviceArray, nbytes: int) -> None:
    check(L.lib().sptrsv_copy_device(dst.ptr, src.ptr, nbytes), "sptrsv_copy_device")


def synchronize() -> None:
    check(L.lib().sptrsv_synchronize(), "synchronize")


def set_backoff(mat: SpMat, level: int) -> None:
    check(L.lib().sptrsv_set_backoff(byref(mat), level), "sptrsv_set_backoff")


def set_tolerance(mat: SpMat, rel_tol: float) -> None:
    check(L.lib().sptrsv_set_tolerance(byref(mat), rel_tol), "sptrsv_set_tolerance")


def dominance_margin(mat: SpMat) -> float:
    d = ctypes.c_double()
    check(L.lib().sptrsv_dominance_margin(byref(mat), byref(d)), "sptrsv_dominance_margin")
    return d.value


def set_lds_pad(mat: SpMat, nbytes: int) -> None:
    check(L.lib().sptrsv_set_lds_pad(byref(mat), nbytes), "sptrsv_set_lds_pad")


def trace(mat: SpMat, enable: bool, k: int = 1) -> None:
    check(L.lib().sptrsv_trace(byref(mat), k, 1 if enable else 0), "sptrsv_trace")


def trace_read(mat: SpMat, k: int = 1) -> tuple[np.ndarray, float]:
    """(n*k, 3) wall-clock stamps (start, reduced, stored) and the clock rate in kHz."""
    cnt = 3 * mat.nr * k
    out = np.zeros(cnt, np.int64)
    rate = ctypes.c_double()
    check(L.lib().sptrsv_trace_read(byref(mat), out.ctypes.data_as(ctypes.POINTER(c_int64)), cnt,
                                    byref(rate)), "sptrsv_trace_read")
    return out.reshape(-1, 3), rate.value


def kernel_timer(enable, kind: str = "all") -> None:
    """HIP events around solve kernels: kind "all", "sync-free" or "sweeps"."""
    mode = {"all": 1, "sync-free": 2, "sweeps": 3}[kind] if enable else 0
    check(L.lib().sptrsv_kernel_timer(mode), "kernel_timer")


def kernel_timer_read() -> tuple[float, int]:
    ms, n = c_float(), c_int()
    check(L.lib().sptrsv_kernel_timer_read(byref(ms), byref(n)), "kernel_timer_read")
    return ms.value, n.value


def dep_latency_ns(rounds: int = 2000) -> float:
    ns = c_float()
    check(L.lib().sptrsv_measure_dep_latency(rounds, byref(ns)), "measure_dep_latency")
    return ns.value


def flush_cache(stream=None) -> None:
    check(L.lib().sptrsv_flush_cache(stream), "flush_cache")


class Timer:
    """hip-event timer on a stream (sptrsv_timer_start/stop)."""

    def __init__(self, stream=None):
        self.stream = stream

    def start(self) -> None:
        check(L.lib().sptrsv_timer_start(self.stream), "timer_start")

    def stop(self) -> float:
        ms = c_float()
        check(L.lib().sptrsv_timer_stop(self.stream, byref(ms)), "timer_stop")
        return ms.value
