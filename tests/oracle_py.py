"""ctypes wrapper of oracle/liboracle.so -- the CPU checker (test use only)."""
from __future__ import annotations

import contextlib
import ctypes
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_void_p
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
SO = ROOT / "oracle" / "liboracle.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not SO.exists():
            subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
                            str(ROOT / "oracle" / "oracle.c"), "-o", str(SO)], check=True)
        l = ctypes.CDLL(str(SO))
        P = c_void_p
        protos = {
            "oracle_set_mutation": (None, [c_int]),
            "oracle_get_mutation": (c_int, []),
            "oracle_check_lower_csr": (c_int, [c_int, P, P]),
            "oracle_fwd_f64": (None, [c_int, P, P, P, P, P, c_int]),
            "oracle_fwd_f32": (None, [c_int, P, P, P, P, P, c_int]),
            "oracle_levels": (c_int, [c_int, P, P, P]),
            "oracle_row_class": (c_int, [c_int]),
            "oracle_count_warps": (c_int, [c_int, P, P, c_int]),
            "oracle_analysis": (c_int, [c_int, P, P, c_int, P, P, P, P, P, P, P]),
            "oracle_coo_csr_nnz": (c_int, [c_int, P, P, c_int]),
            "oracle_coo_to_csr": (None, [c_int, c_int, P, P, P, c_int, P, P, P]),
            "oracle_extract_lower": (c_int, [c_int, P, P, P, c_int, P, P, P]),
            "oracle_validate_x": (c_int, [P, c_int]),
        }
        for k, (r, a) in protos.items():
            f = getattr(l, k)
            f.restype, f.argtypes = r, a
        _lib = l
    return _lib


def _p(a: np.ndarray):
    return c_void_p(a.ctypes.data)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


@contextlib.contextmanager
def mutation(m: int):
    lib().oracle_set_mutation(m)
    try:
        yield
    finally:
        lib().oracle_set_mutation(0)


def check_lower_csr(ia, ja) -> int:
    ia, ja = _i32(ia), _i32(ja)
    return lib().oracle_check_lower_csr(len(ia) - 1, _p(ia), _p(ja))


def fwd_f64(ia, ja, a, b, k: int = 1) -> np.ndarray:
    ia, ja = _i32(ia), _i32(ja)
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b)
    lib().oracle_fwd_f64(len(ia) - 1, _p(ia), _p(ja), _p(a), _p(b), _p(x), k)
    return x


def fwd_f32(ia, ja, a, b, k: int = 1) -> np.ndarray:
    ia, ja = _i32(ia), _i32(ja)
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    x = np.zeros_like(b)
    lib().oracle_fwd_f32(len(ia) - 1, _p(ia), _p(ja), _p(a), _p(b), _p(x), k)
    return x


def levels(ia, ja) -> tuple[np.ndarray, int]:
    ia, ja = _i32(ia), _i32(ja)
    n = len(ia) - 1
    lev = np.zeros(max(n, 1), np.int32)
    nl = lib().oracle_levels(n, _p(ia), _p(ja), _p(lev))
    return lev[:n], nl


def row_class(nz: int) -> int:
    return lib().oracle_row_class(nz)


def analysis(ia, ja, lanes: int = 32) -> dict:
    ia, ja = _i32(ia), _i32(ja)
    n = len(ia) - 1
    _, nl = levels(ia, ja)
    nw = lib().oracle_count_warps(n, _p(ia), _p(ja), lanes)
    lev = np.zeros(max(n, 1), np.int32)
    key = np.zeros(max(n, 1), np.int32)
    iorder = np.zeros(max(n, 1), np.int32)
    bo = np.zeros(7 * nl + 1, np.int32)
    wo = np.zeros(7 * nl + 1, np.int32)
    base = np.zeros(nw + 1, np.int32)
    vs = np.zeros(max(nw, 1), np.int32)
    got = lib().oracle_analysis(n, _p(ia), _p(ja), lanes, _p(lev), _p(key), _p(iorder), _p(bo), _p(wo),
                                _p(base), _p(vs))
    assert got == nw
    return dict(level=lev[:n], key=key[:n], iorder=iorder[:n], nlevs=nl, bucket_off=bo, warp_off=wo,
                n_warps=nw, ibase_row=base, ivect_size_warp=vs[:nw])


def coo_to_csr(m: int, ri, ci, v, symmetric: bool):
    ri, ci = _i32(ri), _i32(ci)
    v = np.ascontiguousarray(v, np.float64)
    nz = lib().oracle_coo_csr_nnz(len(ri), _p(ri), _p(ci), int(symmetric))
    rp = np.zeros(m + 1, np.int32)
    col = np.zeros(max(nz, 1), np.int32)
    val = np.zeros(max(nz, 1), np.float64)
    lib().oracle_coo_to_csr(m, len(ri), _p(ri), _p(ci), _p(v), int(symmetric), _p(rp), _p(col), _p(val))
    return rp, col[:nz], val[:nz]


def extract_lower(m: int, rp, col, val, unit: bool):
    rp, col = _i32(rp), _i32(col)
    val = np.ascontiguousarray(val, np.float64)
    cap = len(col) + m
    lp = np.zeros(m + 1, np.int32)
    lc = np.zeros(max(cap, 1), np.int32)
    lv = np.zeros(max(cap, 1), np.float64)
    nz = lib().oracle_extract_lower(m, _p(rp), _p(col), _p(val), int(unit), _p(lp), _p(lc), _p(lv))
    return lp, lc[:nz], lv[:nz]


def validate_x(x) -> int:
    x = np.ascontiguousarray(x, np.float64)
    return lib().oracle_validate_x(_p(x), len(x))


def read_mtx_coo(path):
    """Plain Python reader of a coordinate .mtx file, entries in file order
    (1-based converted to 0-based)."""
    lines = Path(path).read_text().splitlines()
    banner = lines[0].lower().split()
    body = [l for l in lines[1:] if l.strip() and not l.startswith("%")]
    m, n, nz = (int(t) for t in body[0].split())
    ri, ci, v = [], [], []
    for l in body[1:1 + nz]:
        t = l.split()
        ri.append(int(t[0]) - 1)
        ci.append(int(t[1]) - 1)
        v.append(float(t[2]) if len(t) > 2 and banner[3] != "pattern" else 1.0)
    sym = banner[4] in ("symmetric", "hermitian")
    return m, np.array(ri, np.int32), np.array(ci, np.int32), np.array(v, np.float64), sym
