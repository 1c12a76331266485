"""Upper-triangular solves U x = b through allocate_memory_upper.

U is stored as the lower triangle of the reversed index order, so each check
compares the device result with two independent CPU answers: scipy's
spsolve_triangular(lower=False) on U itself, and the oracle's forward
substitution on the explicitly reversed matrix (P U P, P the reversal) with
b reversed and the result reversed back.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve_triangular

import oracle_py as O
from conftest import lower_from_rows
from sptrsv_amd import _lib as L
from sptrsv_amd import interface as I

pytestmark = pytest.mark.gpu

KERNELS = (I.KERNEL_SIMPLE, I.KERNEL_ORDER, I.KERNEL_MULTIROW)


def relerr(x, ref):
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


def upper_of(s: I.System) -> sp.csr_matrix:
    """U = L^T of a generated lower system (sorted columns: diagonal first)."""
    Lm = sp.csr_matrix((s.a, s.ja, s.ia), shape=(s.n, s.n))
    U = sp.csr_matrix(Lm.T)
    U.sort_indices()
    return U


def reversed_of(s: I.System) -> sp.csr_matrix:
    """U = P L P (P the reversal): keeps the row dominance Jacobi needs."""
    Lm = sp.csr_matrix((s.a, s.ja, s.ia), shape=(s.n, s.n))
    U = sp.csr_matrix(Lm[::-1][:, ::-1])
    U.sort_indices()
    return U


def references(U: sp.csr_matrix, B: np.ndarray, k: int) -> np.ndarray:
    n = U.shape[0]
    Bm = B.reshape(n, k)
    ref = np.column_stack([spsolve_triangular(U, Bm[:, r], lower=False) for r in range(k)])
    Lp = sp.csr_matrix(U[::-1][:, ::-1])
    Lp.sort_indices()
    rev = O.fwd_f64(Lp.indptr.astype(np.int32), Lp.indices.astype(np.int32), Lp.data, Bm[::-1].reshape(-1),
                    k=k).reshape(n, k)[::-1]
    assert relerr(rev, ref) <= 1e-11  # the two CPU answers agree
    return ref.reshape(-1)


class Up:
    def __init__(self, U: sp.csr_matrix):
        self.n = U.shape[0]
        self.mat = I.allocate_memory_upper(U.indptr, U.indices, U.data)
        self.info = I.multirow_analysis_base_GPU(self.mat, I.ALL)

    def solve(self, B, precision=I.PREC_F64, k=1, cfg=None, graph=False, **kw):
        dt = np.float32 if precision == I.PREC_F32 else np.float64
        db = I.DeviceArray.from_numpy(np.asarray(B, dt))
        dx = I.DeviceArray(self.n * k, dt)
        if graph:
            I.solve_graph(self.mat, self.info, precision, db, dx, cfg, k=k, refine_steps=kw.get("refine", 1))
        else:
            I.solve(self.mat, self.info, precision, db, dx, k=k, refine_steps=kw.pop("refine", 1), cfg=cfg, **kw)
        assert I.check_watchdog(self.mat) == 0, L.last_error()
        x = dx.numpy()
        db.free()
        dx.free()
        return x

    def close(self):
        I.free_device_analysis(self.mat, self.info)


@pytest.fixture(scope="module")
def device():
    print("device:", I.device_name(0), flush=True)
    return True


@pytest.mark.parametrize("cfg,param", [(2, 30000), (3, 30000), (4, 20)])
def test_upper_all_kernels(device, cfg, param):
    s = I.generate(cfg, param)
    U = upper_of(s)
    k = len(s.b) // s.n
    ref = references(U, s.b, k)
    d = Up(U)
    for kern in KERNELS:
        print("kernel", kern, flush=True)
        assert relerr(d.solve(s.b, k=k, kernel=kern), ref) <= 1e-12, kern
        assert relerr(d.solve(s.b, precision=I.PREC_MIXED, k=k, kernel=kern), ref) <= 1e-10, kern
        assert relerr(d.solve(s.b, precision=I.PREC_F32, k=k, kernel=kern), ref) <= 1e-4, kern
    d.close()


def test_upper_jacobi_multi_rhs_graph_and_autotune(device):
    s = I.generate(2, 50000)
    U = reversed_of(s)
    d = Up(U)
    assert I.dominance_margin(d.mat) > 0
    for k in (1, 3):
        B = np.random.default_rng(k).uniform(-1, 1, s.n * k)
        ref = references(U, B, k)
        print("jacobi k", k, flush=True)
        x = d.solve(B, precision=I.PREC_MIXED, k=k, kernel=I.KERNEL_JACOBI, sweeps=12, outer=4)
        assert relerr(x, ref) <= 1e-10
        assert relerr(d.solve(B, k=k, kernel=I.KERNEL_MULTIROW), ref) <= 1e-12
    cfg, _ = I.autotune(d.mat, d.info, I.PREC_MIXED)
    B = np.random.default_rng(7).uniform(-1, 1, s.n)
    ref = references(U, B, 1)
    assert relerr(d.solve(B, precision=I.PREC_MIXED, cfg=cfg), ref) <= 1e-10
    for rep in range(3):
        assert relerr(d.solve(B, precision=I.PREC_MIXED, cfg=cfg, graph=True), ref) <= 1e-10, rep
    d.close()


def test_upper_unsorted_columns_and_host_interface(device):
    # diagonal first, the other columns of a row in any order
    rng = np.random.default_rng(3)
    n = 300
    rows = [sorted(rng.choice(np.arange(i + 1, n), size=min(n - 1 - i, int(rng.integers(0, 6))),
                              replace=False).tolist()) for i in range(n)]
    ia, ja, a = [0], [], []
    for i, cols in enumerate(rows):
        vals = rng.uniform(-1, 1, len(cols))
        perm = rng.permutation(len(cols))
        ja += [i] + [cols[p] for p in perm]
        a += [1.0 + np.abs(vals).sum()] + [vals[p] for p in perm]
        ia.append(len(ja))
    ia, ja, a = np.array(ia, np.int32), np.array(ja, np.int32), np.array(a)
    U = sp.csr_matrix((a, ja, ia), shape=(n, n))
    b = rng.uniform(-1, 1, n)
    ref = references(U, b, 1)
    d = Up(U)
    assert relerr(I.run_solver(d.mat, d.info, b), ref) <= 1e-12
    assert relerr(I.run_solver(d.mat, None, b), ref) <= 1e-12
    d.close()
    # a lower matrix uploaded the ordinary way is unaffected by earlier upper solves
    li, lj, la = lower_from_rows([[], [0], [0, 1]])
    m = I.allocate_memory(li, lj, la)
    bl = np.array([1.0, 2.0, 3.0])
    assert relerr(I.run_solver(m, None, bl), O.fwd_f64(li, lj, la, bl)) <= 1e-12
    I.free_device_analysis(m, None)


# ---- transposed solves: a handle built on the device from an uploaded one ------------

def stored_reversed_transpose(ia, ja, a, n):
    """P A^T P in CSR with sorted columns: what a transpose handle must store for lower A."""
    Am = sp.csr_matrix((a, ja, ia), shape=(n, n))
    S = sp.csr_matrix(sp.csr_matrix(Am.T)[::-1][:, ::-1])
    S.sort_indices()
    return S


@pytest.mark.parametrize("cfg,param", [(1, 0), (2, 30000), (3, 30000), (4, 20)])
def test_transpose_structure_is_exact(device, cfg, param):
    s = I.generate(cfg, param)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    t = I.transpose(mat)
    got = I.matrix_arrays(t)
    S = stored_reversed_transpose(s.ia, s.ja, s.a, s.n)
    assert np.array_equal(got["ia"], S.indptr)
    assert np.array_equal(got["ja"], S.indices)
    assert np.array_equal(got["a"], S.data)
    assert np.array_equal(got["a32"], S.data.astype(np.float32))
    I.free_device_analysis(t, None)
    I.free_device_analysis(mat, None)


@pytest.mark.parametrize("cfg,param", [(2, 30000), (3, 30000)])
def test_transpose_solves(device, cfg, param):
    s = I.generate(cfg, param)
    Lm = sp.csr_matrix((s.a, s.ja, s.ia), shape=(s.n, s.n))
    LT = sp.csr_matrix(Lm.T)
    LT.sort_indices()
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    t = I.transpose(mat)
    info = I.multirow_analysis_base_GPU(t, I.ALL)
    ref = references(LT, s.b, 1)
    for kern in KERNELS:
        for prec, tol in ((I.PREC_F64, 1e-12), (I.PREC_MIXED, 1e-10)):
            db = I.DeviceArray.from_numpy(s.b)
            dx = I.DeviceArray(s.n, np.float64)
            I.solve(t, info, prec, db, dx, kernel=kern)
            assert I.check_watchdog(t) == 0, L.last_error()
            assert relerr(dx.numpy(), ref) <= tol, (kern, prec)
    # the transpose of the transpose solves L again
    tt = I.transpose(t)
    assert np.array_equal(I.matrix_arrays(tt)["ja"], s.ja)
    assert relerr(I.run_solver(tt, None, s.b), O.fwd_f64(s.ia, s.ja, s.a, s.b)) <= 1e-12
    for m, inf in ((tt, None), (t, info), (mat, None)):
        I.free_device_analysis(m, inf)


def test_transpose_of_upper_and_f32_handles(device):
    s = I.generate(2, 20000)
    U = reversed_of(s)                       # upper; U^T is lower
    up = I.allocate_memory_upper(U.indptr, U.indices, U.data)
    t = I.transpose(up)
    UT = sp.csr_matrix(U.T)
    UT.sort_indices()
    assert np.array_equal(I.matrix_arrays(t)["ja"], UT.indices)  # stored as is: no reversal
    b = np.random.default_rng(11).uniform(-1, 1, s.n)
    ref = O.fwd_f64(UT.indptr, UT.indices, UT.data, b)
    assert relerr(I.run_solver(t, None, b), ref) <= 1e-12
    # an FP32 upload carries only FP32 values; its transpose solves in FP32
    m32 = I.allocate_memory_f32(s.ia, s.ja, s.a.astype(np.float32))
    t32 = I.transpose(m32)
    got = I.matrix_arrays(t32)
    assert got["a"] is None
    S = stored_reversed_transpose(s.ia, s.ja, s.a.astype(np.float32).astype(np.float64), s.n)
    assert np.array_equal(got["a32"], S.data.astype(np.float32))
    Lm = sp.csr_matrix((s.a, s.ja, s.ia), shape=(s.n, s.n))
    ref = references(sp.csr_matrix(Lm.T), b, 1)
    db = I.DeviceArray.from_numpy(b.astype(np.float32))
    dx = I.DeviceArray(s.n, np.float32)
    I.solve(t32, None, I.PREC_F32, db, dx, kernel=I.KERNEL_SIMPLE)
    assert relerr(dx.numpy().astype(np.float64), ref) <= 1e-4
    for m in (t32, m32, t, up):
        I.free_device_analysis(m, None)
