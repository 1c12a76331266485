"""New values on an unchanged pattern (sptrsv_update_values): every kind of
handle solves with the new values while keeping its analysis, tuning and
derived Jacobi data consistent."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve_triangular

import oracle_py as O
from sptrsv_amd import _lib as L
from sptrsv_amd import interface as I

pytestmark = pytest.mark.gpu


def relerr(x, ref):
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.fixture(scope="module")
def device():
    print("device:", I.device_name(0), flush=True)
    return True


def new_values(ia, a, rng, dominant=True):
    """Fresh off-diagonal values, diagonal 1 + sum|off| (or 1.0) on the same pattern."""
    out = rng.uniform(-1, 1, len(a))
    for i in range(len(ia) - 1):
        s, e = ia[i], ia[i + 1]
        out[e - 1] = 1.0 + np.abs(out[s:e - 1]).sum() if dominant else 1.0
    return out


def solve(mat, info, b, precision=I.PREC_F64, cfg=None, graph=False, **kw):
    db = I.DeviceArray.from_numpy(b)
    dx = I.DeviceArray(len(b), np.float64)
    if graph:
        I.solve_graph(mat, info, precision, db, dx, cfg)
    else:
        I.solve(mat, info, precision, db, dx, cfg=cfg, **kw)
    assert I.check_watchdog(mat) == 0, L.last_error()
    return dx.numpy()


def test_lower_handle_keeps_analysis_tuning_and_graphs(device):
    s = I.generate(2, 100000)
    rng = np.random.default_rng(1)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    levels = I.analysis_arrays(info)["level"].copy()
    cfg, _ = I.autotune(mat, info, I.PREC_MIXED)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    assert relerr(solve(mat, info, s.b, I.PREC_MIXED, cfg, graph=True), ref) <= 1e-10
    jac = dict(kernel=I.KERNEL_JACOBI, sweeps=12, outer=4)
    assert relerr(solve(mat, info, s.b, I.PREC_MIXED, **jac), ref) <= 1e-10
    margin0 = I.dominance_margin(mat)
    for rep in range(2):
        a2 = new_values(s.ia, s.a, rng)
        I.update_values(mat, a2)
        ref2 = O.fwd_f64(s.ia, s.ja, a2, s.b)
        print("update", rep, flush=True)
        for kern in (I.KERNEL_SIMPLE, I.KERNEL_ORDER, I.KERNEL_MULTIROW):
            assert relerr(solve(mat, info, s.b, kernel=kern), ref2) <= 1e-12, kern
        assert relerr(solve(mat, info, s.b, I.PREC_MIXED, **jac), ref2) <= 1e-10
        assert relerr(solve(mat, info, s.b, I.PREC_MIXED, cfg, graph=True), ref2) <= 1e-10
        assert relerr(solve(mat, info, s.b, I.PREC_MIXED), ref2) <= 1e-10
        assert I.dominance_margin(mat) != margin0
    assert np.array_equal(I.analysis_arrays(info)["level"], levels)  # the analysis is reused as is
    # a device-resident update, and a non-dominant one: the automatic choice stays correct
    a3 = new_values(s.ia, s.a, rng, dominant=False) * 0.5
    a3[s.ia[1:] - 1] = 1.0
    d3 = I.DeviceArray.from_numpy(a3)
    I.update_values(mat, d3)
    ref3 = O.fwd_f64(s.ia, s.ja, a3, s.b)
    assert I.dominance_margin(mat) <= 0
    x = solve(mat, info, s.b, I.PREC_MIXED)
    assert relerr(x, ref3) <= 1e-9
    with pytest.raises(L.SptrsvError) as e:
        I.update_values(mat, a3[:-1])
    assert e.value.status == L.E_INVALID_ARG
    I.free_device_analysis(mat, info)


def test_upper_transpose_ingest_and_spsv_handles(device):
    rng = np.random.default_rng(2)
    s = I.generate(2, 30000)
    n = s.n
    Lm = sp.csr_matrix((s.a, s.ja, s.ia), shape=(n, n))
    U = sp.csr_matrix(Lm.T)
    U.sort_indices()
    b = rng.uniform(-1, 1, n)
    # upper upload: new values in U's own layout
    up = I.allocate_memory_upper(U.indptr, U.indices, U.data)
    U2 = U.copy()
    U2.data = rng.uniform(-1, 1, U.nnz)
    U2.data[U2.indptr[:-1]] = 4.0 + rng.uniform(0, 1, n)   # diagonal first in every row
    I.update_values(up, U2.data)
    assert relerr(solve(up, None, b), spsolve_triangular(U2, b, lower=False)) <= 1e-12
    # the transpose of a plain upload takes values in the source's layout
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    t = I.transpose(mat)
    a2 = new_values(s.ia, s.a, rng)
    I.update_values(t, a2)
    L2 = sp.csr_matrix((a2, s.ja, s.ia), shape=(n, n))
    assert relerr(solve(t, None, b), spsolve_triangular(sp.csr_matrix(L2.T), b, lower=False)) <= 1e-12
    # the transpose of the upper handle composes both maps
    tu = I.transpose(up)
    I.update_values(tu, U2.data * 0.5)
    UT = sp.csr_matrix((U2.data * 0.5, U.indices, U.indptr), shape=(n, n)).T
    assert relerr(solve(tu, None, b), spsolve_triangular(sp.csr_matrix(UT), b, lower=True)) <= 1e-12
    for m in (tu, t, mat, up):
        I.free_device_analysis(m, None)
    # a handle built from COO takes new COO values
    rows, ci = np.repeat(np.arange(n, dtype=np.int32), np.diff(s.ia)), s.ja
    m = I.coo_to_lower(rows, ci, s.a, n, False, False)
    v2 = new_values(s.ia, s.a, rng)
    I.update_values(m, v2)
    rp, col, val = O.coo_to_csr(n, rows, ci, v2, False)
    lp, lc, lv = O.extract_lower(n, rp, col, val, unit=False)
    assert np.array_equal(I.matrix_arrays(m)["a"], lv)
    I.free_device_analysis(m, None)
    # an spsv handle (upper, transposed) takes values in its CSR's layout
    A = sp.csr_matrix(Lm + sp.csr_matrix(Lm.T))
    A.sort_indices()
    dia, dja = I.DeviceArray.from_numpy(A.indptr.astype(np.int32)), I.DeviceArray.from_numpy(A.indices.astype(np.int32))
    h = I.Spsv(n, dia, dja, I.DeviceArray.from_numpy(A.data), I.FILL_UPPER, I.DIAG_NON_UNIT, I.OP_TRANSPOSE)
    h.analysis(I.PREC_F64)
    A2 = A.copy()
    A2.data = rng.uniform(-1, 1, A.nnz)
    A2.setdiag(6.0)
    A2 = sp.csr_matrix(A2)
    A2.sort_indices()
    assert np.array_equal(A2.indices, A.indices)
    I.update_values(h.matrix(), A2.data)
    T = sp.csr_matrix(sp.triu(A2).T)
    db = I.DeviceArray.from_numpy(b)
    dx = I.DeviceArray(n, np.float64)
    h.solve(db, dx)
    assert relerr(dx.numpy(), spsolve_triangular(T, b, lower=True)) <= 1e-12
    h.close()
