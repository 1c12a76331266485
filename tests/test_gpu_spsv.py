"""The spsv-style handle: every fill mode, diagonal type and operation over a
general device CSR, checked against scipy's spsolve_triangular on the
triangle that the options name (and the oracle for the plain lower case)."""
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


def general_matrix(n, per_row, rng):
    """Both triangles populated, one diagonal entry per row, columns shuffled."""
    ia, ja, a = [0], [], []
    for i in range(n):
        cols = np.unique(rng.integers(0, n, per_row))
        cols = cols[cols != i]
        vals = rng.uniform(-1, 1, len(cols))
        cols = np.append(cols, i)
        vals = np.append(vals, 1.0 + np.abs(vals).sum())
        perm = rng.permutation(len(cols))
        ja += cols[perm].tolist()
        a += vals[perm].tolist()
        ia.append(len(ja))
    return np.array(ia, np.int32), np.array(ja, np.int32), np.array(a)


def triangle(A, fill, diag):
    T = sp.tril(A, -1) if fill == I.FILL_LOWER else sp.triu(A, 1)
    D = sp.identity(A.shape[0]) if diag == I.DIAG_UNIT else sp.diags(A.diagonal())
    return sp.csr_matrix(T + D)


def dev(arr):
    return I.DeviceArray.from_numpy(arr)


@pytest.mark.parametrize("fill", [0, 1])
@pytest.mark.parametrize("diag", [0, 1])
@pytest.mark.parametrize("op", [0, 1])
def test_spsv_options(device, fill, diag, op):
    rng = np.random.default_rng(100 + 4 * fill + 2 * diag + op)
    n = 20000
    ia, ja, a = general_matrix(n, 9, rng)
    A = sp.csr_matrix((a, ja, ia), shape=(n, n))
    T = triangle(A, fill, diag)
    M = sp.csr_matrix(T.T) if op == I.OP_TRANSPOSE else T
    lower = (fill == I.FILL_LOWER) != (op == I.OP_TRANSPOSE)
    k = 3
    B = rng.uniform(-1, 1, (n, k))
    alpha = 2.5
    ref = np.column_stack([spsolve_triangular(M, alpha * B[:, r], lower=lower) for r in range(k)]).reshape(-1)
    if fill == I.FILL_LOWER and op == I.OP_NONE:  # the oracle agrees on the plain lower case
        Ms = sp.csr_matrix(M)
        Ms.sort_indices()
        assert relerr(alpha * O.fwd_f64(Ms.indptr, Ms.indices, Ms.data, B.reshape(-1), k=k), ref) <= 1e-12
    dia, dja, da = dev(ia), dev(ja), dev(a)
    h = I.Spsv(n, dia, dja, da, fill, diag, op)
    for x in (dia, dja, da):
        x.free()                                   # the handle keeps its own copies
    h.analysis(I.PREC_MIXED, k=k, tune=(fill == 0 and diag == 0 and op == 0))
    for prec, tol in ((I.PREC_F64, 1e-12), (I.PREC_MIXED, 1e-10), (I.PREC_F32, 1e-4)):
        dt = np.float32 if prec == I.PREC_F32 else np.float64
        db = dev(B.reshape(-1).astype(dt))
        dx = I.DeviceArray(n * k, dt)
        h.solve(db, dx, prec, alpha=alpha, k=k)
        assert I.check_watchdog(h.matrix()) == 0, L.last_error()
        assert relerr(dx.numpy().astype(np.float64), ref) <= tol, prec
        db.free()
        dx.free()
    h.close()


def test_spsv_diagonal_rules_and_errors(device):
    # row 1 holds two diagonal entries: the first one is used; the upper
    # entry (0, 2) is ignored by a lower solve
    ia = np.array([0, 2, 5, 6], np.int32)
    ja = np.array([0, 2, 1, 0, 1, 2], np.int32)
    a = np.array([2.0, 9.0, 4.0, 1.0, 7.0, 5.0])
    b = np.array([2.0, 5.0, 10.0])
    h = I.Spsv(3, dev(ia), dev(ja), dev(a), I.FILL_LOWER, I.DIAG_NON_UNIT, I.OP_NONE)
    db, dx = dev(b), I.DeviceArray(3, np.float64)
    h.solve(db, dx)
    x0 = 1.0
    x1 = (5.0 - 1.0 * x0) / 4.0
    x2 = 10.0 / 5.0
    assert np.allclose(dx.numpy(), [x0, x1, x2], rtol=0, atol=1e-15)
    h.close()
    # unit diagonal: stored diagonal entries are ignored
    h = I.Spsv(3, dev(ia), dev(ja), dev(a), I.FILL_LOWER, I.DIAG_UNIT, I.OP_NONE)
    h.solve(db, dx)
    assert np.allclose(dx.numpy(), [2.0, 5.0 - 2.0, 10.0], rtol=0, atol=1e-15)
    h.close()
    # a row without a diagonal entry cannot be a non-unit triangle
    ia2 = np.array([0, 1, 2], np.int32)
    ja2 = np.array([0, 0], np.int32)
    with pytest.raises(L.SptrsvError) as e:
        I.Spsv(2, dev(ia2), dev(ja2), dev(np.ones(2)), I.FILL_LOWER, I.DIAG_NON_UNIT, I.OP_NONE)
    assert e.value.status == L.E_INVALID_CSR
    # ... but it is fine with a unit diagonal
    h = I.Spsv(2, dev(ia2), dev(ja2), dev(np.array([1.0, 3.0])), I.FILL_LOWER, I.DIAG_UNIT, I.OP_NONE)
    h.solve(dev(np.array([1.0, 4.0])), dx2 := I.DeviceArray(2, np.float64))
    assert dx2.numpy().tolist() == [1.0, 1.0]
    h.close()
    # a column outside [0, n)
    with pytest.raises(L.SptrsvError) as e:
        I.Spsv(2, dev(ia2), dev(np.array([0, 2], np.int32)), dev(np.ones(2)), I.FILL_LOWER, I.DIAG_UNIT,
               I.OP_NONE)
    assert e.value.status == L.E_INVALID_CSR
