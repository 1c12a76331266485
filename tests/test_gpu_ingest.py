"""The driver's matrix preparation on the device (sptrsv_coo_to_lower).

Each case compares the handle's stored CSR, bit for bit, with the oracle's
host pipeline (coo_to_csr then extract_lower, pinned on the hand fixtures in
test_host.py), then solves with the handle.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as O
from sptrsv_amd import _lib as L
from sptrsv_amd import interface as I

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    print("device:", I.device_name(0), flush=True)
    return True


def expected(n, ri, ci, v, sym, unit):
    rp, col, val = O.coo_to_csr(n, ri, ci, v, sym)
    return O.extract_lower(n, rp, col, val, unit=unit)


def check_handle(m, n, ri, ci, v, sym, unit):
    lp, lc, lv = expected(n, ri, ci, v, sym, unit)
    got = I.matrix_arrays(m)
    assert m.nr == n and m.nnz == len(lc)
    assert np.array_equal(got["ia"], lp)
    assert np.array_equal(got["ja"], lc)
    assert np.array_equal(got["a"], lv)
    assert np.array_equal(got["a32"], lv.astype(np.float32))
    return lp, lc, lv


@pytest.mark.parametrize("name", ["tiny_general", "tiny_symmetric", "tiny_pattern"])
@pytest.mark.parametrize("unit", [True, False])
def test_fixtures(device, golden, name, unit):
    n, ri, ci, v, sym = I.read_mtx_coo(str(golden["dir"] / f"{name}.mtx"))
    m = I.coo_to_lower(ri, ci, v, n, sym, unit)
    lp, lc, lv = check_handle(m, n, ri, ci, v, sym, unit)
    assert lp.tolist() == golden["hand"][name]["L_ia"] and lc.tolist() == golden["hand"][name]["L_ja"]
    I.free_device_analysis(m, None)


@pytest.mark.parametrize("n,nz,sym", [(5000, 40000, False), (5000, 40000, True), (300000, 3000000, True)])
def test_random_coo_with_duplicates_and_missing_diagonals(device, n, nz, sym):
    # uniform rows and columns: duplicates, rows with no diagonal, upper
    # entries ahead of lower ones within a row, and empty rows all occur
    rng = np.random.default_rng(n + nz + sym)
    ri = rng.integers(0, n, nz).astype(np.int32)
    ci = rng.integers(0, n, nz).astype(np.int32)
    if sym:
        ri, ci = np.maximum(ri, ci), np.minimum(ri, ci)
    v = rng.uniform(-2, 2, nz)
    diag_rows = rng.choice(n, n // 2, replace=False).astype(np.int32)   # half the rows get a diagonal
    ri = np.concatenate([ri, diag_rows])
    ci = np.concatenate([ci, diag_rows])
    v = np.concatenate([v, 4.0 + rng.uniform(0, 1, len(diag_rows))])
    perm = rng.permutation(len(ri))
    ri, ci, v = ri[perm], ci[perm], v[perm]
    for unit in (False, True):
        print("unit", unit, flush=True)
        m = I.coo_to_lower(ri, ci, v, n, sym, unit)
        lp, lc, lv = check_handle(m, n, ri, ci, v, sym, unit)
        if unit:  # the reference's known answer: integer b = L*1 gives x = 1 exactly
            b = np.add.reduceat(lv, lp[:-1])
            assert O.validate_x(I.run_solver(m, None, b)) == 0
        I.free_device_analysis(m, None)


def test_edge_cases(device):
    # no entries: the identity
    m = I.coo_to_lower(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0), 4, False, False)
    got = I.matrix_arrays(m)
    assert got["ia"].tolist() == [0, 1, 2, 3, 4] and got["ja"].tolist() == [0, 1, 2, 3]
    assert got["a"].tolist() == [1.0] * 4
    I.free_device_analysis(m, None)
    # an empty matrix
    m = I.coo_to_lower(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0), 0, False, True)
    assert m.nr == 0 and m.nnz == 0
    I.free_device_analysis(m, None)
    # an index outside [0, n) is rejected on the device
    with pytest.raises(L.SptrsvError) as e:
        I.coo_to_lower(np.array([0, 5], np.int32), np.array([0, 1], np.int32), np.ones(2), 3, False, True)
    assert e.value.status == L.E_INVALID_CSR


def test_mtx_file_to_device_matches_host_loader(device, tmp_path):
    from test_host import random_coo, write_mtx
    rng = np.random.default_rng(5)
    n, nz = 20000, 200000
    ri, ci, v = random_coo(n, nz, rng, sym=True)
    path = tmp_path / "m.mtx"
    write_mtx(path, n, ri, ci, v, sym=True)
    for unit in (True, False):
        nn, r, c, x, sym = I.read_mtx_coo(str(path))
        m = I.coo_to_lower(r, c, x, nn, sym, unit)
        _, ia, ja, a = I.read_mtx_lower(str(path), unit=unit)
        got = I.matrix_arrays(m)
        assert np.array_equal(got["ia"], ia) and np.array_equal(got["ja"], ja) and np.array_equal(got["a"], a)
        I.free_device_analysis(m, None)
