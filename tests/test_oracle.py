"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each function is checked against something that does not share its code:
scipy's triangular solve, exact known answers, closed-form level counts,
Kahn peeling / memoised longest path, Python's stable sort, hand-derived
fixtures (tests/golden/hand_expected.json) and scipy.io.mmread.  The last
test plants each of the oracle's switchable mistakes and checks that the
pins catch every one of them.
"""
from __future__ import annotations

import functools

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve_triangular

import oracle_py as O
from conftest import lower_from_rows


def grid_lower(side: int, dims: int, diag: float, off: float):
    """Lower factor of the 2D 5-point / 3D 7-point Laplacian, natural order."""
    n = side ** dims
    idx = np.arange(n)
    x = idx % side
    y = (idx // side) % side
    z = idx // (side * side) if dims == 3 else np.zeros(n, int)
    cols = []
    if dims == 3:
        cols.append(np.where(z > 0, idx - side * side, -1))
    cols.append(np.where(y > 0, idx - side, -1))
    cols.append(np.where(x > 0, idx - 1, -1))
    cols.append(idx)
    C = np.stack(cols, 1)
    mask = C >= 0
    ja = C[mask].astype(np.int32)
    ia = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
    a = np.where(ja == np.repeat(idx, mask.sum(1)), diag, off).astype(np.float64)
    return ia, ja, a, (x, y, z)


def random_rows(n, dens, rng, maxrow=40):
    rows = []
    for i in range(n):
        k = min(i, int(rng.integers(0, dens + 1)) if rng.random() > 0.05 else int(rng.integers(0, maxrow)))
        rows.append(sorted(rng.choice(i, size=k, replace=False).tolist()) if k else [])
    return rows


def longest_path_levels(n, ia, ja):
    @functools.lru_cache(maxsize=None)
    def lev(i):
        deps = ja[ia[i]:ia[i + 1] - 1]
        return 1 + max((lev(int(j)) for j in deps), default=0)
    import sys
    sys.setrecursionlimit(100000)
    return np.array([lev(i) for i in range(n)], np.int32)


# ---- individual pins -----------------------------------------------------------

def pin_class_table():
    table = {0: 6, 1: 0, 2: 1, 3: 2, 4: 2, 5: 3, 6: 3, 8: 3, 9: 4, 16: 4, 17: 5, 2000: 5}
    return all(O.row_class(k) == v for k, v in table.items())


def pin_fwd_scipy(golden):
    r = golden["random"]
    for c in range(3):
        x = O.fwd_f64(r[f"c{c}_ia"], r[f"c{c}_ja"], r[f"c{c}_a"], r[f"c{c}_b"])
        ref = r[f"c{c}_x"]
        if not np.max(np.abs(x - ref)) <= 1e-12 * np.max(np.abs(ref)):
            return False
    return True


def pin_fwd_known_answer():
    ia, ja, a, _ = grid_lower(64, 2, 4.0, -1.0)
    b = np.add.reduceat(a, ia[:-1])            # b = L * ones (test.cu:13-16)
    if O.validate_x(O.fwd_f64(ia, ja, a, b)) != 0:
        return False
    x32 = O.fwd_f32(ia, ja, a.astype(np.float32), b.astype(np.float32))
    return bool(np.all(x32 == 1.0))


def fwd_f32_rowloop(ia, ja, a, b):
    """Serial FP32 forward substitution with every operation rounded to
    float32 (numpy float32 scalars): piv = 1/a_ii, left -= a_ij * x_j in
    storage order, x_i = left * piv."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    x = np.zeros(len(ia) - 1, np.float32)
    one = np.float32(1.0)
    for i in range(len(ia) - 1):
        s, e = int(ia[i]), int(ia[i + 1])
        piv = one / a[e - 1]
        left = b[i]
        for p in range(s, e - 1):
            left = left - a[p] * x[ja[p]]
        x[i] = left * piv
    return x


def pin_fwd_f32_rowloop():
    rng = np.random.default_rng(11)
    for n in (1, 40, 400):
        ia, ja, a = lower_from_rows(random_rows(n, 7, rng))
        b = rng.uniform(-1, 1, n)
        got = O.fwd_f32(ia, ja, a.astype(np.float32), b.astype(np.float32))
        if not np.array_equal(got.view(np.uint32), fwd_f32_rowloop(ia, ja, a, b).view(np.uint32)):
            return False
    return True


def pin_levels_closed_form():
    ia, ja, _, (x, y, _) = grid_lower(64, 2, 4.0, -1.0)
    lev, nl = O.levels(ia, ja)
    if nl != 127 or not np.array_equal(lev, x + y + 1):
        return False
    ia, ja, _, (x, y, z) = grid_lower(12, 3, 1.0, -1 / 6)
    lev, nl = O.levels(ia, ja)
    return nl == 34 and np.array_equal(lev, x + y + z + 1)


def pin_levels_independent(golden):
    r = golden["random"]
    for c in range(3):
        lev, nl = O.levels(r[f"c{c}_ia"], r[f"c{c}_ja"])
        if not np.array_equal(lev, r[f"c{c}_level"]) or nl != r[f"c{c}_level"].max():
            return False
    ia, ja, _ = lower_from_rows(random_rows(200, 5, np.random.default_rng(3)))
    lev, _ = O.levels(ia, ja)
    return np.array_equal(lev, longest_path_levels(200, ia, ja))


def pin_iorder_sorted(golden):
    r = golden["random"]
    for c in range(3):
        an = O.analysis(r[f"c{c}_ia"], r[f"c{c}_ja"])
        if not (np.array_equal(an["key"], r[f"c{c}_key"]) and np.array_equal(an["iorder"], r[f"c{c}_iorder"])):
            return False
    return True


def pin_hand_analysis(golden):
    h = golden["hand"]["analysis8"]
    ia, ja, _ = lower_from_rows(h["offdiag"])
    an = O.analysis(ia, ja, 32)
    ok = (an["level"].tolist() == h["level"] and an["key"].tolist() == h["key"]
          and an["nlevs"] == h["nlevs"] and an["iorder"].tolist() == h["iorder"]
          and an["bucket_off"].tolist() == h["bucket_off"] and an["n_warps"] == h["n_warps32"]
          and an["ibase_row"].tolist() == h["ibase_row32"] and an["ivect_size_warp"].tolist() == h["ivect_size32"])
    h48 = golden["hand"]["analysis48"]
    ia, ja, _ = lower_from_rows(h["offdiag"] + [[]] * 40)
    a32 = O.analysis(ia, ja, 32)
    a64 = O.analysis(ia, ja, 64)
    ok = ok and (a32["n_warps"] == h48["n_warps32"] and a32["ibase_row"].tolist() == h48["ibase_row32"]
                 and a32["ivect_size_warp"].tolist() == h48["ivect_size32"]
                 and a32["bucket_off"].tolist() == h48["bucket_off"]
                 and a64["n_warps"] == h48["n_waves64"] and a64["ibase_row"].tolist() == h48["ibase_row64"]
                 and a64["ivect_size_warp"].tolist() == h48["ivect_size64"])
    return ok


def warp_invariants(ia, ja, lanes):
    an = O.analysis(ia, ja, lanes)
    n = len(ia) - 1
    base, vs, iorder = an["ibase_row"], an["ivect_size_warp"], an["iorder"]
    if base[-1] != n or np.any(np.diff(base) <= 0) or sorted(iorder.tolist()) != list(range(n)):
        return False
    key = an["key"]
    for w in range(an["n_warps"]):
        rows = iorder[base[w]:base[w + 1]]
        if len(set(key[rows].tolist())) != 1:
            return False
        cap = lanes // max(vs[w], 1)
        if len(rows) > cap:
            return False
        cls = key[rows[0]] % 7
        if vs[w] != (0 if cls == 6 else min(2 ** cls, lanes)):
            return False
    return True


def pin_warp_invariants():
    rng = np.random.default_rng(11)
    for lanes in (32, 64):
        ia, ja, _ = lower_from_rows(random_rows(400, 20, rng, maxrow=60))
        if not warp_invariants(ia, ja, lanes):
            return False
    return True


def pin_loader(golden):
    g = golden["hand"]
    for name in ("tiny_general", "tiny_symmetric"):
        path = golden["dir"] / f"{name}.mtx"
        m, ri, ci, v, sym = O.read_mtx_coo(path)
        rp, col, val = O.coo_to_csr(m, ri, ci, v, sym)
        e = g[name]
        if rp.tolist() != e["csr_rp"] or col.tolist() != e["csr_col"] or val.tolist() != e["csr_val"]:
            return False
        # entry multiset agrees with scipy's reader (mirrored for symmetric)
        A = scipy.io.mmread(str(path)).tocoo()
        ref = sorted(zip(A.row.tolist(), A.col.tolist(), A.data.tolist()))
        mine = sorted((r, int(c), float(x)) for r in range(m) for c, x in
                      zip(col[rp[r]:rp[r + 1]], val[rp[r]:rp[r + 1]]))
        if ref != mine:
            return False
        lp, lc, lv = O.extract_lower(m, rp, col, val, unit=False)
        if lp.tolist() != e["L_ia"] or lc.tolist() != e["L_ja"] or lv.tolist() != e["L_val_keep"]:
            return False
        lp, lc, lv = O.extract_lower(m, rp, col, val, unit=True)
        if lc.tolist() != e["L_ja"] or not np.all(lv == 1.0):
            return False
    return True


PINS = {
    "fwd_f32_rowloop": lambda g: pin_fwd_f32_rowloop(),
    "class_table": lambda g: pin_class_table(),
    "fwd_scipy": pin_fwd_scipy,
    "fwd_known_answer": lambda g: pin_fwd_known_answer(),
    "levels_closed_form": lambda g: pin_levels_closed_form(),
    "levels_independent": pin_levels_independent,
    "iorder_sorted": pin_iorder_sorted,
    "hand_analysis": pin_hand_analysis,
    "warp_invariants": lambda g: pin_warp_invariants(),
    "loader": pin_loader,
}


@pytest.mark.parametrize("pin", sorted(PINS))
def test_pin(pin, golden):
    assert PINS[pin](golden), f"oracle pin {pin} failed"


def test_full_size_closed_forms():
    # 128^3 3D 7-point factor: nlevs = 3*127 + 1 = 382; 64x64 2D: 127
    ia, ja, _, (x, y, z) = grid_lower(128, 3, 1.0, -1 / 6)
    lev, nl = O.levels(ia, ja)
    assert nl == 382
    assert np.array_equal(lev, x + y + z + 1)


def test_fwd_random_live_scipy():
    rng = np.random.default_rng(5)
    for n in (1, 2, 50, 700):
        ia, ja, a = lower_from_rows(random_rows(n, 8, rng))
        b = rng.uniform(-1, 1, n)
        x = O.fwd_f64(ia, ja, a, b)
        ref = spsolve_triangular(sp.csr_matrix((a, ja, ia), shape=(n, n)), b, lower=True)
        assert np.max(np.abs(x - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_fwd_multi_rhs_matches_columns():
    rng = np.random.default_rng(6)
    ia, ja, a = lower_from_rows(random_rows(300, 6, rng))
    B = rng.uniform(-1, 1, (300, 5))
    X = O.fwd_f64(ia, ja, a, B.ravel(), k=5).reshape(300, 5)
    for r in range(5):
        assert np.array_equal(X[:, r], O.fwd_f64(ia, ja, a, B[:, r].copy()))


def test_unit_values_known_answer(golden):
    # the reference's own check: unit-valued L, b = L*1, x == 1 exactly
    rng = np.random.default_rng(7)
    ia, ja, _ = lower_from_rows(random_rows(500, 10, rng))
    a = np.ones(len(ja))
    b = np.add.reduceat(a, ia[:-1])
    assert O.validate_x(O.fwd_f64(ia, ja, a, b)) == 0


def test_validate_x():
    assert O.validate_x(np.ones(10)) == 0
    x = np.ones(10)
    x[3] = 1 + 2 ** -52
    x[7] = 0.0
    assert O.validate_x(x) == 2


def test_check_lower_csr():
    assert O.check_lower_csr([0, 1, 3], [0, 0, 1]) == 0
    assert O.check_lower_csr([0, 1, 3], [0, 1, 0]) == 2      # diagonal not last
    assert O.check_lower_csr([0, 1, 3], [0, 1, 1]) == 2      # column == row before the diagonal


def test_pattern_loader(golden):
    m, ri, ci, v, sym = O.read_mtx_coo(golden["dir"] / "tiny_pattern.mtx")
    rp, col, val = O.coo_to_csr(m, ri, ci, v, sym)
    lp, lc, _ = O.extract_lower(m, rp, col, val, unit=True)
    assert lp.tolist() == golden["hand"]["tiny_pattern"]["L_ia"]
    assert lc.tolist() == golden["hand"]["tiny_pattern"]["L_ja"]


MUTATIONS = {
    1: "diagonal read from the first entry",
    2: "sign of the off-diagonal products",
    3: "dropped term in the level max",
    4: "unstable tie-break in iorder",
    5: "class boundary shifted",
    6: "warps per bucket rounded down",
    7: "lower extraction ignores the break rule",
    8: "symmetric entries not mirrored",
}


@pytest.mark.parametrize("m", sorted(MUTATIONS))
def test_pins_catch_planted_mistake(m, golden):
    with O.mutation(m):
        failed = [name for name, pin in PINS.items() if not pin(golden)]
    assert failed, f"planted mistake {m} ({MUTATIONS[m]}) passed every pin"


@pytest.mark.parametrize("m", [1, 2])
def test_fwd_f32_pin_catches_planted_mistakes(m):
    # diagonal from the wrong entry, products added: the float32 row loop
    # must see either
    with O.mutation(m):
        assert not pin_fwd_f32_rowloop()
