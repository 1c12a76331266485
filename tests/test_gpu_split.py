"""GPU tests of the long-row split of the multirow walks.

A row of more than 256 off-diagonal entries keeps its first 128; every
further run of 256 entries is a segment item folded by a wave of its own,
placed in the walk right before the wave that holds the row, and the row adds
its segments' partial sums in a fixed order (lane e of its 32-lane group
takes partials e, e + 32, ...).  The reference walks such rows with at most
32 lanes in one warp (solve_csr_multirow.cu:21-114).

Bars: the number of split rows and segments equals the count the split rule
gives on the host; FP64 within 1e-12 relative (max-norm) of the oracle's
serial substitution, FP32 within 1e-5, fused mixed within 1e-10; repeated
solves bit-identical; whole-row and split walks within the same bars of each
other.  Row lengths straddle every boundary of the rule: 256 off-diagonals
(not split), 257 (one segment), 384 and 385 (one and two), rows needing more
segments than the row's 32 lanes (two and five partials per lane), and chains
of long rows so that the critical path runs through split rows.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as O
from sptrsv_amd import _lib as L
from sptrsv_amd import interface as I

pytestmark = pytest.mark.gpu

SEG_MIN, SEG_OWN, SEG_ENTRIES = 256, 128, 256


def seg_count(off: int) -> int:
    return -(-(off - SEG_OWN) // SEG_ENTRIES) if off > SEG_MIN else 0


def relerr(x, ref):
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


def long_row_matrix(n: int, long_rows: dict, seed: int):
    """Lower CSR, diagonal last: Poisson(6) short rows over [i-400, i), plus
    the rows in long_rows (row -> off-diagonal count), each of which also
    depends on the previous long row (a chain through split rows)."""
    rng = np.random.default_rng(seed)
    cols = []
    prev = None
    for i in range(n):
        if i in long_rows:
            m = min(long_rows[i], i)
            c = rng.choice(i, size=m, replace=False)
            if prev is not None and prev not in c:
                c[0] = prev
            prev = i
        else:
            lo = max(0, i - 400)
            m = min(rng.poisson(6), i - lo)
            c = rng.choice(np.arange(lo, i), size=m, replace=False) if m else np.empty(0, np.int64)
        cols.append(np.sort(c))
    ia = np.zeros(n + 1, np.int32)
    ia[1:] = np.cumsum([len(c) + 1 for c in cols])
    ja = np.empty(ia[-1], np.int32)
    a = np.empty(ia[-1], np.float64)
    for i, c in enumerate(cols):
        s, e = ia[i], ia[i + 1]
        ja[s:e - 1] = c
        ja[e - 1] = i
        v = rng.uniform(-1.0, 1.0, len(c)) / max(1, len(c))
        a[s:e - 1] = v
        a[e - 1] = 1.0 + np.abs(v).sum()
    b = rng.uniform(-1.0, 1.0, n)
    return ia, ja, a, b


LONG = {5000: 256, 5100: 257, 5200: 384, 5300: 385, 9000: 2000, 25000: 20000, 44000: 40000,
        44001: 1000, 44002: 300, 20000: 640, 30000: 9000}
for r in range(31000, 31400, 7):
    LONG[r] = 300 + (r % 500)


@pytest.fixture(scope="module")
def system():
    ia, ja, a, b = long_row_matrix(45000, LONG, seed=11)
    mat = I.allocate_memory(ia, ja, a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    yield ia, ja, a, b, mat, info
    I.free_device_analysis(mat, info)


def solve(mat, info, prec, b, **cfg):
    dt = np.float32 if prec == I.PREC_F32 else np.float64
    db = I.DeviceArray.from_numpy(b.astype(dt))
    dx = I.DeviceArray(len(b), dt)
    # lanes given, so variant 0 (packed waves) is not replaced by the default shape's
    I.solve(mat, info, prec, db, dx, cfg=I._cfg(kernel=I.KERNEL_MULTIROW, lanes=32, **cfg))
    assert I.check_watchdog(mat) == 0, L.last_error()
    x = dx.numpy()
    db.free()
    dx.free()
    return x


def test_segment_table_counts(system):
    ia, ja, a, b, mat, info = system
    off = np.diff(ia) - 1
    expect_segs = sum(seg_count(int(o)) for o in off)
    expect_rows = int(sum(1 for o in off if seg_count(int(o)) > 0))
    rows, segs = I.analysis_segments(info)
    print(f"split rows {rows}, segments {segs}", flush=True)
    assert (rows, segs) == (expect_rows, expect_segs)
    # the boundaries of the rule
    assert seg_count(256) == 0 and seg_count(257) == 1 and seg_count(384) == 1 and seg_count(385) == 2
    assert seg_count(40000) == 156


@pytest.mark.parametrize("variant,wg,cu_waves", [(0, 256, 0), (1, 256, 0), (2, 256, 4), (2, 1024, 8)])
def test_split_f64_matches_oracle(system, variant, wg, cu_waves):
    ia, ja, a, b, mat, info = system
    print(f"split FP64, variant {variant}, wg {wg}", flush=True)
    ref = O.fwd_f64(ia, ja, a, b)
    xs = solve(mat, info, I.PREC_F64, b, wg=wg, variant=variant, cu_waves=cu_waves, split_rows=1)
    xs2 = solve(mat, info, I.PREC_F64, b, wg=wg, variant=variant, cu_waves=cu_waves, split_rows=1)
    xw = solve(mat, info, I.PREC_F64, b, wg=wg, variant=variant, cu_waves=cu_waves, split_rows=-1)
    assert relerr(xs, ref) <= 1e-12
    assert relerr(xw, ref) <= 1e-12
    assert np.array_equal(xs.view(np.uint64), xs2.view(np.uint64))
    # the default configuration splits
    xd = solve(mat, info, I.PREC_F64, b, wg=wg, variant=variant, cu_waves=cu_waves)
    assert np.array_equal(xs.view(np.uint64), xd.view(np.uint64))


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_split_f32_and_mixed(system, variant):
    ia, ja, a, b, mat, info = system
    print(f"split FP32 and mixed, variant {variant}", flush=True)
    ref = O.fwd_f64(ia, ja, a, b)
    x32 = solve(mat, info, I.PREC_F32, b, wg=256, variant=variant, cu_waves=4 if variant == 2 else 0, split_rows=1)
    assert relerr(x32.astype(np.float64), ref) <= 1e-5
    xm = solve(mat, info, I.PREC_MIXED, b, wg=256, variant=variant, cu_waves=4 if variant == 2 else 0,
               split_rows=1)
    xm2 = solve(mat, info, I.PREC_MIXED, b, wg=256, variant=variant, cu_waves=4 if variant == 2 else 0,
                split_rows=1)
    assert relerr(xm, ref) <= 1e-10
    assert np.array_equal(xm.view(np.uint64), xm2.view(np.uint64))


def test_split_noalloc_matches_allocating_entry(system):
    ia, ja, a, b, mat, info = system
    print("split rows through the allocation-free entry point", flush=True)
    db = I.DeviceArray.from_numpy(b)
    dx1 = I.DeviceArray(len(b), np.float64)
    dx2 = I.DeviceArray(len(b), np.float64)
    cfg = I._cfg(kernel=I.KERNEL_MULTIROW, lanes=32, wg=256, variant=0, fused=1)
    I.solve(mat, info, I.PREC_MIXED, db, dx1, cfg=cfg)
    I.solve_mixed_noalloc(mat, info, db, dx2, k=1, cfg=cfg)
    assert I.check_watchdog(mat) == 0
    assert np.array_equal(dx1.numpy().view(np.uint64), dx2.numpy().view(np.uint64))
    for d in (db, dx1, dx2):
        d.free()


def test_split_on_the_power_law_workload():
    print("split rows, power-law n=300000", flush=True)
    s = I.generate(3, 300000)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    rows, segs = I.analysis_segments(info)
    off = np.diff(s.ia) - 1
    assert segs == sum(seg_count(int(o)) for o in off) and rows > 0
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    for variant in (0, 1, 2):
        x = solve(mat, info, I.PREC_F64, s.b, wg=256, variant=variant, cu_waves=4 if variant == 2 else 0)
        assert relerr(x, ref) <= 1e-12
        xm = solve(mat, info, I.PREC_MIXED, s.b, wg=256, variant=variant, cu_waves=4 if variant == 2 else 0)
        assert relerr(xm, ref) <= 1e-10
    # an ORDER-only analysis has no wave layout and so no segments
    info2 = I.multirow_analysis_base_GPU(mat, I.ORDER)
    assert I.analysis_segments(info2) == (0, 0)
    I.free_device_analysis(None, info2)
    I.free_device_analysis(mat, info)


def test_captured_solve_follows_the_analysis(system):
    """A captured solve holds the analysis arrays it was recorded with: a new
    analysis of the same matrix (the old one freed) gets a solve of its own."""
    ia, ja, a, b, _, _ = system
    print("captured solve after re-analysis", flush=True)
    mat = I.allocate_memory(ia, ja, a)
    db = I.DeviceArray.from_numpy(b)
    dx = I.DeviceArray(len(b), np.float64)
    cfg = I._cfg(kernel=I.KERNEL_MULTIROW, lanes=32, wg=256, variant=0)
    ref = O.fwd_f64(ia, ja, a, b)
    info1 = I.multirow_analysis_base_GPU(mat, I.ALL)
    I.solve_graph(mat, info1, I.PREC_F64, db, dx, cfg)
    I.synchronize()
    x1 = dx.numpy()
    I.free_device_analysis(None, info1)
    info2 = I.multirow_analysis_base_GPU(mat, I.ALL)
    for _ in range(3):
        I.solve_graph(mat, info2, I.PREC_F64, db, dx, cfg)
        I.synchronize()
        assert I.check_watchdog(mat) == 0
        x2 = dx.numpy()
        assert np.array_equal(x1.view(np.uint64), x2.view(np.uint64))
    assert relerr(x1, ref) <= 1e-12
    db.free()
    dx.free()
    I.free_device_analysis(mat, info2)
