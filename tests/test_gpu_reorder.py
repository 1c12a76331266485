"""GPU tests of the level-ordered copy (config reorder = 1): the solve walks a
copy of L whose row p is row iorder[p] (entries in storage order, columns
renamed), with the copy's own analysis, reading b and writing x in the
caller's order.

Bars: bit-identical to reorder = 0 with the same configuration (ORDER and
MULTIROW, variants 0..2, FP64 and fused mixed, one and several right-hand
sides, split long rows) and within 1e-12 (FP64) / 1e-10 (mixed) of the
oracle; exact x == 1 on the 2D Laplacian factor; an upper-triangular handle;
captured solves replayed bit-identical; the copy rebuilt after new values and
for another analysis; tiny matrices (one row, chains, diagonal-only rows);
an autotuned copy serving automatic solves; and the refusals (SIMPLE, FP32,
unfused mixed, the in-degree walk, the allocation-free solve).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as O
from conftest import lower_from_rows
from sptrsv_amd import _lib as L
from sptrsv_amd import interface as I

pytestmark = pytest.mark.gpu


def relerr(x, ref):
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


def solve(mat, info, prec, b, k=1, **cfg):
    db = I.DeviceArray.from_numpy(b)
    dx = I.DeviceArray(len(b), np.float64)
    I.solve(mat, info, prec, db, dx, k=k, cfg=I._cfg(**cfg))
    assert I.check_watchdog(mat) == 0, L.last_error()
    x = dx.numpy()
    db.free()
    dx.free()
    return x


def bits(x):
    return x.view(np.uint64)


def same_as_original(mat, info, prec, b, k=1, **shape):
    x0 = solve(mat, info, prec, b, k=k, reorder=0, **shape)
    x1 = solve(mat, info, prec, b, k=k, reorder=1, **shape)
    assert np.array_equal(bits(x1), bits(x0)), (prec, k, shape)
    return x1


@pytest.mark.parametrize("gen,param", [(2, 100000), (3, 100000)])
def test_copy_equals_the_original_one_rhs(gen, param):
    print(f"level-ordered copy, generator {gen}", flush=True)
    s = I.generate(gen, param)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    shapes = [dict(kernel=I.KERNEL_ORDER, lanes=16, wg=256, variant=v) for v in (0, 1, 2)]
    shapes += [dict(kernel=I.KERNEL_MULTIROW, lanes=32, wg=wg, variant=v, cu_waves=8)
               for v in (0, 1, 2) for wg in (256, 1024)]
    shapes += [dict(kernel=I.KERNEL_MULTIROW, lanes=32, wg=1024, variant=2, split_rows=-1)]
    for shape in shapes:
        for prec, tol in ((I.PREC_F64, 1e-12), (I.PREC_MIXED, 1e-10)):
            x = same_as_original(mat, info, prec, s.b, **shape)
            assert relerr(x, ref) <= tol, (shape, prec)
    I.free_device_analysis(mat, info)


def test_copy_equals_the_original_several_rhs():
    print("level-ordered copy, 3D Laplacian factor, 4 right-hand sides", flush=True)
    s = I.generate(4, 24)
    k = 4
    rng = np.random.default_rng(3)
    b = rng.standard_normal(s.n * k)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    ref = O.fwd_f64(s.ia, s.ja, s.a, b, k=k)
    for shape in (dict(kernel=I.KERNEL_ORDER, lanes=4, wg=256, variant=0),
                  dict(kernel=I.KERNEL_ORDER, lanes=1, wg=1024, variant=2, cu_waves=6),
                  dict(kernel=I.KERNEL_MULTIROW, lanes=16, wg=256, variant=2)):
        for prec, tol in ((I.PREC_F64, 1e-12), (I.PREC_MIXED, 1e-10)):
            x = same_as_original(mat, info, prec, b, k=k, **shape)
            assert relerr(x, ref) <= tol, (shape, prec)
    I.free_device_analysis(mat, info)


def test_copy_known_answer_upper_and_replays():
    print("level-ordered copy: x == 1, upper handle, captured replays", flush=True)
    s = I.generate(1)                        # 2D Laplacian factor, b = L * ones
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ORDER)
    x = solve(mat, info, I.PREC_F64, s.b, kernel=I.KERNEL_ORDER, lanes=4, wg=256, variant=1, reorder=1)
    assert np.array_equal(x, np.ones(s.n))
    # the same system as an upper-triangular handle (rows and columns reversed)
    n = s.n
    rows = np.repeat(np.arange(n), np.diff(s.ia))
    ur, uc = n - 1 - rows, n - 1 - s.ja
    order = np.lexsort((uc, ur))              # diagonal first in each upper row
    uia = np.zeros(n + 1, np.int32)
    np.add.at(uia, ur + 1, 1)
    uia = np.cumsum(uia).astype(np.int32)
    up = I.allocate_memory_upper(uia, uc[order].astype(np.int32), s.a[order])
    uinfo = I.multirow_analysis_base_GPU(up, I.ALL)
    ub = s.b[::-1].copy()
    for kern in (I.KERNEL_ORDER, I.KERNEL_MULTIROW):
        xu = same_as_original(up, uinfo, I.PREC_F64, ub, kernel=kern, lanes=4, wg=256, variant=2)
        assert np.array_equal(xu, np.ones(n))
    I.free_device_analysis(up, uinfo)
    # captured replays with eager work in between
    t = I.generate(3, 60000)
    m3 = I.allocate_memory(t.ia, t.ja, t.a)
    i3 = I.multirow_analysis_base_GPU(m3, I.ALL)
    db = I.DeviceArray.from_numpy(t.b)
    dx = I.DeviceArray(t.n, np.float64)
    for prec in (I.PREC_F64, I.PREC_MIXED):
        cfg = I._cfg(I.KERNEL_MULTIROW, 32, 1024, 1, 2, 8, reorder=1)
        I.solve(m3, i3, prec, db, dx, cfg=cfg)
        I.synchronize()
        eager = dx.numpy()
        for _ in range(3):
            dx.fill_bytes(0)
            I.solve_graph(m3, i3, prec, db, dx, cfg)
            I.synchronize()
            assert I.check_watchdog(m3) == 0
            assert np.array_equal(bits(dx.numpy()), bits(eager)), prec
            I.solve(m3, i3, I.PREC_F64, db, dx, cfg=I._cfg(I.KERNEL_ORDER, 16, 256, 1, 1))
            I.synchronize()
        assert relerr(eager, O.fwd_f64(t.ia, t.ja, t.a, t.b)) <= 1e-10
    db.free()
    dx.free()
    I.free_device_analysis(m3, i3)
    I.free_device_analysis(mat, info)


def test_copy_follows_new_values_and_analyses():
    print("level-ordered copy: new values, another analysis", flush=True)
    t = I.generate(2, 50000)
    mat = I.allocate_memory(t.ia, t.ja, t.a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    shape = dict(kernel=I.KERNEL_MULTIROW, lanes=32, wg=256, variant=2)
    x = same_as_original(mat, info, I.PREC_F64, t.b, **shape)
    assert relerr(x, O.fwd_f64(t.ia, t.ja, t.a, t.b)) <= 1e-12
    # a captured solve on the copy, then new values: the capture (which holds
    # the copy's arrays) is dropped with the copy and made again
    db = I.DeviceArray.from_numpy(t.b)
    dx = I.DeviceArray(t.n, np.float64)
    gcfg = I._cfg(**shape, reorder=1)
    I.solve_graph(mat, info, I.PREC_MIXED, db, dx, gcfg)
    I.synchronize()
    assert relerr(dx.numpy(), O.fwd_f64(t.ia, t.ja, t.a, t.b)) <= 1e-10
    a2 = t.a.copy()
    off = np.ones(len(a2), bool)
    off[t.ia[1:] - 1] = False
    a2[off] *= -0.5
    I.update_values(mat, a2)
    dx.fill_bytes(0)
    I.solve_graph(mat, info, I.PREC_MIXED, db, dx, gcfg)
    I.synchronize()
    assert I.check_watchdog(mat) == 0
    assert relerr(dx.numpy(), O.fwd_f64(t.ia, t.ja, a2, t.b)) <= 1e-10
    db.free()
    dx.free()
    x2 = same_as_original(mat, info, I.PREC_F64, t.b, **shape)
    assert relerr(x2, O.fwd_f64(t.ia, t.ja, a2, t.b)) <= 1e-12
    # an ORDER analysis of the same matrix: the copy is made again for it
    oinfo = I.multirow_analysis_base_GPU(mat, I.ORDER)
    x3 = same_as_original(mat, oinfo, I.PREC_MIXED, t.b, kernel=I.KERNEL_ORDER, lanes=8, wg=256, variant=1)
    assert relerr(x3, O.fwd_f64(t.ia, t.ja, a2, t.b)) <= 1e-10
    x4 = same_as_original(mat, info, I.PREC_F64, t.b, **shape)
    assert np.array_equal(bits(x4), bits(x2))
    I.free_device_analysis(None, oinfo)
    I.free_device_analysis(mat, info)


def test_copy_refusals():
    print("level-ordered copy refusals", flush=True)
    s = I.generate(2, 20000)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    db = I.DeviceArray.from_numpy(s.b)
    dx = I.DeviceArray(s.n, np.float64)
    bad = [(I.PREC_F64, I._cfg(I.KERNEL_SIMPLE, 16, 256, 1, 1, reorder=1)),
           (I.PREC_MIXED, I._cfg(I.KERNEL_ORDER, 16, 256, 0, 1, reorder=1)),      # unfused
           (I.PREC_F64, I._cfg(I.KERNEL_ORDER, 16, 256, 1, 3, reorder=1)),        # in-degree walk
           (I.PREC_F64, I._cfg(I.KERNEL_ORDER, 16, 256, 1, 1, reorder=2))]
    for prec, cfg in bad:
        with pytest.raises(L.SptrsvError) as e:
            I.solve(mat, info, prec, db, dx, cfg=cfg)
        assert e.value.status == L.E_INVALID_ARG, (prec, cfg.kernel, cfg.variant, cfg.reorder)
    d32 = I.DeviceArray.from_numpy(s.b.astype(np.float32))
    x32 = I.DeviceArray(s.n, np.float32)
    with pytest.raises(L.SptrsvError) as e:
        I.solve(mat, info, I.PREC_F32, d32, x32, cfg=I._cfg(I.KERNEL_ORDER, 16, 256, 1, 1, reorder=1))
    assert e.value.status == L.E_INVALID_ARG
    with pytest.raises(L.SptrsvError) as e:
        I.solve_mixed_noalloc(mat, info, db, dx, k=1, cfg=I._cfg(I.KERNEL_ORDER, 16, 256, 1, 1, reorder=1))
    assert e.value.status == L.E_INVALID_ARG
    for a in (db, dx, d32, x32):
        a.free()
    I.free_device_analysis(mat, info)


def test_tuned_copy_serves_auto_solves():
    print("level-ordered copy: autotune, then automatic solves (fused and unfused)", flush=True)
    s = I.generate(3, 200000)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    for prec, tol in ((I.PREC_F64, 1e-12), (I.PREC_MIXED, 1e-10)):
        cfg, ms = I.autotune(mat, info, prec, reps=2)
        print(f"  tuned precision {prec}: kernel {cfg.kernel} variant {cfg.variant} reorder {cfg.reorder}", flush=True)
        assert ms > 0 and cfg.reorder in (0, 1)
        fused = (1, 0) if prec == I.PREC_MIXED else (1,)
        for f in fused:                          # an unfused call drops the copy, never fails on it
            x = solve(mat, info, prec, s.b, kernel=I.KERNEL_AUTO, fused=f)
            assert relerr(x, ref) <= tol, (prec, f)
    I.free_device_analysis(mat, info)


@pytest.mark.parametrize("rows", [[[]], [[], [0]], [[], [], [0, 1], [2], []],
                                  [[], [0], [1], [2], [3], [0, 2, 4]]])
def test_copy_on_tiny_matrices(rows):
    print(f"level-ordered copy, {len(rows)} rows", flush=True)
    ia, ja, a = lower_from_rows(rows)
    n = len(rows)
    b = np.linspace(1.0, 2.0, n)
    mat = I.allocate_memory(ia, ja, a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    ref = O.fwd_f64(ia, ja, a, b)
    for kern in (I.KERNEL_ORDER, I.KERNEL_MULTIROW):
        for var in (0, 1, 2):
            for prec, tol in ((I.PREC_F64, 1e-14), (I.PREC_MIXED, 1e-12)):
                x = same_as_original(mat, info, prec, b, kernel=kern, lanes=4, wg=256, variant=var)
                assert relerr(x, ref) <= tol, (kern, var, prec)
    I.free_device_analysis(mat, info)
