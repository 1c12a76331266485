"""GPU tests of the in-degree walk (variant 3): every row waits on a counter
of its unsolved dependencies, decremented by each solved dependency through
the column pattern of L, then folds its entries in storage order exactly as
the lean walk (variant 1) does.

Bars: bit-identical to the lean walk with the same lanes, workgroup and
window, in FP64, FP32 and fused mixed precision, for natural and level order;
within 1e-12 (FP64) / 1e-10 (mixed) of the oracle; exact x == 1 on the 2D
Laplacian factor's known answer; captured solves replayed bit-identical;
counters rebuilt per solve (repeat solves and new values); and the refusals
(multirow, several right-hand sides, the allocation-free solve).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as O
from sptrsv_amd import _lib as L
from sptrsv_amd import interface as I

pytestmark = pytest.mark.gpu


def relerr(x, ref):
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


def solve(mat, info, prec, b, **cfg):
    dt = np.float32 if prec == I.PREC_F32 else np.float64
    db = I.DeviceArray.from_numpy(b.astype(dt))
    dx = I.DeviceArray(len(b), dt)
    I.solve(mat, info, prec, db, dx, cfg=I._cfg(**cfg))
    assert I.check_watchdog(mat) == 0, L.last_error()
    x = dx.numpy()
    db.free()
    dx.free()
    return x


def bits(x):
    return x.view(np.uint64 if x.dtype == np.float64 else np.uint32)


@pytest.mark.parametrize("gen,param", [(2, 100000), (3, 100000)])
@pytest.mark.parametrize("kernel", [I.KERNEL_SIMPLE, I.KERNEL_ORDER])
def test_indeg_walk_equals_the_lean_walk(gen, param, kernel):
    print(f"in-degree walk, generator {gen}, kernel {kernel}", flush=True)
    s = I.generate(gen, param)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    for lanes, wg, window in ((16, 256, 0), (32, 1024, 4), (8, 64, 1)):
        for prec, tol in ((I.PREC_F64, 1e-12), (I.PREC_F32, 1e-4), (I.PREC_MIXED, 1e-10)):
            shape = dict(kernel=kernel, lanes=lanes, wg=wg, window=window)
            x3 = solve(mat, info, prec, s.b, variant=3, **shape)
            x1 = solve(mat, info, prec, s.b, variant=1, **shape)
            assert np.array_equal(bits(x3), bits(x1)), (lanes, wg, window, prec)
            assert relerr(x3.astype(np.float64), ref) <= tol, (lanes, wg, window, prec)
    I.free_device_analysis(mat, info)


def test_indeg_walk_known_answer_and_captured_replays():
    print("in-degree walk: x == 1 and captured replays", flush=True)
    s = I.generate(1)                        # 2D Laplacian factor, b = L * ones
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    x = solve(mat, None, I.PREC_F64, s.b, kernel=I.KERNEL_SIMPLE, lanes=4, wg=256, variant=3)
    assert np.array_equal(x, np.ones(s.n))
    t = I.generate(3, 60000)
    m3 = I.allocate_memory(t.ia, t.ja, t.a)
    info = I.multirow_analysis_base_GPU(m3, I.ORDER)
    cfg = I._cfg(I.KERNEL_ORDER, 32, 256, 1, 3, 0, window=4)
    db = I.DeviceArray.from_numpy(t.b)
    dx = I.DeviceArray(t.n, np.float64)
    I.solve(m3, info, I.PREC_F64, db, dx, cfg=cfg)
    I.synchronize()
    eager = dx.numpy()
    for _ in range(3):                       # replays, with eager work in between
        dx.fill_bytes(0)
        I.solve_graph(m3, info, I.PREC_F64, db, dx, cfg)
        I.synchronize()
        assert I.check_watchdog(m3) == 0
        assert np.array_equal(bits(dx.numpy()), bits(eager))
        I.solve(m3, info, I.PREC_MIXED, db, dx, cfg=cfg)
        I.synchronize()
    assert relerr(eager, O.fwd_f64(t.ia, t.ja, t.a, t.b)) <= 1e-12
    # new values on the same pattern: the counters are structure only
    a2 = t.a.copy()
    diag = t.ia[1:] - 1
    off = np.ones(len(a2), bool)
    off[diag] = False
    a2[off] *= 0.5
    I.update_values(m3, a2)
    x2 = solve(m3, info, I.PREC_F64, t.b, kernel=I.KERNEL_ORDER, lanes=32, wg=256, variant=3)
    assert relerr(x2, O.fwd_f64(t.ia, t.ja, a2, t.b)) <= 1e-12
    db.free()
    dx.free()
    I.free_device_analysis(m3, info)
    I.free_device_analysis(mat, None)


def test_indeg_walk_refusals():
    print("in-degree walk refusals", flush=True)
    s = I.generate(2, 20000)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    db = I.DeviceArray.from_numpy(np.tile(s.b, 2))
    dx = I.DeviceArray(2 * s.n, np.float64)
    with pytest.raises(L.SptrsvError) as e:     # multirow waves have no in-degree form
        I.solve(mat, info, I.PREC_F64, db, dx, cfg=I._cfg(I.KERNEL_MULTIROW, 32, 256, 1, 3))
    assert e.value.status == L.E_INVALID_ARG
    with pytest.raises(L.SptrsvError) as e:     # one right-hand side only
        I.solve(mat, info, I.PREC_F64, db, dx, k=2, cfg=I._cfg(I.KERNEL_SIMPLE, 16, 256, 1, 3))
    assert e.value.status == L.E_INVALID_ARG
    with pytest.raises(L.SptrsvError) as e:     # its counters are a workspace
        I.solve_mixed_noalloc(mat, info, db, dx, k=1, cfg=I._cfg(I.KERNEL_SIMPLE, 16, 256, 1, 3))
    assert e.value.status == L.E_INVALID_ARG
    db.free()
    dx.free()
    I.free_device_analysis(mat, info)
