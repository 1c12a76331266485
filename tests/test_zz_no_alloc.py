"""GPU test of sptrsv_solve_mixed_noalloc, the allocation-free variant of the
fused mixed-precision solve (its published words live in the library's fixed
per-device scratch array instead of the matrix's workspace).

Each call is compared with sptrsv_solve on the same matrix, right-hand side
and configuration: the two share their kernels, so the results must be
bit-identical, and both must be within 1e-10 (relative, max-norm) of the
oracle's serial FP64 substitution.  Two calls in a row, the second on a
smaller system, check that the shared scratch carries nothing from one call
to the next.
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


def both(mat, info, b, k, cfg):
    """(x from sptrsv_solve, x from sptrsv_solve_mixed_noalloc)."""
    n = mat.nr
    db = I.DeviceArray.from_numpy(np.asarray(b, np.float64))
    xa = I.DeviceArray(n * k, np.float64)
    xb = I.DeviceArray(n * k, np.float64)
    I.solve(mat, info, I.PREC_MIXED, db, xa, k=k, refine_steps=1, cfg=cfg)
    assert I.check_watchdog(mat) == 0, L.last_error()
    I.solve_mixed_noalloc(mat, info, db, xb, k=k, cfg=cfg)
    assert I.check_watchdog(mat) == 0, L.last_error()
    out = xa.numpy(), xb.numpy()
    for d in (db, xa, xb):
        d.free()
    return out


CFGS = {
    "simple-lean": dict(kernel=I.KERNEL_SIMPLE, lanes=16, wg=256, variant=1),
    "simple-packed": dict(kernel=I.KERNEL_SIMPLE, lanes=8, wg=256, variant=0),
    "order-lean": dict(kernel=I.KERNEL_ORDER, lanes=16, wg=64, variant=1),
    "multirow": dict(kernel=I.KERNEL_MULTIROW, lanes=0, wg=256, variant=0),
    "order-persistent": dict(kernel=I.KERNEL_ORDER, lanes=16, wg=64, variant=2, cu_waves=4),
}


@pytest.mark.parametrize("name", list(CFGS))
def test_noalloc_equals_sptrsv_solve_on_two_calls(name):
    print(f"no-alloc {name}: random n=200000 then 2D Laplacian n=4096", flush=True)
    cfg = I._cfg(fused=1, **CFGS[name])
    # first call: random lower system (the 1e6-row generator at 2e5 rows)
    s = I.generate(2, 200000)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    xa, xb = both(mat, info, s.b, 1, cfg)
    assert np.array_equal(xa.view(np.uint64), xb.view(np.uint64))
    assert relerr(xb, O.fwd_f64(s.ia, s.ja, s.a, s.b)) <= 1e-10
    # second call, smaller system on the same scratch: the reference's
    # known answer x == 1 exactly
    t = I.generate(1)
    mat2 = I.allocate_memory(t.ia, t.ja, t.a)
    info2 = I.multirow_analysis_base_GPU(mat2, I.ALL)
    ya, yb = both(mat2, info2, t.b, 1, cfg)
    assert np.array_equal(ya.view(np.uint64), yb.view(np.uint64))
    assert O.validate_x(yb) == 0
    I.free_device_analysis(mat, info)
    I.free_device_analysis(mat2, info2)


def test_noalloc_several_rhs():
    print("no-alloc k=4 then k=2 on the 3D Laplacian side 24", flush=True)
    s = I.generate(4, 24)
    k = s.k
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    for kk in (4, 2):
        b = s.b.reshape(s.n, k)[:, :kk].copy()
        cfg = I._cfg(kernel=I.KERNEL_ORDER, fused=1)
        xa, xb = both(mat, info, b.ravel(), kk, cfg)
        assert np.array_equal(xa.view(np.uint64), xb.view(np.uint64))
        ref = O.fwd_f64(s.ia, s.ja, s.a, b.ravel(), kk)
        assert relerr(xb, ref) <= 1e-10
    I.free_device_analysis(mat, info)


def test_noalloc_after_tuning():
    """After autotuning, an automatic call takes the tuned shape; a tuned
    level-ordered copy (a per-matrix workspace) is walked as the matrix as
    given, with the same bits, and a tuned JACOBI choice becomes the default
    sync-free shape."""
    print("no-alloc after autotune: 16 random batch members as one matrix", flush=True)
    mem = [I.generate_batch_member(m, 2000, 200000, seed=5) for m in range(1, 32, 2)]
    n = sum(s.n for s in mem)
    ia = np.zeros(n + 1, np.int64)
    ja, a, b = [], [], []
    r0 = e0 = 0
    for s in mem:
        ia[r0 + 1:r0 + s.n + 1] = s.ia[1:].astype(np.int64) - s.ia[0] + e0
        ja.append(s.ja + r0)
        a.append(s.a)
        b.append(s.b)
        r0 += s.n
        e0 += s.nnz
    ia, ja, a, b = ia.astype(np.int32), np.concatenate(ja), np.concatenate(a), np.concatenate(b)
    assert n <= L.NOALLOC_MAX_WORDS
    mat = I.allocate_memory(ia, ja, a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    cfg, _ = I.autotune(mat, info, I.PREC_MIXED, reps=2)
    print(f"  tuned kernel {cfg.kernel} variant {cfg.variant} reorder {cfg.reorder}", flush=True)
    ref = O.fwd_f64(ia, ja, a, b)
    for _ in range(2):
        xa, xb = both(mat, info, b, 1, None)
        if cfg.kernel != I.KERNEL_JACOBI:    # same sync-free shape: same bits
            assert np.array_equal(xa.view(np.uint64), xb.view(np.uint64))
        assert relerr(xb, ref) <= 1e-10      # (a tuned JACOBI becomes the default sync-free shape)
    I.free_device_analysis(mat, info)


def test_noalloc_rejects_what_it_cannot_do():
    print("no-alloc argument checks", flush=True)
    n = L.NOALLOC_MAX_WORDS // 64 + 1
    ia = np.arange(n + 1, dtype=np.int32)
    ja = np.arange(n, dtype=np.int32)
    mat = I.allocate_memory(ia, ja, np.full(n, 2.0))
    db = I.DeviceArray(n * 64, np.float64)
    dx = I.DeviceArray(n * 64, np.float64)
    with pytest.raises(L.SptrsvError) as e:  # n * k over the scratch
        I.solve_mixed_noalloc(mat, None, db, dx, k=64)
    assert e.value.status == L.E_CAPACITY
    with pytest.raises(L.SptrsvError) as e:  # JACOBI keeps per-matrix workspaces
        I.solve_mixed_noalloc(mat, None, db, dx, k=1, cfg=I._cfg(kernel=I.KERNEL_JACOBI, sweeps=4, outer=1))
    assert e.value.status == L.E_INVALID_ARG
    up = I.allocate_memory_upper(ia, ja, np.full(n, 2.0))
    with pytest.raises(L.SptrsvError) as e:  # reversed storage needs reversal buffers
        I.solve_mixed_noalloc(up, None, db, dx, k=1)
    assert e.value.status == L.E_INVALID_ARG
    # the checks ran before any kernel: a valid call on the same handle works
    b = np.full(n, 2.0)
    db1 = I.DeviceArray.from_numpy(b)
    dx1 = I.DeviceArray(n, np.float64)
    I.solve_mixed_noalloc(mat, None, db1, dx1, k=1, cfg=I._cfg(kernel=I.KERNEL_SIMPLE))
    assert I.check_watchdog(mat) == 0
    assert np.all(dx1.numpy() == 1.0)
    for d in (db, dx, db1, dx1):
        d.free()
    I.free_device_analysis(mat, None)
    I.free_device_analysis(up, None)
