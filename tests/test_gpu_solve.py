"""GPU parity tests: the HIP library, called through the C-ABI, against the
CPU oracle on the same inputs.

Bars: levels / iorder / warp arrays bit-exact; FP64 solves within 1e-12
relative (max-norm) of the oracle's serial forward substitution; FP32
solves within 1e-5 on diagonally dominant systems; mixed precision (FP32
bulk + FP64 refinement) within 1e-10; exact x == 1 on the reference's
known-answer inputs.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as O
from conftest import lower_from_rows
from sptrsv_amd import _lib as L
from sptrsv_amd import interface as I

pytestmark = pytest.mark.gpu

KERNELS = {"simple": I.KERNEL_SIMPLE, "order": I.KERNEL_ORDER, "multirow": I.KERNEL_MULTIROW}


def relerr(x, ref):
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


def random_rows(n, dens, rng, maxrow=40):
    rows = []
    for i in range(n):
        k = min(i, int(rng.integers(0, dens + 1)) if rng.random() > 0.05 else int(rng.integers(0, maxrow)))
        rows.append(sorted(rng.choice(i, size=k, replace=False).tolist()) if k else [])
    return rows


class Dev:
    """A matrix on the device with its ALL analysis."""

    def __init__(self, ia, ja, a, analyse=True):
        self.ia, self.ja, self.a = ia, ja, a
        self.n = len(ia) - 1
        self.mat = I.allocate_memory(ia, ja, a)
        self.info = I.multirow_analysis_base_GPU(self.mat, I.ALL) if analyse else None

    def solve(self, b, precision=I.PREC_F64, kernel=I.KERNEL_SIMPLE, k=1, lanes=0, wg=0, fused=1, refine=1,
              variant=0, cu_waves=0, sweeps=0, outer=0, expect=0, window=0):
        dt = np.float32 if precision == I.PREC_F32 else np.float64
        db = I.DeviceArray.from_numpy(np.asarray(b, dt))
        dx = I.DeviceArray(self.n * k, dt)
        I.solve(self.mat, self.info, precision, db, dx, k=k, refine_steps=refine, kernel=kernel, lanes=lanes,
                wg=wg, fused=fused, variant=variant, cu_waves=cu_waves, sweeps=sweeps, outer=outer,
                window=window)
        if expect is not None:  # None: the caller reads (and clears) the status itself
            assert I.check_watchdog(self.mat) == expect, L.last_error()
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


# ---- the reference's known-answer check -------------------------------------------

def test_known_answer_laplacian_2d(device):
    s = I.generate(1)
    d = Dev(s.ia, s.ja, s.a)
    for name, kern in KERNELS.items():
        x = d.solve(s.b, kernel=kern)
        assert O.validate_x(x) == 0, name
        x32 = d.solve(s.b.astype(np.float32), precision=I.PREC_F32, kernel=kern)
        assert np.all(x32 == 1.0), name
        xm = d.solve(s.b, precision=I.PREC_MIXED, kernel=kern)
        assert O.validate_x(xm) == 0, name
    d.close()


def test_run_solver_host_interface(device):
    s = I.generate(1)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    assert O.validate_x(I.run_solver(mat, None, s.b)) == 0          # SIMPLE, no analysis
    for mode in (I.ORDER, I.MULTIROW, I.ALL):
        info = I.AnalysisInfo()
        info.mode = mode
        assert I.run_analysis(info, mat) == 0
        assert O.validate_x(I.run_solver(mat, info, s.b)) == 0
        I.free_device_analysis(None, info)
    info = I.AnalysisInfo()
    info.mode = I.SIMPLE
    assert I.run_analysis(info, mat) == L.E_NO_ANALYSIS_NEEDED
    I.free_device_analysis(mat, None)


# ---- analysis: bit-exact ---------------------------------------------------------

def check_analysis(ia, ja, info):
    got = I.analysis_arrays(info)
    ref32 = O.analysis(ia, ja, 32)
    ref64 = O.analysis(ia, ja, 64)
    assert got["nlevs"] == ref32["nlevs"]
    assert np.array_equal(got["level"], ref32["level"])
    assert np.array_equal(got["iorder"], ref32["iorder"])
    assert np.array_equal(got["bucket_off"], ref32["bucket_off"])
    assert got["n_warps"] == ref32["n_warps"]
    assert np.array_equal(got["ibase_row"], ref32["ibase_row"])
    assert np.array_equal(got["ivect_size_warp"], ref32["ivect_size_warp"])
    assert got["n_waves"] == ref64["n_warps"]
    assert np.array_equal(got["ibase_row64"], ref64["ibase_row"])
    assert np.array_equal(got["ivect_size_wave"], ref64["ivect_size_warp"])
    n = len(ia) - 1
    assert np.array_equal(got["inv_iorder"][got["iorder"]], np.arange(n))
    assert np.array_equal(got["lev_size"], np.bincount(ref32["level"] - 1, minlength=ref32["nlevs"]))
    assert np.array_equal(got["warp_lev"], ref32["warp_off"][0::7])


def test_analysis_hand_fixture(device, golden):
    h = golden["hand"]["analysis8"]
    for rows in (h["offdiag"], h["offdiag"] + [[]] * 40):
        ia, ja, a = lower_from_rows(rows)
        d = Dev(ia, ja, a)
        check_analysis(ia, ja, d.info)
        d.close()


@pytest.mark.parametrize("case", [0, 1, 2])
def test_analysis_golden_random(device, golden, case):
    r = golden["random"]
    ia, ja, a = r[f"c{case}_ia"], r[f"c{case}_ja"], r[f"c{case}_a"]
    d = Dev(ia, ja, a)
    got = I.analysis_arrays(d.info)
    assert np.array_equal(got["level"], r[f"c{case}_level"])
    assert np.array_equal(got["iorder"], r[f"c{case}_iorder"])
    check_analysis(ia, ja, d.info)
    d.close()


@pytest.mark.parametrize("cfg,param", [(1, 64), (2, 30000), (3, 30000), (4, 24)])
def test_analysis_configs(device, cfg, param):
    s = I.generate(cfg, param)
    d = Dev(s.ia, s.ja, s.a)
    check_analysis(s.ia, s.ja, d.info)
    d.close()


def test_analysis_order_mode_has_no_warps(device):
    s = I.generate(1, 32)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ORDER)
    got = I.analysis_arrays(info)
    assert info.n_warps == 0 and "ibase_row" not in got
    assert np.array_equal(got["iorder"], O.analysis(s.ia, s.ja)["iorder"])
    I.free_device_analysis(mat, info)


# ---- solves vs the oracle ------------------------------------------------------------

@pytest.mark.parametrize("case", [0, 1, 2])
def test_fp64_all_kernels_golden(device, golden, case):
    r = golden["random"]
    ia, ja, a, b = r[f"c{case}_ia"], r[f"c{case}_ja"], r[f"c{case}_a"], r[f"c{case}_b"]
    ref = O.fwd_f64(ia, ja, a, b)
    assert relerr(ref, r[f"c{case}_x"]) <= 1e-12
    d = Dev(ia, ja, a)
    for name, kern in KERNELS.items():
        for lanes in ([0] if kern == I.KERNEL_MULTIROW else [1, 2, 8, 64]):
            for wg in ([256, 1024] if kern == I.KERNEL_SIMPLE else [256]):
                for variant in (0, 1):
                    ln = lanes if lanes or variant == 0 else 4
                    x = d.solve(b, kernel=kern, lanes=ln, wg=wg, variant=variant)
                    assert relerr(x, ref) <= 1e-12, (name, lanes, wg, variant)
                    xm = d.solve(b, precision=I.PREC_MIXED, kernel=kern, lanes=ln, wg=wg, variant=variant)
                    assert relerr(xm, ref) <= 1e-10, (name, lanes, wg, variant)
    d.close()


def test_residency_caps_and_small_workgroups(device):
    s = I.generate(2, 100000)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    d = Dev(s.ia, s.ja, s.a)
    for name, kern in KERNELS.items():
        for wg in (64, 256):
            for cw in (1, 2, 8):
                for variant in (0, 1, 2):
                    x = d.solve(s.b, kernel=kern, lanes=16, wg=wg, variant=variant, cu_waves=cw)
                    assert relerr(x, ref) <= 1e-12, (name, wg, cw, variant)
                    xm = d.solve(s.b, precision=I.PREC_MIXED, kernel=kern, lanes=16, wg=wg, variant=variant,
                                 cu_waves=cw)
                    assert relerr(xm, ref) <= 1e-10, (name, wg, cw, variant)
    d.close()


def test_persistent_edge_cases(device):
    # more waves than rows, a single row, long chains, long rows
    rng = np.random.default_rng(2)
    cases = {"one_row": [[]], "chain": [[i - 1] if i else [] for i in range(3000)],
             "long_row": [[]] * 2000 + [list(range(2000))], "diag": [[]] * 100}
    for name, rows in cases.items():
        ia, ja, a = lower_from_rows(rows, rng)
        b = rng.uniform(-1, 1, len(rows))
        ref = O.fwd_f64(ia, ja, a, b)
        d = Dev(ia, ja, a)
        for kern in KERNELS.values():
            for cw in (1, 16):
                assert relerr(d.solve(b, kernel=kern, lanes=8, wg=64, variant=2, cu_waves=cw), ref) <= 1e-12, name
        d.close()


def test_fast_and_general_paths_agree_bitwise(device):
    # same per-lane folding order and the same reduction tree for equal lanes
    s = I.generate(3, 60000)
    d = Dev(s.ia, s.ja, s.a)
    for lanes in (4, 16, 64):
        x0 = d.solve(s.b, kernel=I.KERNEL_ORDER, lanes=lanes, variant=0)
        x1 = d.solve(s.b, kernel=I.KERNEL_ORDER, lanes=lanes, variant=1)
        x2 = d.solve(s.b, kernel=I.KERNEL_SIMPLE, lanes=lanes, variant=1)
        assert np.array_equal(x0, x1) and np.array_equal(x1, x2), lanes
    d.close()


@pytest.mark.parametrize("cfg,param", [(2, 200000), (3, 200000)])
def test_fp64_fp32_mixed_configs(device, cfg, param):
    s = I.generate(cfg, param)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    d = Dev(s.ia, s.ja, s.a)
    for name, kern in KERNELS.items():
        for variant in (0, 1):
            lanes = 0 if variant == 0 else 16
            assert relerr(d.solve(s.b, kernel=kern, lanes=lanes, variant=variant), ref) <= 1e-12, name
            x32 = d.solve(s.b.astype(np.float32), precision=I.PREC_F32, kernel=kern, lanes=lanes,
                          variant=variant)
            assert relerr(x32.astype(np.float64), ref) <= 1e-5, name
            for fused in (1, 0):
                xm = d.solve(s.b, precision=I.PREC_MIXED, kernel=kern, fused=fused, lanes=lanes,
                             variant=variant)
                assert relerr(xm, ref) <= 1e-10, (name, fused, variant)
    d.close()


def test_long_row_window(device):
    """16 entries per lane in flight (power-law rows up to 2000 entries)."""
    s = I.generate(3, 100000)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    d = Dev(s.ia, s.ja, s.a)
    for name, kern in KERNELS.items():
        for variant in (0, 1, 2):
            x = d.solve(s.b, kernel=kern, lanes=32 if variant else 0, wg=256, variant=variant, window=16)
            assert relerr(x, ref) <= 1e-12, (name, variant)
            xm = d.solve(s.b, precision=I.PREC_MIXED, kernel=kern, lanes=32 if variant else 0, wg=256,
                         variant=variant, window=16)
            assert relerr(xm, ref) <= 1e-10, (name, variant)
    d.close()


def test_mixed_refine_steps(device):
    s = I.generate(2, 100000)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    d = Dev(s.ia, s.ja, s.a)
    errs = [relerr(d.solve(s.b, precision=I.PREC_MIXED, refine=k, fused=0), ref) for k in (1, 2, 3)]
    assert errs[0] <= 1e-10 and errs[1] <= errs[0] * 1.01 + 1e-15 and errs[2] <= 1e-13
    xf = d.solve(s.b, precision=I.PREC_MIXED, refine=2, fused=1)
    assert relerr(xf, ref) <= 1e-13
    d.close()


def test_multi_rhs(device):
    s = I.generate(4, 24)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b, k=16)
    d = Dev(s.ia, s.ja, s.a)
    for name, kern in KERNELS.items():
        assert relerr(d.solve(s.b, kernel=kern, k=16), ref) <= 1e-12, name
        assert relerr(d.solve(s.b, precision=I.PREC_MIXED, kernel=kern, k=16), ref) <= 1e-10, name
    for k in (1, 3, 5, 64):
        rng = np.random.default_rng(k)
        B = rng.uniform(-1, 1, s.n * k)
        refk = O.fwd_f64(s.ia, s.ja, s.a, B, k=k)
        assert relerr(d.solve(B, k=k), refk) <= 1e-12, k
        assert relerr(d.solve(B, precision=I.PREC_MIXED, k=k, fused=0), refk) <= 1e-10, k
    d.close()


def test_multi_rhs_persistent_walk(device):
    # several right-hand sides on the persistent walk (variant 2): natural
    # order, level order and multirow waves, with and without residency caps
    for cfg, param in ((4, 24), (2, 60000), (3, 60000)):
        s = I.generate(cfg, param)
        d = Dev(s.ia, s.ja, s.a)
        for k in (3, 16):
            B = np.random.default_rng(k + cfg).uniform(-1, 1, s.n * k)
            ref = O.fwd_f64(s.ia, s.ja, s.a, B, k=k)
            for name, kern in KERNELS.items():
                for lanes in ((1, 2) if kern != I.KERNEL_MULTIROW else (0,)):
                    for cw in (0, 8):
                        print(cfg, k, name, lanes, cw, flush=True)
                        x = d.solve(B, kernel=kern, k=k, lanes=lanes, wg=256, variant=2, cu_waves=cw)
                        assert relerr(x, ref) <= 1e-12, (cfg, k, name, lanes, cw)
                x = d.solve(B, precision=I.PREC_MIXED, kernel=kern, k=k, wg=256, variant=2, cu_waves=4)
                assert relerr(x, ref) <= 1e-10, (cfg, k, name, "mixed")
        d.close()


def test_bit_reproducible(device):
    s = I.generate(3, 100000)
    d = Dev(s.ia, s.ja, s.a)
    for kern in KERNELS.values():
        x1 = d.solve(s.b, kernel=kern)
        x2 = d.solve(s.b, kernel=kern)
        assert np.array_equal(x1, x2)
    d.close()


# ---- edge cases ------------------------------------------------------------------

def test_edge_cases(device):
    rng = np.random.default_rng(1)
    cases = {
        "one_row": [[]],
        "diagonal_only": [[]] * 1000,
        "bidiagonal_chain": [[i - 1] if i else [] for i in range(5000)],
        "dense_lower_64": [list(range(i)) for i in range(64)],
        "one_long_row": [[]] * 3000 + [list(range(3000))],
        "star": [[]] + [[0]] * 2000,
    }
    for name, rows in cases.items():
        ia, ja, a = lower_from_rows(rows, rng)
        b = rng.uniform(-1, 1, len(rows))
        ref = O.fwd_f64(ia, ja, a, b)
        d = Dev(ia, ja, a)
        check_analysis(ia, ja, d.info)
        for kname, kern in KERNELS.items():
            assert relerr(d.solve(b, kernel=kern), ref) <= 1e-12, (name, kname)
            assert relerr(d.solve(b, precision=I.PREC_MIXED, kernel=kern), ref) <= 1e-10, (name, kname)
        d.close()


def test_empty_matrix(device):
    mat = I.allocate_memory(np.array([0], np.int32), np.zeros(0, np.int32), np.zeros(0))
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    assert info.nlevs == 0
    I.free_device_analysis(mat, info)


def test_unsorted_columns_and_nonzero_base(device):
    # columns need not be sorted (main.cu keeps file order); ia may start > 0
    rows = [[], [0], [1, 0], [2, 0, 1]]
    ia, ja, a = lower_from_rows(rows)
    b = np.array([1.0, 2.0, 3.0, 4.0])
    ref = O.fwd_f64(ia, ja, a, b)
    d = Dev(ia, ja, a)
    assert relerr(d.solve(b), ref) <= 1e-12
    d.close()
    pad_ja = np.concatenate([[7, 7], ja]).astype(np.int32)
    pad_a = np.concatenate([[9.0, 9.0], a])
    mat = I.allocate_memory(ia + 2, pad_ja[0:], pad_a, n=4)
    assert relerr(I.run_solver(mat, None, b), ref) <= 1e-12
    I.free_device_analysis(mat, None)


def test_watchdog_reports_instead_of_hanging(device):
    # a tiny wait bound forced from the host: waits give up, rows become NaN,
    # the kernel ends and the status reports it; then a normal bound works again
    rows = [[i - 1] if i else [] for i in range(20000)]
    ia, ja, a = lower_from_rows(rows)
    b = np.ones(len(rows))
    d = Dev(ia, ja, a, analyse=False)
    I.set_watchdog_ms(d.mat, 1e-4)
    db = I.DeviceArray.from_numpy(b)
    dx = I.DeviceArray(len(rows), np.float64)
    I.csr_L_solve_simple(d.mat, db, dx)
    assert I.check_watchdog(d.mat) == L.E_WATCHDOG
    assert np.isnan(dx.numpy()).any()
    I.set_watchdog_ms(d.mat, 20000.0)
    I.csr_L_solve_simple(d.mat, db, dx)
    assert I.check_watchdog(d.mat) == 0
    assert relerr(dx.numpy(), O.fwd_f64(ia, ja, a, b)) <= 1e-12
    d.close()


def test_fp32_input_matrix(device):
    s = I.generate(2, 50000)
    mat = I.allocate_memory_f32(s.ia, s.ja, s.a.astype(np.float32))
    ref = O.fwd_f32(s.ia, s.ja, s.a.astype(np.float32), s.b.astype(np.float32))
    db = I.DeviceArray.from_numpy(s.b.astype(np.float32))
    dx = I.DeviceArray(s.n, np.float32)
    I.csr_L_solve_simple_f32(mat, db, dx)
    assert relerr(dx.numpy().astype(np.float64), ref.astype(np.float64)) <= 1e-5
    with pytest.raises(L.SptrsvError):
        I.csr_L_solve_simple(mat, db, dx)            # no FP64 values
    I.free_device_analysis(mat, None)


# ---- autotune, batch, vendor baseline ----------------------------------------------

def test_autotune_choice_is_correct(device):
    s = I.generate(2, 100000)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    d = Dev(s.ia, s.ja, s.a)
    for prec, tol in ((I.PREC_F64, 1e-12), (I.PREC_MIXED, 1e-10)):
        cfg, ms = I.autotune(d.mat, d.info, prec, reps=2)
        allowed = set(KERNELS.values()) | ({I.KERNEL_JACOBI} if prec == I.PREC_MIXED else set())
        assert ms > 0 and cfg.kernel in allowed
        x = d.solve(s.b, precision=prec, kernel=I.KERNEL_AUTO)
        assert relerr(x, ref) <= tol
    d.close()


def test_batch_graph_replays(device):
    systems = [I.generate_batch_member(m, 500, 6000, seed=5) for m in range(10)]
    bt = I.Batch(systems)
    B = np.concatenate([s.b for s in systems])
    refs = np.concatenate([O.fwd_f64(s.ia, s.ja, s.a, s.b) for s in systems])
    db = I.DeviceArray.from_numpy(B)
    dx = I.DeviceArray(bt.rows, np.float64)
    for prec, tol in ((I.PREC_F64, 1e-12), (I.PREC_MIXED, 1e-10)):
        bt.solve(prec, db, dx, graph=False)
        I.synchronize()
        assert relerr(dx.numpy(), refs) <= tol
        for rep in range(3):                     # replays with eager work in between
            dx.fill_bytes(0)
            bt.solve(prec, db, dx, graph=True)
            I.synchronize()
            assert relerr(dx.numpy(), refs) <= tol, rep
    bt.tune(I.PREC_MIXED, tune=True)
    bt.solve(I.PREC_MIXED, db, dx, graph=True)
    I.synchronize()
    assert relerr(dx.numpy(), refs) <= 1e-10
    assert all(bt.choice(m).lanes >= 1 for m in range(len(systems)))
    bt.close()


def test_batch_patterns_share_tuning(device):
    """Members on the same structure with other values: found as one
    structure, timed once, given the same lanes, and each solved right."""
    import types
    rng = np.random.default_rng(17)
    base = [I.generate_batch_member(m, 500, 6000, seed=5) for m in range(3)]

    def revalued(s):
        a = s.a.copy()
        diag = s.ia[1:] - 1
        off = np.ones(len(a), bool)
        off[diag] = False
        a[off] *= rng.uniform(0.5, 1.5, off.sum())
        rows = np.repeat(np.arange(s.n), np.diff(s.ia))
        a[diag] = 1.0 + np.bincount(rows[off], weights=np.abs(a[off]), minlength=s.n)
        return types.SimpleNamespace(n=s.n, ia=s.ia, ja=s.ja.copy(), a=a, b=rng.uniform(-1, 1, s.n))

    systems = [base[0], base[1], revalued(base[0]), revalued(base[1]), revalued(base[0]), base[2]]
    bt = I.Batch(systems)
    count, of = bt.patterns()
    assert count == 3 and of.tolist() == [0, 1, 0, 1, 0, 5]
    B = np.concatenate([s.b for s in systems])
    db = I.DeviceArray.from_numpy(B)
    dx = I.DeviceArray(bt.rows, np.float64)
    for prec, tol in ((I.PREC_F64, 1e-12), (I.PREC_MIXED, 1e-10)):
        bt.tune(prec, tune=True)
        lanes = [bt.choice(m).lanes for m in range(len(systems))]
        assert lanes[0] == lanes[2] == lanes[4] and lanes[1] == lanes[3], lanes
        bt.solve(prec, db, dx, graph=True)
        I.synchronize()
        assert bt.check() == 0
        x = dx.numpy()
        for m, s in enumerate(systems):
            lo, hi = bt.row_off[m], bt.row_off[m + 1]
            assert relerr(x[lo:hi], O.fwd_f64(s.ia, s.ja, s.a, s.b)) <= tol, (prec, m)
    bt.close()


def test_rocsparse_baseline(device):
    s = I.generate(2, 50000)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    try:
        rb = I.RocsparseBaseline(mat)
    except L.SptrsvError as e:
        pytest.skip(f"rocSPARSE unavailable: {e}")
    db = I.DeviceArray.from_numpy(s.b)
    dx = I.DeviceArray(s.n, np.float64)
    rb.solve(db, dx)
    I.synchronize()
    assert relerr(dx.numpy(), ref) <= 1e-12
    rb.close()
    # csritsv: FP64 Jacobi to |b - L x|_inf <= 1e-12 (delta = 1 for this matrix)
    ri = I.RocsparseBaseline(mat, itsv=True)
    iters = ri.solve_itsv(db, dx, 1e-12, max_iter=2000)
    I.synchronize()
    assert 20 <= iters < 2000 * 21 and (iters - 20) % 21 == 0  # (k + 1) * 20 + k iterations
    assert relerr(dx.numpy(), ref) <= 1e-10
    # testing after every iteration finds the first iteration that meets the
    # bound; testing every 21st finds it at most 20 iterations later
    every = ri.solve_itsv(db, dx, 1e-12, max_iter=20000, free_iter=0)
    I.synchronize()
    assert relerr(dx.numpy(), ref) <= 1e-10
    print(f"csritsv: {every} iterations when testing each, {iters} when testing every 21st", flush=True)
    assert every <= iters <= every + 21
    ri.close()
    db.free()
    dx.free()
    I.free_device_analysis(mat, None)


def test_bench_solve_and_power(device):
    s = I.generate(1)
    d = Dev(s.ia, s.ja, s.a)
    db = I.DeviceArray.from_numpy(s.b)
    dx = I.DeviceArray(s.n, np.float64)
    r = I.bench_solve(d.mat, d.info, I.PREC_F64, db, dx, reps=5, check_ones=True)
    assert r.reps == 5 and r.mean_ms > 0 and r.errors == 0
    assert r.power_mw < 0 or 0 < r.power_mw / 1000 < 1400
    d.close()


def test_power_sampling_around_a_solve_loop(device):
    """amdsmiAPIRun / amdsmiAPIEnd (nvmlAPIRun / nvmlAPIEnd, nvmlPower.cpp:65-158)
    around a loop of solves: at least one sample, and a mean board power
    between 0 and 1400 W (the MI355X board limit).  Each run initialises and
    each end shuts down amdsmi, so a second run works too; an end without a
    run reports no reading."""
    import time
    s = I.generate(2, 300000)
    d = Dev(s.ia, s.ja, s.a, analyse=False)
    db = I.DeviceArray.from_numpy(s.b)
    dx = I.DeviceArray(s.n, np.float64)
    for rnd in range(2):
        assert I.power_start("solve loop", 0, 1000), L.last_error()
        t0 = time.time()
        while time.time() - t0 < 0.3:
            I.solve(d.mat, None, I.PREC_F64, db, dx, kernel=I.KERNEL_SIMPLE)
            I.synchronize()
        mw = I.power_end()
        samples = L.lib().amdsmiAPISamples()
        print(f"power run {rnd}: {samples} samples, {mw / 1000:.1f} W", flush=True)
        assert samples >= 1
        assert 0 < mw / 1000 < 1400
    assert I.power_end() < 0
    assert I.check_watchdog(d.mat) == 0
    db.free()
    dx.free()
    d.close()
