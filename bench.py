"""Benchmark: sparse lower-triangular solve on one MI355X.

Workload (default): random lower-triangular CSR, n = 1,000,000, Poisson(8)
off-diagonals per row in [max(0,i-5000), i), values U(-1,1), diagonal
1 + sum|off|, one right-hand side U(-1,1) (seed fixed).  Inputs are resident
in HBM before timing.  One step = one full solve; the headline is the
mixed-precision method (FP32 bulk + FP64 residual/update refinement) in
GNNZ/s, with the plain FP64 sync-free solve measured beside it.  Both are
timed as replays of the solve captured in a hipGraph (one launch per solve).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--workload 2]

Prints ONE JSON line.  With --gpus > 1 it exits: this solver runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "GNNZ/s + HBM GB/s per solve; FP32+refine vs plain FP64 at rel err <=1e-10"

WORKLOADS = {
    2: ("random lower CSR n=1e6, Poisson(8) off-diagonals in [i-5000,i), U(-1,1), diag 1+sum|off|, 1 RHS", 1000000),
    3: ("power-law lower CSR n=2e6, Zipf(1.8) off-diagonals (cap 2000) in [0,i), diag 1+sum|off|, 1 RHS", 2000000),
    4: ("3D 7-point Laplacian lower factor 128^3 (n=2097152), unit diag, off -1/6, 16 RHS", 128),
    5: ("batch of 256 independent lower CSR systems, n~U[2000,200000]; even members banded (bandwidth U[1,32]), "
        "odd members Poisson(6) off-diagonals in [0,i); U(-1,1), diag 1+sum|off|, 1 RHS each", 256),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_f64(n, nnz, k):
    """Plain FP64 sync-free pass: ia + ja + a64 + b + x."""
    return 4 * (n + 1) + 12 * nnz + 16 * k * n


def bytes_mixed_fused(n, nnz, k, reorder=0):
    """Fused mixed pass: ia + ja + a32 + a64 + b (f64) + x (f64) + packed (x0,d) words
    (+ the caller's row of each row when a level-ordered copy is walked)."""
    return 4 * (n + 1) + 16 * nnz + 24 * k * n + (4 * n if reorder else 0)


def bytes_method_convention(n, nnz, k):
    """Separate-pass convention: FP32 solve + FP64 residual + FP32 solve + FP64 update."""
    f32 = 4 * (n + 1) + 8 * nnz + 8 * k * n
    res = 4 * (n + 1) + 12 * nnz + 24 * k * n
    upd = 20 * k * n
    return 2 * f32 + res + upd


def _jacobi_sweep_totals(n, nnz, k, sweeps, outer, sweeps_next=0):
    """(bytes, launches) of all sweep launches of a solve.

    A sweep reads ia, ja, a32 (strict lower part and diagonal), the previous
    iterate d (4kn, each entry once) and r32 (4kn).  Inner sweeps write d
    (4kn); the last sweep of a step writes x in FP64 (8kn) and, after the
    first step, also reads it (8kn).  Step 1 runs sweeps-1 launches, later
    steps sweeps_next-1 (sweeps_next 0: the same as step 1)."""
    per_step = [sweeps] + [sweeps_next or sweeps] * (outer - 1)
    base = 4 * (n + 1) + 8 * nnz + 8 * k * n
    inner = sum(s - 2 for s in per_step) * (base + 4 * k * n)
    last = (base + 8 * k * n) + (outer - 1) * (base + 16 * k * n)
    return inner + last, sum(s - 1 for s in per_step)


def bytes_jacobi_sweeps(n, nnz, k, sweeps, outer, sweeps_next=0):
    """Average algorithmic bytes of one Jacobi sweep launch over a solve."""
    total, launches = _jacobi_sweep_totals(n, nnz, k, sweeps, outer, sweeps_next)
    return total / launches


def bytes_jacobi_method(n, nnz, k, sweeps, outer, sweeps_next=0):
    """Whole Jacobi-refinement solve: per step one FP64 residual (the first
    reads no x) and its sweeps, then the certifying FP64 residual."""
    res = 4 * (n + 1) + 12 * nnz + 16 * k * n + 4 * k * n + 4 * k * n   # b, x, r32, d0
    first_res = res - 8 * k * n
    sweeps_b, _ = _jacobi_sweep_totals(n, nnz, k, sweeps, outer, sweeps_next)
    cert = 4 * (n + 1) + 12 * nnz + 16 * k * n + 4 * k * n
    return first_res + (outer - 1) * res + sweeps_b + cert


KERNEL_NAMES = {1: "simple", 2: "order", 3: "multirow", 4: "jacobi"}

# Configurations the autotuner chose on MI355X for each workload (sptrsv_config_t
# fields), used with --preset so that profiler runs time exactly the tuned
# kernels without running the tuner.
PRESETS = {
    2: {"f64": dict(kernel=2, lanes=16, wg=256, fused=1, variant=2, cu_waves=4),
        "mixed": dict(kernel=4, lanes=2, wg=256, fused=1, variant=1, sweeps=12, outer=2, window=16)},
    3: {"f64": dict(kernel=3, lanes=32, wg=1024, fused=1, variant=2, cu_waves=8),
        "mixed": dict(kernel=3, lanes=32, wg=1024, fused=1, variant=2, cu_waves=8, reorder=1)},
    4: {"f64": dict(kernel=2, lanes=1, wg=1024, fused=1, variant=2, cu_waves=6),
        "mixed": dict(kernel=2, lanes=1, wg=1024, fused=1, variant=2, cu_waves=12, reorder=1),
        "f64_per_rhs": dict(kernel=3, lanes=4, wg=256, fused=1, variant=2, cu_waves=6)},
}


def cfg_dict(c) -> dict:
    d = {"kernel": KERNEL_NAMES.get(c.kernel, c.kernel), "lanes": c.lanes, "wg": c.wg}
    if c.kernel == 4:
        d = {"kernel": "jacobi", "tile_products": 1024 * max(1, c.lanes), "wg": c.wg, "sweeps": c.sweeps,
             "sweeps_next": c.sweeps_next or c.sweeps, "outer": c.outer, "entries_per_lane": c.window or 8,
             "column_sorted_tiles": bool(c.variant == 1),
             "block_inner_iterations": (c.inner or 4) if c.variant == 2 else 0}
    else:
        d.update(fused=c.fused, variant={0: "multi-row-wave", 1: "one-row-per-wave", 2: "persistent"}.get(
            c.variant, c.variant), cu_waves=c.cu_waves, window=c.window)
        if c.variant == 2:
            d["claim_rounds"] = c.claim_rounds or 1
        if c.kernel == 3:
            d["split_rows"] = c.split_rows >= 0
        if c.reorder:
            d["level_ordered_copy"] = True
    return d


def cfg_fields(c) -> dict:
    """Every sptrsv_config_t field (the form PRESETS records)."""
    return {name: getattr(c, name) for name, _ in c._fields_}


def sweep_layout(n, tw, variant):
    """Entry layout of the point sweep kernel the library launches (its third
    template argument): 0 storage order, 1 column-sorted with a separate
    2-byte position, 2 column-sorted with the position packed above the
    column, used when n <= 2^(32 - p), p the bits of a position below the
    tile's LDS capacity of tw + tw/4 products."""
    if variant != 1:
        return 0
    pbits = (tw + tw // 4 - 1).bit_length()
    return 2 if n <= (1 << (32 - pbits)) else 1


def profiled_traffic(kernel_key: str):
    """HBM-side bytes per launch of `kernel_key` from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summaries committed under profiles/ (calibrated there against a
    copy kernel of known size), or None when no summary covers that kernel."""
    import re
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            m = re.match(r"traffic_bytes_per_launch\s+(\S+)\s+(\d+)", line.strip("` "))
            if m and m.group(1) == kernel_key:
                return int(m.group(2))
    return None


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def _native_substitution():
    """The oracle's serial FP64 forward substitution compiled on this host with
    gcc -O3 -march=native (the CPU baseline's build), or None without gcc."""
    import ctypes
    import subprocess
    import tempfile
    out = Path(tempfile.mkdtemp(prefix="sptrsv_cpu_")) / "liboracle_native.so"
    try:
        subprocess.run(["gcc", "-O3", "-march=native", "-fPIC", "-shared", "-ffp-contract=off",
                        str(ROOT / "oracle" / "oracle.c"), "-o", str(out)], check=True, capture_output=True,
                       timeout=120)
        return ctypes.CDLL(str(out))
    except (OSError, subprocess.SubprocessError):
        return None


def cpu_baseline(systems, budget_s: float):
    """Serial FP64 forward substitution, one core, median of >= 5 timed solves
    after a warm-up, over the given systems in turn until the time budget is
    spent; GNNZ/s over the solves timed.  Built with -O3 -march=native on this
    host; its results are checked bit for bit against the oracle."""
    import ctypes
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_py as O
    lib = _native_substitution()
    flags = "gcc -O3 -march=native -ffp-contract=off, built on this host" if lib else \
        "gcc -O2 -ffp-contract=off (the checker's build; no compiler on this host)"

    def p(a):
        return a.ctypes.data_as(ctypes.c_void_p)

    def solve(sy):
        if lib is None:
            return O.fwd_f64(sy.ia, sy.ja, sy.a, sy.b[: sy.n])
        x = np.zeros(sy.n)
        b = np.ascontiguousarray(sy.b[: sy.n])
        lib.oracle_fwd_f64(sy.n, p(sy.ia), p(sy.ja), p(sy.a), p(b), p(x), 1)
        return x
    first = systems[0]
    x = solve(first)                                        # warm-up, and the bitwise check
    assert np.array_equal(x, O.fwd_f64(first.ia, first.ja, first.a, first.b[: first.n]))
    rates, secs, solves, entries = [], 0.0, 0, 0
    t_end = time.perf_counter() + budget_s
    i = 0
    while time.perf_counter() < t_end or len(rates) < 5:
        sy = systems[i % len(systems)]
        i += 1
        t0 = time.perf_counter()
        solve(sy)
        dt = time.perf_counter() - t0
        rates.append(sy.nnz / dt)
        secs += dt
        solves += 1
        entries += sy.nnz
    what = "the full matrix" if len(systems) == 1 else f"{min(solves, len(systems))} of the {len(systems)} systems"
    return {"value": round(float(np.median(rates)) / 1e9, 4), "unit": "GNNZ/s", "cores": 1, "kind": "port",
            "sample": f"serial FP64 forward substitution (oracle/oracle.c, {flags}) on {what}, 1 RHS, "
                      f"median of {solves} solves ({secs:.1f} s, {entries / 1e9:.2f} G entries) on "
                      f"{_cpu_model()}, 1 of {os.cpu_count()} cores used"}


def calibrate_copy(I) -> None:
    """Five copies of 256 MiB with the library's k_copy16: the known-size
    launches the FETCH_SIZE / WRITE_SIZE counters are scaled against."""
    src = I.DeviceArray(32 << 20, np.float64)                       # 256 MiB
    dst = I.DeviceArray(32 << 20, np.float64)
    for _ in range(5):
        I.copy_device(dst, src, 256 << 20)
    I.synchronize()
    src.free()
    dst.free()
    log("[bench] calibration: 5 k_copy16 copies of 262144 KiB (read 262144 KiB, write 262144 KiB each)")


def run_batch(args, I, O) -> int:
    """Workload 5: 256 independent systems in one captured graph (MIXED, per-matrix
    kernel choice) against the same systems solved one after another in plain FP64."""
    desc, nmat = WORKLOADS[5]
    t0 = time.time()
    systems = [I.generate_batch_member(m, 2000, 200000, seed=5) for m in range(nmat)]
    rows = sum(x.n for x in systems)
    nnz = sum(x.nnz for x in systems)
    log(f"[bench] generated {nmat} systems: rows={rows} nnz={nnz} ({time.time() - t0:.1f}s)")
    bt = I.Batch(systems)
    B = np.concatenate([x.b for x in systems])
    db = I.DeviceArray.from_numpy(B)
    dx = I.DeviceArray(rows, np.float64)
    t0 = time.time()
    bt.tune(I.PREC_MIXED, tune=not args.no_tune)
    kinds = [bt.choice(m).kernel for m in range(nmat)]
    jc = next((bt.choice(m) for m in range(nmat) if kinds[m] == I.KERNEL_JACOBI), None)
    log(f"[bench] tuned in {time.time() - t0:.1f}s: {kinds.count(I.KERNEL_JACOBI)} Jacobi members"
        + (f" (sweeps {jc.sweeps}/{jc.sweeps_next or jc.sweeps}, steps {jc.outer}"
           + (f", block sweeps with {jc.inner or 4} inner iterations" if jc.variant == 2 else "") + ")" if jc else ""))
    for _ in range(args.warmup):
        bt.solve(I.PREC_MIXED, db, dx, graph=True)
    I.synchronize()
    timer = I.Timer()
    timer.start()
    for _ in range(args.steps):
        bt.solve(I.PREC_MIXED, db, dx, graph=True)
    msm = timer.stop() / args.steps
    I.synchronize()
    assert bt.check() == 0
    # the two halves' kernels timed apart (HIP events on their own streams, in
    # as many eager solves): Jacobi sweeps of the Jacobi members, and the
    # sync-free pass of the others
    per_kind = {}
    for kind in ("sweeps", "sync-free"):
        I.kernel_timer(True, kind)
        for _ in range(args.steps):
            bt.solve(I.PREC_MIXED, db, dx, graph=False)
        I.synchronize()
        per_kind[kind] = I.kernel_timer_read()
        I.kernel_timer(False)
    assert bt.check() == 0
    kms = sum(v[0] for v in per_kind.values())
    launches = sum(v[1] for v in per_kind.values())
    x = dx.numpy()
    errs = []
    for m, sm in enumerate(systems):
        ref = O.fwd_f64(sm.ia, sm.ja, sm.a, sm.b)
        xm = x[bt.row_off[m]:bt.row_off[m + 1]]
        errs.append(float(np.max(np.abs(xm - ref)) / np.max(np.abs(ref))))
    log(f"[bench] batched MIXED graph: {msm:.3f} ms, max member rel err {max(errs):.2e}")
    ms64b = ms_seq = None
    if not args.no_extras:
        # every member at once, plain FP64 sync-free (heuristic lanes), one launch
        bt.tune(I.PREC_F64, tune=False)
        bt.solve(I.PREC_F64, db, dx, graph=False)
        I.synchronize()
        timer = I.Timer()
        timer.start()
        for _ in range(max(1, args.steps // 4)):
            bt.solve(I.PREC_F64, db, dx, graph=False)
        ms64b = timer.stop() / max(1, args.steps // 4)
        assert bt.check() == 0
        # sequential plain FP64: one system at a time, each its own matrix and launch
        mats = [I.allocate_memory(sm.ia, sm.ja, sm.a) for sm in systems]
        bufs = [(I.DeviceArray.from_numpy(sm.b), I.DeviceArray(sm.n, np.float64)) for sm in systems]
        I.solve(mats[0], None, I.PREC_F64, bufs[0][0], bufs[0][1], kernel=I.KERNEL_SIMPLE)
        I.synchronize()
        timer = I.Timer()
        timer.start()
        for mt, (bb, xx) in zip(mats, bufs):
            I.solve(mt, None, I.PREC_F64, bb, xx, kernel=I.KERNEL_SIMPLE)
        ms_seq = timer.stop()
        for mt in mats:
            assert I.check_watchdog(mt) == 0
            I.free_device_analysis(mt, None)
        for bb, xx in bufs:
            bb.free()
            xx.free()
        log(f"[bench] FP64 batched one launch: {ms64b:.3f} ms; FP64 sequential: {ms_seq:.1f} ms")
    if args.calibrate_copy:
        calibrate_copy(I)
    bytes_jac_rows = sum(sm.nnz for m, sm in enumerate(systems) if kinds[m] == I.KERNEL_JACOBI)
    # roofline of the dominant kernel: the kind with the larger summed time
    jac_n = sum(sm.n for m, sm in enumerate(systems) if kinds[m] == I.KERNEL_JACOBI)
    sf_n, sf_nnz = rows - jac_n, nnz - bytes_jac_rows
    t_sw, n_sw = per_kind["sweeps"]
    t_sf, n_sf = per_kind["sync-free"]
    traffic = None
    if jc is not None and t_sw >= t_sf and n_sw > 0:
        b_launch = bytes_jacobi_sweeps(jac_n, bytes_jac_rows, 1, jc.sweeps, jc.outer, jc.sweeps_next)
        k_ms = t_sw / n_sw
        if jc.variant == 2:
            traffic = profiled_traffic(f"k_jac_bsweep<{1024 * max(1, jc.lanes)}>@workload5")
        kname = (f"Jacobi {'block ' if jc.variant == 2 else ''}sweep (FP32) over the {kinds.count(I.KERNEL_JACOBI)} "
                 f"Jacobi members, inside FP64 refinement")
        note = ("effective bandwidth: the Jacobi members' arrays stay in the 256 MB Infinity Cache across a "
                "step's sweeps when they fit")
    else:
        b_launch = bytes_mixed_fused(sf_n, sf_nnz, 1)
        k_ms = t_sf / max(n_sf, 1)
        kname = "sync-free fused mixed-precision pass over the sync-free members"
        note = "one sync-free pass, HBM-fed; latency-bound by the members' dependency chains"
    achieved = b_launch / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    log(f"[bench] kernels per solve: sweeps {t_sw / args.steps:.3f} ms ({n_sw // max(args.steps, 1)} launches), "
        f"sync-free {t_sf / args.steps:.3f} ms ({n_sf // max(args.steps, 1)} launches)")
    cpu = None if args.no_cpu else cpu_baseline(systems, args.cpu_seconds)
    out = {
        "metric": METRIC,
        "value": round(nnz / (msm * 1e-3) / 1e9, 4),
        "unit": "GNNZ/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(msm, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {
            "workload": desc, "systems": nmat, "rows": rows, "nnz": nnz,
            "jacobi_members": kinds.count(I.KERNEL_JACOBI),
            "jacobi_schedule": {"sweeps": jc.sweeps, "sweeps_next": jc.sweeps_next or jc.sweeps,
                                "outer": jc.outer,
                                "block_inner_iterations": (jc.inner or 4) if jc.variant == 2 else 0} if jc else None,
            "jacobi_nnz_fraction": round(bytes_jac_rows / nnz, 4),
            "schedule": dict(zip(("serial_halves", "sync_free_members_as_one_analysed_matrix"),
                                 (bool(v) for v in bt.schedule()))),
            "rel_err_mixed_max_member": max(errs),
            "f64_batched_one_launch_ms": None if ms64b is None else round(ms64b, 3),
            "f64_sequential_ms": None if ms_seq is None else round(ms_seq, 2),
            "speedup_vs_sequential_f64": None if ms_seq is None else round(ms_seq / msm, 2),
            "speedup_vs_batched_f64": None if ms64b is None else round(ms64b / msm, 3),
            "kernel_ms_sum_per_solve": round(kms / max(args.steps, 1), 4),
            "kernel_launches_per_solve": launches // max(args.steps, 1),
            "sweeps_ms_per_solve": round(t_sw / max(args.steps, 1), 4),
            "sync_free_ms_per_solve": round(t_sf / max(args.steps, 1), 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "kernel": kname,
            "kernel_ms": round(k_ms, 4),
            "bytes_per_launch": int(b_launch),
            "note": note,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    db.free()
    dx.free()
    bt.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=2, choices=sorted(WORKLOADS),
                    help="2 (default, the headline), 3, 4 (16 RHS) or 5 (batch of 256 systems)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--jacobi", type=int, nargs="+", metavar="N",
                    help="SWEEPS STEPS TILE EPL SORTED [SWEEPS_NEXT]: use the Jacobi-refinement mixed path "
                         "with this schedule (tile width 1 or 2 = 1024 or 2048 products, phase-2 entries per "
                         "lane, 1 = column-sorted tiles, sweeps of the steps after the first) instead of "
                         "tuning it")
    ap.add_argument("--preset", action="store_true",
                    help="use the recorded tuned configurations (PRESETS) instead of tuning (profiler runs)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the rocSPARSE, power and cold-cache legs (for profiler runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run 5 plain device copies of 256 MiB (the library's k_copy16), the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this SpTRSV engine runs on one GPU; --gpus must be 1"}))
        return 2

    from sptrsv_amd import interface as I
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_py as O

    if args.workload == 5:
        return run_batch(args, I, O)
    desc, param = WORKLOADS[args.workload]
    t0 = time.time()
    s = I.generate(args.workload, param)
    k = s.k
    log(f"[bench] generated workload {args.workload}: n={s.n} nnz={s.nnz} k={k} ({time.time() - t0:.1f}s)")
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    analysis_ms = []
    for _ in range(2):                      # GPU level-set analysis (ALL mode), first and repeat call
        t_a = time.perf_counter()
        info = I.multirow_analysis_base_GPU(mat, I.ALL)
        analysis_ms.append((time.perf_counter() - t_a) * 1e3)
        if len(analysis_ms) == 1:
            I.free_device_analysis(None, info)
    nlevs = info.nlevs
    log(f"[bench] device {I.device_name(0)}; analysis nlevs={nlevs}")
    db = I.DeviceArray.from_numpy(s.b)
    dx = I.DeviceArray(s.n * k, np.float64)
    I.reserve(mat, k)

    tuned = {}
    for prec in (I.PREC_F64, I.PREC_MIXED):
        if prec == I.PREC_MIXED and args.jacobi:
            j = args.jacobi
            if len(j) not in (5, 6):
                ap.error("--jacobi takes 5 or 6 integers")
            tuned[prec] = I._cfg(I.KERNEL_JACOBI, j[2], 256, 1, j[4], 0, j[0], j[1], j[3],
                                 j[5] if len(j) == 6 else 0)
        elif args.preset and args.workload in PRESETS:
            tuned[prec] = I._cfg(**PRESETS[args.workload]["f64" if prec == I.PREC_F64 else "mixed"])
        elif args.no_tune:
            tuned[prec] = I.default_config(mat, info, prec, k)
        else:
            tuned[prec], _ = I.autotune(mat, info, prec, k=k, reps=3)
        log(f"[bench] precision {prec}: {cfg_dict(tuned[prec])}")

    x_ref = O.fwd_f64(s.ia, s.ja, s.a, s.b, k=k)

    def run(prec):
        """Timed steps replay the solve captured in a hipGraph (one launch per
        solve); the dominant kernel's duration is then measured live with HIP
        events around its launches (each sync-free pass; for Jacobi one pair
        around each refinement step's consecutive sweeps, divided by their
        count), in the same number of eager solves."""
        c = tuned[prec]
        for _ in range(args.warmup):
            I.solve_graph(mat, info, prec, db, dx, c, k=k)
        I.synchronize()
        assert I.check_watchdog(mat) == 0
        timer = I.Timer()
        timer.start()
        for _ in range(args.steps):
            I.solve_graph(mat, info, prec, db, dx, c, k=k)
        ms = timer.stop() / args.steps
        I.synchronize()
        assert I.check_watchdog(mat) == 0
        x = dx.numpy()
        err = float(np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref)))
        I.kernel_timer(True)
        for _ in range(args.steps):
            I.solve(mat, info, prec, db, dx, k=k, cfg=c)
        I.synchronize()
        kms, launches = I.kernel_timer_read()
        I.kernel_timer(False)
        assert I.check_watchdog(mat) == 0
        return ms, kms / max(launches, 1), err

    # cold-cache (MALL flushed) single solves, reported beside the steady state
    def cold(prec, reps=5):
        c = tuned[prec]
        out = []
        for _ in range(reps):
            I.flush_cache()
            I.synchronize()
            timer = I.Timer()
            timer.start()
            I.solve_graph(mat, info, prec, db, dx, c, k=k)
            out.append(timer.stop())
        return float(np.median(out))

    def jacobi_setup_ms():
        """One-time set-up of the Jacobi path on a fresh handle of the same matrix
        (row-scaled FP32 copy, contiguous diagonal, tile tables, column-sorted
        tile copy): its first solve minus a steady eager solve."""
        c = tuned[I.PREC_MIXED]
        if c.kernel != I.KERNEL_JACOBI:
            return None
        m2 = I.allocate_memory(s.ia, s.ja, s.a)
        I.synchronize()
        timer = I.Timer()
        timer.start()
        I.reserve(m2, k)
        I.solve(m2, None, I.PREC_MIXED, db, dx, k=k, cfg=c)
        first = timer.stop()
        timer = I.Timer()
        timer.start()
        for _ in range(3):
            I.solve(m2, None, I.PREC_MIXED, db, dx, k=k, cfg=c)
        steady = timer.stop() / 3
        assert I.check_watchdog(m2) == 0
        I.free_device_analysis(m2, None)
        return round(max(0.0, first - steady), 3)

    ms64, kms64, err64 = run(I.PREC_F64)
    log(f"[bench] plain FP64: {ms64:.3f} ms/solve (kernel {kms64:.3f} ms) rel err {err64:.2e}")
    msm, kmsm, errm = run(I.PREC_MIXED)
    log(f"[bench] FP32+refine: {msm:.3f} ms/solve (kernel {kmsm:.3f} ms) rel err {errm:.2e}")
    per_rhs = None
    if k > 1:
        # the plain FP64 sync-free solve one right-hand side at a time (k solves)
        b1 = I.DeviceArray.from_numpy(np.ascontiguousarray(s.b.reshape(s.n, k)[:, 0]))
        x1 = I.DeviceArray(s.n, np.float64)
        if args.preset and args.workload in PRESETS:
            c1 = I._cfg(**PRESETS[args.workload]["f64_per_rhs"])
        elif args.no_tune:
            c1 = I.default_config(mat, info, I.PREC_F64, 1)
        else:
            c1, _ = I.autotune(mat, info, I.PREC_F64, k=1, reps=3)
        for _ in range(args.warmup):
            I.solve_graph(mat, info, I.PREC_F64, b1, x1, c1, k=1)
        I.synchronize()
        timer = I.Timer()
        timer.start()
        for _ in range(args.steps):
            I.solve_graph(mat, info, I.PREC_F64, b1, x1, c1, k=1)
        t1 = timer.stop() / args.steps
        assert I.check_watchdog(mat) == 0
        ref1 = O.fwd_f64(s.ia, s.ja, s.a, np.ascontiguousarray(s.b.reshape(s.n, k)[:, 0]))
        e1 = float(np.max(np.abs(x1.numpy() - ref1)) / np.max(np.abs(ref1)))
        per_rhs = {"ms_per_rhs": round(t1, 4), "ms_all_rhs": round(t1 * k, 4),
                   "gnnz_s": round(s.nnz / (t1 * 1e-3) / 1e9, 4), "kernel": cfg_dict(c1), "rel_err": e1,
                   "speedup_mixed_vs_per_rhs_f64": round(t1 * k / msm, 3)}
        log(f"[bench] plain FP64 per RHS: {t1:.3f} ms x {k} = {t1 * k:.3f} ms")
        b1.free()
        x1.free()
    cold64, coldm = (cold(I.PREC_F64), cold(I.PREC_MIXED)) if not args.no_extras else (float("nan"),) * 2
    jac_setup = jacobi_setup_ms()
    tdep = I.dep_latency_ns(2000)
    log(f"[bench] cold-cache: FP64 {cold64:.3f} ms, mixed {coldm:.3f} ms; t_dep {tdep:.0f} ns one-way")

    # vendor baseline on the same device matrix (analysis timed separately)
    rocsparse = None
    try:
        if args.no_extras:
            raise RuntimeError("skipped (--no-extras)")
        # csrsv takes one contiguous right-hand side: with k > 1 it solves
        # RHS 0, split out of the row-major n x k block
        if k == 1:
            rb_b, rb_x, rb_ref = db, dx, x_ref
        else:
            b0 = np.ascontiguousarray(s.b.reshape(s.n, k)[:, 0])
            rb_b, rb_x = I.DeviceArray.from_numpy(b0), I.DeviceArray(s.n, np.float64)
            rb_ref = np.ascontiguousarray(x_ref.reshape(s.n, k)[:, 0])
        rb = I.RocsparseBaseline(mat)
        rb.solve(rb_b, rb_x)
        I.synchronize()
        timer = I.Timer()
        timer.start()
        for _ in range(args.steps):
            rb.solve(rb_b, rb_x)
        rs_ms = timer.stop() / args.steps
        x = rb_x.numpy()
        rs_err = float(np.max(np.abs(x - rb_ref)) / np.max(np.abs(rb_ref)))
        if k > 1:
            rb_b.free()
            rb_x.free()
        rocsparse = {"ms": round(rs_ms, 4), "gnnz_s": round(k * s.nnz / (rs_ms * 1e-3) / 1e9 if k == 1 else
                                                           s.nnz / (rs_ms * 1e-3) / 1e9, 4),
                     "analysis_ms": round(rb.analysis_ms, 3), "rel_err": rs_err,
                     "note": "rocsparse_dcsrsv_solve, 1 RHS" if k == 1 else "rocsparse_dcsrsv_solve on RHS 0 only"}
        rb.close()
        log(f"[bench] rocSPARSE csrsv: {rs_ms:.3f} ms (analysis {rb.analysis_ms:.1f} ms)")
    except Exception as e:  # reported baseline only
        rocsparse = {"error": str(e)[:200]}

    # rocSPARSE's own iterative triangular solve (FP64 Jacobi, host-checked
    # residual), stopped at |r|_inf <= 1e-11, beside the certified FP32 Jacobi
    itsv = None
    if not args.no_extras and k == 1:
        try:
            ri = I.RocsparseBaseline(mat, itsv=True)
            iters = ri.solve_itsv(db, dx, 1e-11)
            I.synchronize()
            timer = I.Timer()
            timer.start()
            reps = 3
            for _ in range(reps):
                iters = ri.solve_itsv(db, dx, 1e-11)
            it_ms = timer.stop() / reps
            x = dx.numpy()
            itsv = {"ms": round(it_ms, 4), "iterations": iters, "analysis_ms": round(ri.analysis_ms, 3),
                    "rel_err": float(np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref))),
                    "note": "rocsparse_dcsritsv_solve_ex, tol 1e-11 on |b - L x|_inf, 20 iterations between tests"}
            ri.close()
            log(f"[bench] rocSPARSE csritsv: {it_ms:.3f} ms, {iters} iterations")
        except Exception as e:  # reported baseline only
            itsv = {"error": str(e)[:200]}

    power = None
    if not args.no_extras and I.power_start("bench", 0, 1000):
        for _ in range(max(args.steps, 50)):
            I.solve_graph(mat, info, I.PREC_MIXED, db, dx, tuned[I.PREC_MIXED], k=k)
        I.synchronize()
        p = I.power_end()
        power = round(p / 1000.0, 1) if p > 0 else None

    n, nnz = s.n, s.nnz
    cm_is_jacobi = tuned[I.PREC_MIXED].kernel == I.KERNEL_JACOBI
    gnnz_m = k * nnz / (msm * 1e-3) / 1e9
    gnnz_64 = k * nnz / (ms64 * 1e-3) / 1e9
    cm = tuned[I.PREC_MIXED]
    traffic = None
    if cm.kernel == I.KERNEL_JACOBI:
        b_launch = bytes_jacobi_sweeps(n, nnz, k, cm.sweeps, cm.outer, cm.sweeps_next)
        b_method = bytes_jacobi_method(n, nnz, k, cm.sweeps, cm.outer, cm.sweeps_next)
        kpow = 1
        while kpow < k:
            kpow *= 2
        tw = 1024 * max(1, cm.lanes)
        key = f"k_jac_sweep<{kpow},{tw},{sweep_layout(n, tw, cm.variant)}>@workload{args.workload}"
        kname = f"Jacobi sweep (FP32, {key.split('@')[0]}) inside FP64 refinement"
        traffic = profiled_traffic(key)
    else:
        b_launch = b_method = bytes_mixed_fused(n, nnz, k, cm.reorder)
        kname = "sync-free fused mixed-precision solve"
        # multirow takes its lanes per row from the analysis: lanes is not part of its key
        kl = 0 if cm.kernel == I.KERNEL_MULTIROW else cm.lanes
        ro = ",reorder=1" if cm.reorder else ""
        key = (f"fused-mixed[kernel={cm.kernel},lanes={kl},wg={cm.wg},variant={cm.variant}{ro},k={k}]"
               f"@workload{args.workload}")
        traffic = profiled_traffic(key)
    achieved = b_launch / (kmsm * 1e-3) / 1e9
    cpu = None if args.no_cpu else cpu_baseline([s], args.cpu_seconds)
    out = {
        "metric": METRIC,
        "value": round(gnnz_m, 4),
        "unit": "GNNZ/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(msm, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": desc, "n": n, "nnz": nnz, "rhs": k, "nlevs": nlevs,
            "kernel_mixed": cfg_dict(tuned[I.PREC_MIXED]),
            "kernel_f64": cfg_dict(tuned[I.PREC_F64]),
            "rel_err_mixed": errm, "rel_err_f64": err64,
            "plain_f64_gnnz_s": round(gnnz_64, 4), "plain_f64_ms": round(ms64, 4),
            "speedup_mixed_vs_f64": round(ms64 / msm, 3),
            "hbm_gbs_mixed": round(b_method / (msm * 1e-3) / 1e9, 2),
            "hbm_gbs_mixed_compulsory": round(bytes_method_convention(n, nnz, k) / (msm * 1e-3) / 1e9, 2),
            "bytes_mixed_method": int(b_method),
            "hbm_gbs_f64": round(bytes_f64(n, nnz, k) / (ms64 * 1e-3) / 1e9, 2),
            "bytes_method_separate_pass_convention": bytes_method_convention(n, nnz, k),
            "cold_cache_ms": None if args.no_extras else {"mixed": round(coldm, 4), "f64": round(cold64, 4)},
            "t_dep_ns": round(tdep, 1),
            "latency_bound_fraction_mixed": (None if cm_is_jacobi else round(nlevs * tdep * 1e-6 / msm, 3)),
            "latency_bound_fraction_f64": round(nlevs * tdep * 1e-6 / ms64, 3),
            "power_w": power,
            "rocsparse_csrsv": rocsparse,
            "rocsparse_csritsv": itsv,
            "analysis_ms": {"first": round(analysis_ms[0], 3), "repeat": round(analysis_ms[1], 3)},
            "jacobi_setup_ms": jac_setup,
            "plain_f64_per_rhs": per_rhs,
            "tuned_fields": {"f64": cfg_fields(tuned[I.PREC_F64]), "mixed": cfg_fields(tuned[I.PREC_MIXED]),
                             **({"f64_per_rhs": cfg_fields(c1)} if k > 1 else {})},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "kernel": kname,
            "kernel_ms": round(kmsm, 4),
            "bytes_per_launch": int(b_launch),
            "note": (f"effective bandwidth: the {b_launch / 1e6:.0f} MB a sweep reads stay in the 256 MB "
                     f"Infinity Cache across a step's sweeps, so achieved counts cache-fed bytes; a whole "
                     f"cold-cache solve takes {coldm:.3f} ms against {msm:.3f} ms warm"
                     if cm_is_jacobi else "one sync-free pass, HBM-fed"),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    db.free()
    dx.free()
    I.free_device_analysis(mat, info)
    if args.calibrate_copy:
        calibrate_copy(I)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
