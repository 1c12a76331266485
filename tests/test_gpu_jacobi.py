"""GPU tests of the certified Jacobi-refinement path (KERNEL_JACOBI, mixed
precision): FP32 Jacobi sweeps inside FP64 iterative refinement, with the
a-posteriori bound |x - x*|_inf <= |b - L x|_inf / delta checked on the
device.

The result is compared with the oracle's serial FP64 forward substitution
(the reference's solution) at the mixed-precision bar, 1e-10 relative
max-norm error.  The certificate itself is pinned by soundness: whenever it
reports success, the true error against the oracle is within the tolerance.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as O
from sptrsv_amd import _lib as L
from sptrsv_amd import interface as I
from test_gpu_solve import Dev, relerr

pytestmark = pytest.mark.gpu

J = dict(precision=I.PREC_MIXED, kernel=I.KERNEL_JACOBI)


def margin_np(ia, a):
    n = len(ia) - 1
    rows = np.repeat(np.arange(n), np.diff(ia))
    diag = a[ia[1:] - 1]
    off = np.bincount(rows, weights=np.abs(a), minlength=n) - np.abs(diag)
    return float(np.min(np.abs(diag) - off))


@pytest.mark.parametrize("cfg,param", [(1, 64), (2, 200000), (3, 200000), (4, 24)])
def test_jacobi_certified_configs(cfg, param):
    s = I.generate(cfg, param)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b, k=s.k)
    d = Dev(s.ia, s.ja, s.a)
    assert I.dominance_margin(d.mat) == pytest.approx(margin_np(s.ia, s.a), rel=1e-12)
    for lanes in (4, 16):
        x = d.solve(s.b, k=s.k, lanes=lanes, sweeps=16, outer=4, **J)
        assert relerr(x, ref) <= 1e-10, (cfg, lanes)
    d.close()


def test_jacobi_known_answer():
    """x = 1 on the reference's known-answer input, to the refinement's
    tolerance: exact equality is a property of direct substitution, which
    the sync-free kernels keep; an iterative solve stops within rounding."""
    s = I.generate(1)
    d = Dev(s.ia, s.ja, s.a)
    x = d.solve(s.b, lanes=4, sweeps=12, outer=3, **J)
    assert np.max(np.abs(x - 1.0)) <= 1e-10
    x = d.solve(s.b, lanes=4, sweeps=12, outer=6, **J)   # more steps: converges to rounding
    assert np.max(np.abs(x - 1.0)) <= 1e-14
    d.close()


def test_jacobi_certificate_is_sound():
    """For every schedule the device accepts, the true error is within tol."""
    s = I.generate(3, 100000)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    d = Dev(s.ia, s.ja, s.a)
    accepted = rejected = 0
    for sweeps in (2, 4, 8):
        for outer in (1, 2, 3, 4, 6):
            db = I.DeviceArray.from_numpy(s.b)
            dx = I.DeviceArray(s.n, np.float64)
            I.solve(d.mat, None, I.PREC_MIXED, db, dx, kernel=I.KERNEL_JACOBI, lanes=8, sweeps=sweeps,
                    outer=outer)
            st = I.check_watchdog(d.mat)
            err = relerr(dx.numpy(), ref)
            db.free()
            dx.free()
            assert st in (0, I.E_NOT_CONVERGED), L.last_error()
            if st == 0:
                accepted += 1
                assert err <= 1e-10, (sweeps, outer, err)
            else:
                rejected += 1
    assert accepted > 0 and rejected > 0        # both verdicts are exercised
    assert I.check_watchdog(d.mat) == 0         # the flag is cleared once read
    d.close()


def test_jacobi_tolerance_setting():
    s = I.generate(2, 50000)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    d = Dev(s.ia, s.ja, s.a)
    I.set_tolerance(d.mat, 1e-4)
    x = d.solve(s.b, lanes=8, sweeps=12, outer=1, **J)  # accepted at the loose tolerance
    assert relerr(x, ref) <= 1e-4
    I.set_tolerance(d.mat, 1e-14)
    d.solve(s.b, lanes=8, sweeps=12, outer=1, expect=I.E_NOT_CONVERGED, **J)
    with pytest.raises(L.SptrsvError):
        I.set_tolerance(d.mat, 0.0)
    # a captured solve certifies against the tolerance in force when it is
    # replayed, not the one it was captured with
    db = I.DeviceArray.from_numpy(s.b)
    dx = I.DeviceArray(s.n, np.float64)
    c = I._cfg(I.KERNEL_JACOBI, 8, 256, 1, 0, 0, 12, 1)
    I.set_tolerance(d.mat, 1e-4)
    I.solve_graph(d.mat, d.info, I.PREC_MIXED, db, dx, c)
    I.synchronize()
    assert I.check_watchdog(d.mat) == 0
    I.set_tolerance(d.mat, 1e-14)
    I.solve_graph(d.mat, d.info, I.PREC_MIXED, db, dx, c)
    I.synchronize()
    assert I.check_watchdog(d.mat) == I.E_NOT_CONVERGED
    I.set_tolerance(d.mat, 1e-4)
    I.solve_graph(d.mat, d.info, I.PREC_MIXED, db, dx, c)
    I.synchronize()
    assert I.check_watchdog(d.mat) == 0
    db.free()
    dx.free()
    d.close()


def test_jacobi_certifies_each_rhs_column_on_its_own():
    """k = 2 with the second column 1e-9 times the first: the certificate
    compares each column's residual with that column's own solution, so it
    either reports a miss or every column is within tol of the oracle on its
    own scale (pooling the columns would accept the small column's error)."""
    s = I.generate(2, 100000)
    k = 2
    rng = np.random.default_rng(3)
    B = np.stack([rng.uniform(-1, 1, s.n), 1e-9 * rng.uniform(-1, 1, s.n)], axis=1).ravel()
    ref = O.fwd_f64(s.ia, s.ja, s.a, B, k=k).reshape(s.n, k)
    d = Dev(s.ia, s.ja, s.a, analyse=False)
    verdicts = set()
    for outer in (1, 2, 3, 4):
        x = d.solve(B, k=k, lanes=8, sweeps=8, outer=outer, expect=None, **J).reshape(s.n, k)
        st = I.check_watchdog(d.mat)
        errs = [relerr(x[:, c], ref[:, c]) for c in range(k)]
        print(f"outer {outer}: status {st}, per-column errors {errs}", flush=True)
        assert st in (0, I.E_NOT_CONVERGED)
        verdicts.add(st)
        if st == 0:
            assert max(errs) <= 1e-10, (outer, errs)
    assert 0 in verdicts  # enough steps certify both columns
    d.close()


@pytest.mark.parametrize("path", ["jacobi", "fused-mixed", "upper"])
def test_captured_solve_survives_workspace_regrowth(path):
    """A solve captured for k = 1, then an eager solve with k = 3 (which
    regrows the workspaces the captured solve used), then the k = 1 graph
    replayed: it must equal the eager k = 1 solve bit for bit."""
    s = I.generate(2, 60000)
    rng = np.random.default_rng(5)
    if path == "upper":
        import scipy.sparse as sp
        Lm = sp.csr_matrix((s.a, s.ja, s.ia), shape=(s.n, s.n))
        U = sp.csr_matrix(Lm.T)
        U.sort_indices()
        mat = I.allocate_memory_upper(U.indptr, U.indices, U.data)
        cfg = I._cfg(I.KERNEL_SIMPLE, 0, 256, 1, 0, 0)
        prec = I.PREC_F64
    else:
        mat = I.allocate_memory(s.ia, s.ja, s.a)
        prec = I.PREC_MIXED
        cfg = (I._cfg(I.KERNEL_JACOBI, 2, 256, 1, 1, 0, 12, 2) if path == "jacobi"
               else I._cfg(I.KERNEL_SIMPLE, 0, 256, 1, 0, 0))
    b1 = I.DeviceArray.from_numpy(s.b)
    x1 = I.DeviceArray(s.n, np.float64)
    b3 = I.DeviceArray.from_numpy(rng.uniform(-1, 1, 3 * s.n))
    x3 = I.DeviceArray(3 * s.n, np.float64)
    I.solve(mat, None, prec, b1, x1, cfg=cfg)
    I.synchronize()
    eager = x1.numpy()
    for rep in range(2):
        x1.fill_bytes(0)
        I.solve_graph(mat, None, prec, b1, x1, cfg)   # captured with k = 1
        I.synchronize()
        assert np.array_equal(x1.numpy(), eager), (path, rep, "before regrowth")
        I.solve(mat, None, prec, b3, x3, k=3, cfg=cfg)  # regrows the k-sized workspaces
        I.synchronize()
        assert I.check_watchdog(mat) == 0, L.last_error()
        x1.fill_bytes(0)
        I.solve_graph(mat, None, prec, b1, x1, cfg)   # replayed (re-captured) after the regrowth
        I.synchronize()
        assert I.check_watchdog(mat) == 0, L.last_error()
        assert np.array_equal(x1.numpy(), eager), (path, rep, "after regrowth")
    for dvec in (b1, x1, b3, x3):
        dvec.free()
    I.free_device_analysis(mat, None)


@pytest.mark.parametrize("k", [1, 3, 16, 64])
def test_jacobi_multi_rhs(k):
    s = I.generate(4, 20)
    rng = np.random.default_rng(k)
    B = rng.uniform(-1, 1, s.n * k)
    ref = O.fwd_f64(s.ia, s.ja, s.a, B, k=k)
    d = Dev(s.ia, s.ja, s.a, analyse=False)
    x = d.solve(B, k=k, lanes=4, sweeps=12, outer=4, **J)
    assert relerr(x, ref) <= 1e-10, k
    d.close()


def test_jacobi_rejects_non_dominant():
    ia = np.array([0, 1, 3, 6], np.int32)
    ja = np.array([0, 0, 1, 0, 1, 2], np.int32)
    a = np.array([1.0, -0.75, 1.0, -0.75, -0.75, 1.0])      # row 2: 1 - 1.5 < 0
    d = Dev(ia, ja, a, analyse=False)
    assert I.dominance_margin(d.mat) == pytest.approx(-0.5)
    db = I.DeviceArray.from_numpy(np.ones(3))
    dx = I.DeviceArray(3, np.float64)
    with pytest.raises(L.SptrsvError) as e:
        I.solve(d.mat, None, I.PREC_MIXED, db, dx, kernel=I.KERNEL_JACOBI, lanes=2, sweeps=4, outer=2)
    assert e.value.status == L.E_INVALID_ARG
    # the sync-free kernels still solve it
    I.solve(d.mat, d.info, I.PREC_MIXED, db, dx, kernel=I.KERNEL_SIMPLE)
    ref = O.fwd_f64(ia, ja, a, np.ones(3))
    assert relerr(dx.numpy(), ref) <= 1e-12
    db.free()
    dx.free()


def test_jacobi_autotune_and_batch_members():
    for m in range(4):
        s = I.generate_batch_member(m, 2000, 20000, seed=7)
        ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
        d = Dev(s.ia, s.ja, s.a)
        cfg, ms = I.autotune(d.mat, d.info, I.PREC_MIXED, reps=2)
        x = d.solve(s.b, precision=I.PREC_MIXED, kernel=I.KERNEL_AUTO)
        assert relerr(x, ref) <= 1e-10, (m, cfg.kernel)
        if cfg.kernel == I.KERNEL_JACOBI:
            assert cfg.sweeps >= 2 and cfg.outer >= 1
        d.close()


def test_batch_per_member_jacobi_choice():
    """Banded members (one dependency level per row) go to Jacobi, the rest
    stay sync-free; both halves run in one captured graph, and every member
    meets the mixed bar on its own (per-member relative error)."""
    systems = [I.generate_batch_member(m, 3000, 12000, seed=11) for m in range(8)]
    bt = I.Batch(systems)
    B = np.concatenate([s.b for s in systems])
    refs = [O.fwd_f64(s.ia, s.ja, s.a, s.b) for s in systems]
    db = I.DeviceArray.from_numpy(B)
    dx = I.DeviceArray(bt.rows, np.float64)
    bt.tune(I.PREC_MIXED, tune=True)
    kinds = [bt.choice(m).kernel for m in range(bt.nmat)]
    print("choices:", kinds, flush=True)
    for graph in (False, True, True, True):
        dx.fill_bytes(0)
        bt.solve(I.PREC_MIXED, db, dx, graph=graph)
        I.synchronize()
        assert bt.check() == 0, L.last_error()
        x = dx.numpy()
        for m, s in enumerate(systems):
            xm = x[bt.row_off[m]:bt.row_off[m + 1]]
            assert relerr(xm, refs[m]) <= 1e-10, (m, kinds[m], graph)
    # every schedule the caller may set gives the same answers: halves
    # concurrent or serial, sync-free members as tasks or as a separate
    # analysed matrix (persistent kernels claim their work in order, so a
    # persistent one runs next to the concurrent Jacobi sweeps too)
    print("tuned schedule:", bt.schedule(), flush=True)
    for serial in (0, 1):
        for separate in (0, 1):
            try:
                bt.set_schedule(serial, separate)
            except L.SptrsvError as e:
                # refused only when there is no sync-free member to separate
                all_jacobi = all(k == I.KERNEL_JACOBI for k in kinds)
                assert e.status == L.E_INVALID_ARG and separate == 1 and all_jacobi, kinds
                continue
            for graph in (False, True):
                dx.fill_bytes(0)
                bt.solve(I.PREC_MIXED, db, dx, graph=graph)
                I.synchronize()
                assert bt.check() == 0, L.last_error()
                x = dx.numpy()
                for m, s in enumerate(systems):
                    xm = x[bt.row_off[m]:bt.row_off[m + 1]]
                    assert relerr(xm, refs[m]) <= 1e-10, (m, serial, separate, graph)
    # FP64 batches stay sync-free for every member
    bt.solve(I.PREC_F64, db, dx, graph=False)
    I.synchronize()
    x = dx.numpy()
    for m in range(bt.nmat):
        assert relerr(x[bt.row_off[m]:bt.row_off[m + 1]], refs[m]) <= 1e-12
    db.free()
    dx.free()
    bt.close()


@pytest.mark.parametrize("cfg,param,k", [(2, 150000, 1), (3, 100000, 1), (4, 20, 16), (4, 20, 3)])
def test_jacobi_tile_shapes(cfg, param, k):
    """Every sweep shape gives the certified answer: entries in storage or
    column-sorted order, 1024- or 2048-product tiles, phase-2 lanes summing
    1..16 entries."""
    s = I.generate(cfg, param)
    rng = np.random.default_rng(cfg + k)
    B = s.b if k == s.k else rng.uniform(-1, 1, s.n * k)
    ref = O.fwd_f64(s.ia, s.ja, s.a, B, k=k)
    d = Dev(s.ia, s.ja, s.a, analyse=False)
    for variant in (0, 1):
        for tile in (1, 2):
            for epl in (1, 8, 16):
                x = d.solve(B, k=k, lanes=tile, variant=variant, window=epl, sweeps=16, outer=4, **J)
                assert relerr(x, ref) <= 1e-10, (variant, tile, epl)
    d.close()


def short_row_matrix(n, seed):
    """Lower CSR whose first 6000 rows are diagonal-only (tiles of exactly
    TW / K rows, the most a tile can hold) and whose other rows hold at most
    one off-diagonal entry; diag 1 + |off|."""
    rng = np.random.default_rng(seed)
    ia, ja, a = [0], [], []
    for i in range(n):
        if i >= 6000 and rng.random() < 0.3:
            j = int(rng.integers(max(0, i - 3000), i))
            v = float(rng.uniform(-1, 1))
            ja += [j, i]
            a += [v, 1.0 + abs(v)]
        else:
            ja.append(i)
            a.append(float(rng.uniform(1, 2)))
        ia.append(len(ja))
    return np.array(ia, np.int32), np.array(ja, np.int32), np.array(a)


@pytest.mark.parametrize("k", [1, 3])
def test_jacobi_tiles_full_of_one_entry_rows(k):
    """Tiles holding the most rows a tile can (one entry each) stage their
    row pointers and right-hand sides in LDS sized for the table's longest
    tile: every width, entry layout and (one right-hand side) the block
    sweep meet the mixed bar."""
    ia, ja, a = short_row_matrix(40000, k)
    n = len(ia) - 1
    rng = np.random.default_rng(10 + k)
    B = rng.uniform(-1, 1, n * k)
    ref = O.fwd_f64(ia, ja, a, B, k=k)
    d = Dev(ia, ja, a, analyse=False)
    for variant in ((0, 1, 2) if k == 1 else (0, 1)):   # block sweeps take one right-hand side
        for tile in (1, 2):
            x = d.solve(B, k=k, lanes=tile, variant=variant, sweeps=12, outer=3, **J)
            assert relerr(x, ref) <= 1e-10, (variant, tile)
    d.close()


@pytest.mark.parametrize("cfg,param,tiles", [(2, 150000, (1, 2)), (3, 60000, (1, 2)), (2, 1_100_000, (2,))])
def test_sorted_tiles_equal_storage_order_bitwise(cfg, param, tiles):
    """Column-sorted tiles only reorder the gathers: every product lands in
    its storage slot and rows are summed in storage order, so the solve is
    bit-identical to the storage-order sweep, for either tile width.  The
    sizes cover the position packed into the column word (n <= 2^21 for
    1024-product tiles, n <= 2^20 for 2048) and the separate position array
    (2048-product tiles at n > 2^20)."""
    s = I.generate(cfg, param)
    d = Dev(s.ia, s.ja, s.a, analyse=False)
    for tile in tiles:
        xs = [d.solve(s.b, lanes=tile, variant=v, window=8, sweeps=12, outer=3, **J) for v in (0, 1)]
        assert np.array_equal(xs[0], xs[1]), tile
    d.close()


def test_graph_captured_solves_replay_exactly():
    """sptrsv_solve_graph: every replay, with eager solves of other
    configurations in between, equals the eager solve bit for bit."""
    s = I.generate(2, 60000)
    d = Dev(s.ia, s.ja, s.a)
    db = I.DeviceArray.from_numpy(s.b)
    dx = I.DeviceArray(s.n, np.float64)
    cfgs = [(I.PREC_F64, I._cfg(I.KERNEL_SIMPLE, 16, 256, 1, 1, 4)),
            (I.PREC_MIXED, I._cfg(I.KERNEL_MULTIROW, 0, 256, 1, 0, 0)),
            (I.PREC_MIXED, I._cfg(I.KERNEL_JACOBI, 2, 256, 1, 0, 0, 12, 2, 8))]
    eager = []
    for prec, c in cfgs:
        I.solve(d.mat, d.info, prec, db, dx, cfg=c)
        I.synchronize()
        eager.append(dx.numpy())
    for rep in range(3):
        for j, (prec, c) in enumerate(cfgs):
            dx.fill_bytes(0)
            I.solve_graph(d.mat, d.info, prec, db, dx, c)
            I.synchronize()
            assert I.check_watchdog(d.mat) == 0
            assert np.array_equal(dx.numpy(), eager[j]), (rep, j)
            I.solve(d.mat, d.info, I.PREC_F64, db, dx, kernel=I.KERNEL_ORDER)   # eager work between replays
    db.free()
    dx.free()
    d.close()


# ---- block sweeps (variant 2): exact solves of each tile's diagonal block ------------

def block_sweep_np(ia, ja, a, d_old, rs, window, cap, inner):
    """One block sweep on the row-scaled system, restated in float64: tile t =
    rows whose first entry lies in [t*window, (t+1)*window);
      base_i = rs_i - sum_{j<r0} N'_ij d_old_j,  z^0 = d_old,
      z^s_i = base_i - sum_{r0<=j<i} N'_ij z^{s-1}_j,  d = z^inner;
    a tile of more than `cap` entries takes the point update."""
    import scipy.sparse as sp
    n = len(ia) - 1
    diag = a[ia[1:] - 1]
    rows = np.repeat(np.arange(n), np.diff(ia))
    off = ja != rows
    tile = ia[:-1] // window
    vals = a[off] / diag[rows[off]]
    Np = sp.csr_matrix((vals, (rows[off], ja[off])), shape=(n, n))
    same = tile[rows[off]] == tile[ja[off]]
    Nin = sp.csr_matrix((vals[same], (rows[off][same], ja[off][same])), shape=(n, n))
    Nout = Np - Nin
    base = rs - Nout @ d_old
    z = d_old.copy()
    for _ in range(inner):
        z = base - Nin @ z
    # tiles beyond LDS capacity: point update
    ent = np.bincount(tile, weights=np.diff(ia), minlength=tile.max() + 1)
    big = ent[tile] > cap
    d = np.where(big, rs - Np @ d_old, z)
    return d


@pytest.mark.parametrize("src", ["banded", "random", "powerlaw"])
@pytest.mark.parametrize("tile", [1, 2])
def test_block_sweep_matches_restatement(src, tile):
    s = {"banded": lambda: I.generate_batch_member(0),
         "random": lambda: I.generate(2, 40000),
         "powerlaw": lambda: I.generate(3, 40000)}[src]()
    d = Dev(s.ia, s.ja, s.a)
    window, cap = 1024 * tile, 1280 * tile
    rs = s.b / s.a[s.ia[1:] - 1]
    for inner in (1, 4, 16):
        d1 = block_sweep_np(s.ia, s.ja, s.a, rs, rs, window, cap, inner)
        d2 = block_sweep_np(s.ia, s.ja, s.a, d1, rs, window, cap, inner)
        for sweeps, want in ((2, d1), (3, d2)):
            # one step, no certification expected: x is the last sweep's iterate
            db = I.DeviceArray.from_numpy(s.b)
            dx = I.DeviceArray(s.n, np.float64)
            I.solve(d.mat, None, I.PREC_MIXED, db, dx, kernel=I.KERNEL_JACOBI, lanes=tile, variant=2,
                    sweeps=sweeps, outer=1, inner=inner)
            I.check_watchdog(d.mat)   # the certificate may fail here: only the iterate is compared
            assert relerr(dx.numpy(), want) <= 2e-5, (src, tile, inner, sweeps)
    d.close()


@pytest.mark.parametrize("member,sweeps", [(0, 6), (2, 6), (3, 16)])
def test_block_jacobi_certified(member, sweeps):
    # banded members (even) need few block sweeps; a random member (odd)
    # couples tiles everywhere and needs about as many as point sweeps
    s = I.generate_batch_member(member)
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    d = Dev(s.ia, s.ja, s.a)
    for tile in (1, 2):
        x = d.solve(s.b, lanes=tile, variant=2, sweeps=sweeps, outer=3, **J)
        assert relerr(x, ref) <= 1e-10, (member, tile)
    d.close()


def test_block_sweeps_beat_point_sweeps_on_banded():
    s = I.generate_batch_member(0)           # banded: nearly every dependency inside a tile
    ref = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    d = Dev(s.ia, s.ja, s.a)
    errs = {}
    for var in (0, 2):
        db = I.DeviceArray.from_numpy(s.b)
        dx = I.DeviceArray(s.n, np.float64)
        I.solve(d.mat, None, I.PREC_MIXED, db, dx, kernel=I.KERNEL_JACOBI, lanes=1, variant=var, sweeps=4, outer=1)
        I.check_watchdog(d.mat)
        errs[var] = relerr(dx.numpy(), ref)
    assert errs[2] < 1e-2 * errs[0], errs   # 4 in-tile iterations per sweep by default
    cfg, _ = I.autotune(d.mat, None, I.PREC_MIXED)
    db = I.DeviceArray.from_numpy(s.b)
    dx = I.DeviceArray(s.n, np.float64)
    I.solve(d.mat, None, I.PREC_MIXED, db, dx, cfg=cfg)
    assert I.check_watchdog(d.mat) == 0
    assert relerr(dx.numpy(), ref) <= 1e-10
    print("tuned", cfg.kernel, cfg.variant, cfg.sweeps, cfg.outer, flush=True)
    d.close()
