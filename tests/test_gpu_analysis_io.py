"""GPU tests of analysis reuse: sptrsv_save_analysis / sptrsv_load_analysis
(a checksummed analysis file) and sptrsv_copy_analysis (the analysis of one
matrix given to another with the same structure).

The reference recomputes its analysis on every run (analysis_csr.cu:272-410
into dfr_syncfree.h:25-41).  Bars: every array of a loaded or copied
analysis is bit-exact with the original (and with the oracle's analysis),
solves through it are bit-identical to solves through the original, and a
damaged, truncated or foreign file is refused.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as O
from sptrsv_amd import _lib as L
from sptrsv_amd import interface as I

pytestmark = pytest.mark.gpu

KEYS = ("level", "iorder", "inv_iorder", "lev_size", "bucket_off")
WARP_KEYS = ("warp_lev", "ibase_row", "ivect_size_warp", "ibase_row64", "ivect_size_wave")


def relerr(x, ref):
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


def same_arrays(a, b, warps):
    assert a["nlevs"] == b["nlevs"] and a["n_warps"] == b["n_warps"] and a["n_waves"] == b["n_waves"]
    for k in KEYS + (WARP_KEYS if warps else ()):
        assert np.array_equal(a[k], b[k]), k


def solve(mat, info, b, kernel):
    db = I.DeviceArray.from_numpy(b)
    dx = I.DeviceArray(len(b), np.float64)
    I.solve(mat, info, I.PREC_F64, db, dx, kernel=kernel)
    assert I.check_watchdog(mat) == 0, L.last_error()
    x = dx.numpy()
    db.free()
    dx.free()
    return x


@pytest.mark.parametrize("mode", [I.ALL, I.ORDER])
def test_save_load_round_trip(tmp_path, mode):
    print(f"analysis file round trip, mode {mode}, power-law n=300000", flush=True)
    s = I.generate(3, 300000)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, mode)
    path = tmp_path / "analysis.bin"
    I.save_analysis(path, mat, info)
    loaded = I.load_analysis(path, mat)
    a, b = I.analysis_arrays(info), I.analysis_arrays(loaded)
    warps = mode != I.ORDER
    same_arrays(a, b, warps)
    ref = O.analysis(s.ia, s.ja, 32)
    assert np.array_equal(b["level"], ref["level"]) and np.array_equal(b["iorder"], ref["iorder"])
    if warps:
        assert np.array_equal(b["ibase_row"], ref["ibase_row"])
    kernels = (I.KERNEL_ORDER, I.KERNEL_MULTIROW) if warps else (I.KERNEL_ORDER,)
    xr = O.fwd_f64(s.ia, s.ja, s.a, s.b)
    for kern in kernels:
        x1 = solve(mat, info, s.b, kern)
        x2 = solve(mat, loaded, s.b, kern)
        assert np.array_equal(x1.view(np.uint64), x2.view(np.uint64)), kern
        assert relerr(x2, xr) <= 1e-12
    # loading into a struct that holds an analysis replaces it
    I.load_analysis(path, mat, loaded)
    same_arrays(a, I.analysis_arrays(loaded), warps)
    I.free_device_analysis(None, loaded)
    I.free_device_analysis(mat, info)


def test_damaged_and_foreign_files_are_refused(tmp_path):
    print("analysis file checks", flush=True)
    s = I.generate(2, 50000)
    mat = I.allocate_memory(s.ia, s.ja, s.a)
    info = I.multirow_analysis_base_GPU(mat, I.ALL)
    path = tmp_path / "a.bin"
    I.save_analysis(path, mat, info)
    raw = path.read_bytes()
    # a flipped bit in the arrays
    bad = bytearray(raw)
    bad[len(raw) // 2] ^= 0x10
    (tmp_path / "flip.bin").write_bytes(bytes(bad))
    with pytest.raises(L.SptrsvError) as e:
        I.load_analysis(tmp_path / "flip.bin", mat)
    assert e.value.status == L.E_IO
    # truncated
    (tmp_path / "cut.bin").write_bytes(raw[:-100])
    with pytest.raises(L.SptrsvError) as e:
        I.load_analysis(tmp_path / "cut.bin", mat)
    assert e.value.status == L.E_IO
    # not an analysis file
    (tmp_path / "junk.bin").write_bytes(b"\x00" * 200)
    with pytest.raises(L.SptrsvError) as e:
        I.load_analysis(tmp_path / "junk.bin", mat)
    assert e.value.status == L.E_IO
    # same size, another structure: the last off-diagonal column of a row moved
    ja2 = s.ja.copy()
    r = next(i for i in range(1000, s.n) if s.ia[i + 1] - s.ia[i] >= 3 and s.ja[s.ia[i + 1] - 2] + 1 < i)
    ja2[s.ia[r + 1] - 2] += 1
    other = I.allocate_memory(s.ia, ja2, s.a)
    with pytest.raises(L.SptrsvError) as e:
        I.load_analysis(path, other)
    assert e.value.status == L.E_INVALID_ARG
    with pytest.raises(L.SptrsvError) as e:
        I.copy_analysis(mat, info, other)
    assert e.value.status == L.E_INVALID_ARG
    # another size
    small = I.generate(1)
    m3 = I.allocate_memory(small.ia, small.ja, small.a)
    with pytest.raises(L.SptrsvError) as e:
        I.load_analysis(path, m3)
    assert e.value.status == L.E_INVALID_ARG
    for m in (other, m3):
        I.free_device_analysis(m, None)
    I.free_device_analysis(mat, info)


def test_copy_analysis_to_a_matrix_on_the_same_pattern():
    print("analysis copied to another system on the same pattern", flush=True)
    s = I.generate(2, 200000)
    rng = np.random.default_rng(9)
    a2 = s.a.copy()
    diag = s.ia[1:] - 1
    off = np.ones(len(a2), bool)
    off[diag] = False
    a2[off] = rng.uniform(-1, 1, off.sum())
    rows = np.repeat(np.arange(s.n), np.diff(s.ia))
    a2[diag] = 1.0 + np.bincount(rows[off], weights=np.abs(a2[off]), minlength=s.n)
    m1 = I.allocate_memory(s.ia, s.ja, s.a)
    m2 = I.allocate_memory(s.ia, s.ja, a2)
    info1 = I.multirow_analysis_base_GPU(m1, I.ALL)
    info2 = I.copy_analysis(m1, info1, m2)
    same_arrays(I.analysis_arrays(info1), I.analysis_arrays(info2), True)
    fresh = I.multirow_analysis_base_GPU(m2, I.ALL)
    same_arrays(I.analysis_arrays(fresh), I.analysis_arrays(info2), True)
    xr = O.fwd_f64(s.ia, s.ja, a2, s.b)
    for kern in (I.KERNEL_ORDER, I.KERNEL_MULTIROW):
        x = solve(m2, info2, s.b, kern)
        assert np.array_equal(x.view(np.uint64), solve(m2, fresh, s.b, kern).view(np.uint64))
        assert relerr(x, xr) <= 1e-12
    # the copy is independent: freeing the source leaves it usable
    I.free_device_analysis(m1, info1)
    x = solve(m2, info2, s.b, I.KERNEL_MULTIROW)
    assert relerr(x, xr) <= 1e-12
    I.free_device_analysis(None, fresh)
    I.free_device_analysis(m2, info2)
