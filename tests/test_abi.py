"""The C-ABI as a compiled C++ caller sees it: tests/abi/caller.cpp is built
with g++ against include/sptrsv_amd.h and linked to libsptrsv_amd.so.

Its static_asserts check that the reference's sp_mat_t and
dfr_analysis_info_t (dfr_syncfree.h:7-41) are a prefix of the library's at
the same offsets; a build with two fields swapped in the restated reference
struct must fail.  On the CPU the program calls only host-side entry points;
on the GPU it runs the reference's test flow (test.cu:4-30) and
test_solve_L_analysis_multirow (test.h:4, as main.cu:259 calls it) on the 2D
Laplacian factor, where x == 1 exactly.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

from sptrsv_amd import _lib as L

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "abi" / "caller.cpp"


def build_caller(out: Path, *defines: str) -> subprocess.CompletedProcess:
    L.lib()  # the library must be built
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    cmd = [gxx, "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", *[f"-D{d}" for d in defines],
           str(SRC), f"-L{L.LIB_PATH.parent}", "-lsptrsv_amd", f"-Wl,-rpath,{L.LIB_PATH.parent}", "-o", str(out)]
    return subprocess.run(cmd, capture_output=True, text=True)


def test_cpp_caller_compiles_links_and_calls_host_entry_points(tmp_path):
    exe = tmp_path / "caller"
    r = build_caller(exe)
    assert r.returncode == 0, r.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "host entry points ok: n = 4096, nnz = 12160" in run.stdout


def test_layout_checks_reject_a_swapped_field(tmp_path):
    r = build_caller(tmp_path / "bad", "PLANT_SWAPPED_FIELDS")
    assert r.returncode != 0
    assert "offset differs" in r.stderr


@pytest.mark.gpu
def test_cpp_caller_runs_the_reference_flow_on_gpu(tmp_path):
    print("C++ caller: reference test flow on the 64x64 2D Laplacian", flush=True)
    exe = tmp_path / "caller"
    r = build_caller(exe)
    assert r.returncode == 0, r.stderr
    run = subprocess.run([str(exe), "run"], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "reference flow: 0 errors; test_solve_L_analysis_multirow: 0; _f32: 0" in run.stdout
    assert run.stdout.count("[PASS!]") == 6, run.stdout
