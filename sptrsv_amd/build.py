"""Builds libsptrsv_amd.so (gfx950) in-tree with hipcc, in parallel.

    python -m sptrsv_amd.build            # library only
    python -m sptrsv_amd.build --oracle   # also the CPU oracle (tests only)

Every HIP translation unit is compiled with packed-FP32 instructions
disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = ROOT / "_build" / "sptrsv_amd"
LIB = PKG / "libsptrsv_amd.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = os.environ.get("SPTRSV_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{CSRC}",
             f"-I{ROCM / 'include'}", "-D__HIP_PLATFORM_AMD__"]

HIP_SOURCES = ["inst_f64.hip", "inst_f32.hip", "inst_mixed.hip", "inst_level.hip",
               "api_solve.hip", "analysis.hip", "interface.hip", "batch.hip", "iterative.hip",
               "transpose.hip", "ingest.hip", "spsv.hip"]
HOST_HIP_SOURCES = ["power.cpp", "rocsparse_baseline.cpp"]   # need HIP headers, no kernels
CXX_SOURCES = ["mmio.cpp", "generate.cpp"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{err}")
    if err.strip():
        sys.stderr.write(err + "\n")


def build_library(jobs: int | None = None) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    steps = []
    for src in HIP_SOURCES:
        steps.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(BUILD / (src + ".o"))])
    for src in HOST_HIP_SOURCES:
        steps.append([HIPCC, "-x", "hip", "--cuda-host-only", *HIP_FLAGS, "-c", str(CSRC / src),
                      "-o", str(BUILD / (src + ".o"))])
    for src in CXX_SOURCES:
        steps.append(["g++", *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(BUILD / (src + ".o"))])
    jobs = jobs or min(16, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, s) for s in steps]:
            f.result()
    objs = [str(BUILD / (s + ".o")) for s in HIP_SOURCES + HOST_HIP_SOURCES + CXX_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp),
          "-ldl", "-lpthread", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    """CPU restatement used only by tests/, smoke() and bench's cpu_baseline."""
    out = ROOT / "oracle" / "liboracle.so"
    cc = shutil.which("gcc") or "cc"
    _run([cc, "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-Wall", "-Wextra",
          str(ROOT / "oracle" / "oracle.c"), "-o", str(out)])
    return out


def main(argv: list[str]) -> int:
    build_library()
    print(f"built {LIB}")
    if "--oracle" in argv:
        print(f"built {build_oracle()}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main(sys.argv[1:]))
