"""Builds the native libraries in-tree.

* ``atrc_amd/libatrc_gpu.so`` — the gfx950 path tracer behind the C ABI of
  ``include/atrc_gpu.h`` (hipcc, ``--offload-arch=gfx950``).
* ``atrc_amd/libatrc_host.so`` and ``atrc_amd/atrc_cli`` — the C++ host: Atrc
  JSON config, factory and the ``gpu_pt`` renderer (g++; loads libatrc_gpu.so).
* ``oracle/liboracle.so`` — the CPU restatement the tests compare the GPU path with.

Every build compiles from scratch (no incremental skipping), so ``build()``
is also the "does it build" check.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "atrc_amd"
CSRC = PKG / "csrc"
HOST = PKG / "host"
ORACLE = ROOT / "oracle"

GPU_LIB = PKG / "libatrc_gpu.so"
HOST_LIB = PKG / "libatrc_host.so"
CLI = PKG / "atrc_cli"
ORACLE_LIB = ORACLE / "liboracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AGX_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: FP32 arithmetic exactly as written (bit parity with Atrc's
# operation order); packed FP32 instructions disabled for correctness.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-parameter", "-Wno-unused-function",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.stdout:
        print(res.stdout, end="", flush=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}")


def build_gpu() -> Path:
    srcs = [CSRC / "agx_kernels.hip", CSRC / "agx_guide.hip", CSRC / "agx_api.cpp", CSRC / "agx_build.cpp",
            CSRC / "agx_bvh4.cpp"]
    tmp = GPU_LIB.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-pthread", "-Wl,--no-undefined", "-o", str(tmp), *map(str, srcs)])
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_host() -> list[Path]:
    """C++ host: config, factory, gpu_pt renderer (linked against libatrc_gpu.so)."""
    srcs = sorted(str(p) for p in HOST.glob("*.cpp") if p.name != "cli_main.cpp")
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-parameter", f"-I{ROOT / 'include'}"]
    link = [f"-L{PKG}", "-latrc_gpu", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-pthread"]
    tmp = HOST_LIB.with_suffix(".so.tmp")
    _run(["g++", *flags, "-shared", "-o", str(tmp), *srcs, *link])
    os.replace(tmp, HOST_LIB)
    _run(["g++", *flags, "-o", str(CLI), str(HOST / "cli_main.cpp"), f"-L{PKG}", "-latrc_host", *link])
    return [HOST_LIB, CLI]


def build_oracle() -> Path:
    _run(["make", "-B", "-C", str(ORACLE)])
    return ORACLE_LIB


def build_all() -> None:
    build_gpu()
    build_host()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"gpu": build_gpu, "host": build_host, "oracle": build_oracle, "all": build_all}[w]()
