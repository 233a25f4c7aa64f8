"""Build recipe of the HIP backend library (libdpcull_hip.so) and of the C++
host mirror linked into it.

Every HIP translation unit is compiled for gfx950 with:
  -ffp-contract=off                 no fused multiply-add (the reference builds without FMA)
  -fno-gpu-flush-denormals-to-zero  IEEE float denormals, as on the host
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                    no packed FP32 instructions
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIBNAME = "libdpcull_hip.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")

HIP_SOURCES = ["runtime.hip", "cull.hip", "tree.hip", "batch.hip"]
CXX_SOURCES = ["host/Manager.cpp", "host/Tree.cpp", "host/Serving.cpp"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wextra",
             f"-I{ROOT / 'include'}", f"-I{ROCM / 'include'}", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}")
    # hipcc's host pass does not know the packed-fp32 feature and says so; that is harmless.
    noise = ("packed-fp32-ops", "not a recognized feature")
    for line in (res.stdout + res.stderr).splitlines():
        if line.strip() and not any(n in line for n in noise):
            sys.stderr.write(line + "\n")


def build(verbose: bool = False) -> Path:
    """Compile every HIP and C++ source and link libdpcull_hip.so in-tree."""
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = objdir / (Path(src).stem + ".hip.o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in CXX_SOURCES:
        if not (CSRC / src).exists():
            continue
        obj = objdir / (Path(src).stem + ".cpp.o")
        jobs.append(["g++", *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()
    objs = [j[-1] for j in jobs]
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs,
          f"-L{ROCM / 'lib'}", "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM / 'lib'}", "-Wl,--no-undefined"])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
