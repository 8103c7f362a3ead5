"""Build the gfx950 shared library libmps_sph.so in-tree with hipcc.

Usage: python -m mps_cuda_amd.build   (or mps_cuda_amd.build.build())
Compiles every translation unit every time it is called (in parallel), then links.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmps_sph.so")

SOURCES = ["hash.hip", "sph.hip", "halo.hip", "sampling.hip", "adjoint.hip", "capi.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPS_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: never emit v_pk_{add,mul,fma}_f32.
# -ffp-contract=off + correctly rounded div/sqrt: FP32 results follow the reference's
# expression order (the kernels also say `#pragma clang fp contract(off)`).
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _compile(src: str, csrc: str = CSRC, out_dir: str = OUT_DIR) -> str:
    obj = os.path.join(out_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", os.path.join(csrc, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True, csrc: str = CSRC, lib: str = LIB, out_dir: str = OUT_DIR) -> str:
    """Compile `csrc` (default: the package's sources) into the shared library `lib`."""
    os.makedirs(out_dir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda f: _compile(f, csrc, out_dir), SOURCES))
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"[mps_cuda_amd] built {lib}", file=sys.stderr)
    if lib == LIB:
        build_facade_driver(verbose)
    return lib


REPO = os.path.dirname(HERE)
FACADE_SRC = os.path.join(REPO, "tests", "cpp", "facade_scene.cpp")
FACADE_BIN = os.path.join(OUT_DIR, "facade_scene")


def build_facade_driver(verbose: bool = True) -> str:
    """Host-only C++ program that drives the library through include/mps_system.hpp (the
    reference's System/SPHMaterial/MeshObject/... names); found next to the library at run time."""
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-Wextra", "-Wpedantic", "-Wshadow", "-Werror",
           "-I", os.path.join(REPO, "include"),
           FACADE_SRC, "-L", HERE, "-lmps_sph", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib",
           "-o", FACADE_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade driver build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"[mps_cuda_amd] built {FACADE_BIN}", file=sys.stderr)
    return FACADE_BIN


if __name__ == "__main__":
    build()
