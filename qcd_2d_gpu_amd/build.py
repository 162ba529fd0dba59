"""Build recipe of libsu2mc.so (gfx950) and the su2mc command-line driver.

Runs hipcc directly, in-tree: objects go to qcd_2d_gpu_amd/_build/, the
library and the executable to qcd_2d_gpu_amd/.  Every translation unit is
compiled with automatic floating-point contraction off, so the only fused
multiply-adds are the ones the update bodies write out (the bit-exact
contract with the CPU checker), and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libsu2mc.so")
EXE = os.path.join(PKG, "su2mc")

ARCH = os.environ.get("SU2MC_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

SOURCES = ["su2_sweep.hip", "su2_misc.hip", "su2_batch.hip", "su2_probe.hip", "su2_capi.cpp", "su2_handle.cpp",
           "su2_engine.cpp"]
C_SOURCES = ["su2_driver.c"]  # host driver in C, on the C ABI
CC = os.environ.get("CC", shutil.which("gcc") or "cc")
CFLAGS = ["-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-Wall", "-I", INCLUDE, "-I", CSRC]

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", INCLUDE,
    "-I", CSRC,
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")


def _obj(src: str) -> str:
    return os.path.join(BUILD, os.path.splitext(src)[0] + ".o")


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every source and link libsu2mc.so and su2mc. Returns the
    library path."""
    os.makedirs(BUILD, exist_ok=True)
    gen = os.path.join(CSRC, "su2_mrg32k3a_jump.h")
    with open(gen, "w") as f:
        subprocess.run([sys.executable, os.path.join(CSRC, "gen_mrg32k3a_jump.py")], stdout=f, check=True)

    def compile_one(src: str) -> None:
        cmd = [HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", _obj(src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)

    def compile_c(src: str) -> None:
        cmd = [CC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", _obj(src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, SOURCES))
        list(ex.map(compile_c, C_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *[_obj(s) for s in SOURCES + C_SOURCES]])
    os.replace(tmp, LIB)
    _run([CC, "-O2", "-std=c11", "-Wall", "-I", INCLUDE, os.path.join(CSRC, "su2mc_main.c"), "-o", EXE,
          "-L", PKG, "-lsu2mc", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
