"""Builds libspaqr_gpu.so in-tree: hipcc, gfx950 device code, one shared library.

Objects are compiled in parallel into pspaqr_public_amd/_build/ and linked into
pspaqr_public_amd/libspaqr_gpu.so.  Run ``python -m pspaqr_public_amd.build``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libspaqr_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPAQR_ARCH", "gfx950")

SOURCES = ["kernels.hip", "resident.hip", "learned.hip", "krylov.hip", "rrqr.cpp", "tree.cpp", "tree_rank.cpp",
           "tree_batch.cpp", "status.hip", "capi.cpp", "noalloc.cpp", "mmio.cpp", "symbolic.hip"]

# per-file flags: the resident-accumulator applies keep MFMA accumulators in VGPRs
EXTRA = {"resident.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # packed-FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + EXTRA.get(src, []) + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


CONSUMER_SRC = os.path.join(os.path.dirname(HERE), "tests", "cxx", "ptree_consumer.cpp")
CONSUMER = os.path.join(os.path.dirname(HERE), "tests", "cxx", "ptree_consumer")


def build_consumer(verbose: bool = False) -> str:
    """The C++ test program that drives the library through include/spaqr_ptree.hpp (plain g++,
    linked against libspaqr_gpu.so, found next to it at run time through its rpath)."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(os.path.dirname(HERE), "include"), CONSUMER_SRC,
           "-L", HERE, "-lspaqr_gpu", "-Wl,-rpath,$ORIGIN/../../pspaqr_public_amd", "-o", CONSUMER]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("consumer build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    if verbose:
        print("built", CONSUMER)
    return CONSUMER


if __name__ == "__main__":
    build(verbose=True)
    build_consumer(verbose=True)
    sys.exit(0)
