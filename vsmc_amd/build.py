"""Builds libvsmch.so (all HIP sources of vsmc_amd/csrc) for gfx950, in-tree.

Run: python -m vsmc_amd.build  (or via __graft_entry__.build()).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libvsmch.so")
ARCH = os.environ.get("VSMCH_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a * b + c stays two roundings, as in the reference's x86-64 build:
    # __dmul_rn / __dadd_rn are plain * and + in HIP's headers, which the
    # default contraction would fuse into an FMA
    "-ffp-contract=off",
    "-Wall", "-Wshadow", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


ROOT = os.path.dirname(HERE)
CPP_TESTS = [os.path.join(ROOT, "tests", "cpp", "pf_hip.cpp"),
             os.path.join(ROOT, "tests", "cpp", "rindex_hip.cpp"),
             os.path.join(ROOT, "tests", "cpp", "rw_hip.cpp"),
             os.path.join(ROOT, "tests", "cpp", "plug_hip.cpp")]


def build_cpp_tests(verbose=True):
    """C++ programs written against include/vsmch/sampler.hpp (the drop-in
    layer), linked to the in-tree libvsmch.so. -ffp-contract=off keeps the
    device functors' a * b + c rounded twice, as the reference's host build."""
    outs = []
    for src in CPP_TESTS:
        exe = os.path.splitext(src)[0]
        cmd = [HIPCC, "-O3", "-std=c++17", f"--offload-arch={ARCH}",
               "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-I", os.path.join(ROOT, "include"), src,
               "-L", HERE, "-lvsmch", "-Wl,-rpath,$ORIGIN/../../vsmc_amd",
               "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
        outs.append(exe)
        if verbose:
            print(f"built {exe}")
    return outs


if __name__ == "__main__":
    build()
    build_cpp_tests()
    sys.exit(0)
