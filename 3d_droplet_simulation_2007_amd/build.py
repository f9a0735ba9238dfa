"""Builds libdroplet_gpu.so (HIP, gfx950) in-tree, next to this file.

Every .hip source under csrc/ is compiled by hipcc for --offload-arch=gfx950
and linked into one shared library whose exported C-ABI is
include/droplet_gpu.h.  The compile is a cross-compile: no GPU is needed.

Flags that matter for results:
  -ffp-contract=off   no FMA contraction, so FP64 kernels round every
                      operation like the reference's x86-64 build;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions on gfx950.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(REPO, "include")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdroplet_gpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DG_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=False):
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_fortran_host(verbose)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
HOST_DIR = os.path.join(HERE, "host")
HOST_BIN = os.path.join(HERE, "droplet_host")
DROPIN_BIN = os.path.join(HERE, "dropin_f77")


def build_fortran_host(verbose=False):
    """The Fortran host (host/*.f90) over the C-ABI, linked with an $ORIGIN
    rpath so it finds libdroplet_gpu.so next to it."""
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found: Fortran host not built")
        return None
    os.makedirs(OBJDIR, exist_ok=True)
    mod_o = os.path.join(OBJDIR, "droplet_gpu_mod.o")
    host_o = os.path.join(OBJDIR, "droplet_host.o")
    steps = [
        [FLANG, "-O2", "-J", OBJDIR, "-c", os.path.join(HOST_DIR, "droplet_gpu_mod.f90"), "-o", mod_o],
        [FLANG, "-O2", "-J", OBJDIR, "-c", os.path.join(HOST_DIR, "droplet_host.f90"), "-o", host_o],
        [FLANG, "-o", HOST_BIN + ".tmp", host_o, mod_o, f"-L{HERE}", "-ldroplet_gpu", "-Wl,-rpath,$ORIGIN"],
    ]
    # the Fortran-77 program on the reference's own call forms (implicit
    # interfaces, COMMON blocks of input.h's layout)
    f77_dir = os.path.join(HOST_DIR, "f77")
    f77_o = os.path.join(OBJDIR, "dropin_f77.o")
    steps += [
        [FLANG, "-O2", "-I", f77_dir, "-c", os.path.join(f77_dir, "dropin.f"), "-o", f77_o],
        [FLANG, "-o", DROPIN_BIN + ".tmp", f77_o, f"-L{HERE}", "-ldroplet_gpu", "-Wl,-rpath,$ORIGIN"],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed: {' '.join(cmd)}\n{r.stderr[-4000:]}")
    os.replace(HOST_BIN + ".tmp", HOST_BIN)
    os.replace(DROPIN_BIN + ".tmp", DROPIN_BIN)
    if verbose:
        print(f"built {HOST_BIN} and {DROPIN_BIN}")
    return HOST_BIN


if __name__ == "__main__":
    try:
        build(verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
