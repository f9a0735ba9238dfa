#!/usr/bin/env python3
"""Links the reference program against libdroplet_gpu.so: the drop-in check.

Reads main.f where it lies under /root/reference, leaves out the program
units whose call forms libdroplet_gpu.so defines (the time step's routines
that PROGRAM THREE DIM's loop calls, main.f:19-73, and those only they
call), compiles the rest with ROCm flang next to the reference's unmodified
input.h, and links it against the library.  Nothing else of main.f is
changed.  Output: oracle/_ref/three_dim_dropin (git-ignored), which the
GPU test of printout's files runs.  The temporary copy of main.f is
deleted.

usage: python oracle/link_dropin.py
Exit status 0 when the link succeeds; prints the undefined symbols the
library resolved.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
import tempfile

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
PKG = os.path.join(REPO, "3d_droplet_simulation_2007_amd")
OUT = os.path.join(HERE, "_ref", "three_dim_dropin")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

# program units the library replaces (its exported Fortran call forms)
REPLACED = ("vofmethod", "swp", "calch3d", "eddyvisc", "kepsmodel", "uv2p", "checktime", "cfl", "csf", "sgip",
            "sgip2", "solvelpres", "convective", "mglin")

_UNIT = re.compile(r"^[ \t]+(?:(?:double\s+precision|integer|real|logical)\s+)?"
                   r"(subroutine|function|program)\s+(\w+)", re.IGNORECASE)


def units(lines):
    """(name, first line, end line) of each program unit of a fixed-form
    file: a unit runs from its header to the line before the next header."""
    heads = []
    for q, ln in enumerate(lines):
        if ln[:1] in ("c", "C", "*", "!"):
            continue
        m = _UNIT.match(ln)
        if m:
            heads.append((m.group(2).lower(), q))
    out = []
    for k, (name, q) in enumerate(heads):
        end = heads[k + 1][1] if k + 1 < len(heads) else len(lines)
        out.append((name, q, end))
    return out


def main():
    if not os.path.exists(os.path.join(REF, "main.f")):
        print("link_dropin: /root/reference/main.f not present; skipping", file=sys.stderr)
        return 0
    lib = os.path.join(PKG, "libdroplet_gpu.so")
    if not os.path.exists(lib):
        print("link_dropin: build libdroplet_gpu.so first", file=sys.stderr)
        return 1
    with open(os.path.join(REF, "main.f"), errors="replace") as f:
        lines = f.readlines()
    us = units(lines)
    found = {n for n, _, _ in us}
    missing = [n for n in REPLACED if n not in found]
    if missing:
        print(f"link_dropin: units not found in main.f: {missing}", file=sys.stderr)
        return 1
    drop = set()
    for name, q, end in us:
        if name in REPLACED:
            drop.update(range(q, end))
    kept = [ln for q, ln in enumerate(lines) if q not in drop]
    tmp = tempfile.mkdtemp(prefix="dropin_link.")
    try:
        src = os.path.join(tmp, "main_dropin.f")
        with open(src, "w") as f:
            f.writelines(kept)
        os.symlink(os.path.join(REF, "input.h"), os.path.join(tmp, "input.h"))
        obj = os.path.join(tmp, "main_dropin.o")
        cc = [FLANG, "-O2", "-cpp", "-Ddsind=sind", "-Ddcosd=cosd", "-Ddatand=atand", "-c", src, "-o", obj]
        r = subprocess.run(cc, capture_output=True, text=True, cwd=tmp)
        if r.returncode != 0:
            print(r.stderr[-4000:], file=sys.stderr)
            return 1
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        ld = [FLANG, "-o", OUT + ".tmp", obj, f"-L{PKG}", "-ldroplet_gpu", f"-Wl,-rpath,{PKG}"]
        r = subprocess.run(ld, capture_output=True, text=True, cwd=tmp)
        if r.returncode != 0:
            print(r.stderr[-4000:], file=sys.stderr)
            return 1
        os.replace(OUT + ".tmp", OUT)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    print(f"link_dropin: linked {OUT} with {len(REPLACED)} units of main.f taken from libdroplet_gpu.so: "
          + ", ".join(REPLACED))
    return 0


if __name__ == "__main__":
    sys.exit(main())
