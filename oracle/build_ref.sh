#!/usr/bin/env bash
# Builds the reference Fortran (farzan-t/3D_Droplet_simulation_2007, main.f)
# from its sources where they lie under /root/reference, one shared library
# per grid size, into oracle/_ref/libref_<NX>x<NY>x<NZ>.so.
#
# The reference fixes its grid size at compile time through the PARAMETERs
# NGM/NGMX/NGMY/NGMZ of input.h (input.h:11,26-28; readinput checks
# nx == NMX at main.f:240-242).  For each size this script writes a header
# with those four values changed into a temporary directory, next to a
# symlink to the unmodified main.f (Fortran INCLUDE searches the including
# file's directory first), compiles, and deletes the temporary directory.
# Only the built .so lands in oracle/_ref/.  Nothing of the reference is
# committed.
#
# Flags: -Ddsind=sind etc. map the GNU degree-trig extensions (used only by
# the contact-angle diagnostics, main.f:1026ff) onto the F2023 intrinsics.
# No -ffast-math and no -march: the default x86-64 target has no FMA, so
# every FP64 operation is a separately rounded IEEE operation.
#
# usage: oracle/build_ref.sh NGMX,NGMY,NGMZ [...]     e.g.  4,4,4  6,5,5
set -euo pipefail
REF=/root/reference
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
FLANG=${FLANG:-/opt/rocm/lib/llvm/bin/flang}
if [ ! -f "$REF/main.f" ]; then
  echo "build_ref: $REF/main.f not present; skipping" >&2
  exit 0
fi
if [ ! -x "$FLANG" ]; then
  echo "build_ref: flang not found at $FLANG" >&2
  exit 1
fi
mkdir -p "$OUT"
for spec in "$@"; do
  IFS=, read -r gx gy gz <<< "$spec"
  nx=$(( (1 << gx) + 2 )); ny=$(( (1 << gy) + 2 )); nz=$(( (1 << gz) + 2 ))
  g=$gx; [ "$gy" -gt "$g" ] && g=$gy; [ "$gz" -gt "$g" ] && g=$gz
  lib="$OUT/libref_${nx}x${ny}x${nz}.so"
  if [ -f "$lib" ] && [ "$lib" -nt "$REF/main.f" ] && [ "$lib" -nt "$0" ]; then
    continue
  fi
  tmp=$(mktemp -d /tmp/droplet_ref.XXXXXX)
  ln -s "$REF/main.f" "$tmp/main.f"
  sed -e "s/PARAMETER (NGM=[0-9]*)/PARAMETER (NGM=$g)/" \
      -e "s/NGMX=[0-9]*,/NGMX=$gx,/" \
      -e "s/NGMY=[0-9]*,/NGMY=$gy,/" \
      -e "s/NGMZ=[0-9]*,/NGMZ=$gz,/" "$REF/input.h" > "$tmp/input.h"
  extra=""
  # COMMON blocks above 2 GiB need the medium code model
  if [ "$g" -ge 8 ]; then extra="-mcmodel=medium"; fi
  "$FLANG" -O2 -fPIC -shared $extra -cpp -Ddsind=sind -Ddcosd=cosd -Ddatand=atand \
      -o "$lib.tmp" "$tmp/main.f"
  mv "$lib.tmp" "$lib"
  rm -rf "$tmp"
  echo "build_ref: built $lib"
done
