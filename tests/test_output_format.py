"""printout's text formats on the host (no GPU): the library's formatter
against what flang's runtime writes for the same values.

The reference's output goes through two Fortran edit forms: list-directed
WRITE (outputdata.out, the screen, the .plt headers) and the explicit
(1x,e15.7) of its formats 111 and 122 (main.f:5099-5100, 5150).  A small
Fortran program, compiled here with the same flang as the reference build,
writes a set of values both ways; dg_format_real must give the same text
for every one.  The values cover the special cases (zeros of both signs,
NaN, infinities, subnormals, the largest double), the F/E switch of
list-directed output around 0.1 and 1e15, rounding carries in E15.7
(9.99999995 -> 0.1000000E+02), ties (12345665 -> 0.1234566E+08) and
three-digit exponents, plus 4000 seeded values spread over the whole
exponent range.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

PROGRAM = """
program fmtprobe
  implicit none
  integer :: n, i
  double precision, allocatable :: v(:)
  open(1, file='v.bin', access='stream', form='unformatted', status='old')
  read(1) n
  allocate(v(n))
  read(1) v
  close(1)
  open(2, file='ld.txt')
  open(3, file='e.txt')
  do i = 1, n
    write(2,*) v(i)
    write(3,'(1x,e15.7)') v(i)
  end do
  close(2)
  close(3)
end program
"""


def values():
    special = [0.0, -0.0, 1.0, -1.0, 0.5, 0.1, 0.09999999999999999, 0.099, 0.095, 0.0949999, 0.05, 0.25,
               1e15, 9.99e14, 9.4e14, 9.5e14, 1e16, 9.5e15, 123456789012345678.0, 2.0 / 3.0, 582400.0,
               3e-300, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, -1.7976931348623157e308,
               9.99999995, 9.99999994, 12345665.0, 12345675.0, 1e-100, 1e100, 1e-99, 1e99, 0.99999999,
               float("nan"), float("inf"), float("-inf"), 2.015195741758242, 1.4285714285714284e-4, 735.0,
               0.0728, 7.28e-2, 1e21, 99999999999999999.0, 6.000000089406968e-05]
    rng = np.random.default_rng(20071)
    mant = rng.uniform(1.0, 10.0, 4000)
    expo = rng.integers(-320, 308, 4000)
    rnd = mant * np.power(10.0, expo.astype(np.float64))
    rnd = np.where(rng.random(4000) < 0.5, -rnd, rnd)
    rnd = rnd[np.isfinite(rnd)]
    # values near the list-directed F/E switches and short decimals
    near = np.concatenate([np.nextafter(np.array([0.095, 1e15, 9.5e14, 0.1]), 0.0),
                           np.nextafter(np.array([0.095, 1e15, 9.5e14, 0.1]), 1e300),
                           np.round(rng.uniform(-1000, 1000, 300), 3)])
    return np.concatenate([np.array(special), rnd, near])


def flang_text(tmp, v):
    with open(os.path.join(tmp, "p.f90"), "w") as f:
        f.write(PROGRAM)
    r = subprocess.run([FLANG, "-O1", "p.f90", "-o", "p"], cwd=tmp, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    with open(os.path.join(tmp, "v.bin"), "wb") as f:
        f.write(np.int32(len(v)).tobytes())
        f.write(v.astype("<f8").tobytes())
    subprocess.run(["./p"], cwd=tmp, check=True, timeout=120)
    ld = open(os.path.join(tmp, "ld.txt")).read().splitlines()
    e = open(os.path.join(tmp, "e.txt")).read().splitlines()
    return ld, e


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not present")
def test_formats_match_flang(dg, tmp_path):
    lib = dg.lib()
    v = values()
    ld, e = flang_text(str(tmp_path), v)
    assert len(ld) == len(v) and len(e) == len(v)
    buf = ctypes.create_string_buffer(64)
    n = ctypes.c_int(64)
    bad = []
    for q, x in enumerate(v):
        xd = ctypes.c_double(float(x))
        for mode, want in ((0, ld[q]), (1, e[q])):
            assert lib.dg_format_real(ctypes.byref(xd), ctypes.byref(ctypes.c_int(mode)), buf, ctypes.byref(n)) == 0
            got = " " + buf.value.decode()
            if got != want:
                bad.append((mode, repr(float(x)), got, want))
    assert not bad, bad[:20]


def test_fast_e15_7_equals_exact_digits(dg):
    """The writers' E15.7 (mode 1, integer rounding of a scaled value where
    it is certain) against the same edit through printf's correctly rounded
    digits (mode 2) on 200000 values: log-uniform magnitudes, values on and
    next to 7-digit rounding ties, powers of ten and their neighbours."""
    lib = dg.lib()
    rng = np.random.default_rng(7)
    mags = 10.0 ** rng.uniform(-40, 40, 100000) * rng.choice([-1.0, 1.0], 100000)
    ties = (rng.integers(1000000, 9999999, 40000) + 0.5) * 10.0 ** rng.integers(-30, 30, 40000).astype(float)
    near = np.concatenate([np.nextafter(ties[:20000], 0.0), np.nextafter(ties[:20000], np.inf)])
    p10 = 10.0 ** np.arange(-30, 31, dtype=float)
    edge = np.concatenate([p10, np.nextafter(p10, 0.0), np.nextafter(p10, np.inf), 9.9999995 * p10,
                           np.nextafter(9.9999995 * p10, 0.0), 0.99999995 * p10])
    vals = np.concatenate([mags, ties, near, edge, rng.standard_normal(20000)])
    buf1, buf2 = ctypes.create_string_buffer(32), ctypes.create_string_buffer(32)
    n = ctypes.c_int(32)
    m1, m2 = ctypes.c_int(1), ctypes.c_int(2)
    bad = []
    for x in vals:
        xd = ctypes.c_double(float(x))
        lib.dg_format_real(ctypes.byref(xd), ctypes.byref(m1), buf1, ctypes.byref(n))
        lib.dg_format_real(ctypes.byref(xd), ctypes.byref(m2), buf2, ctypes.byref(n))
        if buf1.value != buf2.value:
            bad.append((repr(float(x)), buf1.value, buf2.value))
    assert not bad, bad[:10]
