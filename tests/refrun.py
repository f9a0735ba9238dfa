"""Runs the flang-built reference (oracle/_ref/libref_*.so) through ctypes.

Test infrastructure only.  The libraries are compiled from the reference's
own main.f by oracle/build_ref.sh; they exist only where /root/reference
exists.  Each run happens in a fresh subprocess because the reference keeps
all state in zero-initialised COMMON blocks (input.h:19-95) and a second run
in the same process would start from the first run's leftovers.

The driver loop restates PROGRAM THREE DIM (main.f:19-73) with printout
(output only) left out.
"""
from __future__ import annotations

import ctypes
import multiprocessing as mp
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF_DIR = os.path.join(REPO, "oracle", "_ref")

FIELD_NAMES = ("cc", "u", "v", "w", "p", "s11", "s22", "s33", "s12", "s13", "s23",
               "rhon", "mun", "mut", "ktur", "eps")
# interface heights of CalcH3d, and their order in COMMON /NEWW/ (input.h:90-95)
HEIGHT_NAMES = ("hxmin", "hxmax", "hymin", "hymax", "hzmin", "hzmax")
_NEWW_ORDER = ("hxmax", "hxmin", "hymax", "hymin", "hzmax", "hzmin")


def ref_lib_path(nx, ny, nz):
    return os.path.join(REF_DIR, f"libref_{nx}x{ny}x{nz}.so")


def have_ref(nx, ny, nz):
    return os.path.exists(ref_lib_path(nx, ny, nz))


def _ngm(n):
    g = 0
    while (1 << g) + 2 < n:
        g += 1
    assert (1 << g) + 2 == n, n
    return g


class _Layout:
    """COMMON block layouts from input.h (sizes fixed at compile time)."""

    def __init__(self, nx, ny, nz):
        self.nx, self.ny, self.nz = nx, ny, nz
        gx, gy, gz = _ngm(nx), _ngm(ny), _ngm(nz)
        self.NGM = max(gx, gy, gz)
        NGM = self.NGM
        nn3 = (2 ** (3 * NGM) - 1) // 7
        self.MEMLEN = 64 * nn3 + 64 * 2 ** (2 * NGM) + 192 * 2 ** NGM + 72 * NGM
        self.n = nx * ny * nz


def _child(lib_path, inputdata, nsteps, init_fields, want_fields, conn, every_step_fields, workdir=None,
           printout_ts=(), loop_printout=False):
    try:
        tmp = workdir or tempfile.mkdtemp(prefix="refrun_")
        # stdout (printout's list-directed lines) into the work directory
        out = os.open(os.path.join(tmp, "stdout.txt"), os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
        os.dup2(out, 1)
        with open(os.path.join(tmp, "inputdata.in"), "w") as f:
            f.write(inputdata)
        os.chdir(tmp)
        lib = ctypes.CDLL(lib_path)
        nx, ny, nz = (int(l.split()[0]) for l in inputdata.splitlines()[:3])
        L = _Layout(nx, ny, nz)
        lib.openfiles_()
        lib.initialize_()
        NGM = L.NGM
        ptrs = np.frombuffer((ctypes.c_int * (8 * NGM + 1)).in_dll(lib, "pointers_"), dtype=np.int32)
        icc, iu, ires, irhs, ia, ib, ic, iei = [ptrs[k * NGM:(k + 1) * NGM].copy() for k in range(8)]
        ng = int(ptrs[8 * NGM])
        zbuf = (ctypes.c_double * L.MEMLEN).in_dll(lib, "memory_")
        z = np.frombuffer(zbuf, dtype=np.float64)
        zaddr = ctypes.addressof(zbuf)
        n = L.n

        def zslot(idx):
            return z[int(idx) - 1:int(idx) - 1 + n].reshape(nz, ny, nx)

        def zp(idx):
            return ctypes.c_void_p(zaddr + 8 * (int(idx) - 1))

        surf = np.frombuffer((ctypes.c_double * (9 * n)).in_dll(lib, "surfcom_"), dtype=np.float64)
        names9 = ("u", "v", "w", "s11", "s22", "s33", "s12", "s13", "s23")
        arrs = {nm: surf[q * n:(q + 1) * n].reshape(nz, ny, nx) for q, nm in enumerate(names9)}
        # /PHYSICS/: 19 doubles, then rhon, mun
        phys = np.frombuffer((ctypes.c_double * (19 + 2 * n)).in_dll(lib, "physics_"), dtype=np.float64)
        arrs["rhon"] = phys[19:19 + n].reshape(nz, ny, nx)
        arrs["mun"] = phys[19 + n:19 + 2 * n].reshape(nz, ny, nx)
        # /TURBULENCE/: mut, lmix(NMX), turbst, ktur, eps, cmu, ...
        turb = np.frombuffer((ctypes.c_double * (11 * n + nx)).in_dll(lib, "turbulence_"), dtype=np.float64)
        arrs["mut"] = turb[0:n].reshape(nz, ny, nx)
        arrs["lmix"] = turb[n:n + nx]
        for q, nm in enumerate(("ktur", "eps", "cmu", "u2p", "v2p", "w2p", "uvp", "uwp", "vwp")):
            arrs[nm] = turb[n + nx + (q + 1) * n:n + nx + (q + 2) * n].reshape(nz, ny, nx)
        arrs["cc"] = zslot(icc[ng - 1])
        arrs["p"] = zslot(iu[ng - 1])
        # /CYCLING/ MAXCFL,OMEGA (2 doubles = 4 ints), then CYCLES,NPRE,NPOST,NTOP,
        # restart,icase,ijet,idrop,ibubble,isheet,iturb,nfilter,ibc,tswap,
        # convectype,hswitch,surfacet (input.h:47-48)
        cyc_i = np.frombuffer((ctypes.c_int * 21).in_dll(lib, "cycling_"), dtype=np.int32)
        physd = phys  # tau=0, time=1, radius=2, h=3
        diff = np.frombuffer((ctypes.c_double * 4).in_dll(lib, "diff_"), dtype=np.float64)
        # /DIFF/ ANGLEDIF,EPSC,EPSURR,tauh (double) then tauh2 (INTEGER)
        diff_i = np.frombuffer((ctypes.c_int * 9).in_dll(lib, "diff_"), dtype=np.int32)
        iturb = int(cyc_i[14])
        hswitch = int(cyc_i[19])
        surfacet = int(cyc_i[20])
        neww = np.frombuffer((ctypes.c_double * (6 * n)).in_dll(lib, "neww_"), dtype=np.float64)
        for q, nm in enumerate(_NEWW_ORDER):
            arrs[nm] = neww[q * n:(q + 1) * n].reshape(nz, ny, nx)
        for nm, val in (init_fields or {}).items():
            arrs[nm][...] = val
        hist = []
        snaps = []
        tim = np.frombuffer((ctypes.c_int * 6).in_dll(lib, "times_"), dtype=np.int32)
        t0 = int(tim[0])
        for t in range(t0, t0 + nsteps):
            ti = ctypes.c_int(t)
            if loop_printout:
                lib.printout_(ctypes.byref(ti), zp(icc[ng - 1]), zp(iu[ng - 1]))
            lib.vofmethod_(ctypes.byref(ti))
            if hswitch == 1:
                lib.calch3d_(zp(icc[ng - 1]))
            if iturb in (1, 2):
                lib.eddyvisc_(zp(icc[ng - 1]), zp(iu[ng - 1]))
            if iturb in (3, 4):
                lib.kepsmodel_(ctypes.byref(ti), zp(iu[ng - 1]), zp(ires[ng - 1]))
            lib.checktime_(ctypes.byref(ti))
            if surfacet == 2:
                lib.csf_(zp(icc[ng - 1]), zp(iu[ng - 1]), zp(ires[ng - 1]), zp(irhs[ng - 1]))
            elif surfacet == 5:
                lib.sgip_(zp(icc[ng - 1]), zp(iu[ng - 1]), zp(ires[ng - 1]), zp(irhs[ng - 1]))
            elif surfacet == 6:
                lib.sgip2_(zp(icc[ng - 1]), zp(iu[ng - 1]), zp(ires[ng - 1]), zp(irhs[ng - 1]))
            else:
                raise RuntimeError(f"surfacet={surfacet} not on this path")
            lib.solvelpres_(zp(icc[ng - 1]), zp(iu[ng - 1]), zp(ires[ng - 1]), zp(irhs[ng - 1]),
                            zp(ia[ng - 1]), zp(ib[ng - 1]), zp(ic[ng - 1]), zp(iei[ng - 1]), ctypes.byref(ti))
            hist.append(dict(t=t, tau=float(physd[0]), time=float(physd[1]),
                             tauh=float(diff[3]), tauh2=int(diff_i[8]),
                             npre=int(cyc_i[5]), npost=int(cyc_i[6]), ntop=int(cyc_i[7]),
                             vol=float(arrs["cc"][1:-1, 1:-1, 1:-1].sum()), tswap=int(cyc_i[17])))
            if every_step_fields:
                snaps.append({nm: arrs[nm].copy() for nm in every_step_fields})
        out = {nm: arrs[nm].copy() for nm in want_fields}
        scal = dict(tau=float(physd[0]), time=float(physd[1]), h=float(physd[3]), tauh=float(diff[3]),
                    tauh2=int(diff_i[8]), npre=int(cyc_i[5]), npost=int(cyc_i[6]), ntop=int(cyc_i[7]),
                    tswap=int(cyc_i[17]), t0=t0)
        # printout (main.f:5020-5100) of the final state at the given steps
        for t in printout_ts:
            ti = ctypes.c_int(t)
            lib.printout_(ctypes.byref(ti), zp(icc[ng - 1]), zp(iu[ng - 1]))
        _flush_units(lib, (2, 6, 23, 24, 26, 27, 28))
        conn.send(dict(fields=out, hist=hist, scalars=scal, snaps=snaps))
    except Exception as e:  # pragma: no cover - reported to the parent
        conn.send(dict(error=repr(e)))
    finally:
        conn.close()


def _flush_units(lib, units):
    """FLUSH of the reference's open units through the Fortran runtime that
    the library carries (the child ends without the runtime's exit handler,
    which would otherwise write the buffers)."""
    begin = lib._FortranAioBeginFlush
    begin.restype = ctypes.c_void_p
    begin.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    end = lib._FortranAioEndIoStatement
    end.argtypes = [ctypes.c_void_p]
    for u in units:
        end(begin(u, b"refrun.py", 0))


def run_reference(inputdata: str, nsteps: int, init_fields=None, want_fields=FIELD_NAMES,
                  every_step_fields=None, timeout=600, workdir=None, printout_ts=(), loop_printout=False):
    """Runs initialize and nsteps of the loop from t0 (0, or save.dat's t0
    with restart = 1: put save.dat into workdir).  workdir (default: a fresh
    temporary directory) receives inputdata.in, stdout.txt and the files
    the reference writes; loop_printout calls printout at the start of each
    step as the reference's loop does, printout_ts calls it on the final
    state for each listed step."""
    nx, ny, nz = (int(l.split()[0]) for l in inputdata.splitlines()[:3])
    lib_path = ref_lib_path(nx, ny, nz)
    if not os.path.exists(lib_path):
        raise FileNotFoundError(lib_path)
    ctx = mp.get_context("fork")
    a, b = ctx.Pipe(duplex=False)
    proc = ctx.Process(target=_child, args=(lib_path, inputdata, nsteps, init_fields, want_fields, b,
                                            every_step_fields, workdir, tuple(printout_ts), loop_printout))
    proc.start()
    b.close()
    if not a.poll(timeout):
        proc.kill()
        raise TimeoutError("reference run timed out")
    res = a.recv()
    proc.join(30)
    if "error" in res:
        raise RuntimeError(res["error"])
    return res


# ---- pure subroutines (no COMMON state): called in-process ---------------
class RefKernels:
    """Direct bindings to the reference's explicit-interface subroutines."""

    def __init__(self, nx, ny, nz):
        self.lib = ctypes.CDLL(ref_lib_path(nx, ny, nz))
        self.lib.al3d_.restype = ctypes.c_double
        self.lib.fl3d_.restype = ctypes.c_double

    @staticmethod
    def _d(x):
        return ctypes.byref(ctypes.c_double(float(x)))

    @staticmethod
    def _i(x):
        return ctypes.byref(ctypes.c_int(int(x)))

    @staticmethod
    def _p(arr):
        assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"]
        return arr.ctypes.data_as(ctypes.c_void_p)

    def al3d(self, b1, b2, b3, cc):
        return self.lib.al3d_(self._d(b1), self._d(b2), self._d(b3), self._d(cc))

    def fl3d(self, m1, m2, m3, alpha, r0, dr0):
        return self.lib.fl3d_(self._d(m1), self._d(m2), self._d(m3), self._d(alpha), self._d(r0), self._d(dr0))

    def relax(self, u, lengthsize, rhs, a, b, c, ei, omega):
        nz, ny, nx = u.shape
        self.lib.relax_(self._p(u), self._d(lengthsize), self._p(rhs), self._p(a), self._p(b), self._p(c),
                        self._p(ei), self._i(2), self._i(nx), self._i(ny), self._i(nz), self._d(omega))

    def resid(self, lengthsize, res, u, rhs, a, b, c):
        nz, ny, nx = u.shape
        self.lib.resid_(self._d(lengthsize), self._p(res), self._p(u), self._p(rhs), self._p(a), self._p(b),
                        self._p(c), self._i(2), self._i(nx), self._i(ny), self._i(nz))

    def rstrct(self, uc, uf):
        ncz, ncy, ncx = uc.shape
        self.lib.rstrct_(self._p(uc), self._p(uf), self._i(2), self._i(ncx), self._i(ncy), self._i(ncz))

    def interp(self, uf, uc):
        nz, ny, nx = uf.shape
        self.lib.interp_(self._p(uf), self._p(uc), self._i(2), self._i(nx), self._i(ny), self._i(nz))

    def addint(self, uf, uc, res):
        nz, ny, nx = uf.shape
        self.lib.addint_(self._p(uf), self._p(uc), self._p(res), self._i(2), self._i(nx), self._i(ny), self._i(nz))

    def acfromac(self, a, b, c, ei, a2, b2, c2):
        nz, ny, nx = a.shape
        self.lib.acfromac_(self._p(a), self._p(b), self._p(c), self._p(ei), self._p(a2), self._p(b2), self._p(c2),
                           self._i(nx), self._i(ny), self._i(nz))

    def abcei(self, a, b, c, ei, cc, r1, r2):
        nz, ny, nx = a.shape
        self.lib.abcei_(self._p(a), self._p(b), self._p(c), self._p(ei), self._p(cc), self._d(r1), self._d(r2),
                        self._i(2), self._i(nx), self._i(ny), self._i(nz))

    def slvsml(self, u, rhs, a, b, c, ei, ntop, omega, lengthsize):
        nz, ny, nx = u.shape
        self.lib.slvsml_(self._p(u), self._p(rhs), self._p(a), self._p(b), self._p(c), self._p(ei), self._i(2),
                         self._i(nx), self._i(ny), self._i(nz), self._i(ntop), self._d(omega), self._d(lengthsize))

    def bc_scalar(self, s):
        nz, ny, nx = s.shape
        self.lib.bc_scalar_(self._p(s), self._i(2), self._i(nx), self._i(ny), self._i(nz))

    def filterc27(self, cc, cf):
        nz, ny, nx = cc.shape
        self.lib.filterc27_(self._p(cc), self._p(cf), self._i(nx), self._i(ny), self._i(nz))
