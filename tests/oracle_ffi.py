"""ctypes binding of the CPU restatement (oracle/_build/liboracle.so).

Test infrastructure only: the checker for the GPU product and the CPU
baseline of bench.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
ORACLE_SO = os.path.join(REPO, "oracle", "_build", "liboracle.so")

FIELD_IDS = dict(cc=0, u=1, v=2, w=3, p=4, s11=5, s22=6, s33=7, s12=8, s13=9, s23=10,
                 rhon=11, mun=12, mut=13, ktur=14, eps=15, a=16, b=17, c=18, ei=19, rhs=20, res=21,
                 pprev=22, hxmin=23, hxmax=24, hymin=25, hymax=26, hzmin=27, hzmax=28,
                 u2p=29, v2p=30, w2p=31, uvp=32, uwp=33, vwp=34, cmu=35)

_lib = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = ctypes.CDLL(ORACLE_SO)
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_field.restype = ctypes.POINTER(ctypes.c_double)
        L.or_field.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_level_array.restype = ctypes.POINTER(ctypes.c_double)
        L.or_level_array.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.or_level_dims.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        for nm in ("or_step", "or_vofmethod", "or_kepsmodel", "or_checktime", "or_solvelpres"):
            getattr(L, nm).argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_csf.argtypes = [ctypes.c_void_p]
        for nm in ("or_calch3d", "or_sgip", "or_sgip2", "or_eddyvisc", "or_uv2p"):
            getattr(L, nm).argtypes = [ctypes.c_void_p]
        L.or_calcrho.argtypes = [ctypes.c_void_p]
        L.or_cfl.argtypes = [ctypes.c_void_p]
        L.or_cfl.restype = ctypes.c_double
        L.or_al3d.restype = ctypes.c_double
        L.or_al3d.argtypes = [ctypes.c_double] * 4
        L.or_fl3d.restype = ctypes.c_double
        L.or_fl3d.argtypes = [ctypes.c_double] * 6
        L.or_mg_solve_tol.argtypes = [ctypes.c_void_p]
        L.or_set_block_smoother.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4
        L.or_set_block_overlap.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_set_block_max.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_set_block_sweeps.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_set_block_large.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4
        L.or_mglin.argtypes = [ctypes.c_void_p]
        L.or_resid_norm.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.or_resid_norm.restype = ctypes.c_double
        dp, i_, d_ = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.or_relax.argtypes = [dp, d_, dp, dp, dp, dp, dp, i_, i_, i_, d_]
        L.or_resid.argtypes = [d_, dp, dp, dp, dp, dp, dp, i_, i_, i_]
        L.or_rstrct.argtypes = [dp, dp, i_, i_, i_]
        L.or_interp.argtypes = [dp, dp, i_, i_, i_]
        L.or_addint.argtypes = [dp, dp, dp, i_, i_, i_]
        L.or_acfromac.argtypes = [dp, dp, dp, dp, dp, dp, dp, i_, i_, i_]
        L.or_slvsml.argtypes = [dp, dp, dp, dp, dp, dp, i_, i_, i_, i_, d_, d_]
        L.or_abcei.argtypes = [dp, dp, dp, dp, dp, d_, d_, i_, i_, i_]
        L.or_bc_scalar.argtypes = [dp, i_, i_, i_]
        L.or_filterc27.argtypes = [dp, dp, i_, i_, i_]
        _lib = L
    return _lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleKernels:
    """The restatement's explicit-argument kernels, same call shapes as
    refrun.RefKernels and the GPU package's host-pointer kernels."""

    @staticmethod
    def al3d(b1, b2, b3, cc):
        return lib().or_al3d(b1, b2, b3, cc)

    @staticmethod
    def fl3d(m1, m2, m3, alpha, r0, dr0):
        return lib().or_fl3d(m1, m2, m3, alpha, r0, dr0)

    @staticmethod
    def relax(u, lengthsize, rhs, a, b, c, ei, omega):
        nz, ny, nx = u.shape
        lib().or_relax(_ptr(u), lengthsize, _ptr(rhs), _ptr(a), _ptr(b), _ptr(c), _ptr(ei), nx, ny, nz, omega)

    @staticmethod
    def resid(lengthsize, res, u, rhs, a, b, c):
        nz, ny, nx = u.shape
        lib().or_resid(lengthsize, _ptr(res), _ptr(u), _ptr(rhs), _ptr(a), _ptr(b), _ptr(c), nx, ny, nz)

    @staticmethod
    def rstrct(uc, uf):
        ncz, ncy, ncx = uc.shape
        lib().or_rstrct(_ptr(uc), _ptr(uf), ncx, ncy, ncz)

    @staticmethod
    def interp(uf, uc):
        nz, ny, nx = uf.shape
        lib().or_interp(_ptr(uf), _ptr(uc), nx, ny, nz)

    @staticmethod
    def addint(uf, uc, res):
        nz, ny, nx = uf.shape
        lib().or_addint(_ptr(uf), _ptr(uc), _ptr(res), nx, ny, nz)

    @staticmethod
    def acfromac(a, b, c, ei, a2, b2, c2):
        nz, ny, nx = a.shape
        lib().or_acfromac(_ptr(a), _ptr(b), _ptr(c), _ptr(ei), _ptr(a2), _ptr(b2), _ptr(c2), nx, ny, nz)

    @staticmethod
    def slvsml(u, rhs, a, b, c, ei, ntop, omega, lengthsize):
        nz, ny, nx = u.shape
        lib().or_slvsml(_ptr(u), _ptr(rhs), _ptr(a), _ptr(b), _ptr(c), _ptr(ei), nx, ny, nz, ntop, omega, lengthsize)

    @staticmethod
    def abcei(a, b, c, ei, cc, r1, r2):
        nz, ny, nx = a.shape
        lib().or_abcei(_ptr(a), _ptr(b), _ptr(c), _ptr(ei), _ptr(cc), r1, r2, nx, ny, nz)

    @staticmethod
    def bc_scalar(s):
        nz, ny, nx = s.shape
        lib().or_bc_scalar(_ptr(s), nx, ny, nz)

    @staticmethod
    def filterc27(cc, cf):
        nz, ny, nx = cc.shape
        lib().or_filterc27(_ptr(cc), _ptr(cf), nx, ny, nz)


class OrState(ctypes.Structure):
    """Mirror of or_state's leading scalar members (droplet_oracle.h)."""


def _state_fields():
    d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p

    class Level(ctypes.Structure):
        _fields_ = [("nx", i), ("ny", i), ("nz", i)] + [(n, p) for n in ("rhs", "res", "u", "cc", "a", "b", "c", "ei")]

    f = [("nx", i), ("ny", i), ("nz", i), ("tmax", i), ("nfilter", i), ("cycles", i), ("omega", d),
         ("npre", i), ("ntop", i), ("npost", i), ("restart", i), ("maxcfl", d), ("iturb", i), ("convectype", i),
         ("hswitch", i)]
    f += [(n, d) for n in ("xvel1_si", "yvel1_si", "zvel1_si", "xvel2_si", "yvel2_si", "zvel2_si", "rho2_si",
                           "rho1_si", "sigma_si", "visc2_si", "visc1_si", "grav_si", "grav1_si")]
    f += [("icase", i), ("ibc", i)]
    f += [(n, d) for n in ("length_si", "radius_si", "y0_si", "z0_si")]
    f += [(n, i) for n in ("tprint", "ntecplot", "tsave", "surfacet", "ijet", "idrop", "ibubble", "isheet", "ng")]
    f += [("lev", Level * 16)]
    f += [(n, d) for n in ("tau_si", "timescale", "lengthscale", "lengthsize", "vscale", "tau", "time", "radius",
                           "h", "rho1", "rho2", "mu1", "mu2", "sigma", "g", "g1", "y0", "z0", "xvel1", "yvel1",
                           "zvel1", "xvel2", "yvel2", "zvel2", "tauh")]
    f += [("tauh2", i), ("tswap", i)]
    f += [(n, p) for n in ("u", "v", "w", "s11", "s22", "s33", "s12", "s13", "s23", "rhon", "mun", "mut", "ktur",
                           "eps", "cmu")]
    f += [("mg_mode", i), ("mg_nu1", i), ("mg_nu2", i), ("mg_maxcycles", i), ("mg_guess", i), ("mg_tol", d),
          ("mg_cycles_last", i), ("mg_relres_last", d), ("mg_bnorm_last", d), ("p_prev", p), ("cfl_last", d), ("mg_outer", i), ("mg_prolong", i), ("vof_guard", i), ("mg_omega", d), ("mg_nuc", i), ("ref_grid", i)]
    f += [(n, p) for n in ("hxmin", "hxmax", "hymin", "hymax", "hzmin", "hzmax")]
    f += [("stop_code", i)]
    f += [(n, p) for n in ("u2p", "v2p", "w2p", "uvp", "uwp", "vwp", "lmix")]
    return f


OrState._fields_ = _state_fields()


class Oracle:
    """One CPU-restatement instance created from inputdata.in text."""

    def __init__(self, inputdata: str):
        L = lib()
        with tempfile.NamedTemporaryFile("w", suffix=".in", delete=False) as f:
            f.write(inputdata)
            path = f.name
        err = ctypes.create_string_buffer(256)
        try:
            self.h = L.or_create(path.encode(), err, 256)
        finally:
            os.unlink(path)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.st = ctypes.cast(self.h, ctypes.POINTER(OrState)).contents
        self.nx, self.ny, self.nz = self.st.nx, self.st.ny, self.st.nz

    def close(self):
        if self.h:
            lib().or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, name):
        ptr = lib().or_field(self.h, FIELD_IDS[name])
        n = self.nx * self.ny * self.nz
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self.nz, self.ny, self.nx)

    def level(self, lev, which):
        dims = (ctypes.c_int * 3)()
        lib().or_level_dims(self.h, lev, dims)
        ptr = lib().or_level_array(self.h, lev, which)
        nx, ny, nz = dims
        return np.ctypeslib.as_array(ptr, shape=(nx * ny * nz,)).reshape(nz, ny, nx)

    def step(self, t):
        lib().or_step(self.h, int(t))

    def run(self, nsteps, t0=0):
        hist = []
        for t in range(t0, t0 + nsteps):
            self.step(t)
            hist.append(dict(t=t, tau=self.st.tau, time=self.st.time, tauh=self.st.tauh, tauh2=self.st.tauh2,
                             npre=self.st.npre, npost=self.st.npost, ntop=self.st.ntop,
                             vol=float(self.field("cc")[1:-1, 1:-1, 1:-1].sum()), tswap=self.st.tswap,
                             mg_cycles=self.st.mg_cycles_last, relres=self.st.mg_relres_last))
        return hist


def inputdata_text(nx=18, ny=18, nz=18, tmax=5, cycles=2, omega=1.5, npre=2, ntop=4, npost=2, maxcfl=0.25,
                   iturb=0, convectype=2, hswitch=0, xvel1=0.0, yvel1=-1.0, zvel1=0.0, xvel2=0.0, yvel2=0.0,
                   zvel2=0.0, rho2=1.2, rho1=1000.0, sigma=0.072, visc2=1.8e-5, visc1=1.0e-3, grav=-9.81,
                   grav1=0.0, length=0.004, radius=0.0008, y0=0.0025, z0=1000.0, surfacet=2, restart=0,
                   tprint=100000, ntecplot=100000, tsave=100000):
    """inputdata.in in the reference's 38-line order (main.f:188-234)."""
    vals = [nx, ny, nz, tmax, 0, cycles, omega, npre, ntop, npost, restart, maxcfl, iturb, convectype, hswitch,
            xvel1, yvel1, zvel1, xvel2, yvel2, zvel2, rho2, rho1, sigma, visc2, visc1, grav, grav1, 2, 2,
            length, radius, y0, z0, tprint, ntecplot, tsave, surfacet]
    return "".join(f"{v!r}\n" if isinstance(v, float) else f"{v}\n" for v in vals)


# the values of the reference's shipped inputdata.in (66x34x34, water and a
# heavier second fluid, inflow 0.5 m/s, no gravity, a droplet centred just
# below the wall, SGIP surface tension with CalcH3d heights)
SHIPPED = dict(nx=66, ny=34, nz=34, tmax=400, cycles=2, omega=1.5, npre=5, ntop=4, npost=5, maxcfl=0.1, iturb=0,
               convectype=2, hswitch=1, xvel1=0.0, yvel1=0.0, zvel1=0.0, xvel2=0.5, yvel2=0.0, zvel2=0.0,
               rho2=500.0, rho1=1000.0, sigma=0.0728, visc2=0.001, visc1=0.001, grav=0.0, grav1=0.0,
               length=0.008, radius=0.001, y0=-0.00017365, z0=1000.0, surfacet=5)


def shipped_text(**over):
    kw = dict(SHIPPED)
    kw.update(over)
    return inputdata_text(**kw)


def clipped_cc(txt):
    """smooth2's initial volume fraction clipped to [0, 1]: unclipped, 8
    cells hold 1 + 4e-13, the first VOFmethod turns 100 cells to NaN (AL3D's
    cube root of a negative number) and CalcH3d STOPs with NO MATCH CASES
    FOUND at the first step."""
    o = Oracle(txt)
    cc = o.field("cc").clip(0.0, 1.0).copy()
    o.close()
    return cc


def sampled_cc(shape, h, drops):
    """The sub-point count of smooth2 (main.f:5693-5731) for several
    droplets: volume fractions of the interior, numpy restatement."""
    nz, ny, nx = shape
    index = 30
    inc = 1.0 / float(index ** 3)
    table = np.zeros(index ** 3 + 1)
    acc = 0.0
    for q in range(1, index ** 3 + 1):
        acc = acc + inc
        table[q] = acc
    cc = np.zeros(shape)
    off = (np.arange(1, index + 1, dtype=np.float64) - 0.5) / float(index) * h
    for k in range(2, nz):
        for j in range(2, ny):
            for i in range(2, nx):
                xa, xb = ((i - 1) - 1.5) * h, (i - 1.5) * h
                ya, yb = ((j - 1) - 1.5) * h, (j - 1.5) * h
                za, zb = ((k - 1) - 1.5) * h, (k - 1.5) * h
                cand = []
                full = False
                for (x0, y0, z0, r, *_v) in drops:
                    dmin = min((xa - x0) ** 2, (xb - x0) ** 2) + min((ya - y0) ** 2, (yb - y0) ** 2) + \
                        min((za - z0) ** 2, (zb - z0) ** 2)
                    if not dmin < r * r:
                        continue
                    dmax = max((xa - x0) ** 2, (xb - x0) ** 2) + max((ya - y0) ** 2, (yb - y0) ** 2) + \
                        max((za - z0) ** 2, (zb - z0) ** 2)
                    if not dmax > r * r:
                        full = True
                        break
                    cand.append((x0, y0, z0, r))
                if full:
                    cc[k - 1, j - 1, i - 1] = 1.0
                    continue
                if not cand:
                    continue
                tx = (xa + off)[:, None, None]
                ty = (ya + off)[None, :, None]
                tz = (za + off)[None, None, :]
                inside = np.zeros((index, index, index), dtype=bool)
                for (x0, y0, z0, r) in cand:
                    inside |= (tx - x0) ** 2 + (ty - y0) ** 2 + (tz - z0) ** 2 < r * r
                cc[k - 1, j - 1, i - 1] = table[int(inside.sum())]
    return cc
