"""Python host side of libdroplet_gpu.so.

Mirrors the reference's driver (PROGRAM THREE DIM, main.f:2-76) and its
subroutine interfaces over the C-ABI declared in include/droplet_gpu.h.  The
GPU library is required: nothing here computes the physics on the CPU, and
every entry point raises if the library is missing or reports an error.
"""
from __future__ import annotations

import ctypes
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libdroplet_gpu.so")

FIELD_IDS = dict(cc=0, u=1, v=2, w=3, p=4, s11=5, s22=6, s33=7, s12=8, s13=9, s23=10,
                 rhon=11, mun=12, mut=13, ktur=14, eps=15, a=16, b=17, c=18, ei=19, rhs=20, res=21,
                 pprev=22, hxmin=23, hxmax=24, hymin=25, hymax=26, hzmin=27, hzmax=28,
                 u2p=29, v2p=30, w2p=31, uvp=32, uwp=33, vwp=34, cmu=35)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class DgState(ctypes.Structure):
    _fields_ = [("tau", ctypes.c_double), ("time", ctypes.c_double), ("tauh", ctypes.c_double),
                ("tauh2", ctypes.c_int), ("tswap", ctypes.c_int), ("npre", ctypes.c_int),
                ("ntop", ctypes.c_int), ("npost", ctypes.c_int),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("ng", ctypes.c_int),
                ("mg_iters", ctypes.c_int), ("mg_relres", ctypes.c_double), ("cfl", ctypes.c_double),
                ("epsdiv", ctypes.c_double)]


class DropletError(RuntimeError):
    pass


_lib = None


def lib():
    """Loads the in-tree HIP library (fails loudly when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise DropletError(f"{LIBPATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIBPATH)
        L.dg_last_error.restype = ctypes.c_char_p
        L.dg_init.argtypes = [ctypes.c_char_p]
        L.dg_get_field.argtypes = [_ip, _dp]
        L.dg_set_field.argtypes = [_ip, _dp]
        L.dg_get_state.argtypes = [ctypes.POINTER(DgState)]
        L.dg_set_state.argtypes = [ctypes.POINTER(DgState)]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = lib().dg_last_error().decode(errors="replace")
        raise DropletError(f"{what}: {msg}")


def _i(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _d(v):
    return ctypes.byref(ctypes.c_double(float(v)))


def _p(a):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise DropletError("arrays must be C-contiguous float64 (reference layout (nz,ny,nx) in numpy order)")
    return a.ctypes.data_as(_dp)


def _shapes(what, shape, *arrays):
    """Every array must have `shape`: the library reads and writes that many
    doubles through each pointer."""
    for q, a in enumerate(arrays):
        if a.ndim != 3 or tuple(a.shape) != tuple(shape):
            raise DropletError(f"{what}: array {q + 1} has shape {a.shape}, expected {tuple(shape)}")


def _coarse(shape):
    """Extents of the next-coarser level: n -> n/2 + 1 (memoryaloc, main.f:274-322)."""
    return tuple(n // 2 + 1 for n in shape)


def make_inputdata(nx=66, ny=66, nz=66, tmax=49, cycles=2, omega=1.5, npre=2, ntop=4, npost=2, maxcfl=0.25,
                   iturb=0, convectype=2, hswitch=0, xvel1=0.0, yvel1=-1.0, zvel1=0.0, xvel2=0.0, yvel2=0.0,
                   zvel2=0.0, rho2=1.2, rho1=1000.0, sigma=0.072, visc2=1.8e-5, visc1=1.0e-3, grav=-9.81,
                   grav1=0.0, length=0.004, radius=0.0008, y0=0.0025, z0=1000.0, tprint=100000,
                   ntecplot=100000, tsave=100000, surfacet=2):
    """Text of an inputdata.in file in the reference's 38-line order
    (readinput, main.f:188-234).  Defaults: a water droplet in air on a 4 mm
    cube falling at 1 m/s along the gravity axis (the reference's y)."""
    vals = [nx, ny, nz, tmax, 0, cycles, omega, npre, ntop, npost, 0, maxcfl, iturb, convectype, hswitch,
            xvel1, yvel1, zvel1, xvel2, yvel2, zvel2, rho2, rho1, sigma, visc2, visc1, grav, grav1, 2, 2,
            length, radius, y0, z0, tprint, ntecplot, tsave, surfacet]
    return "".join(f"{v!r}\n" if isinstance(v, float) else f"{v}\n" for v in vals)


class DropletGPU:
    """The resident device state of one run (the reference's COMMON blocks).

    Only one instance is active per process, as in the reference."""

    def __init__(self, inputdata_path=None, inputdata_text=None):
        L = lib()
        if inputdata_text is not None:
            with tempfile.NamedTemporaryFile("w", suffix=".in", delete=False) as f:
                f.write(inputdata_text)
                path = f.name
            try:
                _check(L.dg_init(path.encode()), "dg_init")
            finally:
                os.unlink(path)
        else:
            _check(L.dg_init(os.fspath(inputdata_path).encode()), "dg_init")
        st = self.state()
        self.nx, self.ny, self.nz = st["nx"], st["ny"], st["nz"]

    def close(self):
        lib().dg_finalize()

    # ---- the reference's per-step operators
    def step(self, t):
        _check(lib().dg_step(_i(t)), "dg_step")

    def run(self, nsteps, t0=0):
        _check(lib().dg_run(_i(t0), _i(nsteps)), "dg_run")

    def vofmethod(self, t):
        _check(lib().dg_vofmethod(_i(t)), "dg_vofmethod")

    def kepsmodel(self, t):
        _check(lib().dg_kepsmodel(_i(t)), "dg_kepsmodel")

    def checktime(self, t):
        _check(lib().dg_checktime(_i(t)), "dg_checktime")

    def csf(self):
        _check(lib().dg_csf(), "dg_csf")

    def eddyvisc(self):
        """eddyvisc (main.f:1828-1885), the mixing-length model of iturb 1, 2."""
        _check(lib().dg_eddyvisc(), "dg_eddyvisc")

    def uv2p(self):
        """uv2p (main.f:2032-2155), the anisotropic stresses of iturb 4."""
        _check(lib().dg_uv2p(), "dg_uv2p")

    def smooth2_droplet(self):
        """The droplet dg_init placed (smooth2, main.f:5678-5752) as one row of
        init_droplets' parameters: x0, y0, z0, r, u, v, w in solver units."""
        p = np.zeros(7)
        _check(lib().dg_smooth2_droplet(_p(p)), "dg_smooth2_droplet")
        return p

    def init_droplets(self, params):
        """smooth2 generalised to several droplets: params is (n, 7) with rows
        (x0, y0, z0, r, u, v, w) in solver units (see smooth2_droplet)."""
        a = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, 7)
        _check(lib().dg_init_droplets(_i(a.shape[0]), _p(a)), "dg_init_droplets")

    def calch3d(self):
        _check(lib().dg_calch3d(), "dg_calch3d")

    def sgip(self):
        _check(lib().dg_sgip(), "dg_sgip")

    def sgip2(self):
        _check(lib().dg_sgip2(), "dg_sgip2")

    def solvelpres(self, t):
        _check(lib().dg_solvelpres(_i(t)), "dg_solvelpres")

    def calcrho(self):
        _check(lib().dg_calcrho(), "dg_calcrho")

    def cfl(self):
        """CFL() (main.f:4848): max |u|,|v|,|w| over the interior times tauh."""
        v = ctypes.c_double()
        _check(lib().dg_cfl(ctypes.byref(v)), "dg_cfl")
        return v.value

    def clip_cc(self):
        _check(lib().dg_clip_cc(), "dg_clip_cc")

    def set_solver(self, mode=0, tol=1e-8, nu1=2, nu2=2, maxit=200, guess=0):
        _check(lib().dg_set_solver(_i(mode), _d(tol), _i(nu1), _i(nu2), _i(maxit), _i(guess)), "dg_set_solver")

    def set_ordered_sums(self, on):
        _check(lib().dg_set_ordered_sums(_i(1 if on else 0)), "dg_set_ordered_sums")

    def set_vof_guard(self, on):
        """Advect partly filled cells whose Youngs normal vanishes with a plane
        normal to the sweep (the reference evaluates 0/0 there)."""
        _check(lib().dg_set_vof_guard(_i(1 if on else 0)), "dg_set_vof_guard")

    # ---- state exchange
    def get(self, name):
        out = np.empty((self.nz, self.ny, self.nx), dtype=np.float64)
        _check(lib().dg_get_field(ctypes.byref(ctypes.c_int(FIELD_IDS[name])), _p(out)), "dg_get_field")
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.nz, self.ny, self.nx)
        _check(lib().dg_set_field(ctypes.byref(ctypes.c_int(FIELD_IDS[name])), _p(a)), "dg_set_field")

    def state(self):
        s = DgState()
        _check(lib().dg_get_state(ctypes.byref(s)), "dg_get_state")
        return {k: getattr(s, k) for k, _ in DgState._fields_}

    def set_state(self, **kw):
        s = DgState()
        _check(lib().dg_get_state(ctypes.byref(s)), "dg_get_state")
        for k, v in kw.items():
            setattr(s, k, v)
        _check(lib().dg_set_state(ctypes.byref(s)), "dg_set_state")

    # ---- output, checkpoint and restart
    def run_limits(self):
        """(t0, tmax) of the run: 0 and inputdata.in's tmax, or save.dat's
        t0 and t0 + its tmax after a restart (iniget, main.f:698-793)."""
        t0, tmax = ctypes.c_int(), ctypes.c_int()
        _check(lib().dg_run_limits(ctypes.byref(t0), ctypes.byref(tmax)), "dg_run_limits")
        return t0.value, tmax.value

    def output_open(self, directory=".", echo_stdout=False):
        """openfiles (main.f:105-118) in `directory`; starts the writer thread."""
        _check(lib().dg_output_open(os.fspath(directory).encode(), _i(1 if echo_stdout else 0)), "dg_output_open")

    def printout(self, t):
        """printout(t, cc, p) (main.f:5020-5100): queues the downloads and the
        files of step t; returns before they are written."""
        _check(lib().dg_printout(_i(t)), "dg_printout")

    def output_flush(self):
        _check(lib().dg_output_flush(), "dg_output_flush")

    def output_close(self):
        _check(lib().dg_output_close(), "dg_output_close")

    def output_stats(self):
        jobs = ctypes.c_int()
        b, w, t = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().dg_output_stats(ctypes.byref(jobs), ctypes.byref(b), ctypes.byref(w), ctypes.byref(t)),
               "dg_output_stats")
        return dict(jobs=jobs.value, bytes=b.value, wait_ms=w.value, write_ms=t.value)

    def savedata(self, path):
        """savedata (main.f:5414-5467): the restart records iniget reads."""
        _check(lib().dg_savedata(os.fspath(path).encode()), "dg_savedata")

    def save_state(self, path):
        """The full resident state, for a bit-identical resume."""
        _check(lib().dg_save_state(os.fspath(path).encode()), "dg_save_state")

    def load_state(self, path):
        _check(lib().dg_load_state(os.fspath(path).encode()), "dg_load_state")


def format_real(x, mode=0):
    """x as flang writes it list-directed (mode 0) or with E15.7 (mode 1);
    host code, no GPU."""
    buf = ctypes.create_string_buffer(64)
    _check(lib().dg_format_real(_d(x), _i(mode), buf, _i(64)), "dg_format_real")
    return buf.value.decode()


# ---- host-pointer kernels with the reference's argument lists -------------

def relax(u, lengthsize, rhs, a, b, c, ei, omega):
    _shapes("relax", u.shape, u, rhs, a, b, c, ei)
    nz, ny, nx = u.shape
    _check(lib().dg_relax(_p(u), _d(lengthsize), _p(rhs), _p(a), _p(b), _p(c), _p(ei), _i(2), _i(nx), _i(ny),
                          _i(nz), _d(omega)), "dg_relax")


def resid(lengthsize, res, u, rhs, a, b, c):
    _shapes("resid", u.shape, res, u, rhs, a, b, c)
    nz, ny, nx = u.shape
    _check(lib().dg_resid(_d(lengthsize), _p(res), _p(u), _p(rhs), _p(a), _p(b), _p(c), _i(2), _i(nx), _i(ny),
                          _i(nz)), "dg_resid")


def rstrct(uc, uf):
    _shapes("rstrct", uc.shape, uc)
    _shapes("rstrct", tuple(2 * n - 2 for n in uc.shape), uf)
    ncz, ncy, ncx = uc.shape
    _check(lib().dg_rstrct(_p(uc), _p(uf), _i(2), _i(ncx), _i(ncy), _i(ncz)), "dg_rstrct")


def interp(uf, uc):
    _shapes("interp", uf.shape, uf)
    _shapes("interp", _coarse(uf.shape), uc)
    nz, ny, nx = uf.shape
    _check(lib().dg_interp(_p(uf), _p(uc), _i(2), _i(nx), _i(ny), _i(nz)), "dg_interp")


def addint(uf, uc, res):
    _shapes("addint", uf.shape, uf, res)
    _shapes("addint", _coarse(uf.shape), uc)
    nz, ny, nx = uf.shape
    _check(lib().dg_addint(_p(uf), _p(uc), _p(res), _i(2), _i(nx), _i(ny), _i(nz)), "dg_addint")


def acfromac(a, b, c, ei, a2, b2, c2):
    _shapes("acfromac", a.shape, a, b, c, ei)
    _shapes("acfromac", tuple(2 * n - 2 for n in a.shape), a2, b2, c2)
    nz, ny, nx = a.shape
    _check(lib().dg_acfromac(_p(a), _p(b), _p(c), _p(ei), _p(a2), _p(b2), _p(c2), _i(nx), _i(ny), _i(nz)),
           "dg_acfromac")


def slvsml(u, rhs, a, b, c, ei, ntop, omega, lengthsize):
    _shapes("slvsml", u.shape, u, rhs, a, b, c, ei)
    nz, ny, nx = u.shape
    _check(lib().dg_slvsml(_p(u), _p(rhs), _p(a), _p(b), _p(c), _p(ei), _i(2), _i(nx), _i(ny), _i(nz), _i(ntop),
                           _d(omega), _d(lengthsize)), "dg_slvsml")


def abcei(a, b, c, ei, cc, r1, r2):
    _shapes("abcei", a.shape, a, b, c, ei, cc)
    nz, ny, nx = a.shape
    _check(lib().dg_abcei(_p(a), _p(b), _p(c), _p(ei), _p(cc), _d(r1), _d(r2), _i(2), _i(nx), _i(ny), _i(nz)),
           "dg_abcei")


def bc_scalar(s):
    _shapes("bc_scalar", s.shape, s)
    nz, ny, nx = s.shape
    _check(lib().dg_bc_scalar(_p(s), _i(2), _i(nx), _i(ny), _i(nz)), "dg_bc_scalar")


def filterc27(cc, cf):
    _shapes("filterc27", cc.shape, cc, cf)
    nz, ny, nx = cc.shape
    _check(lib().dg_filterc27(_p(cc), _p(cf), _i(nx), _i(ny), _i(nz)), "dg_filterc27")


def al3d(b1, b2, b3, cc):
    b1, b2, b3, cc = (np.ascontiguousarray(x, dtype=np.float64) for x in (b1, b2, b3, cc))
    if not (b1.size == b2.size == b3.size == cc.size):
        raise DropletError("al3d: inputs of different sizes")
    out = np.empty_like(b1)
    _check(lib().dg_al3d_batch(_i(b1.size), _p(b1), _p(b2), _p(b3), _p(cc), _p(out)), "dg_al3d_batch")
    return out


def fl3d(m1, m2, m3, alpha, r0, dr0):
    m1, m2, m3, alpha, r0, dr0 = (np.ascontiguousarray(x, dtype=np.float64) for x in (m1, m2, m3, alpha, r0, dr0))
    if len({x.size for x in (m1, m2, m3, alpha, r0, dr0)}) != 1:
        raise DropletError("fl3d: inputs of different sizes")
    out = np.empty_like(m1)
    _check(lib().dg_fl3d_batch(_i(m1.size), _p(m1), _p(m2), _p(m3), _p(alpha), _p(r0), _p(dr0), _p(out)),
           "dg_fl3d_batch")
    return out


def poisson_solve(rhs, a, b, c, p0=None, lengthsize=1.0, mode=2, tol=1e-8, maxit=200):
    """Tolerance-mode solve of resid's operator; returns (p, iterations, relres)."""
    nz, ny, nx = rhs.shape
    p = np.zeros_like(rhs) if p0 is None else np.array(p0, dtype=np.float64, copy=True)
    _shapes("poisson_solve", rhs.shape, rhs, a, b, c, p)
    it = ctypes.c_int(0)
    rel = ctypes.c_double(0.0)
    _check(lib().dg_poisson_solve(_i(nx), _i(ny), _i(nz), _d(lengthsize), _p(rhs), _p(a), _p(b), _p(c), _p(p),
                                  _i(mode), _d(tol), _i(maxit), ctypes.byref(it), ctypes.byref(rel)),
           "dg_poisson_solve")
    return p, it.value, rel.value


def set_smoother(omega=1.0, nu_coarse=0):
    """Tolerance-mode smoother: SOR weight of the red-black passes and the
    sweep count on the levels below the finest (0: the solver's nu1/nu2)."""
    _check(lib().dg_set_smoother(_d(omega), _i(nu_coarse)), "dg_set_smoother")


def set_tile_smoother(on=0, nu=0):
    """Tile smoother on the V-cycle levels between the finest and the tail
    (one launch per level visit, restricted additive Schwarz with red-black
    SOR inside each 16x8x8 tile): 1 (or True) on those of at most 2^18
    cells, 2 on all of them, 0 (False): global red-black passes.  nu: the
    sweeps before and after the coarse correction there (0: nu_coarse)."""
    _check(lib().dg_set_tile_smoother(_i(int(on)), _i(int(nu))), "dg_set_tile_smoother")


def set_tuning(graphs=True, split=True, tail="lds"):
    """Tolerance-mode execution switches:
    graphs -- replay each PCG iteration as a captured HIP graph (bit-identical);
    split -- mode 2's FP32 V-cycle on red-black split storage with FP16
             coefficients where the densities allow (same tolerance reached);
             False natural storage, "fp32" / "fp16" split with that
             coefficient type regardless of the densities;
    tail -- with split storage, the levels of at most 4096 cells and the
            coarsest solve: None one launch per operation, "global" one
            single-workgroup kernel, "lds" that kernel on LDS copies when
            they fit (bit-identical)."""
    sv = {False: 0, True: 1, "fp32": 2, "fp16": 3}[split]
    tv = {None: 0, "global": 1, "lds": 2}[tail]
    _check(lib().dg_set_tuning(_i(1 if graphs else 0), _i(sv), _i(tv)), "dg_set_tuning")


def set_lookahead(ratio=1e3):
    """Queue PCG iteration n+1 before iteration n's residual reaches the host
    while the last residual seen exceeds ratio * tol (<= 0: never);
    bit-identical either way."""
    _check(lib().dg_set_lookahead(_d(ratio)), "dg_set_lookahead")


# kernel families dg_profile_start accepts
PROFILE_KERNELS = {"gs_colour": 1, "resid": 2, "pcg_dir": 3, "pcg_update": 4, "relax": 5, "sweep": 6,
                   "rb_colour": 7, "conv_fwd": 8}


def profile_start(kernel):
    """Start HIP-event timing of every launch of one kernel family on the
    library's stream (graph replay is off meanwhile)."""
    kid = PROFILE_KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
    _check(lib().dg_profile_start(ctypes.byref(ctypes.c_int(kid))), "dg_profile_start")


def profile_stop():
    """Stop timing; returns dict(total_ms, launches, bytes) where bytes is the
    algorithmic byte count of the timed launches (for conv_fwd: their
    algorithmic FLOPs)."""
    ms, nl, by = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
    _check(lib().dg_profile_stop(ctypes.byref(ms), ctypes.byref(nl), ctypes.byref(by)), "dg_profile_stop")
    return dict(total_ms=ms.value, launches=nl.value, bytes=by.value)
