"""Writes the golden fixtures under tests/golden/ from the flang-built
reference (oracle/_ref/libref_*.so, built by oracle/build_ref.sh from
/root/reference/main.f).  Run in the build container, where the reference
exists:  python tests/golden/make_golden.py

Fixtures are data only: seeded inputs and the reference's outputs.

  plic_vectors.npz   AL3D (main.f:4562) / FL3D (main.f:4639) on random and
                     edge-case inputs
  kernels_18.npz     relax, resid, rstrct, interp, addint, acfromac, abcei,
                     filterc27, bc_scalar, slvsml on seeded 18^3 / 10^3 /
                     6x4x4 inputs
  run_34.npz         10 steps at 34^3 (droplet, CSF, central convection):
                     per-step history and final fields
  run_66x34x34.npz   6 steps on the shipped input.h grid: history and fields
  run_18_upwind_keps.npz  upwind convection + k-epsilon from k0/eps0, 5 steps
  run_66_hist.npz    17 steps at 66^3: per-step history and field checksums
  run_shipped_sgip.npz  the reference's shipped parameters (66x34x34, hswitch=1
                     CalcH3d heights, surfacet=5 SGIP with its contact-angle
                     wall row; initial cc clipped to [0, 1], see below): 4
                     steps, history, fields and heights; then SGIP2 and CSF
                     with CalcH3d, 3 steps each: history and checksums
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle_ffi import inputdata_text  # noqa: E402
from refrun import RefKernels, run_reference  # noqa: E402

HIST_KEYS = ("tau", "time", "tauh", "tauh2", "npre", "ntop", "npost", "vol", "tswap")


def hist_arrays(hist):
    return {f"h_{k}": np.array([h[k] for h in hist]) for k in HIST_KEYS}


def plic():
    K = RefKernels(18, 18, 18)
    rng = np.random.default_rng(20240611)
    n = 4000
    m = rng.uniform(0, 1, size=(n, 3))
    m[:300, 0] = 0.0                      # one zero component
    m[300:500, :2] = 0.0                  # axis-aligned normals
    m[500:700, 1] = m[500:700, 0]         # ties
    m /= m.sum(axis=1, keepdims=True)
    cc = rng.uniform(0, 1, size=n)
    cc[700:900] = rng.uniform(0, 1e-7, size=200)
    cc[900:1100] = 1 - rng.uniform(0, 1e-7, size=200)
    cc[1100:1200] = 0.5
    al = np.array([K.al3d(m[q, 0], m[q, 1], m[q, 2], cc[q]) for q in range(n)])
    mm = rng.uniform(-1, 1, size=(n, 3))
    alpha = rng.uniform(-0.5, 1.5, size=n)
    r0 = rng.uniform(-0.3, 1.0, size=n)
    dr0 = rng.uniform(0.01, 1.0, size=n)
    fl = np.array([K.fl3d(mm[q, 0], mm[q, 1], mm[q, 2], alpha[q], r0[q], dr0[q]) for q in range(n)])
    np.savez_compressed(os.path.join(HERE, "plic_vectors.npz"), al_m=m, al_cc=cc, al_out=al, fl_m=mm,
                        fl_alpha=alpha, fl_r0=r0, fl_dr0=dr0, fl_out=fl)


def coef(shape, rng):
    return np.where(rng.uniform(size=shape) < 0.5, 1.0, 833.3333) * rng.uniform(0.9, 1.1, size=shape)


def kernels():
    K = RefKernels(18, 18, 18)
    rng = np.random.default_rng(7)
    out = {}
    s = (18, 18, 18)  # (nz,ny,nx)
    u, rhs = rng.uniform(-1, 1, s), rng.uniform(-1, 1, s)
    a, b, c = coef(s, rng), coef(s, rng), coef(s, rng)
    ei = 1.0 / (a + b + c)
    out.update(in_u=u, in_rhs=rhs, in_a=a, in_b=b, in_c=c, in_ei=ei)
    u1 = u.copy()
    K.relax(u1, 1.0, rhs, a, b, c, ei, 1.5)
    out["relax"] = u1
    res = np.zeros(s)
    K.resid(1.0, res, u1, rhs, a, b, c)
    out["resid"] = res
    uc = np.zeros((10, 10, 10))
    K.rstrct(uc, u)
    out["rstrct"] = uc
    uf = rhs.copy()
    K.interp(uf, uc)
    out["interp"] = uf
    g, r2 = u.copy(), np.zeros(s)
    K.addint(g, uc, r2)
    out["addint_uf"], out["addint_res"] = g, r2
    ca = [np.zeros((10, 10, 10)) for _ in range(4)]
    K.acfromac(*ca, a, b, c)
    out["acfromac_a"], out["acfromac_b"], out["acfromac_c"], out["acfromac_ei"] = ca
    cc = rng.uniform(0, 1, s)
    cc[rng.uniform(size=s) < 0.3] = 1.0
    out["in_cc"] = cc
    ab = [np.zeros(s) for _ in range(4)]
    K.abcei(*ab, cc, 0.4994, 0.0012)
    out["abcei_a"], out["abcei_b"], out["abcei_c"], out["abcei_ei"] = ab
    cf = np.zeros(s)
    K.filterc27(cc, cf)
    out["filterc27"] = cf
    sc = u.copy()
    K.bc_scalar(sc)
    out["bc_scalar"] = sc
    s6 = (4, 4, 6)
    u6, r6 = rng.uniform(-1, 1, s6), rng.uniform(-1, 1, s6)
    a6, b6, c6 = coef(s6, rng), coef(s6, rng), coef(s6, rng)
    e6 = 1.0 / (a6 + b6 + c6)
    out.update(in6_u=u6, in6_rhs=r6, in6_a=a6, in6_b=b6, in6_c=c6, in6_ei=e6)
    u6b = u6.copy()
    K.slvsml(u6b, r6, a6, b6, c6, e6, 7, 1.5, 0.004)
    out["slvsml"] = u6b
    np.savez_compressed(os.path.join(HERE, "kernels_18.npz"), **out)


def runs():
    txt = inputdata_text(nx=34, ny=34, nz=34)
    r = run_reference(txt, 10, want_fields=("cc", "u", "v", "w", "p", "s12", "s13", "s23"))
    np.savez_compressed(os.path.join(HERE, "run_34.npz"), inputdata=np.array(txt), **hist_arrays(r["hist"]),
                        **{f"f_{k}": v for k, v in r["fields"].items()})
    txt = inputdata_text(nx=66, ny=34, nz=34, y0=0.0015, radius=0.0005)
    r = run_reference(txt, 6, want_fields=("cc", "u", "v", "w", "p"))
    np.savez_compressed(os.path.join(HERE, "run_66x34x34.npz"), inputdata=np.array(txt), **hist_arrays(r["hist"]),
                        **{f"f_{k}": v for k, v in r["fields"].items()})
    txt = inputdata_text(nx=18, ny=18, nz=18, iturb=3, convectype=1)
    init = dict(ktur=1e-4, eps=1e-3 * 0.004)
    r = run_reference(txt, 5, init_fields=init, want_fields=("cc", "u", "v", "w", "p", "ktur", "eps", "mut"))
    np.savez_compressed(os.path.join(HERE, "run_18_upwind_keps.npz"), inputdata=np.array(txt),
                        init_ktur=init["ktur"], init_eps=init["eps"], **hist_arrays(r["hist"]),
                        **{f"f_{k}": v for k, v in r["fields"].items()})
    txt = inputdata_text(nx=66, ny=66, nz=66, radius=0.00079)
    r = run_reference(txt, 17, want_fields=("cc", "u", "v", "w", "p"))
    chk = {}
    for k, v in r["fields"].items():
        inner = v[1:-1, 1:-1, 1:-1]
        chk[f"sum_{k}"] = inner.sum()
        chk[f"sumsq_{k}"] = (inner * inner).sum()
        chk[f"maxabs_{k}"] = np.abs(inner).max()
    np.savez_compressed(os.path.join(HERE, "run_66_hist.npz"), inputdata=np.array(txt), **hist_arrays(r["hist"]),
                        **chk)


from oracle_ffi import clipped_cc, shipped_text  # noqa: E402


def sgip_runs():
    from refrun import FIELD_NAMES, HEIGHT_NAMES
    out = {}
    txt = shipped_text()
    cc0 = clipped_cc(txt)
    r = run_reference(txt, 4, init_fields={"cc": cc0}, want_fields=FIELD_NAMES + HEIGHT_NAMES)
    out.update(inputdata=np.array(txt), init_cc=cc0, **hist_arrays(r["hist"]))
    out.update({f"f_{k}": r["fields"][k] for k in ("cc", "u", "v", "w", "p", "s11", "s22", "s33") + HEIGHT_NAMES})
    for tag, over in (("sgip2", dict(surfacet=6)), ("csfh", dict(surfacet=2))):
        t2 = shipped_text(**over)
        r = run_reference(t2, 3, init_fields={"cc": cc0}, want_fields=FIELD_NAMES + HEIGHT_NAMES)
        out[f"{tag}_inputdata"] = np.array(t2)
        out.update({f"{tag}_{k}": v for k, v in hist_arrays(r["hist"]).items()})
        for k, v in r["fields"].items():
            out[f"{tag}_sum_{k}"] = v[1:-1, 1:-1, 1:-1].sum()
            out[f"{tag}_sumsq_{k}"] = (v[1:-1, 1:-1, 1:-1] ** 2).sum()
    np.savez_compressed(os.path.join(HERE, "run_shipped_sgip.npz"), **out)


if __name__ == "__main__":
    which = sys.argv[1:] or ["plic", "kernels", "runs", "sgip"]
    if "plic" in which:
        plic()
    if "kernels" in which:
        kernels()
    if "runs" in which:
        runs()
    if "sgip" in which:
        sgip_runs()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
