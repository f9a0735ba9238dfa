"""Benchmark: time steps/s of the droplet solver at 256^3 with every step's
pressure Poisson solve taken to a 1e-8 FP64 relative residual.

Workload (one MI355X): 258^3 grid with ghosts (256^3 cells) on a 6 mm cube,
a water droplet (r = 1 mm) in air 1.5 mm above the no-slip wall (reference
axis y is vertical), impact velocity -3 m/s, gravity -9.81, CSF surface
tension, laminar.  Synthetic: the droplet initial condition of the
reference's smooth2, no dataset.  Two deviations from the reference's
arithmetic keep this workload finite: the initial volume fraction is clipped
to [0, 1] (smooth2 leaves values above 1 that AL3D turns into NaN), and the
VOF guard advects partly filled cells whose Youngs normal vanishes with a
plane normal to the sweep (the reference evaluates 0/0 there; on this
workload that first happens within the first ten steps).  A timed step whose
Poisson solve does not reach the tolerance, or non-finite fields after the
timed steps, abort the run.

Measured paths, each timed over K steps after W warm-up steps:
  target       FP64 fields, PCG with an FP32 V-cycle preconditioner (FP64
               residual/update; red-black split storage, FP16 face
               coefficients), initial guess: the learned network's pressure
               prediction (BF16, on the 128^3 level) over a zero start,
               accepted only when it lowers the residual, with the fallback
               controller;
  mixed        the same solver without the network, with the faster of the
               zero and previous-step guesses;
  denominator  the same PCG with an FP64 V-cycle (the plain FP64 multigrid
               solver) on the same red-black split storage, launch shapes and
               single-workgroup tail, FP64 coefficients, with the faster of
               the zero and previous-step guesses;
  fp64 natural the FP64 V-cycle on the natural layout (reported beside).
Every path uses the same smoother: V(2,2), SOR weight 1.15, 4 sweeps on the
levels in the single-workgroup tail, and (--tiles 2, the default) the levels
between the finest and the tail (128^3, 64^3 and 32^3 here) smoothed as
tiles, 5 sweeps before and after the coarse correction in one launch each
(the natural-layout FP64 path has no tiles and 4 sweeps on every level
below the finest).
The target and the denominator are also timed over a later window of the
same run (steps 150..169 by default), where the splash has developed.
Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "time steps/s at 256^3 to 1e-8 Poisson residual: learned+mixed MG vs FP64 MG"
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense BF16 MFMA of one MI355X
# HBM bytes per launch of the dominant kernel family at 258^3, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this script scaled by
# the same counters on a plain 256 MiB copy (fetch x1.9998, write x1.0):
# of the 8 finest passes of a V(2,2) cycle, 232.1 MB for each of the 5
# plain passes, 167.7 MB for the zero-iterate pass and 251.9 + 246.6 MB for
# the two passes that take the prolongation in, averaged as bench.py times
# them; keyed by kernel and grid extent, re-measured whenever the kernel
# changes
ROOF_KERNEL = "k_rb_colour4<_Float16,*,true>"
PMC_TRAFFIC = {(ROOF_KERNEL, 258): (5 * 230.0e6 + 167.7e6 + 246.0e6 + 245.1e6) / 8}


TILE_LABEL = {0: "", 1: "; levels of 4096 < cells <= 2^18 as 16x8x8 tiles (restricted additive Schwarz, "
                        "one launch per level visit)",
              2: "; every level of more than 4096 cells below the finest as tiles of 16x8x8 cells "
                 "(32x16x8 above 2^18 cells; restricted additive Schwarz, one launch per level visit)"}


def splash_inputdata(dg, n, **kw):
    # 6 mm cube; droplet radius 1 mm, its lowest point 1.5 mm above the wall
    # (centre 2.5 mm); the reference centres z and forces x0 = L/4.
    return dg.make_inputdata(nx=n, ny=n, nz=n, tmax=200, maxcfl=0.25, iturb=0, convectype=2, yvel1=-3.0,
                             rho2=1.2, rho1=1000.0, sigma=0.072, visc2=1.8e-5, visc1=1.0e-3, grav=-9.81,
                             length=0.006, radius=0.001, y0=0.0025, z0=1000.0, cycles=2, npre=2, ntop=4, npost=2,
                             **kw)


def run_path(dg, txt, mode, steps, warmup, tol, guess=1, profile_kernel=None, profile_steps=2, late_start=0,
             late_steps=0, output_dir=None):
    """W warm-up steps, K timed steps (headline, no instrumentation), then,
    when profile_kernel is given, profile_steps more steps of the same run
    with HIP events around every launch of that kernel on the library's
    stream (graph replay is off while events are recorded), then, when
    late_steps > 0, untimed steps up to late_start and late_steps timed
    steps from there.  With output_dir, printout runs before every step (the
    reference's loop, main.f:21-23) into that directory; the time to finish
    the queued files after the timed steps is measured apart."""
    import numpy as np
    gs = importlib.import_module("3d_droplet_simulation_2007_amd.guess")
    g = dg.DropletGPU(inputdata_text=txt)
    g.clip_cc()
    g.set_ordered_sums(False)
    g.set_vof_guard(True)
    g.set_solver(mode=mode, tol=tol, nu1=2, nu2=2, maxit=200, guess=guess)
    vol0 = float(g.get("cc")[1:-1, 1:-1, 1:-1].sum())
    if output_dir:
        g.output_open(output_dir)
    for t in range(warmup):
        if output_dir:
            g.printout(t)
        g.step(t)
    g.state()  # synchronises the stream
    if output_dir:
        g.output_flush()
    gs0 = gs.stats()
    o0 = g.output_stats()
    t0 = time.perf_counter()
    timed = []
    for t in range(warmup, warmup + steps):
        if output_dir:
            g.printout(t)
        g.step(t)
        s = g.state()
        timed.append((s["mg_iters"], s["mg_relres"]))
    st = g.state()  # synchronises
    t1 = time.perf_counter()
    output = None
    if output_dir:
        g.output_flush()
        t2 = time.perf_counter()
        o1 = g.output_stats()
        g.output_close()
        output = dict(drain_ms=1e3 * (t2 - t1), **{k: o1[k] - o0[k] for k in o1})
    gs1 = gs.stats()
    guess_stats = {k: gs1[k] - gs0[k] for k in ("accepted", "rejected", "skipped")}
    ran = guess_stats["accepted"] + guess_stats["rejected"]
    if ran:
        tot = gs1["mean_gain"] * (gs1["accepted"] + gs1["rejected"]) - gs0["mean_gain"] * (gs0["accepted"] + gs0["rejected"])
        guess_stats["mean_residual_gain"] = tot / ran
    # a step whose solve did not reach the tolerance (NaN residual, zero or
    # exhausted iterations) makes the measurement invalid
    bad = [(warmup + q, it, rr) for q, (it, rr) in enumerate(timed) if not (it >= 1 and rr < tol)]
    if bad:
        raise RuntimeError(f"mode {mode}: timed steps whose Poisson solve did not converge: {bad}")
    vol1 = float(g.get("cc")[1:-1, 1:-1, 1:-1].sum())
    p = g.get("p")
    for name in ("u", "v", "w", "cc"):
        if not np.isfinite(g.get(name)).all():
            raise RuntimeError(f"mode {mode}: non-finite {name} after the timed steps")
    timed_iters = [it for it, _ in timed]
    prof = None
    if profile_kernel is not None:
        dg.profile_start(profile_kernel)
        for t in range(warmup + steps, warmup + steps + profile_steps):
            g.step(t)
        prof = dg.profile_stop()
        if prof["launches"] == 0:
            raise RuntimeError(f"profiling recorded no launch of kernel family {profile_kernel!r}")
    late = None
    if late_steps > 0:
        t = warmup + steps + (profile_steps if profile_kernel is not None else 0)
        while t < late_start:
            g.step(t)
            t += 1
        g.state()
        lt0 = time.perf_counter()
        its = []
        for q in range(late_steps):
            g.step(t + q)
            s = g.state()
            its.append(s["mg_iters"])
            if not s["mg_relres"] < tol:
                raise RuntimeError(f"mode {mode}: late-window step {t + q} did not converge")
        lt1 = time.perf_counter()
        late = dict(first_step=t, steps=late_steps, ms_per_step=1e3 * (lt1 - lt0) / late_steps,
                    mg_iters_per_step=float(np.mean(its)))
    g.close()
    print(f"bench: mode {mode} guess {guess}: {1e3 * (t1 - t0) / steps:.2f} ms/step over {steps} steps, "
          f"{float(np.mean([q[0] for q in timed])):.2f} PCG iterations/step"
          + (f", later window {late['ms_per_step']:.2f} ms/step" if late else ""), file=sys.stderr, flush=True)
    return dict(seconds=t1 - t0, steps=steps, iters=timed_iters, relres=st["mg_relres"], tau=st["tau"],
                vol0=vol0, vol1=vol1, prof=prof, p=p, guess_stats=guess_stats, late=late, output=output)


def host_facts():
    """CPU model, logical cores and the restatement's compiler and flags."""
    import platform
    import subprocess
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cc = subprocess.run(["gcc", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        cc = "gcc (version unknown)"
    flags = ""
    try:
        with open(os.path.join(HERE, "oracle", "Makefile")) as f:
            for ln in f:
                if ln.startswith("CFLAGS"):
                    flags = ln.split("=", 1)[1].strip()
    except OSError:
        pass
    return dict(cpu_model=model, host_logical_cores=os.cpu_count(), compiler=cc, flags=flags)


def cpu_baseline(n, steps, tol, omega=1.0, nu_coarse=0):
    """The CPU restatement (oracle), one thread, FP64 PCG with an FP64 V-cycle
    (the same algorithm as the denominator), on `steps` steps of the same
    workload; initialisation excluded."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import oracle_ffi
    dg = importlib.import_module("3d_droplet_simulation_2007_amd")
    txt = splash_inputdata(dg, n)
    o = oracle_ffi.Oracle(txt)
    cc = o.field("cc")
    cc[...] = cc.clip(0.0, 1.0)
    o.st.mg_mode, o.st.mg_outer, o.st.mg_prolong, o.st.vof_guard = 1, 1, 0, 1
    o.st.mg_tol, o.st.mg_nu1, o.st.mg_nu2, o.st.mg_maxcycles, o.st.mg_guess = tol, 2, 2, 200, 1
    o.st.mg_omega, o.st.mg_nuc = omega, nu_coarse
    t0 = time.perf_counter()
    for t in range(steps):
        o.step(t)
    dt = time.perf_counter() - t0
    o.close()
    return steps / dt, dt


def copy_calibration(mib):
    """Five plain device-to-device copies of `mib` MiB: under rocprofv3 --pmc
    their FETCH_SIZE / WRITE_SIZE per dispatch, divided by `mib` MiB, give the
    factor that turns those counters into HBM bytes."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.rand(n, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=258, help="extent with ghosts (2^k+2)")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-denominator", action="store_true")
    ap.add_argument("--omega", type=float, default=1.15, help="SOR weight of the V-cycle smoother (all paths)")
    ap.add_argument("--nu-coarse", type=int, default=4, help="smoothing sweeps below the finest level (all paths)")
    ap.add_argument("--tiles", type=int, choices=(0, 1, 2), default=2,
                    help="tile smoother on the V-cycle levels between the finest and the tail (all paths): "
                         "1 on those of at most 2^18 cells, 2 on all")
    ap.add_argument("--tile-sweeps", type=int, default=5,
                    help="sweeps before and after the coarse correction on the tile levels (0: --nu-coarse)")
    ap.add_argument("--late-start", type=int, default=150, help="first step of the later timed window (0: none)")
    ap.add_argument("--late-steps", type=int, default=20, help="steps in the later timed window")
    ap.add_argument("--weights", default=os.path.join(HERE, "3d_droplet_simulation_2007_amd", "weights",
                                                       "guess_unet.npz"), help="learned-guess network weights")
    ap.add_argument("--no-output", action="store_true", help="skip the run with printout's files")
    ap.add_argument("--only-path", choices=("fp64", "mixed"), default=None,
                    help="run one path alone (zero guess) for a profile: the FP64 denominator or the mixed "
                         "solver without the network; prints its ms/step and no JSON line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (scale for HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)
    dg = importlib.import_module("3d_droplet_simulation_2007_amd")
    dg.lib()
    # one smoother for every path (target, mixed and the FP64 denominator)
    dg.set_smoother(args.omega, args.nu_coarse)
    dg.set_tile_smoother(args.tiles, args.tile_sweeps)
    n = args.grid
    txt = splash_inputdata(dg, n)
    if args.only_path:
        run_path(dg, txt, 1 if args.only_path == "fp64" else 2, args.steps, args.warmup, args.tol, guess=0)
        return
    gs = importlib.import_module("3d_droplet_simulation_2007_amd.guess")
    weights_meta = gs.load_weights(args.weights)
    import numpy as np
    # target: learned guess + FP32-preconditioned FP64 PCG; dominant kernel:
    # the finest-level colour pass of the red-black split FP32 V-cycle
    late = dict(late_start=args.late_start, late_steps=args.late_steps if args.late_start > 0 else 0)
    tgt = run_path(dg, txt, 2, args.steps, args.warmup, args.tol, guess=2, profile_kernel="rb_colour", **late)
    out = {}
    den = mixed = None
    if not args.no_denominator:
        # the mixed solver without the network, with the better of its zero
        # and previous-step guesses, as the denominator gets
        mixes = {gname: run_path(dg, txt, 2, args.steps, args.warmup, args.tol, guess=gi)
                 for gi, gname in ((0, "zero"), (1, "previous step"))}
        mguess = min(mixes, key=lambda k: mixes[k]["seconds"])
        mixed = mixes[mguess]
        # the plain FP64 solver with the better of its zero and previous-step
        # guesses, on the split layout the mixed solver uses
        dens = {gname: run_path(dg, txt, 1, args.steps, args.warmup, args.tol, guess=gi)
                for gi, gname in ((0, "zero"), (1, "previous step"))}
        dguess = min(dens, key=lambda k: dens[k]["seconds"])
        den = dens[dguess]
        # its later window, with the faster guess, in a run of its own
        den_late = run_path(dg, txt, 1, args.steps, args.warmup, args.tol, guess=(0 if dguess == "zero" else 1),
                            **late)["late"]
        # the FP64 V-cycle on the natural layout (stride-2 colour passes)
        dg.set_tuning(split=False)
        try:
            nat = run_path(dg, txt, 1, args.steps, args.warmup, args.tol, guess=(0 if dguess == "zero" else 1))
        finally:
            dg.set_tuning()
        a, b = tgt["p"][1:-1, 1:-1, 1:-1], den["p"][1:-1, 1:-1, 1:-1]
        out["p_rel_l2_vs_fp64"] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    value = tgt["steps"] / tgt["seconds"]
    pr = tgt["prof"]
    avg_ms = pr["total_ms"] / max(pr["launches"], 1)
    achieved = (pr["bytes"] / max(pr["launches"], 1)) / (avg_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "time steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tgt["seconds"] / tgt["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 fields and PCG, f32 V-cycle iterate, f16 face coefficients, bf16 guess network",
        "data": "synthetic (reference smooth2 droplet IC)",
        "config": {"workload": f"{n - 2}^3 droplet-on-wall splash, 6 mm cube, r=1 mm, v=-3 m/s, CSF, "
                               f"each step's Poisson solve to {args.tol:g} relative residual",
                   "grid_with_ghosts": [n, n, n], "solver": f"FP64 PCG + FP32 V(2,2) preconditioner, red-black SOR omega {args.omega:g}, "
                                                         f"{args.nu_coarse} sweeps below the finest level"
                                                         + TILE_LABEL[args.tiles]
                                                         + (f", {args.tile_sweeps} sweeps on the tile levels"
                                                            if args.tiles and args.tile_sweeps else ""),
                   "initial_guess": "learned pressure prediction (encoder-decoder on the next-coarser level) "
                                    "over a zero start, kept only when it lowers the residual, fallback zero, "
                                    "with the fallback controller",
                   "parallelism": "single GPU"},
        "guess_weights": {k: weights_meta.get(k) for k in ("grid", "samples", "epochs")},
        "learned_guess": tgt["guess_stats"],
        "mg_iters_per_step": float(np.mean(tgt["iters"])),
        "final_relres": tgt["relres"],
        "volume_drift_rel": (tgt["vol1"] - tgt["vol0"]) / tgt["vol0"],
        "roofline": {"kernel": ROOF_KERNEL + " (finest-level SOR colour passes of the red-black split "
                                           "FP32 V-cycle, FP16 face coefficients; the zero-iterate pass and the two "
                                           "passes that take the prolongation in included)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": PMC_TRAFFIC.get((ROOF_KERNEL, n)),
                     "avg_launch_us": 1e3 * avg_ms, "launches": pr["launches"],
                     "algorithmic_bytes_per_launch": pr["bytes"] / max(pr["launches"], 1)},
    }
    if den is not None:
        dv = den["steps"] / den["seconds"]
        line["denominator_fp64"] = {"value": dv, "ms_per_step": 1e3 * den["seconds"] / den["steps"],
                                    "mg_iters_per_step": float(np.mean(den["iters"])), "guess": dguess,
                                    "solver": "FP64 PCG + FP64 V(2,2) on red-black split storage, FP64 "
                                              "coefficients, the same launch shapes, tile smoother and "
                                              "single-workgroup tail",
                                    "other_guess_ms_per_step": {k: 1e3 * v["seconds"] / v["steps"]
                                                                for k, v in dens.items() if k != dguess},
                                    "late_window": den_late}
        line["speedup_vs_fp64"] = value / dv
        nv = nat["steps"] / nat["seconds"]
        line["fp64_natural_layout"] = {"value": nv, "ms_per_step": 1e3 * nat["seconds"] / nat["steps"],
                                       "mg_iters_per_step": float(np.mean(nat["iters"])),
                                       "speedup_of_target": value / nv}
        if tgt["late"] and den_late:
            line["late_window"] = {"first_step": tgt["late"]["first_step"], "steps": tgt["late"]["steps"],
                                   "target_ms_per_step": tgt["late"]["ms_per_step"],
                                   "target_mg_iters_per_step": tgt["late"]["mg_iters_per_step"],
                                   "denominator_ms_per_step": den_late["ms_per_step"],
                                   "denominator_mg_iters_per_step": den_late["mg_iters_per_step"],
                                   "speedup_vs_fp64": den_late["ms_per_step"] / tgt["late"]["ms_per_step"]}
        mv = mixed["steps"] / mixed["seconds"]
        line["mixed_without_learned_guess"] = {"value": mv, "ms_per_step": 1e3 * mixed["seconds"] / mixed["steps"],
                                               "mg_iters_per_step": float(np.mean(mixed["iters"])), "guess": mguess,
                                               "speedup_vs_fp64": mv / dv}
    line.update(out)
    # the guess network's BF16 MFMA convolutions (forward) on an input of
    # the workload's next-coarser level, timed with HIP events around their
    # launches on the network's stream
    nc = (n - 2) // 2
    feat = np.random.default_rng(5).standard_normal((nc, nc, nc, 4)).astype(np.float32)
    gs.forward(feat)
    dg.profile_start("conv_fwd")
    for _ in range(3):
        gs.forward(feat)
    cp = dg.profile_stop()
    conv_tf = cp["bytes"] / (cp["total_ms"] * 1e-3) / 1e12
    line["mfma_conv"] = {"kernel": "k_conv_fwd_lds (BF16 MFMA 16x16x32, FP32 accumulate, input tile staged in LDS), "
                                   "every layer of the network",
                         "bound": "mfma", "achieved": conv_tf, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": conv_tf / MFMA_BF16_PEAK_TFLOPS, "input": [nc, nc, nc, 4],
                         "flops_per_forward": cp["bytes"] / 3, "ms_per_forward": cp["total_ms"] / 3,
                         "launches_per_forward": cp["launches"] // 3}
    if not args.no_output and mixed is not None:
        # the mixed path again, with printout's files at tprint 5 and
        # ntecplot 10 (a full-field Tecplot file every 10 steps), written to
        # a scratch directory by the writer thread
        import tempfile
        otxt = splash_inputdata(dg, n, tprint=5, ntecplot=10)
        with tempfile.TemporaryDirectory() as d:
            o = run_path(dg, otxt, 2, args.steps, args.warmup, args.tol, guess=(0 if mguess == "zero" else 1),
                         output_dir=d)
        st = o["output"]
        line["output"] = {"cadence": "printout every step: diagnostics every 5 steps, mainNNNN.plt every 10",
                          "ms_per_step": 1e3 * o["seconds"] / o["steps"],
                          "ms_per_step_without_output": 1e3 * mixed["seconds"] / mixed["steps"],
                          "printouts_queued": st["jobs"], "bytes_downloaded": st["bytes"],
                          "step_wait_for_writer_ms": st["wait_ms"], "writer_thread_ms": st["write_ms"],
                          "drain_after_timed_steps_ms": st["drain_ms"]}
    if not args.no_cpu_baseline:
        v, dt = cpu_baseline(n, args.cpu_steps, args.tol, args.omega, args.nu_coarse)
        line["cpu_baseline"] = {"value": v, "unit": "time steps/s", "cores": 1, "kind": "port",
                                "sample": f"{args.cpu_steps} step(s) of the same {n - 2}^3 workload, FP64 PCG + "
                                          f"FP64 V-cycle (the denominator's algorithm), {dt:.1f} s, one thread",
                                **host_facts()}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
