"""Train and evaluate the learned initial guess of the pressure solve.

Runs on one GPU with libdroplet_gpu.so doing all the work: seeded
single-droplet splash cases are stepped with the tolerance-mode solver
(FP64 PCG with an FP32 V-cycle, SOR weight and coarse sweeps as bench.py
uses them) while every solve's network input and target (the converged
pressure on the next-coarser level) are captured in device memory; the
network is then trained there with Adam; finally held-out seeded cases are
evaluated by the PCG iterations needed to reach the tolerance from the
zero, the previous-step and the learned guess on the same systems.

Cases (6 mm cube, water in air, CSF, gravity, no-slip walls): droplet
radius uniform in [0.5, 1.2] mm, centre height uniform between 0.3 mm above
the wall plus the radius and 4.2 mm, impact velocity uniform in
[-3, -0.5] m/s.  The reference fixes the droplet's x at L/4 and centres z.

usage (from the repository root):
  python 3d_droplet_simulation_2007_amd/train_guess.py --grid 258 --cases 40 --steps 50 \\
      --epochs 4 --out 3d_droplet_simulation_2007_amd/weights/guess_unet.npz --eval-cases 20
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def case_params(rng):
    r = rng.uniform(0.5e-3, 1.2e-3)
    y0 = rng.uniform(r + 0.3e-3, 4.2e-3)
    v = rng.uniform(-3.0, -0.5)
    return dict(radius=r, y0=y0, yvel1=v)


def case_inputdata(dg, n, p):
    return dg.make_inputdata(nx=n, ny=n, nz=n, tmax=100000, maxcfl=0.25, iturb=0, convectype=2,
                             yvel1=p["yvel1"], rho2=1.2, rho1=1000.0, sigma=0.072, visc2=1.8e-5, visc1=1.0e-3,
                             grav=-9.81, length=0.006, radius=p["radius"], y0=p["y0"], z0=1000.0, cycles=2,
                             npre=2, ntop=4, npost=2)


def open_case(dg, n, p, tol, guess):
    g = dg.DropletGPU(inputdata_text=case_inputdata(dg, n, p))  # smoother as set by main()
    g.clip_cc()
    g.set_ordered_sums(False)
    g.set_vof_guard(True)
    g.set_solver(mode=2, tol=tol, nu1=2, nu2=2, maxit=200, guess=guess)
    return g


def generate(dg, gs, args, rng):
    """Step the training cases with capture on; returns per-case records."""
    gs.capture(args.cases * args.steps)
    recs = []
    for ci in range(args.cases):
        p = case_params(rng)
        g = open_case(dg, args.grid, p, args.tol, 1)
        its, ok = [], True
        t0 = time.perf_counter()
        for t in range(args.steps):
            g.step(t)
            st = g.state()
            if not (st["mg_relres"] < args.tol):
                ok = False
                break
            its.append(st["mg_iters"])
        g.close()
        recs.append(dict(case=ci, **p, steps=len(its), ok=ok, mean_iters=float(np.mean(its)) if its else None))
        print(f"case {ci}: r={p['radius']*1e3:.2f} mm y0={p['y0']*1e3:.2f} mm v={p['yvel1']:.2f} m/s "
              f"steps={len(its)} ok={ok} iters/step={recs[-1]['mean_iters']} samples={gs.store_count()} "
              f"({time.perf_counter() - t0:.1f} s)", flush=True)
    return recs


def evaluate(dg, gs, args, rng):
    """Held-out cases: warm steps with the previous-step guess, then one
    evaluation step solving the same system from the three guesses."""
    rows = []
    for ci in range(args.eval_cases):
        p = case_params(rng)
        g = open_case(dg, args.grid, p, args.tol, 1)
        ok = True
        for t in range(args.eval_warm):
            g.step(t)
            if not (g.state()["mg_relres"] < args.tol):
                ok = False
                break
        if ok:
            gs.eval_mode(True)
            try:
                g.step(args.eval_warm)
            finally:
                gs.eval_mode(False)
            it = gs.eval_iters()
            ok = g.state()["mg_relres"] < args.tol
            rows.append(dict(case=ci, **p, ok=bool(ok), **it))
            print(f"eval {ci}: r={p['radius']*1e3:.2f} mm y0={p['y0']*1e3:.2f} mm v={p['yvel1']:.2f} m/s "
                  f"iterations zero={it['zero']} previous={it['previous']} learned={it['learned']}", flush=True)
        g.close()
    good = [r for r in rows if r["ok"]]
    summ = {}
    if good:
        for k in ("zero", "previous", "learned"):
            summ[f"mean_{k}"] = float(np.mean([r[k] for r in good]))
        summ["saved_vs_previous"] = summ["mean_previous"] - summ["mean_learned"]
        summ["saved_vs_zero"] = summ["mean_zero"] - summ["mean_learned"]
        # against the better plain guess of each system
        summ["cases_better"] = int(sum(r["learned"] < min(r["zero"], r["previous"]) for r in good))
        summ["cases_worse"] = int(sum(r["learned"] > min(r["zero"], r["previous"]) for r in good))
        summ["cases"] = len(good)
    return rows, summ


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=258)
    ap.add_argument("--cases", type=int, default=40)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--lr-final", type=float, default=3e-5)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--init", default="", help="start from these weights instead of a random init")
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "guess_unet.npz"))
    ap.add_argument("--eval-cases", type=int, default=20)
    ap.add_argument("--eval-warm", type=int, default=3)
    ap.add_argument("--eval-only", action="store_true", help="evaluate the weights in --out, no training")
    ap.add_argument("--omega", type=float, default=1.15, help="SOR weight of the V-cycle smoother")
    ap.add_argument("--nu-coarse", type=int, default=4, help="smoothing sweeps below the finest level")
    args = ap.parse_args()
    dg = importlib.import_module("3d_droplet_simulation_2007_amd")
    gs = importlib.import_module("3d_droplet_simulation_2007_amd.guess")
    dg.lib()
    dg.set_smoother(args.omega, args.nu_coarse)
    report = dict(args=vars(args))
    if args.eval_only:
        report["weights_meta"] = gs.load_weights(args.out)
    else:
        if args.init:
            gs.load_weights(args.init)
        else:
            gs.init_random(args.seed)
        rng = np.random.default_rng(args.seed)
        t0 = time.perf_counter()
        report["cases"] = generate(dg, gs, args, rng)
        report["samples"] = gs.store_count()
        report["generate_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        losses = []
        for e in range(args.epochs):
            # geometric decay from lr to lr_final over the epochs
            f = e / max(args.epochs - 1, 1)
            lr = args.lr * (args.lr_final / args.lr) ** f
            losses.append(float(gs.train(1, lr, seed=args.seed * 1000 + e)[0]))
            print(f"epoch {e}: lr {lr:.2e} mean loss {losses[-1]:.6g} ({time.perf_counter() - t0:.1f} s)", flush=True)
        report["train_s"] = time.perf_counter() - t0
        report["losses"] = losses
        gs.capture(0)  # frees the sample store
        gs.save_weights(args.out, meta=dict(grid=args.grid, cases=args.cases, steps=args.steps, epochs=args.epochs,
                                            lr=args.lr, lr_final=args.lr_final, seed=args.seed,
                                            samples=report["samples"], losses=json.dumps(losses)))
        print("saved", args.out, flush=True)
    if args.eval_cases > 0:
        rng = np.random.default_rng(10_000 + args.seed)  # held-out stream
        rows, summ = evaluate(dg, gs, args, rng)
        report["eval"] = rows
        report["eval_summary"] = summ
        print("evaluation", json.dumps(summ), flush=True)
    rep_path = os.path.splitext(args.out)[0] + ("_eval.json" if args.eval_only else "_train.json")
    with open(rep_path, "w") as f:
        json.dump(report, f, indent=1)
    print("report", rep_path, flush=True)


if __name__ == "__main__":
    main()
