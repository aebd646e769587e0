"""Benchmark: time-steps/s of the channel solver at 512x96x96 on one MI355X.

Workload (the production configuration): 3-D channel 2.5 x 0.41 x 0.41 on a
512x96x96 MAC grid, full-height cylinder r = 0.05 at (0.5, 0.2), Um = 2.25,
nu = 1e-3 (Re ~ 100), dt = 2.5e-4, GMRES restart 30, rtol 1e-8 for the
momentum (Jacobi) and pressure (PFMG-style MG) solves, initial field = inlet
profile + seeded noise of amplitude 1e-3 Um.  FP64 throughout ("value"); the
FP32-multigrid form (FP32 V-cycle inside FP64 flexible GMRES) is timed on the
same steps and reported beside it.

A step is the whole hot path of main.cu:249-376: BC cleanup, convection,
grad p, momentum rhs, three GMRES+Jacobi solves, scatter, divergence, the
GMRES+PFMG pressure solve, projection and the device-side diagnostics.  Inputs
are resident in HBM before the timed region starts.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "ibm-3d-fpc_amd"))

WORKLOAD = dict(nx=512, ny=96, nz=96, Lx=2.5, Ly=0.41, Lz=0.41, xcirc=0.5, ycirc=0.2, Rcirc=0.05, Um=2.25,
                nu=1e-3, rho=1.0, dt=2.5e-4, vel_kdim=30, p_kdim=30, vel_tol=1e-8, p_tol=1e-8, vel_maxit=250,
                p_maxit=250, init_profile=1, noise_amp=1e-3 * 2.25, seed=0)


def run_sim(F, torch, kw, warmup, steps):
    sim = F.Sim(F.default_params(**kw))
    sim.step(warmup)
    torch.cuda.synchronize()
    its = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step(1)
        its.append([int(sim.scalar(k)) for k in ("its_u", "its_v", "its_w", "its_p")])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return sim, dt, its


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import fpc_amd as F
    F.lib()
    torch.cuda.set_device(0)
    sim64, t64, its64 = run_sim(F, torch, WORKLOAD, args.warmup, args.steps)
    res = {
        "metric": "time-steps/s @512x96x96 FP64 (GMRES+Jacobi u/v/w, GMRES+PFMG p, rtol 1e-8)",
        "value": args.steps / t64,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t64 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (inlet profile + seeded noise), geometry built on the host",
        "config": {"workload": "channel-cylinder 512x96x96 FP64, Re~100, dt 2.5e-4, kdim 30, rtol 1e-8",
                   "grid": [512, 96, 96], "mg_precision": 64},
        "iterations_mean": [float(x) for x in np.mean(np.array(its64), axis=0)],
        "timers_s": {k: sim64.timer(k) for k in ("t_build", "t_mom", "t_p", "t_proj", "t_diag")},
        "cd": sim64.scalar("cd"), "cl": sim64.scalar("cl"),
        "roofline": None,
        "cpu_baseline": None,
    }
    if not args.no_fp32:
        sim32, t32, its32 = run_sim(F, torch, dict(WORKLOAD, mg_precision=32), args.warmup, args.steps)
        errs = {}
        for f in ("u", "v", "w", "p"):
            a, b = sim64.get(f), sim32.get(f)
            errs[f] = float(np.linalg.norm(a - b) / max(np.linalg.norm(a), 1e-300))
        res["fp32mg"] = {"value": args.steps / t32, "ms_per_step": 1e3 * t32 / args.steps,
                         "speedup_vs_fp64": t64 / t32, "field_rel_l2_vs_fp64": errs,
                         "cd_rel_err": abs(sim32.scalar("cd") - sim64.scalar("cd")) / abs(sim64.scalar("cd")),
                         "iterations_mean": [float(x) for x in np.mean(np.array(its32), axis=0)]}
        sim32.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
