"""Benchmark: time-steps/s of the channel solver at 512x96x96 on one MI355X.

Workload (the production configuration): 3-D channel 2.5 x 0.41 x 0.41 on a
512x96x96 MAC grid, full-height cylinder r = 0.05 at (0.5, 0.2), Um = 2.25,
nu = 1e-3 (Re ~ 100), dt = 2.5e-4, GMRES restart 30, rtol 1e-8 for the
momentum (Jacobi) and pressure (PFMG-style MG) solves, initial field = inlet
profile + seeded noise of amplitude 1e-3 Um.  FP64 throughout ("value"); the
FP32-multigrid form (FP32 V-cycle inside FP64 flexible GMRES) is timed on the
same steps and reported beside it.

roofline: the kernel class with the largest device time (found on the last
warmup step with every class timed) is timed over the whole timed region with
HIP events on the library's stream, on every 32nd launch (a step has ~37 launches
of the Gram-Schmidt classes, so the sample cycles through the basis lengths, at
a small fraction of the event cost); achieved = its algorithmic bytes per launch / its
mean launch duration, against the 8 TB/s HBM3E peak.  Launches
enqueued ahead of a converged GMRES cycle return at once and are not counted.
traffic = HBM bytes per launch read from profiles/hbm_traffic.txt (FETCH_SIZE +
WRITE_SIZE per dispatch, calibrated on a copy kernel) when that file names the
class, else null.  copy_gbs = a 1 GiB device-to-device copy timed the same way
(read + written bytes / s): the streaming rate reached in practice, and
frac_of_copy = achieved / copy_gbs.

cpu_baseline: the oracle (C++ restatement of the same path, OpenMP over the
OMP_NUM_THREADS host threads) on the same 512x96x96 workload: after the setup
and the warmup steps, three steps timed, with their iteration counts beside the
GPU run's on the same steps.

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

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "ibm-3d-fpc_amd"))

WORKLOAD = dict(nx=512, ny=96, nz=96, Lx=2.5, Ly=0.41, Lz=0.41, xcirc=0.5, ycirc=0.2, Rcirc=0.05, Um=2.25,
                nu=1e-3, rho=1.0, dt=2.5e-4, vel_kdim=30, p_kdim=30, vel_tol=1e-8, p_tol=1e-8, vel_maxit=250,
                p_maxit=250, init_profile=1, noise_amp=1e-3 * 2.25, seed=0)


HBM_PEAK_GBS = 8000.0
ROOFLINE_SAMPLE = 32  # the roofline class's launches timed: every 32nd
# kernel classes whose every launch covers one fixed box (a roofline figure per
# launch is meaningful); the other MG classes mix the coarse levels
ROOFLINE_CLASSES = ("spmv", "spmv_jacobi", "mg_down0", "mg_up0", "multidot", "orth_update", "explicit",
                    "divergence", "projection")


def _classes(rep):
    return {k: {"launches": n, "ms": ms, "bytes": b, "gbs": (b / (ms * 1e-3) / 1e9) if ms > 0 else None}
            for k, (n, ms, b) in rep.items()}


def step_bytes(classes):
    """Algorithmic bytes of one step summed over the timed kernel classes (the
    event-timed warm step; Arnoldi tails, true residuals and reductions are not
    timed classes), their kernel time, and the rate."""
    if not classes:
        return None
    b = sum(v["bytes"] for v in classes.values())
    ms = sum(v["ms"] for v in classes.values())
    return {"bytes": b, "kernel_ms": ms, "achieved_gbs": b / (ms * 1e-3) / 1e9,
            "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def vcycle_summary(classes):
    """The pressure V-cycle from the event-timed warm step: kernel time and algorithmic
    bytes (every level's down, restriction and up launches, the coarsest GEMV
    excluded) per V-cycle, and the HBM rate they imply."""
    n = classes.get("mg_down0", {}).get("launches")
    if not n:
        return None
    mg = [v for k, v in classes.items() if k.startswith("mg")]
    us = 1e3 * sum(v["ms"] for v in mg) / n
    b = sum(v["bytes"] for v in mg) / n
    return {"vcycles_per_step": n, "us_per_vcycle": us, "bytes_per_vcycle": b, "achieved_gbs": b / us / 1e3,
            "frac": b / us / 1e3 / HBM_PEAK_GBS,
            "level0_us": 1e3 * (classes["mg_down0"]["ms"] + classes["mg_up0"]["ms"]) / n}


def _its_totals(sim):
    return [sim.scalar("its_total_" + c) for c in "uvwp"]


def _its_per_step(sim, tot0, steps):
    """Mean iterations per step over the last `steps` steps (one row: the K steps run
    as one library call, so no host work of this script sits between them)."""
    tot = _its_totals(sim)
    return [[(b - a) / steps for a, b in zip(tot0, tot)]]


def run_sim(F, torch, kw, warmup, steps, roofline_timing=True, sample=ROOFLINE_SAMPLE):
    """W untimed steps (the last one with every kernel class timed), then K timed steps
    with the dominant class timed by HIP events."""
    sim = F.Sim(F.default_params(**kw))
    if not roofline_timing:
        sim.step(warmup)
        torch.cuda.synchronize()
        tot0 = _its_totals(sim)
        t0 = time.perf_counter()
        sim.step(steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        return sim, dt, _its_per_step(sim, tot0, steps), None, None, {}
    if warmup > 1:
        sim.step(warmup - 1)
    F.profile_enable("*")
    sim.step(1 if warmup >= 1 else 0)
    torch.cuda.synchronize()
    classes = _classes(F.profile_report()) if warmup >= 1 else {}
    elig = [k for k in classes if k in ROOFLINE_CLASSES]
    dom = max(elig, key=lambda k: classes[k]["ms"]) if elig else ",".join(ROOFLINE_CLASSES)
    # the dominant class timed on every ROOFLINE_SAMPLE-th launch across the timed region:
    # an unbiased mean launch duration at a fraction of the event cost (profile_enable
    # resets the timing, which frees the warmup step's events: left alive, they slow
    # every later step)
    F.profile_enable(dom)
    F.profile_sample(sample)
    torch.cuda.synchronize()
    tot0 = _its_totals(sim)
    t0 = time.perf_counter()
    sim.step(steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    its = _its_per_step(sim, tot0, steps)
    rep = F.profile_report()
    F.profile_sample(1)
    F.profile_enable("")
    if dom not in rep:
        classes = _classes(rep)
        dom = max(classes, key=lambda k: classes[k]["ms"])
    return sim, dt, its, dom, rep[dom], classes


def read_traffic(cls):
    """(fetch_bytes, write_bytes) per launch of a class from profiles/hbm_traffic.txt, or None."""
    path = os.path.join(REPO, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 3 and not line.startswith("#") and f[0] == cls:
            return float(f[1]), float(f[2])
    return None


def roofline(dom, rec):
    n, ms, b = rec
    per_launch = b / n
    mean_s = ms * 1e-3 / n
    achieved = per_launch / mean_s / 1e9
    tr = read_traffic(dom)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": (tr[0] + tr[1]) if tr else None,
            "kernel": dom, "launches": n, "mean_us": mean_s * 1e6, "algorithmic_bytes_per_launch": per_launch}


def copy_gbs(torch, nbytes=1 << 30, reps=10):
    """Device-to-device copy bandwidth (bytes read + written per second, GB/s): the
    streaming rate this GPU reaches in practice, beside the 8 TB/s nominal peak."""
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    return 2.0 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9


def gpu_step_iterations(F, warmup, nsteps):
    """Iteration counts (u, v, w, p) of steps warmup+1 .. warmup+nsteps of the
    workload on the GPU, one step per call (outside any timed region)."""
    sim = F.Sim(F.default_params(**WORKLOAD))
    sim.step(warmup)
    out = []
    for _ in range(nsteps):
        sim.step(1)
        out.append([int(sim.scalar("its_" + c)) for c in "uvwp"])
    sim.close()
    return out


def cpu_baseline(warmup=2, nsteps=3):
    """The oracle (the C++ restatement of the reference path, OpenMP over the host
    threads OMP_NUM_THREADS allows) on the benchmark workload itself: setup,
    `warmup` untimed steps, then nsteps timed steps (the GPU run's first timed
    steps, whose iteration counts are reported beside these)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    O.set_threads(threads)
    t0 = time.perf_counter()
    sim = O.Sim(O.default_params(**WORKLOAD))
    t_setup = time.perf_counter() - t0
    print(f"cpu_baseline: oracle setup {t_setup:.1f} s", file=sys.stderr, flush=True)
    sim.step(warmup)
    its = []
    t0 = time.perf_counter()
    for _ in range(nsteps):
        sim.step(1)
        sc = sim.scalars()
        its.append([int(sc["its_" + q]) for q in "uvwp"])
        print(f"cpu_baseline: timed step {len(its)} its {its[-1]}", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    return {"value": nsteps / dt, "unit": "time-steps/s", "cores": threads, "kind": "port",
            "iterations": its, "setup_s": t_setup,
            "sample": f"oracle (C++ restatement of the reference path, {threads} OpenMP threads) on the same "
                      f"512x96x96 workload: setup {t_setup:.1f} s, steps 1-{warmup} untimed, steps "
                      f"{warmup + 1}-{warmup + nsteps} timed ({dt:.1f} s)"}


def run_ensemble(F, torch, nthreads, n=64, steps=50, modes=("looped", "concurrent", "batched")):
    """Aggregate cell-updates/s of n seeded uneven channels over `steps` steps: the
    instances one after another (the plain path looped), with `nthreads` host
    workers and one HIP stream per instance, and batched (one stream, one launch
    per stage over all instances).  One untimed step first (solver setup)."""
    pars = F.ensemble_workload(n=n, seed=0)
    cells = sum(p.nx * p.ny * p.nz for p in pars)
    out = {"metric": "cell-updates/s, %d-channel ensemble, %d steps each" % (n, steps), "unit": "cell-updates/s",
           "instances": n, "cells_total": cells, "steps": steps, "dtype": "f64",
           "data": "synthetic (seeded geometry, Re 10-100, inlet profile + noise)"}
    for tag, nth in (("looped", 1), ("concurrent", nthreads), ("batched", 0)):
        if tag not in modes:
            continue
        ens = F.Ensemble(pars, nthreads=nth)
        ens.step(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ens.step(steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[tag] = {"value": cells * steps / dt, "seconds": dt, "threads": nth}
        if nth == 0:
            kz, tmin = ens.tiles()
            out[tag]["autotuned_tiles"] = {"kz": kz, "tile_min_points": tmin}
        ens.close()
    best = max((k for k in ("concurrent", "batched", "looped") if k in out), key=lambda k: out[k]["value"])
    out["value"] = out[best]["value"]
    out["mode"] = best
    if "looped" in out:
        out["speedup_vs_looped"] = out[best]["value"] / out["looped"]["value"]
    return out


def run_vcycle_rate(F, nit=12):
    """Contraction of the pressure V-cycle on the benchmark's own pressure operator
    (512x96x96, embedded with identity rows on solid cells): the stationary iteration
    x += M^-1 (b - A x) from x = 0 on a seeded b (zero on solid cells), residual
    norms per sweep, and the rate over the last sweeps; beside it, the mean residual
    reduction per iteration that GMRES(30) + one V-cycle reaches on the same system."""
    nx, ny, nz = WORKLOAD["nx"], WORKLOAD["ny"], WORKLOAD["nz"]
    S = F.Setup(F.default_params(**WORKLOAD))
    vals = S.get("Ap_vals")
    fl = S.get("fluid_p").astype(bool)
    box = F.StructBox(nx, ny, nz, vals)
    b = np.random.default_rng(0).standard_normal(nx * ny * nz)
    b[~fl] = 0.0
    nb = float(np.linalg.norm(b))
    x = np.zeros_like(b)
    res = [1.0]
    for _ in range(nit):
        r = b - box.matvec(x)
        x = x + F.pfmg_apply(box, r)
        res.append(float(np.linalg.norm(b - box.matvec(x))) / nb)
    tail = 6
    rate = (res[-1] / res[-1 - tail]) ** (1.0 / tail)
    sol = F.Solver(box, tol=1e-8, maxit=250, kdim=30, precond="pfmg")
    _, its, rel, rc = sol.solve(b)
    return {"metric": "pressure V-cycle contraction, 512x96x96 channel operator", "residuals": res,
            "stationary_rate": rate, "gmres_its": its, "gmres_relres": rel,
            "gmres_rate_per_iteration": (rel ** (1.0 / its)) if its else None}


MIXED = dict(WORKLOAD, nx=1024, ny=160, nz=160)


def run_mixed(F, torch, steps=100):
    """The 1024x160x160 channel (26M cells, Re ~100), `steps` steps in FP64 and
    with the FP32 V-cycle inside FP64 flexible GMRES at the same 1e-8 residual:
    time-steps/s of both, per-field relative L2 error and drag/lift error of the
    FP32-MG run against the FP64 run."""
    out = {"metric": "time-steps/s @1024x160x160, %d steps: FP64 vs FP32-MG at 1e-8 resid" % steps,
           "unit": "time-steps/s", "steps": steps, "dtype": "f64 (FP32 V-cycle in the mixed run)",
           "data": "synthetic (inlet profile + seeded noise), geometry built on the host"}
    fields = {}
    for bits in (64, 32):
        t0 = time.perf_counter()
        sim = F.Sim(F.default_params(**dict(MIXED, mg_precision=bits)))
        print(f"[mixed] FP{bits}: setup {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        its, rel = [], []
        t0 = time.perf_counter()
        for n in range(steps):
            sim.step(1)
            its.append([sim.scalar(k) for k in ("its_u", "its_v", "its_w", "its_p")])
            rel.append(sim.scalar("relres_p"))
            if (n + 1) % 25 == 0:
                print(f"[mixed] FP{bits}: step {n + 1}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        fields[bits] = {f: sim.get(f) for f in ("u", "v", "w", "p")}
        out["fp%d" % bits] = {"value": steps / dt, "ms_per_step": 1e3 * dt / steps,
                              "iterations_mean": np.mean(its, axis=0).tolist(), "relres_p_max": max(rel),
                              "cd": sim.scalar("cd"), "cl": sim.scalar("cl")}
        sim.close()
    a, b = fields[64], fields[32]
    out["field_rel_l2_vs_fp64"] = {f: float(np.linalg.norm(b[f] - a[f]) / np.linalg.norm(a[f])) for f in a}
    out["cd_rel_err"] = abs(out["fp32"]["cd"] - out["fp64"]["cd"]) / abs(out["fp64"]["cd"])
    out["cl_rel_err"] = abs(out["fp32"]["cl"] - out["fp64"]["cl"]) / abs(out["fp64"]["cd"])
    out["value"] = out["fp32"]["value"]
    out["speedup_vs_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    return out


ADJOINT = dict(nx=256, ny=64, nz=64, Lx=2.5, Ly=0.41, Lz=0.41, xcirc=0.5, ycirc=0.2, Rcirc=0.05, Um=0.45, nu=1e-3,
               rho=1.0, dt=1e-3, vel_kdim=30, p_kdim=30, vel_tol=1e-12, p_tol=1e-12, vel_maxit=500, p_maxit=500,
               init_profile=1, noise_amp=1e-3 * 0.45, seed=0)


def run_adjoint(F, torch, nsteps=100, avg_from=50, every=10, ndir=16, h=1e-6):
    """Gradient of the mean drag coefficient over steps avg_from..nsteps with
    respect to a 64x64 inlet amplitude field (seeded, 1 +- 5%) and nu, by the
    checkpointed adjoint; checked against central differences on ndir seeded
    directions.  Solves to rtol 1e-12 so the differences resolve 1e-7."""
    rng = np.random.default_rng(0)
    nin = ADJOINT["ny"] * ADJOINT["nz"]
    theta = 1.0 + 0.05 * (2.0 * rng.random(nin) - 1.0)
    nu0 = ADJOINT["nu"]

    def J_of(th, nu):
        s = F.Sim(F.default_params(**dict(ADJOINT, nu=nu)))
        s.set("inlet_amplitude", th)
        cds = []
        for n in range(1, nsteps + 1):
            s.step(1)
            if n >= avg_from:
                cds.append(s.scalar("cd"))
        s.close()
        return float(np.mean(cds))

    sim = F.Sim(F.default_params(**ADJOINT))
    sim.set("inlet_amplitude", theta)
    adj = F.Adjoint(sim)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J, g_th, g_nu = adj.drag_gradient(nsteps, avg_from, every)
    t_grad = time.perf_counter() - t0
    adj.close()
    sim.close()
    t0 = time.perf_counter()
    J_plain = J_of(theta, nu0)
    t_fwd = time.perf_counter() - t0
    errs = []
    for _ in range(ndir):
        dth, dnu = rng.standard_normal(nin), float(rng.standard_normal())
        ad = float(g_th @ dth + g_nu * nu0 * dnu)
        fd = (J_of(theta + h * dth, nu0 * (1 + h * dnu)) - J_of(theta - h * dth, nu0 * (1 - h * dnu))) / (2 * h)
        errs.append(abs(fd - ad) / abs(ad))
    return {"metric": "adjoint gradient of mean c_D, 256x64x64, steps %d-%d, checkpoint every %d" % (avg_from, nsteps, every),
            "J": J, "J_forward_run": J_plain, "dJ_dnu": g_nu, "dJ_dtheta_norm": float(np.linalg.norm(g_th)),
            "directions": ndir, "h": h, "fd_rel_err_max": max(errs), "fd_rel_err_median": float(np.median(errs)),
            "gradient_seconds": t_grad, "forward_seconds": t_fwd, "gradient_over_forward": t_grad / t_fwd,
            "dtype": "f64", "data": "synthetic (seeded inlet amplitude 1 +- 5%, inlet profile + noise)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-roofline", action="store_true", help="no kernel timing at all (A/B timing runs)")
    ap.add_argument("--roofline-sample", type=int, default=ROOFLINE_SAMPLE,
                    help="time every n-th launch of the roofline class in the timed region")
    ap.add_argument("--ensemble", action="store_true",
                    help="64 uneven channels, 50 steps: cell-updates/s looped vs concurrent (no headline line)")
    ap.add_argument("--ensemble-threads", type=int, default=8)
    ap.add_argument("--ensemble-steps", type=int, default=50)
    ap.add_argument("--ensemble-modes", default="looped,concurrent,batched",
                    help="comma list of looped, concurrent, batched")
    ap.add_argument("--adjoint", action="store_true",
                    help="256x64x64 drag-gradient run + 16 finite-difference directions (no headline line)")
    ap.add_argument("--mixed", action="store_true",
                    help="1024x160x160, 100 steps: FP64 against FP32-MG, per-field errors (no headline line)")
    ap.add_argument("--vcycle-rate", action="store_true",
                    help="contraction of the pressure V-cycle on the 512x96x96 operator (no headline line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 2 GiB device-to-device copy 5 times (scale for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch

    torch.cuda.set_device(0)
    if args.copy_calibration:
        a = torch.ones(1 << 28, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": a.numel() * 8, "copies": 5}))
        return
    import fpc_amd as F
    F.lib()
    if args.ensemble:
        print(json.dumps(run_ensemble(F, torch, args.ensemble_threads, steps=args.ensemble_steps,
                                      modes=tuple(args.ensemble_modes.split(",")))))
        return
    if args.adjoint:
        print(json.dumps(run_adjoint(F, torch)))
        return
    if args.mixed:
        print(json.dumps(run_mixed(F, torch)))
        return
    if args.vcycle_rate:
        print(json.dumps(run_vcycle_rate(F)))
        return
    sim64, t64, its64, dom, rec, classes = run_sim(F, torch, WORKLOAD, args.warmup, args.steps,
                                                   not args.no_roofline, args.roofline_sample)
    res = {
        "metric": "time-steps/s @512x96x96: FP64 vs FP32-MG at 1e-8 resid; stencil HBM GB/s",
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
        "roofline": roofline(dom, rec) if dom else None,
        "vcycle": vcycle_summary(classes),
        "timed_classes_step": step_bytes(classes),
        "cpu_baseline": None,
        "kernel_classes": classes,
    }
    if not args.no_fp32:
        sim32, t32, its32, dom32, rec32, cls32 = run_sim(F, torch, dict(WORKLOAD, mg_precision=32), args.warmup,
                                                         args.steps, not args.no_roofline, args.roofline_sample)
        errs = {}
        for f in ("u", "v", "w", "p"):
            a, b = sim64.get(f), sim32.get(f)
            errs[f] = float(np.linalg.norm(a - b) / max(np.linalg.norm(a), 1e-300))
        res["fp32mg"] = {"value": args.steps / t32, "ms_per_step": 1e3 * t32 / args.steps,
                         "speedup_vs_fp64": t64 / t32, "field_rel_l2_vs_fp64": errs,
                         "cd_rel_err": abs(sim32.scalar("cd") - sim64.scalar("cd")) / abs(sim64.scalar("cd")),
                         "iterations_mean": [float(x) for x in np.mean(np.array(its32), axis=0)],
                         "roofline": roofline(dom32, rec32) if dom32 else None, "kernel_classes": cls32}
        sim32.close()
    if res["roofline"]:
        cg = copy_gbs(torch)
        res["roofline"]["copy_gbs"] = cg
        res["roofline"]["frac_of_copy"] = res["roofline"]["achieved"] / cg
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(args.warmup)
        res["cpu_baseline"]["gpu_iterations"] = gpu_step_iterations(F, args.warmup, 3)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
