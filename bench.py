#!/usr/bin/env python3
"""Benchmark of the biogeochemistry step on one MI355X.

Workload: LOBSTER (NO3, NH4, P, Z, sPOM, bPOM, DOM) + DIC/Alk on a
512×512×64 periodic box (1 km × 1 km × 200 m), lognormal tracers, TwoBand
PAR with PAR0 = 100 W/m², sinking sPOM/bPOM, CO2 gas exchange at the
surface, Δt = 60 s.  One step = one RK3 step (three stages of update-state,
tendencies and substep) over every interior cell, inputs resident in HBM.

Metric: tracer-cell updates/s (9 written tracers × cells × steps / s) of the
FP32-mixed path; the plain FP64 path is timed beside it, and both are
compared against the FP64 oracle on a sample of columns after the full
200-step run (columns are independent, so a column sample is exact).

Prints ONE JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "oceanbiome.jl_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--nx", type=int, default=512)
    p.add_argument("--ny", type=int, default=512)
    p.add_argument("--nz", type=int, default=64)
    p.add_argument("--check-steps", type=int, default=200, help="steps of the error check run")
    p.add_argument("--sample-columns", type=int, default=256)
    p.add_argument("--cpu-columns", type=int, default=1024, help="columns of the CPU baseline sample")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--no-check", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--precisions", default="fp32,fp64")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(dev, tr, steps, dt, warmup):
    import torch
    if warmup:
        dev.step(tr, warmup, dt)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    dev.step(tr, steps, dt, t0=warmup * dt)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return start.elapsed_time(stop) / 1e3, wall


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the biogeochemistry step "
              f"shards by column tiles with no exchange, and this build measures one GPU's share.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from oceanbiome import abi, workloads
    from oceanbiome.runtime import HIPBiogeochemistry

    log("building workload", args.nx, args.ny, args.nz)
    wl = workloads.lobster_3d(args.nx, args.ny, args.nz, nsteps=args.check_steps)
    g = wl.grid
    ncell = g.n_interior
    nw = wl.n_written
    results = {}
    outputs = {}
    for prec in args.precisions.split(","):
        dev = HIPBiogeochemistry(wl.bgc, prec)
        tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
        secs, wall = time_steps(dev, tr, args.steps, wl.dt, args.warmup)
        sbytes = 4 if prec == "fp32" else 8
        bytes_step = wl.algorithmic_bytes_per_step(sbytes)
        ms = secs / args.steps * 1e3
        upd = nw * ncell * args.steps / secs
        gbs = bytes_step / (secs / args.steps) / 1e9
        results[prec] = {"ms_per_step": ms, "updates_per_s": upd, "achieved_GBs": gbs,
                         "bytes_per_step": bytes_step, "wall_s": wall}
        log(f"{prec}: {ms:.4f} ms/step  {upd:.3e} upd/s  {gbs:.1f} GB/s  ({gbs / HBM_PEAK_GBS * 100:.1f}% HBM)")
        if not args.no_check:
            tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
            dev.step(tr, args.check_steps, wl.dt)
            torch.cuda.synchronize()
            outputs[prec] = [t.double().cpu().numpy() for t in tr]
        del tr, dev
        torch.cuda.empty_cache()

    errors = {}
    cpu = None
    if not args.no_check or not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        cfg = wl.bgc.config(abi.OBME_FP64)
        nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
        if not args.no_check:
            rng = np.random.default_rng(1)
            ns = min(args.sample_columns, g.Nx * g.Ny)
            flat = rng.choice(g.Nx * g.Ny, ns, replace=False)
            ci, cj = (flat % g.Nx).astype(np.int32), (flat // g.Nx).astype(np.int32)
            ref = [a.copy() for a in wl.tracers]
            O.grid_run_columns(cfg, ref, ci, cj, args.check_steps, wl.dt, nthreads=nthreads)
            names = wl.bgc.required_biogeochemical_tracers()
            ks = np.arange(g.Nz) + g.Hz
            for prec, out in outputs.items():
                e = {}
                for n, name in enumerate(names):
                    a = out[n][ks[:, None], (cj + g.Hy)[None, :], (ci + g.Hx)[None, :]]
                    b = ref[n][ks[:, None], (cj + g.Hy)[None, :], (ci + g.Hx)[None, :]]
                    e[name] = {"rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                               "max_rel": float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))}
                errors[prec] = e
                worst = max(v["rel_l2"] for v in e.values())
                log(f"{prec} vs oracle after {args.check_steps} steps ({ns} columns): worst rel L2 {worst:.3e}")
        if not args.no_cpu:
            nc = min(args.cpu_columns, g.Nx * g.Ny)
            ci = (np.arange(nc) % g.Nx).astype(np.int32)
            cj = (np.arange(nc) // g.Nx).astype(np.int32)
            ref = [a.copy() for a in wl.tracers]
            steps = 20
            t0 = time.perf_counter()
            O.grid_run_columns(cfg, ref, ci, cj, steps, wl.dt, nthreads=nthreads)
            secs = time.perf_counter() - t0
            cpu = {"value": nw * g.Nz * nc * steps / secs, "unit": "tracer-cell updates/s", "cores": nthreads,
                   "kind": "port",
                   "sample": f"{nc} columns x {g.Nz} levels x {steps} steps of the same workload, FP64 oracle, "
                             f"OpenMP over columns ({secs:.1f} s)"}
            log(f"cpu baseline: {cpu['value']:.3e} upd/s on {nthreads} threads")

    head = results.get("fp32") or results["fp64"]
    dom = {"bound": "hbm", "achieved": head["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": head["achieved_GBs"] / HBM_PEAK_GBS, "traffic": None}
    line = {
        "metric": "tracer-cell updates/s (FP32-mixed LOBSTER+DIC/Alk step) + % HBM roofline",
        "value": head["updates_per_s"],
        "unit": "tracer-cell updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if "fp32" in results else "f64",
        "data": "synthetic (seeded lognormal/uniform tracers)",
        "config": {"workload": wl.name, "grid": [g.Nx, g.Ny, g.Nz], "tracers_written": nw, "dt": wl.dt,
                   "precision": "fp32-mixed (fp64 carbonate chemistry)" if "fp32" in results else "fp64",
                   "parallelism": "single GPU"},
        "roofline": dom,
        "cpu_baseline": cpu,
        "fp64": results.get("fp64"),
        "fp32_mixed": results.get("fp32"),
        "speedup_fp32_over_fp64": (results["fp32"]["updates_per_s"] / results["fp64"]["updates_per_s"]
                                   if "fp32" in results and "fp64" in results else None),
        "errors_vs_oracle": errors,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
