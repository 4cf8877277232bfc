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

Prints ONE JSON line.  --gpus N with N > 1 exits: the benchmark runs on one GPU.
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
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work of the baseline sample")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--calibrate-copy-mb", type=int, default=0,
                   help="after the timed steps, copy a buffer of this many MiB once (a known-size kernel "
                        "for scaling HBM counters)")
    p.add_argument("--no-check", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--precisions", default="fp32,fp64")
    p.add_argument("--workload", default="lobster", choices=["lobster", "columns", "carbonate", "pisces", "hooks"],
                   help="lobster: the headline LOBSTER+DIC/Alk step; columns: the ragged NPZD column "
                        "ensemble; carbonate: pH and CO2 flux on point sets; hooks: the Oceananigans "
                        "hook pair (update-state, then tendencies added into G) on the headline grid")
    p.add_argument("--oxygen", action="store_true",
                   help="lobster workload with the O2 tracer and its gas exchange (10 written tracers)")
    p.add_argument("--sinking", default="upwind", choices=["upwind", "centered", "weno5"],
                   help="lobster workload: the sinking face value (centered and weno5 step one stage per launch)")
    p.add_argument("--columns", type=int, default=8192)
    p.add_argument("--sediment", action="store_true",
                   help="columns workload: an InstantRemineralisation sediment under every column (P, D into N)")
    p.add_argument("--points", type=int, default=1 << 24)
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(dev, tr, steps, dt, warmup):
    """Time `steps` steps after `warmup` untimed ones, twice: first plain
    (the step time), then with kernel timing on, where every sweep launch is
    bracketed by HIP events on the stream it runs on (the sweep's launch
    time; the events themselves stretch that run's step time, so it is not
    the one reported).  Returns (seconds of the plain region, its wall
    seconds, mean sweep launch ms)."""
    import torch
    if warmup:
        dev.step(tr, warmup, dt)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    dev.step(tr, steps, dt, t0=warmup * dt)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    secs = start.elapsed_time(stop) / 1e3
    dev.set_kernel_timing(True)
    dev.kernel_timing()          # drop anything recorded before
    dev.step(tr, steps, dt, t0=(warmup + steps) * dt)
    torch.cuda.synchronize()
    sweep_ms, n = dev.kernel_timing()
    dev.set_kernel_timing(False)
    return secs, wall, sweep_ms / max(n, 1)


def surface_errors(wl, cfg, out, ref, ci, cj, O):
    """pH (free scale) and the CO2 flux of the sampled columns' surface cells
    after the check run: the GPU's point kernel on the GPU state against the
    FP64 oracle on the oracle's state (carbon_chemistry.jl:168,
    gas_exchange.jl:26-38)."""
    import torch
    from oceanbiome.runtime import HIPBiogeochemistry
    g = wl.grid
    k = g.Nz - 1 + g.Hz
    top = lambda arr: np.ascontiguousarray(arr[k, cj + g.Hy, ci + g.Hx], dtype=np.float64)
    T = np.full(len(ci), float(wl.extra["T_profile"][-1]))
    S = np.full(len(ci), float(wl.extra["S"]))
    u10 = np.full(len(ci), cfg.co2.wind_speed)
    pH_ref, F_ref = O.co2_flux_batch(cfg.co2, cfg.solver, top(ref[7]), top(ref[8]), T, S, u10,
                                     cfg.co2.air_concentration)
    dev = HIPBiogeochemistry(wl.bgc, "fp64")
    cuda = lambda x: torch.from_numpy(x).cuda()
    pH, F = dev.co2_flux(cuda(top(out[7])), cuda(top(out[8])), cuda(T), cuda(S), cuda(u10), mode=0)
    pH, F = pH.cpu().numpy(), F.cpu().numpy()
    # the floor FP32 storage sets: the oracle's own state rounded to FP32
    r32 = lambda x: x.astype(np.float32).astype(np.float64)
    pH_fl, _ = O.co2_flux_batch(cfg.co2, cfg.solver, r32(top(ref[7])), r32(top(ref[8])), T, S, u10,
                                cfg.co2.air_concentration)
    return {"pH (surface)": {"max_abs": float(np.max(np.abs(pH - pH_ref))),
                             "rms": float(np.sqrt(np.mean((pH - pH_ref) ** 2))),
                             "fp32_storage_floor_max_abs": float(np.max(np.abs(pH_fl - pH_ref)))},
            "CO2 flux (surface)": {"rel_l2": float(np.linalg.norm(F - F_ref) / np.linalg.norm(F_ref)),
                                   "max_rel": float(np.max(np.abs(F - F_ref) / np.abs(F_ref)))}}


# CPU threads one GPU's job may use on the GPU hosts (each GPU's share of
# the host's cores); the baseline reports the CPUs it could see beside it
HOST_SHARE = 16


def host_cpus():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_threads(req):
    """--cpu-threads, else every visible CPU up to the host's share per GPU."""
    if req:
        return req
    return max(1, min(HOST_SHARE, host_cpus()))


def read_traffic(profile_dir, workload, precision):
    """HBM bytes per sweep launch from a committed counter summary: a line
    'traffic_bytes_per_launch <workload> <precision> <bytes>'."""
    if not os.path.isdir(profile_dir):
        return None
    for fn in sorted(os.listdir(profile_dir)):
        if not fn.endswith((".md", ".txt")):
            continue
        for line in open(os.path.join(profile_dir, fn)):
            parts = line.strip().strip("`").split()
            if len(parts) == 4 and parts[0] == "traffic_bytes_per_launch" and parts[1] == workload \
                    and parts[2] == precision:
                return float(parts[3])
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the biogeochemistry step "
              f"shards by column tiles with no exchange, and the benchmark measures one GPU's share.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "columns":
        return bench_columns(args)
    if args.workload == "carbonate":
        return bench_carbonate(args)
    if args.workload == "pisces":
        return bench_pisces(args)
    if args.workload == "hooks":
        return bench_hooks(args)
    import torch
    from oceanbiome import abi, workloads
    from oceanbiome.runtime import HIPBiogeochemistry

    log("building workload", args.nx, args.ny, args.nz)
    wl = workloads.lobster_3d(args.nx, args.ny, args.nz, nsteps=args.check_steps, oxygen=args.oxygen,
                              sinking_scheme=args.sinking)
    g = wl.grid
    ncell = g.n_interior
    nw = wl.n_written
    results = {}
    outputs = {}
    for prec in args.precisions.split(","):
        dev = HIPBiogeochemistry(wl.bgc, prec)
        tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
        secs, wall, sweep_ms = time_steps(dev, tr, args.steps, wl.dt, args.warmup)
        sbytes = 4 if prec == "fp32" else 8
        bytes_step = wl.algorithmic_bytes_per_step(sbytes)
        ms = secs / args.steps * 1e3
        upd = nw * ncell * args.steps / secs
        gbs = bytes_step / (secs / args.steps) / 1e9
        sweep_gbs = bytes_step / (sweep_ms * 1e-3) / 1e9
        results[prec] = {"ms_per_step": ms, "updates_per_s": upd, "achieved_GBs": gbs,
                         "sweep_ms_per_launch": sweep_ms, "sweep_achieved_GBs": sweep_gbs,
                         "bytes_per_step": bytes_step, "wall_s": wall}
        log(f"{prec}: {ms:.4f} ms/step  {upd:.3e} upd/s  {gbs:.1f} GB/s  ({gbs / HBM_PEAK_GBS * 100:.1f}% HBM); "
            f"sweep {sweep_ms:.4f} ms/launch = {sweep_gbs:.1f} GB/s")
        if not args.no_check:
            tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
            dev.step(tr, args.check_steps, wl.dt)
            torch.cuda.synchronize()
            outputs[prec] = [t.double().cpu().numpy() for t in tr]
        del tr, dev
        torch.cuda.empty_cache()
    calibration_copy(args)

    errors = {}
    cpu = None
    if not args.no_check or not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        cfg = wl.bgc.config(abi.OBME_FP64)
        nthreads = cpu_threads(args.cpu_threads)
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
                e.update(surface_errors(wl, cfg, out, ref, ci, cj, O))
                errors[prec] = e
                worst = max(v["rel_l2"] for k, v in e.items() if "(surface" not in k)
                log(f"{prec} vs oracle after {args.check_steps} steps ({ns} columns): worst rel L2 {worst:.3e}")
        if not args.no_cpu:
            # FP64 oracle on every column of the same workload, OpenMP over
            # columns; the step count is sized from a short probe so that the
            # sample is about --cpu-seconds of CPU work.
            allc = np.arange(g.Nx * g.Ny)
            ci = (allc % g.Nx).astype(np.int32)
            cj = (allc // g.Nx).astype(np.int32)
            ref = [a.copy() for a in wl.tracers]
            t0 = time.perf_counter()
            O.grid_run_columns(cfg, ref, ci, cj, 1, wl.dt, nthreads=nthreads)
            per_step = time.perf_counter() - t0
            chunk = int(max(1, min(50, round(args.cpu_seconds / per_step / 4))))
            steps, secs = 0, 0.0
            while secs < args.cpu_seconds and steps < 400:
                t0 = time.perf_counter()
                O.grid_run_columns(cfg, ref, ci, cj, chunk, wl.dt, nthreads=nthreads)
                secs += time.perf_counter() - t0
                steps += chunk
            cpu = {"value": nw * g.Nz * len(ci) * steps / secs, "unit": "tracer-cell updates/s",
                   "cores": nthreads, "host_cpus": host_cpus(), "kind": "port",
                   "sample": f"all {len(ci)} columns x {g.Nz} levels x {steps} steps of the same workload, "
                             f"FP64 oracle (restated reference CPU path), OpenMP over columns, {secs:.1f} s"}
            log(f"cpu baseline: {cpu['value']:.3e} upd/s on {nthreads} threads ({secs:.1f} s)")

    head_prec = "fp32" if "fp32" in results else "fp64"
    head = results[head_prec]
    traffic = read_traffic(os.path.join(ROOT, "profiles"), wl.name, head_prec)
    dom = {"bound": "hbm", "achieved": head["sweep_achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": head["sweep_achieved_GBs"] / HBM_PEAK_GBS, "traffic": traffic,
           "kernel": "k_column_lane (column sweep)", "algorithmic_bytes_per_launch": head["bytes_per_step"],
           "launch_ms": head["sweep_ms_per_launch"]}
    line = {
        "metric": "tracer-cell updates/s + % HBM roofline; FP32-mixed vs FP64 at stated error",
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


def bench_hooks(args):
    """The hooks a coupled Oceananigans model calls once per RK3 stage:
    obme_update_state (the PAR field from P) and obme_add_tendencies (the
    transitions, the sinking divergence and the top CO2 flux added into the
    model's Gⁿ), on the headline workload.  One "stage" is the pair; the
    substep itself is the model's.  Each hook is timed on its own (HIP events
    around every sweep launch) and the pair together.  The tendency sweep's
    algorithmic bytes per cell: the 7 tracers the tendencies depend on read
    (DIC and Alk feed none: the surface kernel reads their top cells), the 9
    tendencies read and written, PAR written (FP32: 104 B; FP64: 208 B); the
    update-state sweep's: P read, PAR written."""
    import torch
    from oceanbiome import workloads
    from oceanbiome.runtime import HIPBiogeochemistry

    wl = workloads.lobster_3d(args.nx, args.ny, args.nz, nsteps=1)
    g = wl.grid
    ncell = g.n_interior
    nw = wl.n_written
    res = {}
    for prec in args.precisions.split(","):
        dev = HIPBiogeochemistry(wl.bgc, prec)
        tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
        G = [torch.zeros_like(t) for t in tr]
        PAR = torch.zeros_like(tr[0])
        es = 4 if prec == "fp32" else 8
        stages = max(1, args.steps)

        def pair(t):
            dev.update_biogeochemical_state(tr, PAR, t)
            dev.update_tendencies(tr, G, PAR, t=t)

        for q in range(args.warmup):
            pair(q * wl.dt)
        torch.cuda.synchronize()
        s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s0.record()
        for q in range(stages):
            pair(q * wl.dt)
        s1.record()
        torch.cuda.synchronize()
        pair_ms = s0.elapsed_time(s1) / stages
        dev.set_kernel_timing(True)
        dev.kernel_timing()
        for q in range(stages):
            dev.update_biogeochemical_state(tr, PAR, q * wl.dt)
        torch.cuda.synchronize()
        par_ms, n1 = dev.kernel_timing()
        for q in range(stages):
            dev.update_tendencies(tr, G, PAR, t=q * wl.dt)
        torch.cuda.synchronize()
        tend_ms, n2 = dev.kernel_timing()
        surf_ms, n3 = dev.kernel_timing(which=1)
        dev.set_kernel_timing(False)
        par_ms, tend_ms = par_ms / max(n1, 1), tend_ms / max(n2, 1)
        assert wl.bgc.carbonate_system is not None
        tbytes = ncell * ((nw - 2) + 2 * nw + 1) * es
        pbytes = ncell * 2 * es          # P read, PAR written
        res[prec] = {"ms_per_stage": pair_ms, "tendency_evaluations_per_s": nw * ncell / (pair_ms * 1e-3),
                     "update_state_sweep_ms": par_ms, "update_state_sweep_GBs": pbytes / (par_ms * 1e-3) / 1e9,
                     "tendency_sweep_ms": tend_ms,
                     "tendency_sweep_bytes": tbytes, "tendency_sweep_GBs": tbytes / (tend_ms * 1e-3) / 1e9,
                     "surface_flux_ms": surf_ms / max(n3, 1)}
        log(f"{prec}: {pair_ms:.4f} ms per stage (update-state {par_ms:.4f} ms = "
            f"{res[prec]['update_state_sweep_GBs']:.1f} GB/s, tendencies {tend_ms:.4f} ms "
            f"= {res[prec]['tendency_sweep_GBs']:.1f} GB/s, surface flux {res[prec]['surface_flux_ms']:.4f} ms)")
        del tr, G, PAR, dev
        torch.cuda.empty_cache()
    calibration_copy(args)
    head_prec = "fp32" if "fp32" in res else "fp64"
    head = res[head_prec]
    traffic = read_traffic(os.path.join(ROOT, "profiles"), wl.name + "_hooks", head_prec)
    print(json.dumps({
        "metric": "tracer-cell tendency evaluations/s of the coupled-model hooks (update-state + tendencies)",
        "value": head["tendency_evaluations_per_s"], "unit": "tracer-cell tendency evaluations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_stage"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if head_prec == "fp32" else "f64", "data": "synthetic (seeded lognormal/uniform tracers)",
        "config": {"workload": wl.name + "_hooks", "grid": [g.Nx, g.Ny, g.Nz], "tracers_written": nw,
                   "one_step_is": "one stage: update-state then tendencies"},
        "roofline": {"bound": "hbm", "achieved": head["tendency_sweep_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": head["tendency_sweep_GBs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_column_lane (tendency mode)",
                     "algorithmic_bytes_per_launch": head["tendency_sweep_bytes"],
                     "launch_ms": head["tendency_sweep_ms"]},
        "cpu_baseline": None, "results": res}))


def bench_columns(args):
    """The ragged NPZD column ensemble: 8192 columns, Nz ~ U{16..128},
    per-column parameters and surface PAR, 500 steps of 60 s, FP32-mixed and
    FP64; packing autotuned first; every column checked against the oracle.
    The warm-up steps run on the tracers themselves and the timed call
    continues from them; the oracle runs the same two calls.  With
    --sediment, an InstantRemineralisation sediment under every column, its
    storage checked too."""
    import torch
    from oceanbiome import InstantRemineralisationSediment, abi, workloads
    from oceanbiome.runtime import HIPBatch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    wl = workloads.npzd_columns(n_columns=args.columns)
    if args.sediment:
        wl.bgc.sediment = InstantRemineralisationSediment()
    off = wl.extra["offsets"]
    ncell = int(off[-1])
    ncol = len(off) - 1
    steps = args.steps if args.steps != 200 else wl.nsteps
    nthreads = cpu_threads(args.cpu_threads)
    sed_ref = O.SedimentState(ncol, 1) if args.sediment else None
    ref = [a.copy() for a in wl.tracers]
    if args.warmup:
        O.batch_run(wl.bgc.config(abi.OBME_FP64), off, list(wl.extra["params"]), wl.extra["dz"], wl.extra["T"], ref,
                    args.warmup, wl.dt, nthreads=nthreads, sediment=sed_ref)
    t0 = time.perf_counter()
    O.batch_run(wl.bgc.config(abi.OBME_FP64), off, list(wl.extra["params"]), wl.extra["dz"], wl.extra["T"], ref,
                steps, wl.dt, nthreads=nthreads, sediment=sed_ref)
    cpu_s = time.perf_counter() - t0
    cpu_steps = steps
    while cpu_s < args.cpu_seconds and cpu_steps < 20 * steps:
        # more of the same work (on a copy) until the CPU sample is about --cpu-seconds
        extra = [a.copy() for a in wl.tracers]
        t0 = time.perf_counter()
        O.batch_run(wl.bgc.config(abi.OBME_FP64), off, list(wl.extra["params"]), wl.extra["dz"], wl.extra["T"],
                    extra, steps, wl.dt, nthreads=nthreads,
                    sediment=O.SedimentState(ncol, 1) if args.sediment else None)
        cpu_s += time.perf_counter() - t0
        cpu_steps += steps
    log(f"oracle: {ncell} cells x {cpu_steps} steps in {cpu_s:.1f} s on {nthreads} threads")
    res = {}
    for prec in args.precisions.split(","):
        b = HIPBatch(wl.bgc, off, wl.extra["params"], wl.extra["dz"], wl.extra["T"], prec)
        tr = [torch.from_numpy(a).to(device="cuda", dtype=b.dtype) for a in wl.tracers]
        b.autotune(tr, 20, wl.dt)
        plan = b.describe()
        b.step(tr, args.warmup, wl.dt)
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        b.step(tr, steps, wl.dt)
        e.record()
        torch.cuda.synchronize()
        secs = s.elapsed_time(e) / 1e3
        out = [t.double().cpu().numpy() for t in tr]
        worst = 0.0
        tracer_err = []
        for a, r in zip(out, ref):
            tracer_err.append(float(np.linalg.norm(a - r) / np.linalg.norm(r)))
        for c in range(len(off) - 1):
            sl = slice(off[c], off[c + 1])
            for a, r in zip(out, ref):
                nr = np.linalg.norm(r[sl])
                if nr > 0:
                    worst = max(worst, float(np.linalg.norm(a[sl] - r[sl]) / nr))
        res[prec] = {"updates_per_s": 4 * ncell * steps / secs, "ms_per_step": secs / steps * 1e3, "plan": plan,
                     "rel_l2_per_tracer": tracer_err, "worst_column_rel_l2": worst}
        if args.sediment:
            st = b.sediment_storage.cpu().numpy()
            res[prec]["storage_rel_l2"] = float(np.linalg.norm(st - sed_ref.storage.ravel()) /
                                                np.linalg.norm(sed_ref.storage))
        log(f"{prec}: {res[prec]['updates_per_s']:.3e} upd/s, {plan}, tracer errors {tracer_err}, worst column {worst:.2e}")
        del b, tr
    head = res.get("fp32") or res["fp64"]
    print(json.dumps({
        "metric": "tracer-cell updates/s (NPZD column ensemble)", "value": head["updates_per_s"],
        "unit": "tracer-cell updates/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if "fp32" in res else "f64", "data": "synthetic (seeded)",
        "config": {"workload": wl.name + ("_sediment" if args.sediment else ""), "columns": ncol, "cells": ncell,
                   "dt": wl.dt, "sediment": "InstantRemineralisation (P, D into N)" if args.sediment else None},
        "cpu_baseline": {"value": 4 * ncell * cpu_steps / cpu_s, "unit": "tracer-cell updates/s", "cores": nthreads, "host_cpus": host_cpus(),
                         "kind": "port",
                         "sample": f"the whole ensemble, {cpu_steps} steps, FP64 oracle, OpenMP over columns, "
                                   f"{cpu_s:.1f} s"},
        "results": res}))


def bench_carbonate(args):
    """pH and CO2 flux on 2^24 points: the reference damped Newton in FP64
    (mode 0) and an FP32 guess refined by FP64 Newton (mode 1); errors against
    the oracle on a 65536-point sample."""
    import torch
    from oceanbiome import abi, workloads
    from oceanbiome.runtime import HIPBiogeochemistry
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    pts = workloads.carbonate_points(args.points)
    wl = workloads.lobster_3d(4, 4, 4, nsteps=1)
    dev = HIPBiogeochemistry(wl.bgc, "fp64")
    cfg = wl.bgc.config(abi.OBME_FP64)
    nthreads = cpu_threads(args.cpu_threads)
    ns = min(65536, args.points)
    t0 = time.perf_counter()
    pH_ref, F_ref = O.co2_flux_batch(cfg.co2, cfg.solver, pts["DIC"][:ns], pts["Alk"][:ns], pts["T"][:ns],
                                     pts["S"][:ns], pts["u10"][:ns], pts["pCO2_air"], nthreads=nthreads)
    cpu_s = time.perf_counter() - t0
    cpu_pts = ns
    if not args.no_cpu:
        # the CPU baseline: the whole point set, repeated until about --cpu-seconds
        cpu_s, cpu_pts = 0.0, 0
        while cpu_s < args.cpu_seconds and cpu_pts < 64 * args.points:
            t0 = time.perf_counter()
            O.co2_flux_batch(cfg.co2, cfg.solver, pts["DIC"], pts["Alk"], pts["T"], pts["S"], pts["u10"],
                             pts["pCO2_air"], nthreads=nthreads)
            cpu_s += time.perf_counter() - t0
            cpu_pts += args.points
    dargs = [torch.from_numpy(pts[k]).cuda() for k in ("DIC", "Alk", "T", "S", "u10")]
    pH = torch.empty_like(dargs[0])
    F = torch.empty_like(dargs[0])
    res = {}
    for mode in (0, 1):
        dev.co2_flux(*dargs, pCO2_air=pts["pCO2_air"], mode=mode, pH=pH, flux=F)
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = max(1, args.steps // 50)
        s.record()
        for _ in range(reps):
            dev.co2_flux(*dargs, pCO2_air=pts["pCO2_air"], mode=mode, pH=pH, flux=F)
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / reps
        p = pH[:ns].cpu().numpy()
        f = F[:ns].cpu().numpy()
        res[f"mode{mode}"] = {"ms": ms, "points_per_s": args.points / (ms * 1e-3),
                              "pH_max_abs_err": float(np.max(np.abs(p - pH_ref))),
                              "flux_rel_l2": float(np.linalg.norm(f - F_ref) / np.linalg.norm(F_ref)),
                              "flux_max_rel": float(np.max(np.abs(f - F_ref) / np.maximum(np.abs(F_ref), 1e-30)))}
        log(f"mode {mode}: {ms:.3f} ms, {res[f'mode{mode}']}")
    print(json.dumps({
        "metric": "surface points/s (pH solve + CO2 flux)", "value": res["mode1"]["points_per_s"],
        "unit": "points/s", "n_gpus": 1, "steps": max(1, args.steps // 50), "warmup": 1,
        "ms_per_step": res["mode1"]["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded)",
        "config": {"workload": f"carbonate_points_{args.points}"},
        "cpu_baseline": {"value": cpu_pts / cpu_s, "unit": "points/s", "cores": nthreads, "host_cpus": host_cpus(), "kind": "port",
                         "sample": f"{cpu_pts} points ({cpu_pts // args.points} passes over the set), FP64 oracle "
                                   f"(reference solve from pH 8), OpenMP, {cpu_s:.1f} s"},
        "results": res}))



def _pisces_oracle_inputs(wl, cols=None):
    """Oracle arguments for the PISCES workload, on all columns or on the
    columns `cols` (flat indices j*Nx + i) gathered into one row."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pisces as OP
    g, m = wl.grid, wl.bgc
    q = OP.params_from_vector(m.param_vector())
    light = dict(surface_PAR=m.surface_PAR, kw=m.kw, e=m.e, chi=m.chi, division=m.division)
    F = {}
    for name, a in zip(OP.TRACERS, wl.tracers):
        x = g.interior(a)
        if cols is not None:
            x = x.reshape(g.Nz, -1)[:, cols].reshape(g.Nz, 1, len(cols))
        F[name] = np.ascontiguousarray(x, dtype=np.float64)
    cc = m.carbon_chemistry
    solver = cc.solver.struct(cc.initial_pH_guess)
    return OP, q, light, F, np.asarray(g.zc), np.asarray(g.zf), np.asarray(g.dz), solver


def pisces_surface_errors(wl, tops, F):
    """pH (free scale) and the CO2 flux (the default
    CarbonDioxideGasExchangeBoundaryCondition, as a diagnostic of the top
    cells' DIC and Alk) of the sampled columns' top cells: the device state
    `tops` (top-cell values per tracer, tracer order) against the oracle's F,
    both through the oracle's point function (carbon_chemistry.jl:168,
    gas_exchange.jl:26-38)."""
    import oracle as O
    from oceanbiome import CarbonDioxideGasExchangeBoundaryCondition
    from oceanbiome.pisces import TRACERS
    gas = CarbonDioxideGasExchangeBoundaryCondition().struct()
    cc = wl.bgc.carbon_chemistry
    solver = cc.solver.struct(cc.initial_pH_guess)
    idx = {n: TRACERS.index(n) for n in ("DIC", "Alk", "T", "S")}
    ref = {n: np.ascontiguousarray(F[n][-1].ravel()) for n in ("DIC", "Alk")}
    T = np.ascontiguousarray(tops[idx["T"]], dtype=np.float64)
    S = np.ascontiguousarray(tops[idx["S"]], dtype=np.float64)
    u10 = np.full(T.shape, gas.wind_speed)
    pH_r, F_r = O.co2_flux_batch(gas, solver, ref["DIC"], ref["Alk"], T, S, u10, gas.air_concentration)
    pH_g, F_g = O.co2_flux_batch(gas, solver, np.ascontiguousarray(tops[idx["DIC"]], dtype=np.float64),
                                 np.ascontiguousarray(tops[idx["Alk"]], dtype=np.float64), T, S, u10,
                                 gas.air_concentration)
    return {"pH (surface)": {"max_abs": float(np.max(np.abs(pH_g - pH_r))),
                             "rms": float(np.sqrt(np.mean((pH_g - pH_r) ** 2)))},
            "CO2 flux (surface, diagnostic)": {"rel_l2": float(np.linalg.norm(F_g - F_r) / np.linalg.norm(F_r)),
                                               "max_rel": float(np.max(np.abs(F_g - F_r) / np.abs(F_r)))}}


def calibration_copy(args):
    """A device copy of known size (--calibrate-copy-mb) for the HBM counters."""
    import torch
    if args.calibrate_copy_mb:
        n = args.calibrate_copy_mb * (1 << 20) // 4
        src = torch.ones(n, device="cuda", dtype=torch.float32)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {n * 4} bytes read, {n * 4} bytes written")
        del src, dst


def bench_pisces(args):
    """PISCES on 256x256x64 (MultiBand PAR, Omega, 24 tendencies, sinking,
    RK3), 300 s steps: FP32-mixed (FP64 Omega) and FP64 timed on the device;
    both checked against the FP64 oracle on a sample of columns after the
    check run; the oracle timed on a bounded sample as the CPU baseline."""
    import torch
    from oceanbiome import workloads
    from oceanbiome.runtime import HIPPISCES
    steps = args.steps if args.steps != 200 else 10
    nchk = args.check_steps if args.check_steps != 200 else 100
    wl = workloads.pisces_3d(256, 256, 64, nsteps=nchk, dt=300.0)
    g = wl.grid
    ncell = g.n_interior
    nw = 24
    calibration_copy(args)
    res, outputs = {}, {}
    for prec in args.precisions.split(","):
        dev = HIPPISCES(wl.bgc, prec)
        tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
        if args.warmup:
            dev.step(tr, args.warmup, wl.dt)
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        dev.step(tr, steps, wl.dt, t0=args.warmup * wl.dt)
        e.record()
        torch.cuda.synchronize()
        secs = s.elapsed_time(e) / 1e3
        sb = 4 if prec == "fp32" else 8
        bytes_step = wl.algorithmic_bytes_per_step(sb)
        res[prec] = {"ms_per_step": secs / steps * 1e3, "updates_per_s": nw * ncell * steps / secs,
                     "achieved_GBs": bytes_step / (secs / steps) / 1e9, "bytes_per_step": bytes_step}
        log(f"pisces {prec}: {res[prec]['ms_per_step']:.3f} ms/step  {res[prec]['updates_per_s']:.3e} upd/s  "
            f"{res[prec]['achieved_GBs']:.0f} GB/s")
        if not args.no_check:
            tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
            dev.step(tr, nchk, wl.dt)
            torch.cuda.synchronize()
            outputs[prec] = [t.double().cpu().numpy() for t in tr]
        del tr, dev
        torch.cuda.empty_cache()
    errors, cpu = {}, None
    nthreads = cpu_threads(args.cpu_threads)
    if not args.no_check:
        rng = np.random.default_rng(1)
        cols = np.sort(rng.choice(g.Nx * g.Ny, min(args.sample_columns, g.Nx * g.Ny), replace=False))
        OP, q, light, F, zc, zf, dz, solver = _pisces_oracle_inputs(wl, cols)
        F0 = {k: v.copy() for k, v in F.items()}
        t0 = time.perf_counter()
        OP.step(q, light, F, zc, zf, dz, nchk, wl.dt, 0.0, solver, nthreads=nthreads)
        log(f"oracle check run: {len(cols)} columns x {nchk} steps in {time.perf_counter() - t0:.1f} s")
        # the FP32 storage floor: the oracle with its state rounded to FP32
        # at the start and after every step
        X = {k: v.astype(np.float32).astype(np.float64) for k, v in F0.items()}
        with np.errstate(all="ignore"):
            for st in range(nchk):
                OP.step(q, light, X, zc, zf, dz, 1, wl.dt, st * wl.dt, solver, nthreads=nthreads)
                for k in OP.WRITTEN:
                    X[k] = X[k].astype(np.float32).astype(np.float64)
        floor = {k: float(np.linalg.norm(X[k] - F[k]) / np.linalg.norm(F[k])) for k in OP.WRITTEN}
        for prec, out in outputs.items():
            e = {}
            for n, name in enumerate(OP.WRITTEN):
                a = g.interior(out[n]).reshape(g.Nz, -1)[:, cols].reshape(F[name].shape)
                b = F[name]
                e[name] = {"rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                           "max_rel": float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))}
                if prec == "fp32":
                    e[name]["fp32_storage_floor"] = floor[name]
            tops = [g.interior(out[n]).reshape(g.Nz, -1)[-1, cols] for n in range(26)]
            e.update(pisces_surface_errors(wl, tops, F))
            errors[prec] = e
            worst = max(v["rel_l2"] for k, v in e.items() if "(surface" not in k)
            log(f"pisces {prec} vs oracle after {nchk} steps ({len(cols)} columns): worst rel L2 {worst:.3e}, "
                f"surface pH max abs {e['pH (surface)']['max_abs']:.2e}")
    if not args.no_cpu:
        rng = np.random.default_rng(2)
        ncol = 4096
        cols = np.sort(rng.choice(g.Nx * g.Ny, ncol, replace=False))
        OP, q, light, F, zc, zf, dz, solver = _pisces_oracle_inputs(wl, cols)
        t0 = time.perf_counter()
        OP.step(q, light, F, zc, zf, dz, 1, wl.dt, 0.0, solver, nthreads=nthreads)
        per = time.perf_counter() - t0
        n = int(max(1, min(100, round(args.cpu_seconds / per))))
        t0 = time.perf_counter()
        OP.step(q, light, F, zc, zf, dz, n, wl.dt, wl.dt, solver, nthreads=nthreads)
        secs = time.perf_counter() - t0
        cpu = {"value": nw * g.Nz * ncol * n / secs, "unit": "tracer-cell updates/s", "cores": nthreads, "host_cpus": host_cpus(),
               "kind": "port",
               "sample": f"{ncol} columns x {g.Nz} levels x {n} steps of the same workload, FP64 numpy oracle "
                         f"(calcite saturation in C over {nthreads} threads, the rest single-threaded), {secs:.1f} s"}
        log(f"cpu baseline: {cpu['value']:.3e} upd/s ({secs:.1f} s)")
    head_prec = "fp32" if "fp32" in res else "fp64"
    head = res[head_prec]
    print(json.dumps({
        "metric": "tracer-cell updates/s + % HBM roofline (PISCES)", "value": head["updates_per_s"],
        "unit": "tracer-cell updates/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if head_prec == "fp32" else "f64", "data": "synthetic (seeded lognormal tracers)",
        "config": {"workload": wl.name, "grid": [g.Nx, g.Ny, g.Nz], "tracers_written": nw, "dt": wl.dt},
        "roofline": {"bound": "hbm", "achieved": head["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": head["achieved_GBs"] / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "whole PISCES step", "algorithmic_bytes_per_launch": head["bytes_per_step"],
                     "launch_ms": head["ms_per_step"]},
        "cpu_baseline": cpu, "results": res, "errors_vs_oracle": errors}))


if __name__ == "__main__":
    main()
