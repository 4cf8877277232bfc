"""GPU parity of the batched NPZD column ensemble (obme_batch_*) against the
FP64 oracle's column runs (oracle.batch_run): ragged depths 16..128, every
rate perturbed per column, per-column surface PAR.  FP64: relative L2 <= 1e-12
per tracer; FP32-mixed: <= 1e-5 per tracer over the ensemble and for every
column's own tracers, at the benchmark's full size (8192 columns, 500 steps)
too.  Also: every packing plan, several launches per call replayed from a
hipGraph (bit-identical to one launch: the FP32 residuals pass between
launches), and the autotuner leaving the state alone."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2
from oceanbiome import abi, workloads
from oceanbiome.runtime import HIPBatch

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5
FP32_COLUMN_TOL = 1e-5


def _ensemble(n=96, nsteps=20, seed=7):
    return workloads.npzd_columns(n_columns=n, nsteps=nsteps, seed=seed)


def _oracle(wl, nsteps, nthreads=8):
    U = [a.copy() for a in wl.tracers]
    O.batch_run(wl.bgc.config(abi.OBME_FP64), wl.extra["offsets"], list(wl.extra["params"]), wl.extra["dz"],
                wl.extra["T"], U, nsteps, wl.dt, nthreads=nthreads)
    return U


def _device(wl, precision):
    import torch
    b = HIPBatch(wl.bgc, wl.extra["offsets"], wl.extra["params"], wl.extra["dz"], wl.extra["T"], precision)
    tr = [torch.from_numpy(a).to(device="cuda", dtype=b.dtype) for a in wl.tracers]
    return b, tr


def _host(tr):
    return [t.double().cpu().numpy() for t in tr]


def _column_errors(wl, out, ref):
    off = wl.extra["offsets"]
    worst = 0.0
    for c in range(len(off) - 1):
        s = slice(off[c], off[c + 1])
        for a, b in zip(out, ref):
            worst = max(worst, rel_l2(a[s], b[s]))
    return worst


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_matches_oracle(precision):
    import torch
    print("case batch", precision, flush=True)
    wl = _ensemble()
    ref = _oracle(wl, wl.nsteps)
    b, tr = _device(wl, precision)
    b.step(tr, wl.nsteps, wl.dt)
    torch.cuda.synchronize()
    out = _host(tr)
    errs = [rel_l2(a, r) for a, r in zip(out, ref)]
    col = _column_errors(wl, out, ref)
    print("errors", errs, "worst column", col, b.describe(), flush=True)
    tol = FP64_TOL if precision == "fp64" else FP32_TOL
    assert all(e <= tol for e in errs), errs
    assert col <= (FP64_TOL * 10 if precision == "fp64" else FP32_COLUMN_TOL), col


def test_every_plan_agrees():
    import torch
    print("case batch plans", flush=True)
    wl = _ensemble(n=48, nsteps=5)
    ref = _oracle(wl, wl.nsteps)
    for plan in range(7):
        b, tr = _device(wl, "fp64")
        try:
            b.configure(plan, 0)
        except Exception as e:   # a plan whose largest tile is too small for the deepest column
            print("plan", plan, "rejected:", e, flush=True)
            continue
        b.step(tr, wl.nsteps, wl.dt)
        torch.cuda.synchronize()
        errs = [rel_l2(a, r) for a, r in zip(_host(tr), ref)]
        print("plan", plan, b.describe(), max(errs), flush=True)
        assert max(errs) <= FP64_TOL, (plan, errs)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_match_single_launch(precision):
    """steps_per_launch = 1 records nsteps launches per class into a hipGraph;
    three replays with eager work and synchronisation between them must each
    equal the single-launch run of the same steps, bit for bit."""
    import torch
    print("case batch graph", precision, flush=True)
    wl = _ensemble(n=40, nsteps=4)
    a, ta = _device(wl, precision)
    g, tg = _device(wl, precision)
    g.configure(1, 1)
    a.configure(1, 0)
    for rep in range(3):
        a.step(ta, wl.nsteps, wl.dt)
        g.step(tg, wl.nsteps, wl.dt)
        torch.cuda.synchronize()
        x = torch.zeros(1000, device="cuda").sum()   # eager work between replays
        torch.cuda.synchronize()
        for u, v in zip(ta, tg):
            assert torch.equal(u, v), rep
        del x


def test_autotune_keeps_state_and_result():
    import torch
    print("case batch autotune", flush=True)
    wl = _ensemble(n=64, nsteps=10)
    ref = _oracle(wl, wl.nsteps)
    b, tr = _device(wl, "fp32")
    before = [t.clone() for t in tr]
    b.autotune(tr, 5, wl.dt)
    torch.cuda.synchronize()
    assert all(torch.equal(u, v) for u, v in zip(tr, before))
    print("tuned", b.describe(), flush=True)
    b.step(tr, wl.nsteps, wl.dt)
    torch.cuda.synchronize()
    errs = [rel_l2(a, r) for a, r in zip(_host(tr), ref)]
    assert all(e <= FP32_TOL for e in errs), errs


def test_full_ensemble_every_column_fp32():
    """The benchmark ensemble (8192 columns, Nz ~ U{16..128}, 500 steps of
    60 s) in FP32-mixed: every tracer of every column within 1e-5 relative
    L2 of the FP64 oracle."""
    import torch
    print("case batch full size fp32", flush=True)
    wl = workloads.npzd_columns()
    b, tr = _device(wl, "fp32")
    b.step(tr, wl.nsteps, wl.dt)
    torch.cuda.synchronize()
    out = _host(tr)
    ref = _oracle(wl, wl.nsteps, nthreads=16)
    errs = [rel_l2(a, r) for a, r in zip(out, ref)]
    col = _column_errors(wl, out, ref)
    print("full-size errors", errs, "worst column", col, flush=True)
    assert all(e <= FP32_TOL for e in errs), errs
    assert col <= FP32_COLUMN_TOL, col
