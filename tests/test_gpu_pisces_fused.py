"""The fused FP32-mixed PISCES step (one wave per column, all three stages in
registers) against the numpy FP64 oracle at full 64-level columns and at a
ragged depth, with the gas exchange and the negative scaling on, over two
calls (the second warm-starts [H+] from the first); against the per-stage
kernels (OBME_PISCES_FUSED=0) on a larger grid; several steps per launch
(the state in registers between steps) against one step per launch, bit for
bit; and the per-stage path that columns deeper than 64 levels take.  FP32-mixed bar: relative L2 <= 1e-5
per tracer.  Inputs around the values of the reference's PISCES tests
(test/test_PISCES.jl:6-16)."""
import os

import numpy as np
import pytest

import pisces as OP
from helpers import rel_l2
from oceanbiome import pisces as HP
from oceanbiome import workloads
from oceanbiome.runtime import HIPPISCES

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _oracle_inputs(wl):
    g = wl.grid
    m = wl.bgc
    q = OP.params_from_vector(m.param_vector())
    light = dict(surface_PAR=m.surface_PAR, kw=m.kw, e=m.e, chi=m.chi, division=m.division)
    F = {n: g.interior(a).copy() for n, a in zip(OP.TRACERS, wl.tracers)}
    cc = m.carbon_chemistry
    solver = cc.solver.struct(cc.initial_pH_guess)
    return q, light, F, np.asarray(g.zc), np.asarray(g.zf), np.asarray(g.dz), solver


def _device_run(wl, calls, nsteps, fused=True, steps_per_launch=None, t_start=0.0):
    import torch
    env = {"OBME_PISCES_FUSED": "1" if fused else "0"}
    if steps_per_launch is not None:
        env["OBME_PISCES_FUSED_STEPS"] = str(steps_per_launch)
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        dev = HIPPISCES(wl.bgc, "fp32")
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
    for c in range(calls):
        dev.step(tr, nsteps, wl.dt, t0=t_start + c * nsteps * wl.dt)
    torch.cuda.synchronize()
    return [t.double().cpu().numpy() for t in tr]


@pytest.mark.parametrize("Nz,gas,scale", [(64, False, False), (64, True, True), (37, True, False)])
def test_fused_step_against_oracle(Nz, gas, scale):
    print("case pisces fused", Nz, "gas", gas, "scale", scale, flush=True)
    wl = workloads.pisces_3d(6, 3, Nz, nsteps=2, dt=300.0, seed=17, medians="test")
    wl.bgc = HP.PISCES(wl.grid, co2_exchange=gas, o2_exchange=gas, scale_negatives=scale)
    q, light, F, zc, zf, dz, solver = _oracle_inputs(wl)
    cfg = wl.bgc.config(0)
    kw = {}
    if gas:
        kw["gas"] = {"co2": cfg.co2, "o2": cfg.o2}
    if scale:
        kw["scaling"] = OP.conserved_groups(q)
    for c in range(2):
        OP.step(q, light, F, zc, zf, dz, wl.nsteps, wl.dt, c * wl.nsteps * wl.dt, solver, **kw)
    out = _device_run(wl, 2, wl.nsteps)
    g = wl.grid
    errs = {name: rel_l2(g.interior(out[n]), F[name]) for n, name in enumerate(OP.WRITTEN)}
    print("errors", {k: float(v) for k, v in errs.items()}, flush=True)
    assert all(e <= TOL for e in errs.values()), errs
    np.testing.assert_array_equal(out[24], wl.tracers[24].astype(np.float32).astype(np.float64))


def test_fused_matches_the_per_stage_kernels():
    print("case pisces fused vs per-stage", flush=True)
    wl = workloads.pisces_3d(40, 9, 64, nsteps=4, dt=300.0, seed=23, medians="test")
    wl.bgc = HP.PISCES(wl.grid, co2_exchange=True, o2_exchange=True)
    a = _device_run(wl, 1, wl.nsteps, fused=True)
    b = _device_run(wl, 1, wl.nsteps, fused=False)
    g = wl.grid
    errs = {name: rel_l2(g.interior(a[n]), g.interior(b[n])) for n, name in enumerate(OP.WRITTEN)}
    print("differences", {k: float(v) for k, v in errs.items()}, flush=True)
    assert all(e <= TOL for e in errs.values()), errs
    # halos untouched by the in-place fused step
    for n in range(len(OP.WRITTEN)):
        np.testing.assert_array_equal(a[n][0], wl.tracers[n][0].astype(np.float32).astype(np.float64))


def test_deep_columns_take_the_per_stage_path():
    print("case pisces deep columns", flush=True)
    wl = workloads.pisces_3d(4, 2, 70, nsteps=2, dt=300.0, seed=29, medians="test")
    q, light, F, zc, zf, dz, solver = _oracle_inputs(wl)
    OP.step(q, light, F, zc, zf, dz, wl.nsteps, wl.dt, 0.0, solver)
    out = _device_run(wl, 1, wl.nsteps)
    g = wl.grid
    errs = {name: rel_l2(g.interior(out[n]), F[name]) for n, name in enumerate(OP.WRITTEN)}
    print("errors", {k: float(v) for k, v in errs.items()}, flush=True)
    assert all(e <= TOL for e in errs.values()), errs


def test_fused_step_with_non_default_parameters():
    """A parameter set other than the defaults (a shallower mixed layer)."""
    print("case pisces fused non-default parameters", flush=True)
    wl = workloads.pisces_3d(5, 3, 64, nsteps=2, dt=300.0, seed=31, mixed_layer_depth=-30.0, medians="test")
    q, light, F, zc, zf, dz, solver = _oracle_inputs(wl)
    OP.step(q, light, F, zc, zf, dz, wl.nsteps, wl.dt, 0.0, solver)
    out = _device_run(wl, 1, wl.nsteps)
    g = wl.grid
    errs = {name: rel_l2(g.interior(out[n]), F[name]) for n, name in enumerate(OP.WRITTEN)}
    print("errors", {k: float(v) for k, v in errs.items()}, flush=True)
    assert all(e <= TOL for e in errs.values()), errs


@pytest.mark.parametrize("gas,scale", [(True, False), (False, True)])
def test_steps_per_launch_bit_identical(gas, scale):
    """20 steps: launches of 16 + 4 steps against 20 launches of one step
    (the [H+] between steps rounded to FP32 as its store between launches
    rounds it), with the diurnal surface PAR from 10 h, in daylight, so the
    surface PAR of every stage differs."""
    print("case pisces fused steps per launch gas", gas, "scale", scale, flush=True)
    wl = workloads.pisces_3d(24, 5, 64, nsteps=20, dt=300.0, seed=37, medians="test")
    wl.bgc = HP.PISCES(wl.grid, co2_exchange=gas, o2_exchange=gas, scale_negatives=scale, surface_PAR="default")
    a = _device_run(wl, 1, wl.nsteps, t_start=36000.0)
    b = _device_run(wl, 1, wl.nsteps, steps_per_launch=1, t_start=36000.0)
    # (cells where the growth rate's day length, day_length(φ, t) in the
    # reference's argument order, is zero and PAR has underflowed to zero
    # are NaN, as in the reference: they must coincide)
    print("non-finite", sum(int((~np.isfinite(x)).sum()) for x in a), flush=True)
    for n, name in enumerate(OP.WRITTEN):
        assert np.array_equal(a[n], b[n], equal_nan=True), (name, float(np.nanmax(np.abs(a[n] - b[n]))))
