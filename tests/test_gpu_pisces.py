"""GPU parity of PISCES (obme_pisces_*) against the numpy FP64 oracle
(oracle/pisces.py) on the same inputs: the update-state fields (MultiBand
PAR, euphotic depth, mixed-layer PAR, calcite saturation), the tendencies
and a few RK3 steps.  FP64 bar: relative L2 <= 1e-12 per tracer, for the
raw tendencies and after the steps (the raw tendencies measured at most
3.2e-15, Fe, whose free-iron root moves that much when exp(...) changes by
one ulp); FP32-mixed: 1e-5.  The inputs lie around the values of the
reference's PISCES tests (test/test_PISCES.jl:6-16), whose phytoplankton
iron quotas sit far from the pole of the iron uptake factor (1 - r)/(1.05 - r)
(mixed_mondo.jl:161-181)."""
import numpy as np
import pytest

import pisces as OP
from helpers import rel_l2
from oceanbiome import workloads
from oceanbiome.runtime import HIPPISCES

pytestmark = pytest.mark.gpu

NAMES = OP.TRACERS


def _oracle_inputs(wl):
    g = wl.grid
    m = wl.bgc
    q = OP.params_from_vector(m.param_vector())
    light = dict(surface_PAR=m.surface_PAR, kw=m.kw, e=m.e, chi=m.chi, division=m.division)
    F = {n: g.interior(a).copy() for n, a in zip(NAMES, wl.tracers)}
    cc = m.carbon_chemistry
    solver = cc.solver.struct(cc.initial_pH_guess)
    return q, light, F, np.asarray(g.zc), np.asarray(g.zf), np.asarray(g.dz), solver


def _device(wl, precision):
    import torch
    dev = HIPPISCES(wl.bgc, precision)
    tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
    return dev, tr


def test_update_state_fp64():
    import torch
    print("case pisces update_state fp64", flush=True)
    wl = workloads.pisces_3d(12, 5, 30, nsteps=1, seed=3, medians="test")
    q, light, F, zc, zf, dz, solver = _oracle_inputs(wl)
    A = OP.update_state(q, light, F, zc, zf, 0.0, solver)
    dev, tr = _device(wl, "fp64")
    PAR, bands, Om, zeu, ml = dev.update_biogeochemical_state(tr, 0.0)
    torch.cuda.synchronize()
    g = wl.grid
    assert rel_l2(g.interior(PAR.cpu().numpy()), A["PAR"]) <= 1e-13
    for n in range(3):
        assert rel_l2(g.interior(bands[n].cpu().numpy()), A[f"PAR{n + 1}"]) <= 1e-13
    assert rel_l2(g.interior(Om.cpu().numpy()), A["Omega"]) <= 1e-12
    assert rel_l2(zeu.cpu().numpy(), A["zeu"][0]) <= 1e-13
    assert rel_l2(ml.cpu().numpy(), A["mlPAR"][0]) <= 1e-13


def test_tendencies_fp64():
    import torch
    print("case pisces tendencies fp64", flush=True)
    wl = workloads.pisces_3d(10, 4, 24, nsteps=1, seed=5, medians="test")
    q, light, F, zc, zf, dz, solver = _oracle_inputs(wl)
    G, _ = OP.full_tendencies(q, light, F, zc, zf, dz, 1.0e5, solver)
    dev, tr = _device(wl, "fp64")
    g = wl.grid
    Gd = [torch.zeros(g.parent_shape, dtype=torch.float64, device="cuda") for _ in range(24)]
    dev.update_tendencies(tr, Gd, t=1.0e5)
    torch.cuda.synchronize()
    worst = {}
    for n, name in enumerate(OP.WRITTEN):
        worst[name] = rel_l2(g.interior(Gd[n].cpu().numpy()), G[name])
    print("tendency errors", worst, flush=True)
    assert all(e <= 1e-12 for e in worst.values()), worst


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_step(precision):
    import torch
    print("case pisces step", precision, flush=True)
    wl = workloads.pisces_3d(8, 4, 20, nsteps=3, dt=300.0, seed=7, medians="test")
    q, light, F, zc, zf, dz, solver = _oracle_inputs(wl)
    OP.step(q, light, F, zc, zf, dz, wl.nsteps, wl.dt, 0.0, solver)
    dev, tr = _device(wl, precision)
    dev.step(tr, wl.nsteps, wl.dt)
    torch.cuda.synchronize()
    g = wl.grid
    errs = {name: rel_l2(g.interior(tr[n].double().cpu().numpy()), F[name]) for n, name in enumerate(OP.WRITTEN)}
    print("errors", errs, flush=True)
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert all(e <= tol for e in errs.values()), errs
    # T and S untouched
    assert np.array_equal(tr[24].double().cpu().numpy(), wl.tracers[24].astype(tr[24].cpu().numpy().dtype))


def _scaled_run(precision, flipped, seed=9):
    import torch
    from oceanbiome import pisces as HP
    wl = workloads.pisces_3d(8, 4, 20, nsteps=3, dt=300.0, seed=seed, medians="test")
    wl.bgc = HP.PISCES(wl.grid, scale_negatives=True)
    rng = np.random.default_rng(3)
    g = wl.grid
    for name in flipped:
        a = g.interior(wl.tracers[HP.TRACERS.index(name)])
        flip = rng.random(a.shape) < 0.1
        a[flip] = -0.2 * np.abs(a[flip])
    q, light, F, zc, zf, dz, solver = _oracle_inputs(wl)
    with np.errstate(all="ignore"):
        OP.step(q, light, F, zc, zf, dz, wl.nsteps, wl.dt, 0.0, solver, scaling=OP.conserved_groups(q))
    dev, tr = _device(wl, precision)
    dev.step(tr, wl.nsteps, wl.dt)
    torch.cuda.synchronize()
    return {name: (g.interior(tr[n].double().cpu().numpy()), F[name]) for n, name in enumerate(OP.WRITTEN)}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_step_with_negative_scaling(precision):
    """PISCES(; scale_negatives = true): the five groups applied at the start
    of every stage, against the oracle (negative nutrients and DOC; every
    group total stays positive, so no fill value enters)."""
    print("case pisces scaling", precision, flush=True)
    res = _scaled_run(precision, ("NH₄", "NO₃", "Fe", "DSi", "PO₄", "DOC"))
    errs = {name: rel_l2(a, b) for name, (a, b) in res.items()}
    print("errors", errs, flush=True)
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert all(e <= tol for e in errs.values()), errs


def test_scaling_zeroed_phytoplankton_gives_the_same_nans():
    """Zeroing a negative P leaves PChl and PFe (members of no group with P,
    or of another group) positive, so the ratios PChl/P and PFe/P divide by
    P + eps(0): the reference's arithmetic gives NaN there.  The device
    gives NaN in exactly the same cells and agrees elsewhere (FP64)."""
    print("case pisces scaling nan", flush=True)
    res = _scaled_run("fp64", ("P", "Z", "POC"))
    for name, (a, b) in res.items():
        assert np.array_equal(np.isnan(a), np.isnan(b)), name
        ok = ~np.isnan(b)
        assert rel_l2(a[ok], b[ok]) <= 1e-12, name


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_step_with_gas_exchange(precision):
    """PISCES with the CO2 and O2 top boundary fluxes of
    validation/PISCES/column.jl:63-64 (DIC and O2 lose F/Δz in the top cell)."""
    import torch
    from oceanbiome import pisces as HP
    print("case pisces gas exchange", precision, flush=True)
    wl = workloads.pisces_3d(8, 4, 20, nsteps=3, dt=300.0, seed=13, medians="test")
    wl.bgc = HP.PISCES(wl.grid, co2_exchange=True, o2_exchange=True)
    q, light, F, zc, zf, dz, solver = _oracle_inputs(wl)
    cfg = wl.bgc.config(0)
    gas = {"co2": cfg.co2, "o2": cfg.o2}
    F0 = {k: v.copy() for k, v in F.items()}
    OP.step(q, light, F, zc, zf, dz, wl.nsteps, wl.dt, 0.0, solver, gas=gas)
    dev, tr = _device(wl, precision)
    dev.step(tr, wl.nsteps, wl.dt)
    torch.cuda.synchronize()
    g = wl.grid
    errs = {name: rel_l2(g.interior(tr[n].double().cpu().numpy()), F[name]) for n, name in enumerate(OP.WRITTEN)}
    print("errors", errs, flush=True)
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert all(e <= tol for e in errs.values()), errs
    # the fluxes act: the top-cell O2 moves towards saturation differently from a run without them
    F1 = {k: v.copy() for k, v in F0.items()}
    OP.step(q, light, F1, zc, zf, dz, wl.nsteps, wl.dt, 0.0, solver)
    assert not np.allclose(F1["O2"][-1], F["O2"][-1], rtol=1e-9)


def test_tendencies_with_gas_exchange_fp64():
    import torch
    from oceanbiome import pisces as HP
    print("case pisces gas exchange tendencies", flush=True)
    wl = workloads.pisces_3d(10, 4, 24, nsteps=1, seed=5, medians="test")
    wl.bgc = HP.PISCES(wl.grid, co2_exchange=True, o2_exchange=True)
    q, light, F, zc, zf, dz, solver = _oracle_inputs(wl)
    cfg = wl.bgc.config(0)
    G, _ = OP.full_tendencies(q, light, F, zc, zf, dz, 1.0e5, solver, gas={"co2": cfg.co2, "o2": cfg.o2})
    dev, tr = _device(wl, "fp64")
    g = wl.grid
    Gd = [torch.zeros(g.parent_shape, dtype=torch.float64, device="cuda") for _ in range(24)]
    dev.update_tendencies(tr, Gd, t=1.0e5)
    torch.cuda.synchronize()
    for n, name in enumerate(OP.WRITTEN):
        assert rel_l2(g.interior(Gd[n].cpu().numpy()), G[name]) <= 1e-12, name
