"""ScaleNegativeTracers on the GPU against the oracle: in the fused step
(at the start of every stage; with the CO2 exchange the flux is computed in
the sweep, because the scaled top-cell DIC feeds back into the other
tracers) and in the update-state hook (in place).  FP64 <= 1e-12 relative
L2 per tracer (the group sums are taken in tracer order, a reassociation of
the member order); FP32-mixed <= 1e-5."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, to_device, to_host
from oceanbiome import NPZD, LOBSTER, CarbonateSystem, RectilinearGrid, TwoBandPhotosyntheticallyActiveRadiation, abi
from oceanbiome import CarbonDioxideGasExchangeBoundaryCondition, workloads
from oceanbiome.runtime import HIPBiogeochemistry
from oceanbiome.workloads import Workload

pytestmark = pytest.mark.gpu
TOL = {"fp64": 1e-12, "fp32": 1e-5}


def _with_negatives(wl, names, frac=0.05, seed=11):
    """Flip the sign of a fraction of the cells of the named tracers."""
    rng = np.random.default_rng(seed)
    g = wl.grid
    order = wl.bgc.required_biogeochemical_tracers()
    for name in names:
        a = g.interior(wl.tracers[order.index(name)])
        flip = rng.random(a.shape) < frac
        a[flip] = -0.3 * np.abs(a[flip])
    return wl


def _run(wl, precision, nsteps):
    import torch
    ref = [a.copy() for a in wl.tracers]
    O.grid_run(wl.bgc.config(abi.OBME_FP64), ref, nsteps, wl.dt)
    dev = HIPBiogeochemistry(wl.bgc, precision)
    tr = to_device(wl.tracers, dev.dtype)
    dev.step(tr, nsteps, wl.dt)
    torch.cuda.synchronize()
    out = to_host(tr)
    g = wl.grid
    return {n: rel_l2(g.interior(a), g.interior(b))
            for n, a, b in zip(wl.bgc.required_biogeochemical_tracers(), out, ref)}, out, ref


def _lobster(shape, co2, seed=7):
    Nx, Ny, Nz = shape
    grid = RectilinearGrid(size=shape, extent=(100.0, 100.0, 2.0 * Nz), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))
    bgc = LOBSTER(grid, carbonate_system=CarbonateSystem(),
                  light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=100.0),
                  gas_exchange=CarbonDioxideGasExchangeBoundaryCondition() if co2 else None,
                  temperature=20 - 0.05 * np.abs(grid.zc), salinity=35.0, scale_negatives=True)
    base = workloads.lobster_3d(Nx, Ny, Nz, nsteps=1, seed=seed)
    return Workload(f"lobster_scaled_{Nx}x{Ny}x{Nz}", bgc, grid, [a.copy() for a in base.tracers], 60.0, 6)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("co2", [True, False])
def test_lobster_step_with_scaling(precision, co2):
    print("case scaling lobster", precision, "co2", co2, flush=True)
    wl = _with_negatives(_lobster((20, 3, 24), co2), ("P", "Z", "sPOM", "NH₄"))
    errs, out, _ = _run(wl, precision, wl.nsteps)
    print("errors", {k: float(v) for k, v in errs.items()}, flush=True)
    assert all(e <= TOL[precision] for e in errs.values()), errs
    # the scaled members are not negative any more (the state after the last
    # stage's update may be, as in the reference, but not by the scaled amount)
    assert np.isfinite(np.array([np.sum(o) for o in out])).all()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_npzd_full_columns_with_scaling(precision):
    print("case scaling npzd", precision, flush=True)
    grid = RectilinearGrid(size=(33, 2, 64), extent=(100.0, 100.0, 128.0), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))
    bgc = NPZD(grid, light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=100.0),
               scale_negatives=True)
    rng = np.random.default_rng(5)
    fields = []
    for lo, hi in ((1.0, 10.0), (-0.02, 0.1), (-0.01, 0.1), (8.0, 20.0), (0.0, 0.05)):
        f = grid.new_field()
        grid.interior(f)[...] = rng.uniform(lo, hi, (64, 2, 33))
        fields.append(f)
    wl = Workload("npzd_scaled", bgc, grid, fields, 60.0, 5)
    errs, _, _ = _run(wl, precision, wl.nsteps)
    print("errors", {k: float(v) for k, v in errs.items()}, flush=True)
    assert all(e <= TOL[precision] for e in errs.values()), errs


def test_reference_case_one_cell():
    """test/test_utils.jl:7-22 on the device: N = 2, P = -1 -> N = 1, P = 0."""
    import torch
    print("case scaling reference", flush=True)
    grid = RectilinearGrid(size=(1, 1, 1), extent=(1, 1, 1), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))
    bgc = NPZD(grid, scale_negatives=True)
    vals = (2.0, -1.0, 0.0, 10.0, 0.0)
    fields = []
    for v in vals:
        f = grid.new_field()
        grid.interior(f)[...] = v
        fields.append(f)
    for precision in ("fp64", "fp32"):
        dev = HIPBiogeochemistry(bgc, precision)
        tr = to_device(fields, dev.dtype)
        dev.step(tr, 1, 1e-10)
        torch.cuda.synchronize()
        out = to_host(tr)
        assert grid.interior(out[0])[0, 0, 0] == pytest.approx(1.0, rel=1e-6)
        assert abs(grid.interior(out[1])[0, 0, 0]) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_update_state_hook_scales_in_place(precision):
    """FP64: the scaled state to 1e-15; FP32: the scaling of the FP32
    state (the oracle on the rounded inputs) to 1e-6, the PAR to 1e-5."""
    import torch
    print("case scaling hook", precision, flush=True)
    wl = _with_negatives(_lobster((12, 3, 20), True), ("P", "Z", "DOM"))
    cfg = wl.bgc.config(abi.OBME_FP64)
    dev = HIPBiogeochemistry(wl.bgc, precision)
    tr = to_device(wl.tracers, dev.dtype)
    start = to_host(tr)
    ref = [a.copy() for a in start]
    O.grid_scale_negatives(cfg, ref)
    g = wl.grid
    PAR_ref = np.zeros_like(wl.tracers[0])
    O.grid_update_state(cfg, ref, PAR_ref)
    PAR = torch.zeros_like(tr[0])
    dev.update_biogeochemical_state(tr, PAR, t=0.0)
    torch.cuda.synchronize()
    out = to_host(tr)
    tol, ptol = (1e-15, 1e-13) if precision == "fp64" else (1e-6, TOL["fp32"])
    for n, (a, b) in enumerate(zip(out, ref)):
        assert rel_l2(g.interior(a), g.interior(b)) <= tol, n
        assert np.array_equal(a[0], start[n][0])        # halos untouched
    assert rel_l2(g.interior(PAR.double().cpu().numpy()), g.interior(PAR_ref)) <= ptol
