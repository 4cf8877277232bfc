"""GPU parity: the HIP library through its C ABI against the FP64 oracle on
the same inputs.  FP64 bar: relative L2 <= 1e-12 per tracer (reassociation
only).  FP32-mixed bar: relative L2 <= 1e-5 per tracer."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, to_device, to_host
from oceanbiome import abi, workloads
from oceanbiome.runtime import HIPBiogeochemistry

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def _run_both(wl, precision, nsteps=None):
    import torch
    nsteps = wl.nsteps if nsteps is None else nsteps
    ref = [a.copy() for a in wl.tracers]
    cfg = wl.bgc.config(abi.OBME_FP64)
    O.grid_run(cfg, ref, nsteps, wl.dt)
    dev = HIPBiogeochemistry(wl.bgc, precision)
    tr = to_device(wl.tracers, dev.dtype)
    dev.step(tr, nsteps, wl.dt)
    torch.cuda.synchronize()
    out = to_host(tr)
    g = wl.grid
    errs = {}
    for name, a, b in zip(wl.bgc.required_biogeochemical_tracers(), out, ref):
        errs[name] = rel_l2(g.interior(a), g.interior(b))
    return errs


def test_npzd_2d_fp64():
    print("case npzd_2d fp64", flush=True)
    errs = _run_both(workloads.npzd_2d(), "fp64")
    assert all(e <= FP64_TOL for e in errs.values()), errs


def test_npzd_2d_fp32_mixed():
    print("case npzd_2d fp32", flush=True)
    errs = _run_both(workloads.npzd_2d(), "fp32")
    print("errors", {k: float(v) for k, v in errs.items()}, flush=True)
    assert all(e <= FP32_TOL for e in errs.values()), errs


@pytest.mark.parametrize("shape", [(32, 8, 64), (45, 3, 37), (20, 2, 7)])
def test_lobster_fp64(shape):
    print("case lobster fp64", shape, flush=True)
    wl = workloads.lobster_3d(*shape, nsteps=10)
    errs = _run_both(wl, "fp64")
    assert all(e <= FP64_TOL for e in errs.values()), errs


@pytest.mark.parametrize("shape", [(32, 8, 64), (45, 3, 37)])
def test_lobster_fp32_mixed(shape):
    print("case lobster fp32", shape, flush=True)
    wl = workloads.lobster_3d(*shape, nsteps=20)
    errs = _run_both(wl, "fp32")
    print("errors", {k: float(v) for k, v in errs.items()}, flush=True)
    assert all(e <= FP32_TOL for e in errs.values()), errs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_diurnal_surface_PAR(precision):
    """default_surface_PAR (Light/Light.jl:219) at the three stage clocks."""
    print("case diurnal", precision, flush=True)
    wl = workloads.lobster_3d(24, 3, 40, nsteps=12, dt=1800.0, surface_PAR="default")
    errs = _run_both(wl, precision)
    tol = FP64_TOL if precision == "fp64" else FP32_TOL
    assert all(e <= tol for e in errs.values()), errs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_surface_PAR_array(precision):
    """A 2-D surface PAR field PAR0[i, j]."""
    import torch
    print("case surface array", precision, flush=True)
    wl = workloads.lobster_3d(30, 5, 33, nsteps=6, surface_PAR=np.zeros((5, 30)))
    g = wl.grid
    PAR0 = np.random.default_rng(11).uniform(20.0, 180.0, (g.Ny, g.Nx))
    ref = [a.copy() for a in wl.tracers]
    O.grid_run(wl.bgc.config(abi.OBME_FP64), ref, wl.nsteps, wl.dt, PAR0=np.ascontiguousarray(PAR0.ravel()))
    dev = HIPBiogeochemistry(wl.bgc, precision, surface_array=torch.from_numpy(PAR0.ravel()))
    tr = to_device(wl.tracers, dev.dtype)
    dev.step(tr, wl.nsteps, wl.dt)
    torch.cuda.synchronize()
    tol = FP64_TOL if precision == "fp64" else FP32_TOL
    for name, a, b in zip(wl.bgc.required_biogeochemical_tracers(), to_host(tr), ref):
        assert rel_l2(g.interior(a), g.interior(b)) <= tol, name


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_stretched_grid_closed_bottom(precision):
    """Non-uniform cell thickness and a closed bottom (w damped to 0 at the
    bottom face): the per-level geometry and face-velocity tables."""
    import torch
    from oceanbiome.grid import RectilinearGrid
    from oceanbiome.models import LOBSTER, CarbonateSystem, TwoParticleAndDissolved, \
        TwoBandPhotosyntheticallyActiveRadiation, CarbonDioxideGasExchangeBoundaryCondition
    print("case stretched closed", precision, flush=True)
    Nz = 30
    faces = -200.0 * (1 - np.linspace(0.0, 1.0, Nz + 1) ** 1.5)
    faces[-1] = 0.0
    g = RectilinearGrid(size=(20, 4, Nz), x=(0, 100), y=(0, 20), z=faces)
    bgc = LOBSTER(g, carbonate_system=CarbonateSystem(), detritus=TwoParticleAndDissolved(open_bottom=False),
                  light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=120.0),
                  gas_exchange=CarbonDioxideGasExchangeBoundaryCondition(), temperature=15.0, salinity=34.0)
    base = workloads.lobster_3d(20, 4, Nz, nsteps=1, seed=5)
    wl = workloads.Workload("stretched", bgc, g, [a.copy() for a in base.tracers], 120.0, 8)
    errs = _run_both(wl, precision)
    tol = FP64_TOL if precision == "fp64" else FP32_TOL
    assert all(e <= tol for e in errs.values()), errs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_update_state_par(precision):
    import torch
    print("case update_state", precision, flush=True)
    wl = workloads.lobster_3d(24, 4, 50, nsteps=1)
    g = wl.grid
    PAR_ref = g.new_field()
    O.grid_update_state(wl.bgc.config(abi.OBME_FP64), wl.tracers, PAR_ref, 0.0)
    dev = HIPBiogeochemistry(wl.bgc, precision)
    tr = to_device(wl.tracers, dev.dtype)
    PAR = torch.zeros(g.parent_shape, dtype=dev.dtype, device="cuda")
    dev.update_biogeochemical_state(tr, PAR, 0.0)
    torch.cuda.synchronize()
    tol = 1e-13 if precision == "fp64" else FP32_TOL
    assert rel_l2(g.interior(PAR.double().cpu().numpy()), g.interior(PAR_ref)) <= tol


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_add_tendencies(precision):
    """The tendencies accumulate into G (FP32: G holds FP32 values, the
    comparison subtracts G's rounded start)."""
    import torch
    print("case add_tendencies", precision, flush=True)
    wl = workloads.lobster_3d(24, 4, 50, nsteps=1)
    g = wl.grid
    rng = np.random.default_rng(3)
    G0 = [rng.standard_normal(g.parent_shape) * 1e-6 for _ in wl.tracers]
    Gref = [a.copy() for a in G0]
    PAR_ref = g.new_field()
    O.grid_tendencies(wl.bgc.config(abi.OBME_FP64), wl.tracers, Gref, 0.0, PAR=PAR_ref)
    dev = HIPBiogeochemistry(wl.bgc, precision)
    tr = to_device(wl.tracers, dev.dtype)
    G = to_device(G0, dev.dtype)
    G0r = to_host(G)
    PAR = torch.zeros(g.parent_shape, dtype=dev.dtype, device="cuda")
    dev.update_tendencies(tr, G, PAR, t=0.0)
    torch.cuda.synchronize()
    tol, ptol = (1e-12, 1e-13) if precision == "fp64" else (FP32_TOL, FP32_TOL)
    for name, a, b, a0, a0r in zip(wl.bgc.required_biogeochemical_tracers(), to_host(G), Gref, G0, G0r):
        assert rel_l2(g.interior(a - a0r), g.interior(b - a0)) <= tol, name
    assert rel_l2(g.interior(PAR.double().cpu().numpy()), g.interior(PAR_ref)) <= ptol


@pytest.mark.parametrize("mode", [0, 1])
def test_co2_flux_points(mode):
    import torch
    print("case co2_flux mode", mode, flush=True)
    pts = workloads.carbonate_points(4096, seed=7)
    wl = workloads.lobster_3d(4, 4, 4, nsteps=1)
    dev = HIPBiogeochemistry(wl.bgc, "fp64")
    cfg = wl.bgc.config(abi.OBME_FP64)
    pH_ref, F_ref = O.co2_flux_batch(cfg.co2, cfg.solver, pts["DIC"], pts["Alk"], pts["T"], pts["S"],
                                     pts["u10"], pts["pCO2_air"])
    args = [torch.from_numpy(pts[k]).cuda() for k in ("DIC", "Alk", "T", "S", "u10")]
    pH, F = dev.co2_flux(*args, pCO2_air=pts["pCO2_air"], mode=mode)
    torch.cuda.synchronize()
    pH = pH.cpu().numpy()
    F = F.cpu().numpy()
    assert np.max(np.abs(pH - pH_ref)) <= 1e-12
    assert rel_l2(F, F_ref) <= 1e-10


SWEEP_WAVES = ["1", "2", "4", "8"]


@pytest.mark.parametrize("waves", SWEEP_WAVES)
def test_sweep_waves_fp64(waves, monkeypatch):
    """Every workgroup size of the sweep gives the same answer (FP64 parity)."""
    print("case sweep waves fp64", waves, flush=True)
    monkeypatch.setenv("OBME_SWEEP_WAVES", waves)
    wl = workloads.lobster_3d(40, 3, 29, nsteps=3)
    errs = _run_both(wl, "fp64")
    assert all(e <= FP64_TOL for e in errs.values()), errs


@pytest.mark.parametrize("waves", SWEEP_WAVES)
def test_sweep_waves_fp32(waves, monkeypatch):
    print("case sweep waves fp32", waves, flush=True)
    monkeypatch.setenv("OBME_SWEEP_WAVES", waves)
    wl = workloads.lobster_3d(40, 3, 29, nsteps=5)
    errs = _run_both(wl, "fp32")
    print("errors", {k: float(v) for k, v in errs.items()}, flush=True)
    assert all(e <= FP32_TOL for e in errs.values()), errs
