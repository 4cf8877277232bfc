"""The benchmark configurations at their full sizes on the GPU, through
properties that do not need the whole grid on the CPU: columns are
independent, so a sample of columns stepped by the oracle is an exact check
of those columns; the halos and the read-only fields stay untouched; every
written value stays finite.  LOBSTER + DIC/Alk with the CO2 exchange on
512x512x64 (the headline workload) in both precisions, the same with the
O2 tracer and its gas exchange in FP32-mixed, and PISCES on 256x256x64
for its 100 steps in both precisions, and the carbonate chemistry on all 2^24
surface points of its configuration in both modes.  Bars as everywhere: FP64 <= 1e-12
relative L2 per tracer, FP32-mixed <= 1e-5 (PISCES around the values of
validation/PISCES/column.jl: or three times the FP32 storage floor, the
deviation the FP64 oracle shows when its state is rounded to FP32 after
every step, where that is larger: the reference's iron uptake has a pole
near those iron quotas); the surface pH of the sampled LOBSTER columns within
1e-6 (FP32-mixed) and 1e-12 (FP64) of the oracle's, and their CO2 flux
within 1e-5 and 1e-10 relative L2."""
import numpy as np
import pytest

import oracle as O
import pisces as OP
from helpers import fp32_storage_floor
from oceanbiome import abi, workloads
from oceanbiome.runtime import HIPBiogeochemistry, HIPPISCES

pytestmark = pytest.mark.gpu
TOL = {"fp64": 1e-12, "fp32": 1e-5}
PH_TOL = {"fp64": 1e-12, "fp32": 1e-6}
FLUX_TOL = {"fp64": 1e-10, "fp32": 1e-5}


def _sample(g, n, seed):
    rng = np.random.default_rng(seed)
    flat = rng.choice(g.Nx * g.Ny, n, replace=False)
    return (flat % g.Nx).astype(np.int32), (flat // g.Nx).astype(np.int32)


def _columns(g, arr, ci, cj):
    ks = np.arange(g.Nz) + g.Hz
    return arr[ks[:, None], (cj + g.Hy)[None, :], (ci + g.Hx)[None, :]]


@pytest.mark.parametrize("precision,oxygen", [("fp32", False), ("fp64", False), ("fp32", True)])
def test_lobster_512x512x64_column_sample(precision, oxygen):
    import torch
    print("case full-size lobster", precision, "oxygen", oxygen, flush=True)
    wl = workloads.lobster_3d(512, 512, 64, nsteps=6, oxygen=oxygen)
    g = wl.grid
    ci, cj = _sample(g, 96, 5)
    ref = [a.copy() for a in wl.tracers]
    O.grid_run_columns(wl.bgc.config(abi.OBME_FP64), ref, ci, cj, wl.nsteps, wl.dt)
    dev = HIPBiogeochemistry(wl.bgc, precision)
    tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
    dev.step(tr, wl.nsteps, wl.dt)
    torch.cuda.synchronize()
    names = wl.bgc.required_biogeochemical_tracers()
    errs = {}
    top = {}
    for n, name in enumerate(names):
        out = tr[n].double().cpu().numpy()
        a, b = _columns(g, out, ci, cj), _columns(g, ref[n], ci, cj)
        top[n] = (a[-1].copy(), b[-1].copy())
        errs[name] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        assert np.isfinite(g.interior(out)).all(), name
        # halos untouched (the bottom halo plane and the first halo row)
        np.testing.assert_array_equal(out[0], wl.tracers[n][0].astype(out.dtype) if precision == "fp64"
                                      else wl.tracers[n][0].astype(np.float32).astype(np.float64))
        del out
    print("errors", errs, flush=True)
    assert all(e <= TOL[precision] for e in errs.values()), errs
    # the surface pH and CO2 flux (carbon_chemistry.jl:168, gas_exchange.jl:26-38)
    # of the GPU state and of the oracle's, both by the oracle's point function
    cfg = wl.bgc.config(abi.OBME_FP64)
    T = np.full(len(ci), float(wl.extra["T_profile"][-1]))
    S = np.full(len(ci), float(wl.extra["S"]))
    u10 = np.full(len(ci), cfg.co2.wind_speed)
    pH_g, F_g = O.co2_flux_batch(cfg.co2, cfg.solver, top[7][0], top[8][0], T, S, u10, cfg.co2.air_concentration)
    pH_r, F_r = O.co2_flux_batch(cfg.co2, cfg.solver, top[7][1], top[8][1], T, S, u10, cfg.co2.air_concentration)
    dpH = float(np.max(np.abs(pH_g - pH_r)))
    dF = float(np.linalg.norm(F_g - F_r) / np.linalg.norm(F_r))
    print("surface pH max abs", dpH, "CO2 flux rel L2", dF, flush=True)
    assert dpH <= PH_TOL[precision] and dF <= FLUX_TOL[precision], (dpH, dF)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_pisces_256x256x64_column_sample(precision):
    """The PISCES workload's full length: 100 steps of 300 s."""
    import torch
    print("case full-size pisces", precision, flush=True)
    wl = workloads.pisces_3d(256, 256, 64, nsteps=100, dt=300.0)
    g = wl.grid
    m = wl.bgc
    ci, cj = _sample(g, 48, 7)
    q = OP.params_from_vector(m.param_vector())
    light = dict(surface_PAR=m.surface_PAR, kw=m.kw, e=m.e, chi=m.chi, division=m.division)
    F = {n: _columns(g, a, ci, cj)[:, None, :].copy() for n, a in zip(OP.TRACERS, wl.tracers)}
    cc = m.carbon_chemistry
    F0 = {k: v.copy() for k, v in F.items()}
    args = (np.asarray(g.zc), np.asarray(g.zf), np.asarray(g.dz))
    solver = cc.solver.struct(cc.initial_pH_guess)
    with np.errstate(all="ignore"):
        OP.step(q, light, F, *args, wl.nsteps, wl.dt, 0.0, solver)
    dev = HIPPISCES(m, precision)
    tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
    dev.step(tr, wl.nsteps, wl.dt)
    torch.cuda.synchronize()
    errs = {}
    for n, name in enumerate(OP.WRITTEN):
        out = tr[n].double().cpu().numpy()
        a = _columns(g, out, ci, cj)
        b = F[name][:, 0, :]
        errs[name] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        assert np.isfinite(g.interior(out)).all(), name
    print("errors", errs, flush=True)
    if precision == "fp64":
        bars = {k: TOL["fp64"] for k in OP.WRITTEN}
    else:
        # 1e-5, or three times the FP32 storage floor where the reference's
        # iron uptake pole makes that larger
        floor = fp32_storage_floor(lambda X, s: OP.step(q, light, X, *args, 1, wl.dt, s * wl.dt, solver),
                                   wl.nsteps, F0, F, OP.WRITTEN)
        print("storage floor", floor, flush=True)
        bars = {k: max(TOL["fp32"], 3 * floor[k]) for k in OP.WRITTEN}
    assert all(errs[k] <= bars[k] for k in OP.WRITTEN), (errs, bars)
    # T and S are read only
    dt_np = np.float32 if precision == "fp32" else np.float64
    np.testing.assert_array_equal(tr[24].cpu().numpy(), wl.tracers[24].astype(dt_np))


@pytest.mark.parametrize("mode", [0, 1])
def test_carbonate_points_2pow24(mode):
    """pH and CO2 flux on the 2^24 points of the carbonate configuration:
    mode 0 (the reference damped Newton) and mode 1 (FP32 guess refined by
    FP64 Newton) against the oracle's reference solve on every point: pH
    within 1e-12 absolute, the flux within 1e-10 relative L2."""
    import os
    import torch
    print("case carbonate 2^24 mode", mode, flush=True)
    pts = workloads.carbonate_points()
    wl = workloads.lobster_3d(4, 4, 4, nsteps=1)
    dev = HIPBiogeochemistry(wl.bgc, "fp64")
    cfg = wl.bgc.config(abi.OBME_FP64)
    nthreads = max(1, min(16, os.cpu_count() or 1))
    pH_ref, F_ref = O.co2_flux_batch(cfg.co2, cfg.solver, pts["DIC"], pts["Alk"], pts["T"], pts["S"],
                                     pts["u10"], pts["pCO2_air"], nthreads=nthreads)
    args = [torch.from_numpy(pts[k]).cuda() for k in ("DIC", "Alk", "T", "S", "u10")]
    pH, F = dev.co2_flux(*args, pCO2_air=pts["pCO2_air"], mode=mode)
    torch.cuda.synchronize()
    pH = pH.cpu().numpy()
    F = F.cpu().numpy()
    assert np.all(np.isfinite(pH)) and np.all(np.isfinite(F))
    dpH = float(np.max(np.abs(pH - pH_ref)))
    dF = float(np.linalg.norm(F - F_ref) / np.linalg.norm(F_ref))
    print("pH max abs", dpH, "flux rel L2", dF, flush=True)
    assert dpH <= 1e-12
    assert dF <= 1e-10
