"""PISCES surface PAR on the GPU against the oracle: default_surface_PAR
(Light.jl:47, the PISCES default) at the clock of each stage, and a 2-D
PAR⁰ array, through the update-state hook and the RK3 step (the per-stage
kernels in FP64, the fused step in FP32-mixed).  FP64 <= 1e-13 for the light
fields and 1e-12 per tracer; FP32-mixed <= 1e-5.  Inputs around the values
of the reference's PISCES tests (test/test_PISCES.jl:6-16).

At night the reference's light limitation divides 0 by the day length
evaluated with its arguments swapped (growth_rate.jl:30), which is zero for
some clocks: NaN there is the reference's result, and the device has to
give NaN in exactly the same cells."""
import numpy as np
import pytest

import pisces as OP
from helpers import rel_l2
from oceanbiome import workloads
from oceanbiome.runtime import HIPPISCES

pytestmark = pytest.mark.gpu
TOL = {"fp64": 1e-12, "fp32": 1e-5}
LTOL = {"fp64": 1e-13, "fp32": 1e-5}


def _case(surface, Nx=8, Ny=4, Nz=20, seed=7):
    wl = workloads.pisces_3d(Nx, Ny, Nz, nsteps=3, dt=600.0, seed=seed, surface_PAR=surface, medians="test")
    g = wl.grid
    m = wl.bgc
    q = OP.params_from_vector(m.param_vector())
    light = dict(surface_PAR=m.surface_PAR, kw=m.kw, e=m.e, chi=m.chi, division=m.division)
    F = {n: g.interior(a).copy() for n, a in zip(OP.TRACERS, wl.tracers)}
    cc = m.carbon_chemistry
    solver = cc.solver.struct(cc.initial_pH_guess)
    return wl, q, light, F, solver


def _surface_array(Ny, Nx, seed=2):
    rng = np.random.default_rng(seed)
    a = rng.uniform(20.0, 200.0, (Ny, Nx))
    a[0, 1] = 0.0
    return a


SURFACES = [("diurnal morning", "default", 4 * 3600.0),
            ("diurnal night", "default", 12 * 3600.0),
            ("array", "array", 0.0)]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name,surface,t", SURFACES, ids=[s[0] for s in SURFACES])
def test_update_state_surface_modes(precision, name, surface, t):
    import torch
    print("case pisces light update_state", name, precision, flush=True)
    surf = _surface_array(4, 8) if surface == "array" else surface
    wl, q, light, F, solver = _case(surf)
    g = wl.grid
    A = OP.update_state(q, light, F, np.asarray(g.zc), np.asarray(g.zf), t, solver)
    dev = HIPPISCES(wl.bgc, precision)
    tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
    PAR, bands, Om, zeu, ml = dev.update_biogeochemical_state(tr, t)
    torch.cuda.synchronize()
    tol = LTOL[precision]
    got = g.interior(PAR.double().cpu().numpy())
    if not np.any(A["PAR"]):
        assert not np.any(got)
    else:
        assert rel_l2(got, A["PAR"]) <= tol
    for n in range(3):
        b = g.interior(bands[n].double().cpu().numpy())
        assert (not np.any(b)) if not np.any(A[f"PAR{n + 1}"]) else rel_l2(b, A[f"PAR{n + 1}"]) <= tol
    assert rel_l2(zeu.double().cpu().numpy(), A["zeu"][0]) <= tol
    mlr = A["mlPAR"][0]
    mlg = ml.double().cpu().numpy()
    assert (not np.any(mlg)) if not np.any(mlr) else rel_l2(mlg, mlr) <= tol


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name,surface,t0", [("diurnal dusk", "default", 6 * 3600.0 - 1500.0),
                                             ("diurnal morning", "default", 4 * 3600.0),
                                             ("array", "array", 0.0)],
                         ids=["diurnal dusk", "diurnal morning", "array"])
def test_step_surface_modes(precision, name, surface, t0):
    """Three steps of 600 s; from 6 h − 1500 s the stage clocks cross dusk
    (PAR⁰ = 0 from 6 h on)."""
    import torch
    print("case pisces light step", name, precision, flush=True)
    surf = _surface_array(4, 8) if surface == "array" else surface
    wl, q, light, F, solver = _case(surf)
    g = wl.grid
    with np.errstate(all="ignore"):
        OP.step(q, light, F, np.asarray(g.zc), np.asarray(g.zf), np.asarray(g.dz), wl.nsteps, wl.dt, t0, solver)
    dev = HIPPISCES(wl.bgc, precision)
    tr = [torch.from_numpy(a).to(device="cuda", dtype=dev.dtype) for a in wl.tracers]
    dev.step(tr, wl.nsteps, wl.dt, t0=t0)
    torch.cuda.synchronize()
    errs = {}
    for n, nm in enumerate(OP.WRITTEN):
        a = g.interior(tr[n].double().cpu().numpy())
        b = F[nm]
        assert np.array_equal(np.isnan(a), np.isnan(b)), nm
        ok = ~np.isnan(b)
        errs[nm] = rel_l2(a[ok], b[ok]) if ok.any() else 0.0
    print("errors", errs, "nan cells", int(np.isnan(F["P"]).sum()), flush=True)
    assert all(e <= TOL[precision] for e in errs.values()), errs
