"""The SimpleMultiG sediment (Models/Sediments/simple_multi_G.jl, nitrogen
form) on the GPU against the oracle: LOBSTER + O2 with the fields Ns, Nf, Nr
advanced in the fused step (two calls: the sediment's lazy first step and
the steps after it), with the CO2 exchange beside it; the update-state
hook's QuasiAdamsBashforth2 sequence and the update_tendencies! hook's
coupled NH4, NO3 and O2 from the fields; and the reference's own
integration test (test/test_sediments.jl:36-75) on the GPU: total nitrogen
kept to 2e-7.  FP64 <= 1e-12 relative L2 per tracer and field; FP32-mixed
<= 1e-5."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, to_device, to_host
from oceanbiome import (LOBSTER, CarbonateSystem, CarbonDioxideGasExchangeBoundaryCondition, Oxygen,
                        RectilinearGrid, SimpleMultiGSediment, TwoBandPhotosyntheticallyActiveRadiation, abi)
from oceanbiome.runtime import HIPBiogeochemistry

pytestmark = pytest.mark.gpu
TOL = {"fp64": 1e-12, "fp32": 1e-5}

RANGES = {"NO₃": (5.0, 12.0), "NH₄": (0.1, 1.0), "P": (0.05, 0.5), "Z": (0.05, 0.3), "sPOM": (0.5, 3.0),
          "bPOM": (0.5, 3.0), "DOM": (0.1, 1.0), "DIC": (2000.0, 2200.0), "Alk": (2300.0, 2400.0),
          "O₂": (150.0, 320.0)}


def _lobster(shape, co2=False, seed=31, depth=None):
    Nx, Ny, Nz = shape
    grid = RectilinearGrid(size=shape, extent=(10.0 * Nx, 10.0 * Ny, depth or 4.0 * Nz), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))
    bgc = LOBSTER(grid, carbonate_system=CarbonateSystem() if co2 else None, oxygen=Oxygen(),
                  light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=80.0),
                  gas_exchange=CarbonDioxideGasExchangeBoundaryCondition() if co2 else None,
                  temperature=20 - 0.05 * np.abs(grid.zc), salinity=35.0,
                  sediment_model=SimpleMultiGSediment(grid))
    rng = np.random.default_rng(seed)
    fields = []
    for n in bgc.required_biogeochemical_tracers():
        lo, hi = RANGES[n]
        f = grid.new_field()
        grid.interior(f)[...] = rng.uniform(lo, hi, (Nz, Ny, Nx))
        fields.append(f)
    return bgc, grid, fields


def _errors(bgc, grid, out, ref, dev, state):
    errs = {n: rel_l2(grid.interior(a), grid.interior(b))
            for n, a, b in zip(bgc.required_biogeochemical_tracers(), out, ref)}
    got = dev.sediment_storage.cpu().numpy()
    for q, name in enumerate(("Ns", "Nf", "Nr")):
        errs[name] = rel_l2(got[q], state.storage[q])
    print("errors", {k: float(v) for k, v in errs.items()}, flush=True)
    return errs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape,co2", [((20, 3, 40), False), ((9, 2, 7), True), ((6, 2, 1), False)])
def test_step_with_simple_multi_g(precision, shape, co2):
    import torch
    print("case smg step", precision, shape, "co2", co2, flush=True)
    bgc, grid, fields = _lobster(shape, co2=co2)
    cfg = bgc.config(abi.OBME_FP64)
    ref = [a.copy() for a in fields]
    state = O.SedimentState(grid.Nx, grid.Ny, fields=3)
    dev = HIPBiogeochemistry(bgc, precision)
    tr = to_device(fields, dev.dtype)
    nsteps, dt = 6, 600.0
    for c in range(2):
        t0 = c * nsteps * dt
        O.grid_run_sediment(cfg, ref, state, nsteps, dt, t0=t0)
        dev.step(tr, nsteps, dt, t0=t0)
    torch.cuda.synchronize()
    errs = _errors(bgc, grid, to_host(tr), ref, dev, state)
    assert all(e <= TOL[precision] for e in errs.values()), errs
    assert np.all(state.storage != 0)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_hooks_with_simple_multi_g(precision):
    import torch
    print("case smg hooks", precision, flush=True)
    bgc, grid, fields = _lobster((7, 3, 12), seed=5)
    cfg = bgc.config(abi.OBME_FP64)
    ref = [a.copy() for a in fields]
    state = O.SedimentState(grid.Nx, grid.Ny, fields=3)
    dev = HIPBiogeochemistry(bgc, precision)
    tr = to_device(fields, dev.dtype)
    tol = TOL[precision]
    for c, dt in enumerate((math.inf, 100.0, 100.0, 100.0, 40.0)):
        O.grid_sediment_update_state(cfg, ref, state, dt)
        dev.update_sediment_state(tr, dt)
        torch.cuda.synchronize()
        got = dev.sediment_storage.cpu().numpy()
        if c == 0:
            assert np.all(got == 0)
            continue
        for q in range(3):
            assert rel_l2(got[q], state.storage[q]) <= tol, (c, q, rel_l2(got[q], state.storage[q]))
    # the coupled tendencies from the fields now
    G_ref = [np.zeros_like(a) for a in fields]
    O.grid_tendencies(cfg, ref, G_ref, sediment_fields=state.storage)
    G = [torch.zeros_like(t) for t in tr]
    dev.update_tendencies(tr, G)
    torch.cuda.synchronize()
    out = to_host(G)
    names = bgc.required_biogeochemical_tracers()
    for n, name in enumerate(names):
        assert rel_l2(grid.interior(out[n]), grid.interior(G_ref[n])) <= tol, name
    for name in ("NH₄", "NO₃", "O₂"):
        n = names.index(name)
        bottom = rel_l2(grid.interior(out[n])[0], grid.interior(G_ref[n])[0])
        assert bottom <= tol, (name, bottom)


def test_reference_integration_conserves_nitrogen_fp64():
    import torch
    print("case smg nitrogen", flush=True)
    grid = RectilinearGrid(size=(3, 3, 50), extent=(10.0, 10.0, 500.0), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))
    bgc = LOBSTER(grid, oxygen=Oxygen(), sediment_model=SimpleMultiGSediment(grid))
    names = list(bgc.required_biogeochemical_tracers())
    init = {"NO₃": 10.0, "NH₄": 1.0, "sPOM": 1.0, "bPOM": 1.0, "O₂": 1000.0}
    fields = []
    for n in names:
        f = grid.new_field()
        grid.interior(f)[...] = init.get(n, 0.0)
        fields.append(f)
    dev = HIPBiogeochemistry(bgc, "fp64")
    tr = to_device(fields, dev.dtype)
    dz = np.diff(grid.zf)

    def total(arrs, sed):
        vol = sum(float(np.sum(grid.interior(arrs[names.index(n)]) * dz[:, None, None]))
                  for n in ("NO₃", "NH₄", "P", "Z", "sPOM", "bPOM", "DOM"))
        return vol + float(np.sum(sed))

    before = total(fields, np.zeros(3))
    dev.step(tr, 100, 1.0)
    torch.cuda.synchronize()
    sed = dev.sediment_storage.cpu().numpy()
    after = total(to_host(tr), sed)
    assert abs(after - before) <= 0.2e-6 * abs(before), (before, after)
    assert np.all(sed != 0)
