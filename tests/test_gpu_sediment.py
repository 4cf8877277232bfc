"""InstantRemineralisation sediment on the GPU against the oracle: the fused
step with the storage advanced in the same pass (two calls, so the
sediment's lazy first tendency and the steps after it are both covered),
with negative scaling and the CO2 exchange compiled in beside it, a closed
bottom (no flux), the update_tendencies! hook (the coupled tendency of the
receiver) and the update-state hook's QuasiAdamsBashforth2 sequence, also
after a fused step (its G⁻ comes from the sweep).  FP64 <= 1e-12 relative
L2 per tracer and for the storage; FP32-mixed <= 1e-5."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, to_device, to_host
from oceanbiome import (LOBSTER, NPZD, CarbonateSystem, CarbonDioxideGasExchangeBoundaryCondition,
                        InstantRemineralisationSediment, RectilinearGrid, TwoBandPhotosyntheticallyActiveRadiation, abi)
from oceanbiome.runtime import HIPBiogeochemistry

pytestmark = pytest.mark.gpu
TOL = {"fp64": 1e-12, "fp32": 1e-5}


def _grid(shape, dz=4.0):
    Nx, Ny, Nz = shape
    return RectilinearGrid(size=shape, extent=(10.0 * Nx, 10.0 * Ny, dz * Nz), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))


def _fields(grid, ranges, seed):
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = grid.Nx, grid.Ny, grid.Nz
    out = []
    for lo, hi in ranges:
        f = grid.new_field()
        grid.interior(f)[...] = rng.uniform(lo, hi, (Nz, Ny, Nx))
        out.append(f)
    return out


NPZD_RANGES = ((1.0, 10.0), (0.05, 0.5), (0.05, 0.3), (8.0, 20.0), (0.5, 3.0))
LOBSTER_RANGES = ((5.0, 12.0), (0.1, 1.0), (0.05, 0.5), (0.05, 0.3), (0.5, 3.0), (0.5, 3.0), (0.1, 1.0),
                  (2000.0, 2200.0), (2300.0, 2400.0))


def _npzd(shape, scale=False, open_bottom=True, seed=21):
    grid = _grid(shape)
    bgc = NPZD(grid, light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=80.0),
               sediment_model=InstantRemineralisationSediment(grid), scale_negatives=scale)
    bgc.detritus.open_bottom = open_bottom
    return bgc, grid, _fields(grid, NPZD_RANGES, seed)


def _lobster(shape, co2=True, scale=True, seed=22):
    grid = _grid(shape)
    sed = InstantRemineralisationSediment(sinking_tracers=("sPOM", "bPOM"), remineralisation_reciever="NH₄")
    bgc = LOBSTER(grid, carbonate_system=CarbonateSystem(),
                  light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=80.0),
                  gas_exchange=CarbonDioxideGasExchangeBoundaryCondition() if co2 else None,
                  temperature=20 - 0.05 * np.abs(grid.zc), salinity=35.0, scale_negatives=scale,
                  sediment_model=sed)
    return bgc, grid, _fields(grid, LOBSTER_RANGES, seed)


def _step_and_compare(bgc, grid, fields, precision, nsteps, dt, calls=2):
    import torch
    cfg = bgc.config(abi.OBME_FP64)
    ref = [a.copy() for a in fields]
    state = O.SedimentState(grid.Nx, grid.Ny)
    dev = HIPBiogeochemistry(bgc, precision)
    tr = to_device(fields, dev.dtype)
    for c in range(calls):
        t0 = c * nsteps * dt
        O.grid_run_sediment(cfg, ref, state, nsteps, dt, t0=t0)
        dev.step(tr, nsteps, dt, t0=t0)
    torch.cuda.synchronize()
    out = to_host(tr)
    errs = {n: rel_l2(grid.interior(a), grid.interior(b))
            for n, a, b in zip(bgc.required_biogeochemical_tracers(), out, ref)}
    errs["storage"] = rel_l2(dev.sediment_storage.cpu().numpy(), state.storage)
    print("errors", {k: float(v) for k, v in errs.items()}, flush=True)
    return errs, dev, state


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(20, 3, 50), (17, 2, 64), (9, 2, 7), (5, 2, 1)])
def test_npzd_step_with_sediment(precision, shape):
    print("case sediment npzd", precision, shape, flush=True)
    bgc, grid, fields = _npzd(shape)
    errs, dev, state = _step_and_compare(bgc, grid, fields, precision, 4, 600.0)
    assert all(e <= TOL[precision] for e in errs.values()), errs
    assert np.all(state.storage > 0)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("co2", [True, False])
def test_lobster_step_with_sediment_scaling_co2(precision, co2):
    print("case sediment lobster", precision, "co2", co2, flush=True)
    bgc, grid, fields = _lobster((12, 3, 24), co2=co2, scale=True)
    errs, _, state = _step_and_compare(bgc, grid, fields, precision, 3, 300.0)
    assert all(e <= TOL[precision] for e in errs.values()), errs
    assert np.all(state.storage > 0)


def test_sediment_without_scaling_lobster_fp64():
    print("case sediment lobster no scaling", flush=True)
    bgc, grid, fields = _lobster((10, 2, 16), co2=True, scale=False)
    errs, _, _ = _step_and_compare(bgc, grid, fields, "fp64", 3, 300.0)
    assert all(e <= TOL["fp64"] for e in errs.values()), errs


def test_closed_bottom_buries_nothing():
    print("case sediment closed bottom", flush=True)
    bgc, grid, fields = _npzd((8, 2, 12), open_bottom=False)
    errs, dev, state = _step_and_compare(bgc, grid, fields, "fp64", 3, 600.0)
    assert all(e <= TOL["fp64"] for k, e in errs.items() if k != "storage"), errs
    assert np.all(dev.sediment_storage.cpu().numpy() == 0) and np.all(state.storage == 0)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_add_tendencies_with_sediment(precision):
    import torch
    print("case sediment tendencies", precision, flush=True)
    bgc, grid, fields = _npzd((11, 3, 20))
    cfg = bgc.config(abi.OBME_FP64)
    G_ref = [np.zeros_like(a) for a in fields]
    O.grid_tendencies(cfg, fields, G_ref)
    dev = HIPBiogeochemistry(bgc, precision)
    tr = to_device(fields, dev.dtype)
    G = [torch.zeros_like(t) for t in tr]
    dev.update_tendencies(tr, G)
    torch.cuda.synchronize()
    out = to_host(G)
    for n, name in enumerate(bgc.required_biogeochemical_tracers()):
        if name == "T":
            continue
        assert rel_l2(grid.interior(out[n]), grid.interior(G_ref[n])) <= TOL[precision], name
    # the receiver's bottom cells carry the coupled tendency
    assert rel_l2(grid.interior(out[0])[0], grid.interior(G_ref[0])[0]) <= TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_update_state_hook_sequence(precision):
    import torch
    print("case sediment hook", precision, flush=True)
    bgc, grid, fields = _npzd((7, 3, 10), seed=5)
    cfg = bgc.config(abi.OBME_FP64)
    ref = [a.copy() for a in fields]
    state = O.SedimentState(grid.Nx, grid.Ny)
    dev = HIPBiogeochemistry(bgc, precision)
    tr = to_device(fields, dev.dtype)
    tol = TOL[precision]
    for c, dt in enumerate((math.inf, 10.0, 10.0, 10.0, 4.0)):
        if c == 3:
            # a new state between the calls
            for a in ref:
                grid.interior(a)[...] *= 1.5
            tr = to_device(ref, dev.dtype)
        O.grid_sediment_update_state(cfg, ref, state, dt)
        dev.update_sediment_state(tr, dt)
        torch.cuda.synchronize()
        got = dev.sediment_storage.cpu().numpy()
        if c == 0:
            assert np.all(got == 0)
            continue
        assert rel_l2(got, state.storage) <= tol, (c, rel_l2(got, state.storage))


def test_fused_step_then_hook_takes_an_ab2_step():
    import torch
    print("case sediment fused then hook", flush=True)
    bgc, grid, fields = _npzd((9, 2, 12), seed=9)
    cfg = bgc.config(abi.OBME_FP64)
    ref = [a.copy() for a in fields]
    state = O.SedimentState(grid.Nx, grid.Ny)
    dev = HIPBiogeochemistry(bgc, "fp64")
    tr = to_device(fields, torch.float64)
    dt = 600.0
    O.grid_run_sediment(cfg, ref, state, 2, dt)
    dev.step(tr, 2, dt)
    last = (3.0 / 4.0 + -5.0 / 12.0) * dt       # the last stage's Δt: the hook repeats it
    assert state.last_dt == last
    O.grid_sediment_update_state(cfg, ref, state, last)
    dev.update_sediment_state(tr, last)
    torch.cuda.synchronize()
    assert rel_l2(dev.sediment_storage.cpu().numpy(), state.storage) <= 1e-12
