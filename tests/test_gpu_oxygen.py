"""LOBSTER (and NPZD) with the Oxygen component on the GPU against the oracle: the
fused step with and without the carbonate system, with the CO2 exchange,
negative scaling and the InstantRemineralisation sediment compiled in beside
it, at a ragged depth and at full 64-level columns (the FP32 default-
formulation kernel), with the O2 gas exchange at the top, and the
update_tendencies! hook.  The other tracers
match the same model without O2 (one-way coupling).  FP64 <= 1e-12
relative L2 per tracer; FP32-mixed <= 1e-5."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, to_device, to_host
from oceanbiome import (LOBSTER, CarbonateSystem, CarbonDioxideGasExchangeBoundaryCondition,
                        InstantRemineralisationSediment, Oxygen, OxygenGasExchangeBoundaryCondition, RectilinearGrid,
                        TwoBandPhotosyntheticallyActiveRadiation, abi)
from oceanbiome.runtime import HIPBiogeochemistry

pytestmark = pytest.mark.gpu
TOL = {"fp64": 1e-12, "fp32": 1e-5}
RANGES = ((5.0, 12.0), (0.1, 1.0), (0.05, 0.5), (0.05, 0.3), (0.5, 3.0), (0.5, 3.0), (0.1, 1.0))


def _model(shape, carbonate, co2=False, scale=False, sediment=False, oxygen=True, o2x=False, seed=3):
    Nx, Ny, Nz = shape
    grid = RectilinearGrid(size=shape, extent=(10.0 * Nx, 10.0 * Ny, 4.0 * Nz), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))
    sed = None
    if sediment:
        sed = InstantRemineralisationSediment(sinking_tracers=("sPOM", "bPOM"), remineralisation_reciever="NH₄")
    bgc = LOBSTER(grid, carbonate_system=CarbonateSystem() if carbonate else None,
                  oxygen=Oxygen() if oxygen else None,
                  light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=80.0),
                  gas_exchange=tuple([CarbonDioxideGasExchangeBoundaryCondition()] if co2 else []) +
                  tuple([OxygenGasExchangeBoundaryCondition()] if o2x else []),
                  temperature=20 - 0.05 * np.abs(grid.zc), salinity=35.0, scale_negatives=scale,
                  sediment_model=sed)
    rng = np.random.default_rng(seed)
    ranges = list(RANGES) + ([(2000.0, 2200.0), (2300.0, 2400.0)] if carbonate else [])
    if oxygen:
        ranges.append((150.0, 300.0))
    fields = []
    for lo, hi in ranges:
        f = grid.new_field()
        grid.interior(f)[...] = rng.uniform(lo, hi, (Nz, Ny, Nx))
        fields.append(f)
    return bgc, grid, fields


def _run(bgc, grid, fields, precision, nsteps, dt):
    import torch
    cfg = bgc.config(abi.OBME_FP64)
    ref = [a.copy() for a in fields]
    dev = HIPBiogeochemistry(bgc, precision)
    tr = to_device(fields, dev.dtype)
    storage = None
    if bgc.sediment is not None:
        state = O.SedimentState(grid.Nx, grid.Ny)
        O.grid_run_sediment(cfg, ref, state, nsteps, dt)
        storage = (dev, state)
    else:
        O.grid_run(cfg, ref, nsteps, dt)
    dev.step(tr, nsteps, dt)
    torch.cuda.synchronize()
    out = to_host(tr)
    errs = {n: rel_l2(grid.interior(a), grid.interior(b))
            for n, a, b in zip(bgc.required_biogeochemical_tracers(), out, ref)}
    if storage is not None:
        errs["storage"] = rel_l2(dev.sediment_storage.cpu().numpy(), storage[1].storage)
    print("errors", {k: float(v) for k, v in errs.items()}, flush=True)
    return errs, out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("carbonate", [False, True])
@pytest.mark.parametrize("shape", [(13, 3, 29), (17, 2, 64)])
def test_step_with_oxygen(precision, carbonate, shape):
    print("case oxygen", precision, "carbonate", carbonate, shape, flush=True)
    bgc, grid, fields = _model(shape, carbonate)
    errs, out = _run(bgc, grid, fields, precision, 3, 600.0)
    assert all(e <= TOL[precision] for e in errs.values()), errs
    # O2 moved
    assert not np.array_equal(grid.interior(out[-1]), grid.interior(fields[-1]))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_oxygen_with_co2_scaling_sediment(precision):
    print("case oxygen with co2, scaling, sediment", precision, flush=True)
    bgc, grid, fields = _model((12, 3, 24), True, co2=True, scale=True, sediment=True)
    errs, _ = _run(bgc, grid, fields, precision, 3, 300.0)
    assert all(e <= TOL[precision] for e in errs.values()), errs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("carbonate,co2,scale", [(False, False, False), (True, True, True), (True, False, False)])
@pytest.mark.parametrize("Nz", [23, 64])
def test_oxygen_gas_exchange(precision, carbonate, co2, scale, Nz):
    print("case oxygen gas exchange", precision, carbonate, co2, scale, Nz, flush=True)
    bgc, grid, fields = _model((9, 2, Nz), carbonate, co2=co2, scale=scale, o2x=True, seed=13)
    errs, _ = _run(bgc, grid, fields, precision, 3, 600.0)
    assert all(e <= TOL[precision] for e in errs.values()), errs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_add_tendencies_with_oxygen_gas_exchange(precision):
    import torch
    print("case oxygen gas exchange tendencies", precision, flush=True)
    bgc, grid, fields = _model((7, 3, 18), True, o2x=True, seed=17)
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
        assert rel_l2(grid.interior(out[n]), grid.interior(G_ref[n])) <= TOL[precision], name


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("Nz,scale,sediment", [(31, False, False), (64, False, False), (40, True, True)])
def test_npzd_with_oxygen(precision, Nz, scale, sediment):
    from oceanbiome import NPZD
    print("case npzd oxygen", precision, Nz, scale, sediment, flush=True)
    grid = RectilinearGrid(size=(11, 2, Nz), extent=(110.0, 20.0, 4.0 * Nz), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))
    bgc = NPZD(grid, oxygen=Oxygen(), scale_negatives=scale,
               light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=80.0),
               sediment_model=InstantRemineralisationSediment(grid) if sediment else None)
    rng = np.random.default_rng(19)
    fields = []
    for lo, hi in ((1.0, 10.0), (0.05, 0.5), (0.05, 0.3), (8.0, 20.0), (0.5, 3.0), (150.0, 300.0)):
        f = grid.new_field()
        grid.interior(f)[...] = rng.uniform(lo, hi, (Nz, 2, 11))
        fields.append(f)
    errs, out = _run(bgc, grid, fields, precision, 3, 600.0)
    assert all(e <= TOL[precision] for e in errs.values()), errs
    assert not np.array_equal(grid.interior(out[-1]), grid.interior(fields[-1]))


def test_other_tracers_match_the_model_without_oxygen():
    print("case oxygen one-way coupling", flush=True)
    import torch
    bgc, grid, fields = _model((10, 2, 20), True, co2=True)
    bgc0, _, _ = _model((10, 2, 20), True, co2=True, oxygen=False)
    a = HIPBiogeochemistry(bgc, "fp64")
    b = HIPBiogeochemistry(bgc0, "fp64")
    ta = to_device(fields, torch.float64)
    tb = to_device(fields[:-1], torch.float64)
    a.step(ta, 3, 600.0)
    b.step(tb, 3, 600.0)
    torch.cuda.synchronize()
    for n, (x, y) in enumerate(zip(to_host(ta)[:-1], to_host(tb))):
        assert rel_l2(grid.interior(x), grid.interior(y)) <= 1e-14, n


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_add_tendencies_with_oxygen(precision):
    import torch
    print("case oxygen tendencies", precision, flush=True)
    bgc, grid, fields = _model((11, 3, 20), True)
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
        assert rel_l2(grid.interior(out[n]), grid.interior(G_ref[n])) <= TOL[precision], name
