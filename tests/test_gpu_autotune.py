"""The column sweep's autotuner (obme_autotune) and its launch shapes
(obme_sweep_configure): every shape gives the same state bit for bit, the
autotuner leaves the caller's tracers untouched, and a handle that has been
tuned steps exactly as a fresh handle of the chosen shape does (the warm
starts of the CO2 solve and the sediment stepper's history are left as they
were).  LOBSTER + DIC/Alk with the CO2 exchange in FP32-mixed, with the O2
exchange in FP64, and NPZD with an InstantRemineralisation sediment in both
precisions."""
import numpy as np
import pytest

from helpers import to_device, to_host
from oceanbiome import (LOBSTER, NPZD, CarbonateSystem, CarbonDioxideGasExchangeBoundaryCondition,
                        InstantRemineralisationSediment, Oxygen, OxygenGasExchangeBoundaryCondition,
                        RectilinearGrid, TwoBandPhotosyntheticallyActiveRadiation)
from oceanbiome.runtime import HIPBiogeochemistry, OBMEError

pytestmark = pytest.mark.gpu
WAVES = (1, 2, 4, 8)


def _grid(shape):
    Nx, Ny, Nz = shape
    return RectilinearGrid(size=shape, extent=(10.0 * Nx, 10.0 * Ny, 4.0 * Nz), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))


def _fields(grid, ranges, seed):
    rng = np.random.default_rng(seed)
    out = []
    for lo, hi in ranges:
        f = grid.new_field()
        grid.interior(f)[...] = rng.uniform(lo, hi, (grid.Nz, grid.Ny, grid.Nx))
        out.append(f)
    return out


LOBSTER_RANGES = ((5.0, 12.0), (0.1, 1.0), (0.05, 0.5), (0.05, 0.3), (0.5, 3.0), (0.5, 3.0), (0.1, 1.0),
                  (2000.0, 2200.0), (2300.0, 2400.0))
NPZD_RANGES = ((1.0, 10.0), (0.05, 0.5), (0.05, 0.3), (8.0, 20.0), (0.5, 3.0))


def _lobster(shape, oxygen=False):
    grid = _grid(shape)
    bgc = LOBSTER(grid, carbonate_system=CarbonateSystem(), oxygen=Oxygen() if oxygen else None,
                  light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=80.0),
                  gas_exchange=(CarbonDioxideGasExchangeBoundaryCondition(),) +
                  ((OxygenGasExchangeBoundaryCondition(),) if oxygen else ()),
                  temperature=20 - 0.05 * np.abs(grid.zc), salinity=35.0)
    ranges = LOBSTER_RANGES + (((150.0, 300.0),) if oxygen else ())
    return bgc, grid, _fields(grid, ranges, 31)


def _npzd_sediment(shape):
    grid = _grid(shape)
    bgc = NPZD(grid, light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=80.0),
               sediment_model=InstantRemineralisationSediment(grid))
    return bgc, grid, _fields(grid, NPZD_RANGES, 32)


CASES = [("lobster co2", "fp32", lambda: _lobster((200, 5, 33))),
         ("lobster co2 o2", "fp64", lambda: _lobster((130, 3, 20), oxygen=True)),
         ("npzd sediment", "fp32", lambda: _npzd_sediment((150, 4, 24))),
         ("npzd sediment", "fp64", lambda: _npzd_sediment((70, 3, 17)))]


@pytest.mark.parametrize("name,precision,make", CASES, ids=[f"{c[0]}-{c[1]}" for c in CASES])
def test_autotune_is_invisible_in_the_results(name, precision, make):
    import torch
    print("case autotune", name, precision, flush=True)
    bgc, grid, fields = make()
    nsteps, dt = 3, 300.0

    # one step first, so the handle carries state (warm starts, the
    # sediment's history) into the autotuner
    tuned = HIPBiogeochemistry(bgc, precision)
    state = to_device(fields, tuned.dtype)
    tuned.step(state, 1, dt)
    torch.cuda.synchronize()
    before = to_host(state)
    storage_before = None if tuned.sediment_storage is None else tuned.sediment_storage.cpu().numpy().copy()

    chosen = tuned.autotune(state, 2, dt)
    torch.cuda.synchronize()
    assert chosen in WAVES
    for n, (a, b) in enumerate(zip(to_host(state), before)):
        assert np.array_equal(a, b), f"autotune changed tracer {n}"
    if storage_before is not None:
        assert np.array_equal(tuned.sediment_storage.cpu().numpy(), storage_before)

    # the same history on a handle that was never tuned, for every shape
    results = {}
    for w in WAVES:
        dev = HIPBiogeochemistry(bgc, precision)
        dev.configure_sweep(w)
        assert dev.sweep_waves == w
        st = to_device(fields, dev.dtype)
        dev.step(st, 1, dt)
        dev.step(st, nsteps, dt, t0=dt)
        torch.cuda.synchronize()
        results[w] = (to_host(st), None if dev.sediment_storage is None else dev.sediment_storage.cpu().numpy())
    tuned.step(state, nsteps, dt, t0=dt)
    torch.cuda.synchronize()
    got = to_host(state)
    for w in WAVES:
        out, storage = results[w]
        for n, (a, b) in enumerate(zip(out, results[WAVES[0]][0])):
            assert np.array_equal(a, b), f"waves {w} differ from waves {WAVES[0]} in tracer {n}"
        if storage is not None:
            assert np.array_equal(storage, results[WAVES[0]][1])
    for n, (a, b) in enumerate(zip(got, results[chosen][0])):
        assert np.array_equal(a, b), f"tuned handle differs from a fresh one in tracer {n}"
    if tuned.sediment_storage is not None:
        assert np.array_equal(tuned.sediment_storage.cpu().numpy(), results[chosen][1])


def test_sweep_configure_rejects_other_shapes():
    print("case sweep configure", flush=True)
    bgc, grid, fields = _npzd_sediment((8, 2, 6))
    dev = HIPBiogeochemistry(bgc, "fp64")
    for w in (0, 3, 16):
        with pytest.raises(OBMEError, match="1, 2, 4 or 8"):
            dev.configure_sweep(w)
    state = to_device(fields, dev.dtype)
    with pytest.raises(OBMEError, match="trial_steps"):
        dev.autotune(state, 0, 60.0)
