"""The centred and fifth-order WENO sinking reconstructions on the GPU (the
drift flux of Oceananigans' Centered(order = 2) and WENO(order = 5) schemes,
the `advection = WENO()` setups) against the FP64 oracle: the RK3 step, run
one stage per launch, for NPZD and LOBSTER (with the carbonate system and
the CO2 exchange, with O2 and its exchange), on uniform and stretched grids,
at column depths where the WENO order drops near both ends, with the
diurnal surface PAR; and the update_tendencies! hook.  FP64 <= 1e-12
relative L2 per tracer, FP32-mixed <= 1e-5.  Parity of the reconstructions
with Oceananigans itself is unpinned (the oracle restates the textbook
scheme)."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, to_device, to_host
from oceanbiome import (LOBSTER, NPZD, CarbonateSystem, CarbonDioxideGasExchangeBoundaryCondition, Oxygen,
                        OxygenGasExchangeBoundaryCondition, RectilinearGrid, TwoBandPhotosyntheticallyActiveRadiation,
                        abi)
from oceanbiome.runtime import HIPBiogeochemistry, OBMEError

pytestmark = pytest.mark.gpu
TOL = {"fp64": 1e-12, "fp32": 1e-5}
SCHEMES = ("centered", "weno5")


def _fields(grid, ranges, seed):
    rng = np.random.default_rng(seed)
    out = []
    for lo, hi in ranges:
        f = grid.new_field()
        grid.interior(f)[...] = rng.uniform(lo, hi, (grid.Nz, grid.Ny, grid.Nx))
        out.append(f)
    return out


NPZD_RANGES = ((1.0, 10.0), (0.05, 0.5), (0.05, 0.3), (8.0, 20.0), (0.5, 3.0))
LOBSTER_RANGES = ((5.0, 12.0), (0.1, 1.0), (0.05, 0.5), (0.05, 0.3), (0.5, 3.0), (0.5, 3.0), (0.1, 1.0),
                  (2000.0, 2200.0), (2300.0, 2400.0))


def _grid(shape, stretched=False):
    Nx, Ny, Nz = shape
    if stretched:
        faces = -200.0 * (1 - np.linspace(0.0, 1.0, Nz + 1) ** 1.5)
        faces[-1] = 0.0
        return RectilinearGrid(size=shape, x=(0, 10.0 * Nx), y=(0, 10.0 * Ny), z=faces)
    return RectilinearGrid(size=shape, extent=(10.0 * Nx, 10.0 * Ny, 4.0 * Nz), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))


def _npzd(shape, scheme, surface="default", seed=41):
    grid = _grid(shape)
    bgc = NPZD(grid, sinking_scheme=scheme,
               light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=surface))
    return bgc, grid, _fields(grid, NPZD_RANGES, seed)


def _lobster(shape, scheme, oxygen=False, stretched=False, seed=42):
    grid = _grid(shape, stretched)
    bgc = LOBSTER(grid, carbonate_system=CarbonateSystem(), oxygen=Oxygen() if oxygen else None,
                  sinking_scheme=scheme,
                  light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=90.0),
                  gas_exchange=(CarbonDioxideGasExchangeBoundaryCondition(),) +
                  ((OxygenGasExchangeBoundaryCondition(),) if oxygen else ()),
                  temperature=20 - 0.05 * np.abs(grid.zc), salinity=35.0)
    ranges = LOBSTER_RANGES + (((150.0, 300.0),) if oxygen else ())
    return bgc, grid, _fields(grid, ranges, seed)


def _step_and_compare(bgc, grid, fields, precision, nsteps, dt, t0=0.0, calls=1):
    import torch
    ref = [a.copy() for a in fields]
    cfg = bgc.config(abi.OBME_FP64)
    dev = HIPBiogeochemistry(bgc, precision)
    tr = to_device(fields, dev.dtype)
    for c in range(calls):
        O.grid_run(cfg, ref, nsteps, dt, t0=t0 + c * nsteps * dt)
        dev.step(tr, nsteps, dt, t0=t0 + c * nsteps * dt)
    torch.cuda.synchronize()
    out = to_host(tr)
    errs = {n: rel_l2(grid.interior(a), grid.interior(b))
            for n, a, b in zip(bgc.required_biogeochemical_tracers(), out, ref)}
    print("errors", {k: float(v) for k, v in errs.items()}, flush=True)
    return errs, out


CASES = [("npzd", (45, 3, 37), {}),
         ("npzd short columns", (30, 2, 5), {}),
         ("npzd three levels", (30, 2, 3), {}),
         ("lobster co2", (70, 2, 64), {}),
         ("lobster co2 o2 stretched", (33, 3, 30), {"oxygen": True, "stretched": True})]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("name,shape,kw", CASES, ids=[c[0] for c in CASES])
def test_step_with_sinking_scheme(precision, scheme, name, shape, kw):
    """Three steps of 1800 s from t = 0 (the diurnal surface PAR changes
    between the stage clocks for NPZD), twice (the second call starts
    with fresh warm starts and scratch)."""
    print("case sinking scheme", scheme, name, precision, flush=True)
    if name.startswith("npzd"):
        bgc, grid, fields = _npzd(shape, scheme)
    else:
        bgc, grid, fields = _lobster(shape, scheme, **kw)
    errs, _ = _step_and_compare(bgc, grid, fields, precision, 3, 1800.0, calls=2)
    assert all(e <= TOL[precision] for e in errs.values()), errs


@pytest.mark.parametrize("scheme", SCHEMES)
def test_schemes_differ_from_upwind(scheme):
    """The reconstruction is in effect: the sinking tracers' states differ
    from the upwind step's by far more than the bars (FP64)."""
    print("case sinking scheme differs", scheme, flush=True)
    bgc, grid, fields = _npzd((16, 2, 40), scheme, surface=100.0)
    bgc_up, _, _ = _npzd((16, 2, 40), "upwind", surface=100.0)
    _, out = _step_and_compare(bgc, grid, fields, "fp64", 5, 3600.0)
    _, out_up = _step_and_compare(bgc_up, grid, fields, "fp64", 5, 3600.0)
    # D (index 4) sinks fastest
    assert rel_l2(grid.interior(out[4]), grid.interior(out_up[4])) > 1e-8


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("scheme", SCHEMES)
def test_add_tendencies_with_sinking_scheme(precision, scheme):
    import torch
    print("case sinking scheme tendencies", scheme, precision, flush=True)
    bgc, grid, fields = _lobster((21, 3, 26), scheme, oxygen=True)
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


def test_sinking_scheme_limits():
    """The centred and WENO steps run without negative scaling and without a
    sediment, and not through the allocation-free entry point."""
    from oceanbiome import InstantRemineralisationSediment
    print("case sinking scheme limits", flush=True)
    grid = _grid((8, 2, 10))
    with pytest.raises(OBMEError, match="without negative scaling"):
        HIPBiogeochemistry(NPZD(grid, sinking_scheme="weno5", scale_negatives=True), "fp64")
    with pytest.raises(OBMEError, match="without a sediment"):
        HIPBiogeochemistry(NPZD(grid, sinking_scheme="weno5", sediment_model=InstantRemineralisationSediment(grid)),
                           "fp64")
    bgc, grid, fields = _npzd((8, 2, 10), "weno5")
    dev = HIPBiogeochemistry(bgc, "fp64")
    tr = to_device(fields, dev.dtype)
    with pytest.raises(OBMEError, match="upwind sinking scheme only"):
        dev.step_noalloc(tr, 1, 60.0)
