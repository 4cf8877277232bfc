"""The benchmark / parity workloads as reproducible synthetic inputs.

Each builder returns a Workload: the model description, its grid, host
float64 tracer parent arrays in tracer order, read-only fields, Δt and the
number of steps.  Random draws use numpy's PCG64 with the given seed.

 - npzd_2d:      NPZD on a 160×32 (Bounded, Flat, Bounded) 10 km × 500 m
                 slice, tanh front initial state, 100 steps of 50 s.
 - lobster_3d:   LOBSTER + DIC/Alk on Nx×Ny×Nz (default 512×512×64, 1 km ×
                 1 km × 200 m), lognormal tracers, CO2 exchange, 200 × 60 s.
 - npzd_columns: ensemble of independent NPZD columns of ragged depth with
                 per-column parameters and surface PAR.
 - carbonate_points: DIC/Alk/T/S/u10 point sets for the CO2 flux.
 - pisces_3d:    PISCES on Nx×Ny×Nz (default 256×256×64, 500 m spacing, 1 km
                 deep), tracers lognormal (σ = 0.3) around the initial
                 values of validation/PISCES/column.jl:80-87, T = (2 + 26
                 (j-½)/Ny) exp(z/500), S = 35, surface PAR 100 W/m².
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import abi
from .grid import RectilinearGrid
from .models import (LOBSTER, NPZD, CarbonateSystem, CarbonDioxideGasExchangeBoundaryCondition, Oxygen,
                     OxygenGasExchangeBoundaryCondition,
                     TwoBandPhotosyntheticallyActiveRadiation)
from .units import day


@dataclass
class Workload:
    name: str
    bgc: object
    grid: RectilinearGrid
    tracers: list                 # host float64 parent arrays, tracer order
    dt: float
    nsteps: int
    T3: Optional[np.ndarray] = None
    S3: Optional[np.ndarray] = None
    extra: dict = field(default_factory=dict)

    @property
    def written(self):
        return self.bgc.written_tracers()

    @property
    def n_written(self):
        return len(self.written)

    def algorithmic_bytes_per_step(self, state_bytes, ro_bytes=None):
        """Algorithmic HBM bytes of one step: every written tracer read and
        written once, plus read-only 3-D inputs (NPZD's T tracer)."""
        n = self.grid.n_interior
        n_ro = getattr(self.bgc, "read_only_3d", None)
        if n_ro is None:
            n_ro = 1 if self.bgc.model == abi.OBME_NPZD else 0
        ro = state_bytes if ro_bytes is None else ro_bytes
        return n * (2 * self.n_written * state_bytes + n_ro * ro)


def npzd_2d(nsteps=100, dt=50.0, surface_PAR=100.0):
    """NPZD 160×32 front (the reference's CPU-runnable 2-D case)."""
    grid = RectilinearGrid(size=(160, 32), extent=(10000.0, 500.0), halo=(3, 3, 3),
                           topology=("Bounded", "Flat", "Bounded"))
    bgc = NPZD(grid, light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=surface_PAR))
    x = grid.xc[None, :]
    z = grid.zc[:, None]
    front = np.tanh((x - 7000 + 4 * z) / 500)
    N = 2.5 - 2 * front
    P = np.where(z > -50, 0.03, 0.01) * np.ones_like(x)
    Z = P.copy()
    D = np.zeros_like(N)
    T = 9 + 0.05 * front
    fields = []
    for v in (N, P, Z, T, D):
        f = grid.new_field()
        grid.interior(f)[:, 0, :] = v
        fields.append(f)
    return Workload("npzd_2d_160x32", bgc, grid, fields, dt, nsteps)


def lobster_3d(Nx=512, Ny=512, Nz=64, nsteps=200, dt=60.0, seed=42, surface_PAR=100.0, gas_exchange=True,
               Lx=1000.0, Ly=1000.0, Lz=200.0, oxygen=False, sinking_scheme="upwind"):
    """LOBSTER + CarbonateSystem on a periodic box with CO2 exchange; with
    oxygen, also the O2 tracer (last) and its gas exchange; the sinking face
    value of `sinking_scheme` ("upwind", "centered", "weno5")."""
    grid = RectilinearGrid(size=(Nx, Ny, Nz), extent=(Lx, Ly, Lz), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))
    T_prof = 20 - 0.05 * np.abs(grid.zc)
    ges = [CarbonDioxideGasExchangeBoundaryCondition()] if gas_exchange else []
    if oxygen and gas_exchange:
        ges.append(OxygenGasExchangeBoundaryCondition())
    bgc = LOBSTER(grid, carbonate_system=CarbonateSystem(), oxygen=Oxygen() if oxygen else None,
                  light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=surface_PAR),
                  gas_exchange=tuple(ges), temperature=T_prof, salinity=35.0, sinking_scheme=sinking_scheme)
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (Nz, Ny, Nx)
    medians = (10.0, 0.1, 0.1, 0.1, 0.05, 0.05, 0.05)
    fields = []
    for m in medians:
        f = grid.new_field()
        grid.interior(f)[...] = m * np.exp(0.5 * rng.standard_normal(shape))
        fields.append(f)
    for lo, hi in ((2000.0, 2200.0), (2250.0, 2400.0)) + (((180.0, 320.0),) if oxygen else ()):
        f = grid.new_field()
        grid.interior(f)[...] = rng.uniform(lo, hi, shape)
        fields.append(f)
    name = f"lobster_dic_alk{'_o2' if oxygen else ''}{'' if sinking_scheme == 'upwind' else '_' + sinking_scheme}" \
           f"_{Nx}x{Ny}x{Nz}"
    return Workload(name, bgc, grid, fields, dt, nsteps,
                    extra={"T_profile": T_prof, "S": 35.0})


# per-column NPZD parameters that are perturbed (obme.h OBME_BP_*)
BATCH_PARAM_DEFAULTS = None


def npzd_column_defaults():
    """NPZD defaults (constructors.jl:177-227) in OBME_BP_* order."""
    b = NPZD(RectilinearGrid(size=(1, 1, 16), extent=(1, 1, 32)))
    pl, d = b.plankton, b.detritus
    v = [0.0] * abi_bp_count()
    v[0] = pl.nitrate_half_saturation
    v[1] = pl.light_half_saturation
    v[2] = pl.phytoplankton_maximum_growth_rate
    v[3] = pl.phytoplankton_exudation_fraction
    v[4] = pl.ammonia_fraction_of_exudate
    v[5] = pl.temperature_coefficient
    v[6] = pl.phytoplankton_mortality_rate
    v[7] = pl.zooplankton_mortality_rate
    v[8] = pl.zooplankton_excretion_rate
    v[9] = pl.phytoplankton_solid_waste_fraction
    v[10] = pl.excretion_inorganic_fraction
    v[11] = pl.preference_for_phytoplankton
    v[12] = pl.maximum_grazing_rate
    v[13] = pl.grazing_half_saturation
    v[14] = pl.zooplankton_assimilation_fraction
    v[15] = d.remineralisation_rate
    v[16] = d.small_particle_fraction
    v[17] = pl.phytoplankton_sinking_speed
    v[18] = d.sinking_speed
    v[19] = 100.0
    return np.array(v)


def abi_bp_count():
    return 20


# parameters kept unperturbed (fractions bounded by 1 in the formulation)
_FRACTIONS = (3, 4, 9, 10, 11, 14, 16)


def npzd_columns(n_columns=8192, nz_min=16, nz_max=128, dz=2.0, nsteps=500, dt=60.0, seed=42, T=10.0):
    """Ensemble of independent NPZD columns: Nz ~ U{nz_min..nz_max}, every
    rate × lognormal(0, 0.2) (fractions kept), PAR0 ~ U(50, 200),
    P, Z ~ U(0.01, 0.1), N ~ U(1, 10), D = 0, uniform T."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nz = rng.integers(nz_min, nz_max + 1, n_columns)
    offsets = np.zeros(n_columns + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(nz)
    base = npzd_column_defaults()
    params = np.tile(base[:, None], (1, n_columns))
    fac = np.exp(0.2 * rng.standard_normal((len(base), n_columns)))
    for k in range(len(base)):
        if k in _FRACTIONS or k == 19:
            continue
        params[k] *= fac[k]
    params[19] = rng.uniform(50.0, 200.0, n_columns)
    ncell = int(offsets[-1])
    N = rng.uniform(1.0, 10.0, ncell)
    P = rng.uniform(0.01, 0.1, ncell)
    Z = rng.uniform(0.01, 0.1, ncell)
    D = np.zeros(ncell)
    bgc = NPZD(RectilinearGrid(size=(1, 1, int(nz_max)), z=(-nz_max * dz, 0.0)),
               light_attenuation=TwoBandPhotosyntheticallyActiveRadiation(surface_PAR=100.0))
    return Workload(f"npzd_columns_{n_columns}", bgc, bgc.grid, [N, P, Z, D], dt, nsteps,
                    extra={"offsets": offsets, "params": params, "dz": dz, "T": T})


# the initial values of validation/PISCES/column.jl:80-87 (CaCO3 = 1e-10)
PISCES_MEDIANS = dict(P=0.1, PChl=0.025, PFe=0.005, D=0.01, DChl=0.003, DFe=0.0006, DSi=0.004, Z=0.06, M=0.5,
                      DOC=4.0, POC=5.4, GOC=8.2, SFe=0.34, BFe=0.5, PSi=0.04, CaCO3=1e-10, NO3=10.0, NH4=0.1,
                      PO4=5.0, Fe=0.6, Si=8.6, DIC=2205.0, Alk=2560.0, O2=317.0)
# the tracer values of the reference's PISCES tests (test/test_PISCES.jl:6-16)
PISCES_TEST_MEDIANS = dict(P=0.5, PChl=0.02, PFe=0.005, D=0.1, DChl=0.004, DFe=0.001, DSi=0.01, Z=0.1, M=0.7,
                           DOC=2.1, POC=7.8, GOC=38.0, SFe=0.206, BFe=1.1, PSi=0.1, CaCO3=0.1, NO3=2.3, NH4=0.9,
                           PO4=0.6, Fe=0.13, Si=8.5, DIC=2205.0, Alk=2566.0, O2=317.0)


def pisces_3d(Nx=256, Ny=256, Nz=64, nsteps=100, dt=300.0, seed=42, Lz=1000.0, spacing=500.0,
              surface_PAR=100.0, mixed_layer_depth=-50.0, medians="column"):
    """PISCES on a periodic box: 24 tracers lognormal(σ = 0.3) around the
    initial values of validation/PISCES/column.jl:80-87, T = (2 + 26 (j-½)/Ny)
    exp(z/500) °C, S = 35, a constant surface PAR of 100 W/m² by default
    ("default" for default_surface_PAR, or a 2-D array), prescribed
    mixed-layer depth.  medians="test": around the values of the reference's
    PISCES tests (test/test_PISCES.jl:6-16) instead."""
    from .pisces import PISCES, TRACERS
    grid = RectilinearGrid(size=(Nx, Ny, Nz), extent=(Nx * spacing, Ny * spacing, Lz), halo=(3, 3, 3),
                           topology=("Periodic", "Periodic", "Bounded"))
    model = PISCES(grid, surface_PAR=surface_PAR, mixed_layer_depth=mixed_layer_depth)
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (Nz, Ny, Nx)
    fields = []
    med = {"column": PISCES_MEDIANS, "test": PISCES_TEST_MEDIANS}[medians]
    for name in med:
        f = grid.new_field()
        grid.interior(f)[...] = med[name] * np.exp(0.3 * rng.standard_normal(shape))
        fields.append(f)
    jj = (np.arange(Ny) + 0.5) / Ny
    T = (2 + 26 * jj)[None, :, None] * np.exp(grid.zc / 500)[:, None, None] * np.ones((1, 1, Nx))
    fT = grid.new_field()
    grid.interior(fT)[...] = T
    fS = grid.new_field(fill=0.0)
    grid.interior(fS)[...] = 35.0
    fields += [fT, fS]
    assert len(fields) == len(TRACERS)
    return Workload(f"pisces_{Nx}x{Ny}x{Nz}", model, grid, fields, dt, nsteps)


def carbonate_points(n=1 << 24, seed=42, pCO2_air=420.0):
    """DIC ~ U(1900, 2300), Alk ~ U(2100, 2500), T ~ U(-2, 32), S ~ U(30, 38),
    u10 = 10 × lognormal(0, 0.3)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return {
        "DIC": rng.uniform(1900.0, 2300.0, n),
        "Alk": rng.uniform(2100.0, 2500.0, n),
        "T": rng.uniform(-2.0, 32.0, n),
        "S": rng.uniform(30.0, 38.0, n),
        "u10": 10.0 * np.exp(0.3 * rng.standard_normal(n)),
        "pCO2_air": pCO2_air,
    }
