"""Host-side mirror of OceanBioME's NutrientsPlanktonDetritus constructors.

Each component carries the reference defaults (the `@kwdef` structs and the
NPZD / LOBSTER constructors); arithmetic on the defaults is written in the
reference's order so every parameter is the same Float64.  `NPZD(grid)` and
`LOBSTER(grid; carbonate_system=CarbonateSystem())` return a
`Biogeochemistry` whose `.config()` is the obme_config the HIP library takes.
"""
import ctypes as C
from dataclasses import dataclass, field, fields, replace
from typing import Optional

import numpy as np

from . import abi
from .units import day, hour


class Linear:
    code = abi.OBME_LINEAR


class Quadratic:
    code = abi.OBME_QUADRATIC


class MondoLightLimitation:
    code = abi.OBME_LIGHT_MONDO


class AnalyticalLightLimitation:
    code = abi.OBME_LIGHT_ANALYTICAL


@dataclass
class PhytoZoo:
    """PhytoZoo (NutrientsPlanktonDetritus/plankton.jl:19-76)."""
    nitrate_half_saturation: float = 0.7
    ammonia_half_saturation: float = 0.001
    iron_half_saturation: float = 2e-4
    nitrate_ammonia_inhibition: float = 3.0
    light_half_saturation: float = 33.0
    phytoplankton_maximum_growth_rate: float = 2.42e-5
    iron_ratio: float = 4.6375e-5
    phytoplankton_exudation_fraction: float = 0.05
    ammonia_fraction_of_exudate: float = 0.75
    light_limitation: type = MondoLightLimitation
    temperature_coefficient: Optional[float] = None
    phytoplankton_mortality_rate: float = 5.8e-7
    zooplankton_mortality_rate: float = 2.31e-6
    zooplankton_excretion_rate: float = 5.8e-7
    phytoplankton_mortality_formulation: type = Quadratic
    phytoplankton_solid_waste_fraction: float = 1.0
    excretion_inorganic_fraction: float = 0.5
    preference_for_phytoplankton: float = 0.5
    maximum_grazing_rate: float = 9.26e-6
    grazing_half_saturation: float = 1.0
    zooplankton_assimilation_fraction: float = 0.7
    grazing_concentration_formulation: type = Quadratic
    zooplankton_calcite_dissolution: float = 0.3
    redfield_ratio: float = 6.56
    carbon_calcite_ratio: float = 0.1
    zooplankton_gut_calcite_dissolution: float = 0.3
    phytoplankton_chlorophyll_ratio: float = 1.31
    phytoplankton_sinking_speed: float = 0.0
    zooplankton_sinking_speed: float = 0.0
    open_bottom: bool = True

    tracers = ("P", "Z")


@dataclass
class Nutrient:
    """Nutrient (nutrients.jl:82-92): a single nutrient N."""
    tracers = ("N",)


@dataclass
class NitrateAmmonia:
    """NitrateAmmonia (nutrients.jl:17-21)."""
    nitrification_rate: float = 5.8e-7
    tracers = ("NO₃", "NH₄")


@dataclass
class TwoParticleAndDissolved:
    """TwoParticleAndDissolved (detritus.jl:116-130, ctor :312-323)."""
    remineralisation_inorganic_fraction: float = 0.0
    small_remineralisation_rate: float = 5.88e-7
    large_remineralisation_rate: float = 5.88e-7
    dissolved_remineralisation_rate: float = 3.86e-7
    small_solid_waste_fraction: float = 0.5
    redfield_ratio: float = 6.56
    small_particle_sinking_speed: float = 3.47e-5
    large_particle_sinking_speed: float = 200 / day
    open_bottom: bool = True
    tracers = ("sPOM", "bPOM", "DOM")


@dataclass
class Detritus:
    """Detritus, single class (detritus.jl:355-384)."""
    remineralisation_rate: float = 0.1213 / day
    small_particle_fraction: float = 0.5
    redfield_ratio: float = 6.56
    sinking_speed: float = 2.7489 / day
    open_bottom: bool = True
    tracers = ("D",)


@dataclass
class CarbonateSystem:
    """CarbonateSystem{1} (carbonate_system.jl:39-47)."""
    tracers = ("DIC", "Alk")


@dataclass
class Oxygen:
    """Oxygen (NutrientsPlanktonDetritus/oxygen.jl:14-17): the O₂ tracer of
    LOBSTER or NPZD, one-way coupled (its tendencies, oxygen.jl:21-46, change
    no other tracer)."""
    respiration_oxygen_nitrogen_ratio: float = 10.75
    nitrification_oxygen_nitrogen_ratio: float = 2.0
    tracers = ("O₂",)

    def struct(self):
        return abi.Oxygen(enabled=1, respiration_oxygen_nitrogen_ratio=self.respiration_oxygen_nitrogen_ratio,
                          nitrification_oxygen_nitrogen_ratio=self.nitrification_oxygen_nitrogen_ratio)


@dataclass
class DampedNewtonRaphsonSolver:
    """DampedNewtonRaphsonSolver (Utils/solvers.jl:81-88) with the
    CarbonChemistry bounds (lower = 0) (carbon_chemistry.jl:81)."""
    max_iters: int = 100
    atol: float = 1e-20
    damping: float = 0.5
    armijo_constant: float = 0.5
    min_damping: Optional[float] = None

    def struct(self, initial_pH_guess=8.0):
        md = self.min_damping if self.min_damping is not None else self.damping ** 10
        return abi.CarbonSolver(max_iters=self.max_iters, atol=self.atol, damping=self.damping,
                                armijo_constant=self.armijo_constant, min_damping=md,
                                initial_pH_guess=initial_pH_guess)


@dataclass
class CarbonChemistry:
    """CarbonChemistry (carbon_chemistry.jl:66-87): the Dickson (2007)
    constants with the TEOS-10 polynomial density and the damped Newton."""
    solver: DampedNewtonRaphsonSolver = field(default_factory=DampedNewtonRaphsonSolver)
    initial_pH_guess: float = 8.0


@dataclass
class CarbonDioxideGasExchangeBoundaryCondition:
    """CarbonDioxideGasExchangeBoundaryCondition (GasExchange.jl:104-124):
    Ho06 k660 (gas_transfer_velocity.jl:61) scaled by the Wanninkhof (2014)
    CO2 Schmidt number (gas_solubility.jl:211-212)."""
    air_concentration: float = 413.0
    wind_speed: float = 2.0
    schmidt: tuple = (2116.8, -136.25, 4.7353, -0.092307, 0.0007555)
    transfer_velocity: tuple = (0.0, 0.0, 0.266 / hour / 100)
    carbon_chemistry: CarbonChemistry = field(default_factory=CarbonChemistry)

    def struct(self):
        s = abi.CO2Exchange(enabled=1, air_concentration=float(self.air_concentration),
                            wind_speed=float(self.wind_speed), air_pressure=1.0)
        for i, v in enumerate(self.schmidt):
            s.schmidt[i] = float(v)
        for i, v in enumerate(self.transfer_velocity):
            s.transfer_velocity[i] = float(v)
        return s


@dataclass
class OxygenGasExchangeBoundaryCondition:
    """OxygenGasExchangeBoundaryCondition (GasExchange.jl:127-146): Ho06 k660
    scaled by the Wanninkhof (2014) O2 Schmidt number (schmidt_number.jl:15-16);
    air side 9352.7 mmol/m3 times the Wanninkhof (1992) solubility β/Tk
    (gas_solubility.jl:28-54)."""
    air_concentration: float = 9352.7
    wind_speed: float = 2.0
    schmidt: tuple = (1920.4, -135.6, 5.2122, -0.10939, 0.00093777)
    transfer_velocity: tuple = (0.0, 0.0, 0.266 / hour / 100)
    solubility: tuple = (-58.3877, 85.8079, 23.8439, -0.034892, 0.015578, -0.0019387)

    def struct(self):
        s = abi.O2Exchange(enabled=1, air_concentration=float(self.air_concentration),
                           wind_speed=float(self.wind_speed))
        for i, v in enumerate(self.schmidt):
            s.schmidt[i] = float(v)
        for i, v in enumerate(self.transfer_velocity):
            s.transfer_velocity[i] = float(v)
        for i, v in enumerate(self.solubility):
            s.solubility[i] = float(v)
        return s


@dataclass
class TwoBandPhotosyntheticallyActiveRadiation:
    """TwoBandPhotosyntheticallyActiveRadiation (Light/2band.jl:106-146).
    surface_PAR: a number (constant), "default" (default_surface_PAR,
    Light/Light.jl:217-219) or a 2-D array of PAR0[i, j]."""
    water_red_attenuation: float = 0.225
    water_blue_attenuation: float = 0.0232
    chlorophyll_red_attenuation: float = 0.037
    chlorophyll_blue_attenuation: float = 0.074
    chlorophyll_red_exponent: float = 0.629
    chlorophyll_blue_exponent: float = 0.674
    pigment_ratio: float = 0.7
    phytoplankton_chlorophyll_ratio: float = 1.31
    surface_PAR: object = "default"

    def struct(self):
        s = abi.TwoBandPAR()
        for f in ("water_red_attenuation", "water_blue_attenuation", "chlorophyll_red_attenuation",
                  "chlorophyll_blue_attenuation", "chlorophyll_red_exponent", "chlorophyll_blue_exponent",
                  "pigment_ratio", "phytoplankton_chlorophyll_ratio"):
            setattr(s, f, float(getattr(self, f)))
        if isinstance(self.surface_PAR, str) and self.surface_PAR == "default":
            s.surface_mode = abi.OBME_SURFACE_DIURNAL
        elif np.isscalar(self.surface_PAR):
            s.surface_mode = abi.OBME_SURFACE_CONSTANT
            s.surface_value = float(self.surface_PAR)
        else:
            s.surface_mode = abi.OBME_SURFACE_ARRAY
        return s


class InstantRemineralisationSediment:
    """InstantRemineralisationSediment(grid; sinking_tracers = (:P, :D),
    remineralisation_reciever = :N, ...) (Models/Sediments/instant_remineralisation.jl:60-98):
    the sinking flux f of `sinking_tracers` through the bottom face is
    remineralised into `remineralisation_reciever`'s bottom cell, a fraction
    e = a + b (f / (k + f))^2 buried in the 2-D `storage` field.  The only
    sediment time stepper built is the default QuasiAdamsBashforth2."""

    def __init__(self, grid=None, sinking_tracers=("P", "D"), remineralisation_reciever="N",
                 burial_efficiency_constant1=0.013, burial_efficiency_constant2=0.53,
                 burial_efficiency_half_saturation=7.0 / 6.56, timestepper="QuasiAdamsBashforth2"):
        if timestepper != "QuasiAdamsBashforth2":
            raise NotImplementedError("only the QuasiAdamsBashforth2 sediment time stepper is built")
        self.grid = grid
        self.sinking_tracers = tuple(sinking_tracers)
        self.remineralisation_reciever = remineralisation_reciever
        self.burial_efficiency_constant1 = float(burial_efficiency_constant1)
        self.burial_efficiency_constant2 = float(burial_efficiency_constant2)
        self.burial_efficiency_half_saturation = float(burial_efficiency_half_saturation)
        self.chi = 0.1   # QuasiAdamsBashforth2TimeStepper default χ

    def required_sediment_fields(self):
        return ("storage",)

    def sinking_fluxes(self):
        return self.sinking_tracers

    def coupled_tracers(self):
        return (self.remineralisation_reciever,)

    def struct(self, names):
        """obme_sediment with tracer names resolved against `names`."""
        s = abi.Sediment()
        s.model = abi.OBME_SEDIMENT_INSTANT_REMINERALISATION
        if self.remineralisation_reciever not in names:
            raise ValueError(f"{self.remineralisation_reciever} is not a tracer of the model")
        s.receiver = names.index(self.remineralisation_reciever)
        if len(self.sinking_tracers) > abi.OBME_MAX_SEDIMENT_TRACERS:
            raise ValueError("too many sinking tracers")
        for q, t in enumerate(self.sinking_tracers):
            if t not in names:
                raise ValueError(f"{t} is not a tracer of the model")
            s.sinking[q] = names.index(t)
        s.n_sinking = len(self.sinking_tracers)
        s.timestepper = abi.OBME_SEDIMENT_AB2
        s.burial_efficiency_constant1 = self.burial_efficiency_constant1
        s.burial_efficiency_constant2 = self.burial_efficiency_constant2
        s.burial_efficiency_half_saturation = self.burial_efficiency_half_saturation
        s.ab2_chi = self.chi
        return s


class SimpleMultiGSediment:
    """SimpleMultiGSediment(grid; ...) (Models/Sediments/simple_multi_G.jl:60-134),
    the nitrogen form (sinking_carbon = nothing, the default): a single-layer
    sediment with slow, fast and refractory organic nitrogen (Ns, Nf, Nr)
    fed by the sinking nitrogen flux of `sinking_nitrogen`, whose
    remineralisation enters the bottom cells of NH₄, NO₃ and O₂ through the
    Soetaert et al. (2000) nitrification, denitrification and anoxic
    fractions.  fast_redfield, slow_redfield and solid_dep_params are
    accepted and, as in the reference's tendencies, unused.  The carbon form
    (sinking_carbon given) is not built."""

    def __init__(self, grid, fast_decay_rate=2 / day, slow_decay_rate=0.2 / day, fast_redfield=0.1509,
                 slow_redfield=0.13, fast_fraction=0.74, slow_fraction=0.26, refactory_fraction=0.1,
                 sedimentation_rate=None, anoxia_half_saturation=1.0,
                 nitrate_oxidation_params=(-1.9785, 0.2261, -0.0615, -0.0289, -0.36109, -0.0232),
                 denitrification_params=(-3.0790, 1.7509, 0.0593, -0.1923, 0.0604, 0.0662),
                 anoxic_params=(-3.9476, 2.6269, -0.2426, -1.3349, 0.1826, -0.0143),
                 solid_dep_params=(0.233, 0.336, 982.0, -1.548), sinking_nitrogen=("sPOM", "bPOM"),
                 sinking_carbon=None, sinking_redfield=6.56, timestepper="QuasiAdamsBashforth2"):
        if sinking_carbon is not None:
            raise NotImplementedError("the carbon form of SimpleMultiG (sinking_carbon) is not built")
        if timestepper != "QuasiAdamsBashforth2":
            raise NotImplementedError("only the QuasiAdamsBashforth2 sediment time stepper is built")
        self.grid = grid
        self.fast_decay_rate = float(fast_decay_rate)
        self.slow_decay_rate = float(slow_decay_rate)
        self.fast_redfield = float(fast_redfield)
        self.slow_redfield = float(slow_redfield)
        self.fast_fraction = float(fast_fraction)
        self.slow_fraction = float(slow_fraction)
        self.refactory_fraction = float(refactory_fraction)
        if sedimentation_rate is None:
            # 982 |z of the bottom cell centre|^-1.548 cm/year (simple_multi_G.jl:104)
            sedimentation_rate = 982 * abs(float(grid.zc[0])) ** (-1.548)
        self.sedimentation_rate = float(sedimentation_rate)
        self.anoxia_half_saturation = float(anoxia_half_saturation)
        self.nitrate_oxidation_params = tuple(float(x) for x in nitrate_oxidation_params)
        self.denitrification_params = tuple(float(x) for x in denitrification_params)
        self.anoxic_params = tuple(float(x) for x in anoxic_params)
        self.solid_dep_params = tuple(float(x) for x in solid_dep_params)
        self.sinking_nitrogen = tuple(sinking_nitrogen)
        self.sinking_redfield = float(sinking_redfield)
        self.chi = 0.1   # QuasiAdamsBashforth2TimeStepper default χ

    def required_sediment_fields(self):
        return ("Ns", "Nf", "Nr")

    def required_tracers(self):
        return ("NO₃", "NH₄", "O₂")

    def sinking_fluxes(self):
        return self.sinking_nitrogen

    def coupled_tracers(self):
        return ("NO₃", "NH₄", "O₂")

    def struct(self, names):
        """obme_sediment with tracer names resolved against `names`."""
        s = abi.Sediment()
        s.model = abi.OBME_SEDIMENT_SIMPLE_MULTI_G
        for t in self.required_tracers() + self.sinking_nitrogen:
            if t not in names:
                raise ValueError(f"{t} is not a tracer of the model (SimpleMultiG needs NO₃, NH₄ and O₂)")
        if len(self.sinking_nitrogen) > abi.OBME_MAX_SEDIMENT_TRACERS:
            raise ValueError("too many sinking tracers")
        for q, t in enumerate(self.sinking_nitrogen):
            s.sinking[q] = names.index(t)
        s.n_sinking = len(self.sinking_nitrogen)
        s.receiver = names.index("NH₄")
        s.ammonia, s.nitrate, s.oxygen = names.index("NH₄"), names.index("NO₃"), names.index("O₂")
        s.timestepper = abi.OBME_SEDIMENT_AB2
        s.ab2_chi = self.chi
        s.fast_decay_rate = self.fast_decay_rate
        s.slow_decay_rate = self.slow_decay_rate
        s.fast_fraction = self.fast_fraction
        s.slow_fraction = self.slow_fraction
        s.refactory_fraction = self.refactory_fraction
        s.sedimentation_rate = self.sedimentation_rate
        s.anoxia_half_saturation = self.anoxia_half_saturation
        s.sinking_redfield = self.sinking_redfield
        for q in range(6):
            s.nitrate_oxidation_params[q] = self.nitrate_oxidation_params[q]
            s.denitrification_params[q] = self.denitrification_params[q]
            s.anoxic_params[q] = self.anoxic_params[q]
        return s


def default_surface_PAR(t):
    """default_surface_PAR (Light/Light.jl:219)."""
    return 100 * max(0.0, np.cos(t * np.pi / (12 * hour)))


class Biogeochemistry:
    """The DiscreteBiogeochemistry wrapper (src/OceanBioME.jl:72-120) around a
    NutrientsPlanktonDetritus composition."""

    def __init__(self, grid, nutrients, plankton, detritus, carbonate_system=None, oxygen=None,
                 light_attenuation=None, gas_exchange=None, sinking_scheme="upwind",
                 temperature=None, salinity=35.0, scale_negatives=False, invalid_fill_value=float("nan"),
                 sediment_model=None):
        self.grid = grid
        self.sediment = sediment_model
        self.scale_negatives = bool(scale_negatives)
        self.invalid_fill_value = float(invalid_fill_value)
        self.nutrients = nutrients
        self.plankton = plankton
        self.detritus = detritus
        self.carbonate_system = carbonate_system
        self.oxygen = oxygen
        self.light_attenuation = light_attenuation or TwoBandPhotosyntheticallyActiveRadiation()
        # a CO2 and/or an O2 boundary condition (one or a tuple of both)
        ges = () if gas_exchange is None else (tuple(gas_exchange) if isinstance(gas_exchange, (tuple, list))
                                                else (gas_exchange,))
        self.gas_exchange = next((g for g in ges if isinstance(g, CarbonDioxideGasExchangeBoundaryCondition)), None)
        self.o2_exchange = next((g for g in ges if isinstance(g, OxygenGasExchangeBoundaryCondition)), None)
        self.sinking_scheme = sinking_scheme
        self.temperature = temperature
        self.salinity = salinity
        if isinstance(nutrients, Nutrient) and isinstance(detritus, Detritus):
            self.model = abi.OBME_NPZD
        elif isinstance(nutrients, NitrateAmmonia) and isinstance(detritus, TwoParticleAndDissolved):
            self.model = abi.OBME_LOBSTER
        else:
            raise NotImplementedError("supported compositions: NPZD and LOBSTER")
        if self.model == abi.OBME_NPZD and carbonate_system is not None:
            raise NotImplementedError("the carbonate system is supported with LOBSTER")
        if oxygen is not None and not isinstance(oxygen, Oxygen):
            raise TypeError("oxygen must be Oxygen() or None")
        if len(ges) != (self.gas_exchange is not None) + (self.o2_exchange is not None):
            raise NotImplementedError("gas_exchange takes one CO2 and one O2 boundary condition at most")
        if self.o2_exchange is not None and (self.model != abi.OBME_LOBSTER or oxygen is None):
            # the O2 flux needs the O2 tracer (LOBSTER with Oxygen())
            raise ValueError("the O2 gas exchange needs LOBSTER with oxygen=Oxygen()")
        self._keep = []

    # -- Oceananigans.Biogeochemistry API -------------------------------------
    def required_biogeochemical_tracers(self):
        """(NutrientsPlanktonDetritus.jl:69-74; plankton.jl:80)"""
        t = list(self.nutrients.tracers) + list(self.plankton.tracers)
        if self.plankton.temperature_coefficient is not None:
            t.append("T")
        t += list(self.detritus.tracers)
        if self.carbonate_system is not None:
            t += list(self.carbonate_system.tracers)
        if self.oxygen is not None:
            t += list(self.oxygen.tracers)
        return tuple(t)

    def required_biogeochemical_auxiliary_fields(self):
        return ("PAR",)

    def conserved_tracers(self):
        """conserved_tracers (NutrientsPlanktonDetritus/coupling_utils.jl:1-51):
        [(names, scalefactors)] — nitrogen, then carbon with the carbonate
        system, members in the reference's order."""
        nitrogen = ["P", "Z"] + list(self.nutrients.tracers) + list(self.detritus.tracers)
        groups = [(nitrogen, [1.0] * len(nitrogen))]
        if self.carbonate_system is not None:
            R = self.plankton.redfield_ratio
            rho = self.plankton.carbon_calcite_ratio
            groups.append((["P", "Z", "DIC"] + list(self.detritus.tracers),
                           [(1 + rho) * R, R, 1.0] + [R] * len(self.detritus.tracers)))
        return groups

    def scaling_struct(self):
        """obme_negative_scaling of NPZD/LOBSTER(grid; scale_negatives = true)
        (NutrientsPlanktonDetritus.jl:46-58): one group per conserved set."""
        if not self.scale_negatives:
            return abi.NegativeScaling()
        groups = [(n, f, self.invalid_fill_value) for n, f in self.conserved_tracers()]
        return abi.negative_scaling(groups, self.required_biogeochemical_tracers())

    def written_tracers(self):
        """Tracers the step writes (T is read-only: zero tendency, no drift)."""
        return tuple(t for t in self.required_biogeochemical_tracers() if t != "T")

    def sinking_speeds(self):
        """biogeochemical_drift_velocity per tracer, as speeds (m/s)."""
        sp = {}
        pl = self.plankton
        sp["P"] = pl.phytoplankton_sinking_speed
        sp["Z"] = pl.zooplankton_sinking_speed
        d = self.detritus
        if isinstance(d, Detritus):
            sp["D"] = d.sinking_speed
        else:
            sp["sPOM"] = d.small_particle_sinking_speed
            sp["bPOM"] = d.large_particle_sinking_speed
        return sp

    def biogeochemical_drift_velocity(self, name):
        s = self.sinking_speeds().get(name, 0.0)
        return None if s == 0.0 else -s

    def open_bottom(self):
        d = self.detritus
        return bool(getattr(d, "open_bottom", True))

    # -- C ABI configuration ----------------------------------------------------
    def npd_struct(self):
        p = abi.NPDParams()
        pl = self.plankton
        for n in ("nitrate_half_saturation", "ammonia_half_saturation", "nitrate_ammonia_inhibition",
                  "light_half_saturation", "phytoplankton_maximum_growth_rate",
                  "phytoplankton_exudation_fraction", "ammonia_fraction_of_exudate",
                  "phytoplankton_mortality_rate", "zooplankton_mortality_rate", "zooplankton_excretion_rate",
                  "phytoplankton_solid_waste_fraction", "excretion_inorganic_fraction",
                  "preference_for_phytoplankton", "maximum_grazing_rate", "grazing_half_saturation",
                  "zooplankton_assimilation_fraction", "redfield_ratio", "carbon_calcite_ratio",
                  "zooplankton_gut_calcite_dissolution", "phytoplankton_chlorophyll_ratio"):
            setattr(p, n, float(getattr(pl, n)))
        p.use_temperature_coefficient = 0 if pl.temperature_coefficient is None else 1
        p.temperature_coefficient = 0.0 if pl.temperature_coefficient is None else float(pl.temperature_coefficient)
        p.light_limitation = pl.light_limitation.code
        p.phytoplankton_mortality_formulation = pl.phytoplankton_mortality_formulation.code
        p.grazing_concentration_formulation = pl.grazing_concentration_formulation.code
        if isinstance(self.nutrients, NitrateAmmonia):
            p.nitrification_rate = float(self.nutrients.nitrification_rate)
        d = self.detritus
        if isinstance(d, TwoParticleAndDissolved):
            p.remineralisation_inorganic_fraction = d.remineralisation_inorganic_fraction
            p.small_remineralisation_rate = d.small_remineralisation_rate
            p.large_remineralisation_rate = d.large_remineralisation_rate
            p.dissolved_remineralisation_rate = d.dissolved_remineralisation_rate
            p.small_solid_waste_fraction = d.small_solid_waste_fraction
            p.detritus_redfield_ratio = d.redfield_ratio
        else:
            p.remineralisation_rate = d.remineralisation_rate
            p.small_particle_fraction = d.small_particle_fraction
            p.detritus_redfield_ratio = d.redfield_ratio
        return p

    def config(self, precision=abi.OBME_FP64):
        """The obme_config for this model on self.grid (host arrays are kept
        alive by this object)."""
        g = self.grid
        cfg = abi.Config()
        cfg.model = self.model
        cfg.precision = precision
        cfg.carbonate_system = 1 if self.carbonate_system is not None else 0
        if self.oxygen is not None:
            cfg.oxygen = self.oxygen.struct()
        # host arrays referenced by the struct stay alive with this object
        zc = np.ascontiguousarray(g.zc, dtype=np.float64)
        zf = np.ascontiguousarray(g.zf, dtype=np.float64)
        dz = np.ascontiguousarray(g.dz, dtype=np.float64)
        self._keep.extend([zc, zf, dz])
        cfg.grid = abi.Grid(Nx=g.Nx, Ny=g.Ny, Nz=g.Nz, Hx=g.Hx, Hy=g.Hy, Hz=g.Hz,
                            zc=zc.ctypes.data_as(C.POINTER(C.c_double)),
                            zf=zf.ctypes.data_as(C.POINTER(C.c_double)),
                            dz=dz.ctypes.data_as(C.POINTER(C.c_double)))
        cfg.npd = self.npd_struct()
        cfg.par = self.light_attenuation.struct()
        sk = abi.Sinking()
        sp = self.sinking_speeds()
        for n, name in enumerate(self.required_biogeochemical_tracers()):
            sk.speed[n] = float(sp.get(name, 0.0))
        sk.open_bottom = 1 if self.open_bottom() else 0
        schemes = {"upwind": abi.OBME_UPWIND1, "centered": abi.OBME_CENTERED2, "weno5": abi.OBME_WENO5}
        if self.sinking_scheme not in schemes:
            raise ValueError(f"sinking_scheme: one of {sorted(schemes)}, not {self.sinking_scheme!r}")
        sk.scheme = schemes[self.sinking_scheme]
        sk.smoothing_distance = 2.0
        cfg.sinking = sk
        cc = self.gas_exchange.carbon_chemistry if self.gas_exchange is not None else CarbonChemistry()
        cfg.solver = cc.solver.struct(cc.initial_pH_guess)
        if self.gas_exchange is not None and self.carbonate_system is not None:
            cfg.co2 = self.gas_exchange.struct()
        if self.o2_exchange is not None:
            cfg.o2 = self.o2_exchange.struct()
        cfg.temperature = self._field_source(self.temperature)
        cfg.salinity = self._field_source(self.salinity)
        cfg.scaling = self.scaling_struct()
        if self.sediment is not None:
            cfg.sediment = self.sediment.struct(list(self.required_biogeochemical_tracers()))
        return cfg

    def _field_source(self, v):
        fs = abi.FieldSource()
        if v is None:
            fs.mode = abi.OBME_FIELD_CONSTANT
            fs.value = 0.0
        elif np.isscalar(v):
            fs.mode = abi.OBME_FIELD_CONSTANT
            fs.value = float(v)
        else:
            prof = np.ascontiguousarray(v, dtype=np.float64)
            assert prof.shape == (self.grid.Nz,)
            self._keep.append(prof)
            fs.mode = abi.OBME_FIELD_PROFILE
            fs.profile = prof.ctypes.data_as(C.POINTER(C.c_double))
        return fs

    def summary(self):
        name = "NPZD" if self.model == abi.OBME_NPZD else "LOBSTER"
        return f"{name} model {self.required_biogeochemical_tracers()}"

    __repr__ = summary


def NPZD(grid, nutrients=None, plankton=None, detritus=None, light_attenuation=None,
         surface_photosynthetically_active_radiation="default", **kw):
    """NPZD(grid) (NutrientsPlanktonDetritus/constructors.jl:177-227)."""
    if plankton is None:
        plankton = PhytoZoo(nitrate_half_saturation=2.3868,
                            phytoplankton_maximum_growth_rate=0.6989 / day,
                            phytoplankton_exudation_fraction=0.0,
                            temperature_coefficient=1.88,
                            phytoplankton_mortality_formulation=Linear,
                            phytoplankton_mortality_rate=(0.066 + 0.0101) / day,
                            preference_for_phytoplankton=1.0,
                            grazing_concentration_formulation=Quadratic,
                            grazing_half_saturation=0.5573,
                            zooplankton_mortality_rate=0.3395 / day,
                            zooplankton_excretion_rate=0.0102 / day,
                            zooplankton_assimilation_fraction=0.9116,
                            phytoplankton_sinking_speed=0.2551 / day,
                            excretion_inorganic_fraction=1.0,
                            phytoplankton_solid_waste_fraction=0.0101 / (0.066 + 0.0101),
                            maximum_grazing_rate=2.1522 / day,
                            light_limitation=AnalyticalLightLimitation,
                            light_half_saturation=(0.6989 / day) / (0.1953 / day))
    if light_attenuation is None:
        light_attenuation = TwoBandPhotosyntheticallyActiveRadiation(
            surface_PAR=surface_photosynthetically_active_radiation)
    return Biogeochemistry(grid, nutrients or Nutrient(), plankton, detritus or Detritus(),
                           light_attenuation=light_attenuation, **kw)


def LOBSTER(grid, nutrients=None, plankton=None, detritus=None, carbonate_system=None, oxygen=None,
            light_attenuation=None, surface_photosynthetically_active_radiation="default", **kw):
    """LOBSTER(grid) (NutrientsPlanktonDetritus/constructors.jl:65-96)."""
    if light_attenuation is None:
        light_attenuation = TwoBandPhotosyntheticallyActiveRadiation(
            surface_PAR=surface_photosynthetically_active_radiation)
    return Biogeochemistry(grid, nutrients or NitrateAmmonia(), plankton or PhytoZoo(),
                           detritus or TwoParticleAndDissolved(), carbonate_system=carbonate_system,
                           oxygen=oxygen, light_attenuation=light_attenuation, **kw)


def required_biogeochemical_tracers(bgc):
    return bgc.required_biogeochemical_tracers()


def biogeochemical_drift_velocity(bgc, name):
    return bgc.biogeochemical_drift_velocity(name)
