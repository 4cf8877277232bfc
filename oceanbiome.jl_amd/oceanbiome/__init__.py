"""oceanbiome — host side of the MI355X biogeochemistry step.

Mirrors the OceanBioME.jl interface for the NutrientsPlanktonDetritus models
(NPZD, LOBSTER + CarbonateSystem + Oxygen) and PISCES (oceanbiome.pisces), TwoBand and
MultiBand PAR, sinking, CarbonChemistry, the CO2 and O2 GasExchange boundary
conditions, ScaleNegativeTracers and the InstantRemineralisation and
SimpleMultiG sediments;
the compute runs in libobme_hip.so.
"""
from . import abi, units
from .grid import ColumnGrid, RectilinearGrid
from .models import (LOBSTER, NPZD, AnalyticalLightLimitation, Biogeochemistry, CarbonateSystem,
                     CarbonChemistry, CarbonDioxideGasExchangeBoundaryCondition, DampedNewtonRaphsonSolver,
                     Detritus, InstantRemineralisationSediment, Linear, MondoLightLimitation, NitrateAmmonia, Nutrient,
                     Oxygen, OxygenGasExchangeBoundaryCondition, PhytoZoo, Quadratic, SimpleMultiGSediment,
                     TwoBandPhotosyntheticallyActiveRadiation, TwoParticleAndDissolved,
                     biogeochemical_drift_velocity, default_surface_PAR, required_biogeochemical_tracers)

__all__ = [
    "abi", "units", "RectilinearGrid", "ColumnGrid", "NPZD", "LOBSTER", "Biogeochemistry", "PhytoZoo",
    "Nutrient", "NitrateAmmonia", "Detritus", "TwoParticleAndDissolved", "CarbonateSystem", "CarbonChemistry",
    "CarbonDioxideGasExchangeBoundaryCondition", "OxygenGasExchangeBoundaryCondition", "DampedNewtonRaphsonSolver",
    "TwoBandPhotosyntheticallyActiveRadiation", "Linear", "Quadratic", "MondoLightLimitation",
    "AnalyticalLightLimitation", "default_surface_PAR", "required_biogeochemical_tracers",
    "biogeochemical_drift_velocity", "InstantRemineralisationSediment", "SimpleMultiGSediment", "Oxygen",
]


def HIPBiogeochemistry(*args, **kwargs):
    """Device biogeochemistry (see oceanbiome.runtime.HIPBiogeochemistry)."""
    from .runtime import HIPBiogeochemistry as _H
    return _H(*args, **kwargs)
