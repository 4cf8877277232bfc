"""oceanbiome — host side of the MI355X biogeochemistry step.

Mirrors the OceanBioME.jl interface for the NutrientsPlanktonDetritus models
(NPZD, LOBSTER + CarbonateSystem), TwoBand PAR, sinking, CarbonChemistry and
the CO2 GasExchange boundary condition; the compute runs in libobme_hip.so.
"""
from . import abi, units
from .grid import ColumnGrid, RectilinearGrid
from .models import (LOBSTER, NPZD, AnalyticalLightLimitation, Biogeochemistry, CarbonateSystem,
                     CarbonChemistry, CarbonDioxideGasExchangeBoundaryCondition, DampedNewtonRaphsonSolver,
                     Detritus, Linear, MondoLightLimitation, NitrateAmmonia, Nutrient, PhytoZoo, Quadratic,
                     TwoBandPhotosyntheticallyActiveRadiation, TwoParticleAndDissolved,
                     biogeochemical_drift_velocity, default_surface_PAR, required_biogeochemical_tracers)

__all__ = [
    "abi", "units", "RectilinearGrid", "ColumnGrid", "NPZD", "LOBSTER", "Biogeochemistry", "PhytoZoo",
    "Nutrient", "NitrateAmmonia", "Detritus", "TwoParticleAndDissolved", "CarbonateSystem", "CarbonChemistry",
    "CarbonDioxideGasExchangeBoundaryCondition", "DampedNewtonRaphsonSolver",
    "TwoBandPhotosyntheticallyActiveRadiation", "Linear", "Quadratic", "MondoLightLimitation",
    "AnalyticalLightLimitation", "default_surface_PAR", "required_biogeochemical_tracers",
    "biogeochemical_drift_velocity",
]


def HIPBiogeochemistry(*args, **kwargs):
    """Device biogeochemistry (see oceanbiome.runtime.HIPBiogeochemistry)."""
    from .runtime import HIPBiogeochemistry as _H
    return _H(*args, **kwargs)
