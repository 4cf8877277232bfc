"""ctypes mirror of include/obme.h (the C ABI of libobme_hip.so).

The structs here must match the header field for field; the library exports
obme_abi_sizeof() so the test-suite can check every size against the compiled
layout.
"""
import ctypes as C

OBME_OK, OBME_ERR_INVALID, OBME_ERR_HIP, OBME_ERR_UNSUPPORTED, OBME_ERR_NOMEM = 0, 1, 2, 3, 4
OBME_NPZD, OBME_LOBSTER = 1, 2
OBME_FP64, OBME_FP32_MIXED = 0, 1
OBME_LIGHT_MONDO, OBME_LIGHT_ANALYTICAL = 0, 1
OBME_LINEAR, OBME_QUADRATIC = 0, 1
OBME_UPWIND1, OBME_CENTERED2, OBME_WENO5 = 0, 1, 2
OBME_SURFACE_CONSTANT, OBME_SURFACE_DIURNAL, OBME_SURFACE_ARRAY = 0, 1, 2
OBME_FIELD_CONSTANT, OBME_FIELD_PROFILE, OBME_FIELD_3D = 0, 1, 2
OBME_MAX_TRACERS = 32

NPD_PARAM_NAMES = (
    "nitrate_half_saturation",
    "ammonia_half_saturation",
    "nitrate_ammonia_inhibition",
    "light_half_saturation",
    "phytoplankton_maximum_growth_rate",
    "phytoplankton_exudation_fraction",
    "ammonia_fraction_of_exudate",
    "temperature_coefficient",
    "phytoplankton_mortality_rate",
    "zooplankton_mortality_rate",
    "zooplankton_excretion_rate",
    "phytoplankton_solid_waste_fraction",
    "excretion_inorganic_fraction",
    "preference_for_phytoplankton",
    "maximum_grazing_rate",
    "grazing_half_saturation",
    "zooplankton_assimilation_fraction",
    "redfield_ratio",
    "carbon_calcite_ratio",
    "zooplankton_gut_calcite_dissolution",
    "phytoplankton_chlorophyll_ratio",
    "nitrification_rate",
    "remineralisation_inorganic_fraction",
    "small_remineralisation_rate",
    "large_remineralisation_rate",
    "dissolved_remineralisation_rate",
    "small_solid_waste_fraction",
    "detritus_redfield_ratio",
    "remineralisation_rate",
    "small_particle_fraction",
)


class NPDParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in NPD_PARAM_NAMES] + [
        ("light_limitation", C.c_int32),
        ("use_temperature_coefficient", C.c_int32),
        ("phytoplankton_mortality_formulation", C.c_int32),
        ("grazing_concentration_formulation", C.c_int32),
    ]


class TwoBandPAR(C.Structure):
    _fields_ = [
        ("water_red_attenuation", C.c_double),
        ("water_blue_attenuation", C.c_double),
        ("chlorophyll_red_attenuation", C.c_double),
        ("chlorophyll_blue_attenuation", C.c_double),
        ("chlorophyll_red_exponent", C.c_double),
        ("chlorophyll_blue_exponent", C.c_double),
        ("pigment_ratio", C.c_double),
        ("phytoplankton_chlorophyll_ratio", C.c_double),
        ("surface_mode", C.c_int32),
        ("pad0", C.c_int32),
        ("surface_value", C.c_double),
        ("surface_array", C.c_void_p),
    ]


class Sinking(C.Structure):
    _fields_ = [
        ("speed", C.c_double * OBME_MAX_TRACERS),
        ("open_bottom", C.c_int32),
        ("scheme", C.c_int32),
        ("smoothing_distance", C.c_double),
    ]


class CarbonSolver(C.Structure):
    _fields_ = [
        ("max_iters", C.c_int32),
        ("pad0", C.c_int32),
        ("atol", C.c_double),
        ("damping", C.c_double),
        ("armijo_constant", C.c_double),
        ("min_damping", C.c_double),
        ("initial_pH_guess", C.c_double),
    ]


class CO2Exchange(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32),
        ("pad0", C.c_int32),
        ("air_concentration", C.c_double),
        ("wind_speed", C.c_double),
        ("air_pressure", C.c_double),
        ("schmidt", C.c_double * 5),
        ("transfer_velocity", C.c_double * 3),
    ]


class Grid(C.Structure):
    _fields_ = [
        ("Nx", C.c_int32), ("Ny", C.c_int32), ("Nz", C.c_int32),
        ("Hx", C.c_int32), ("Hy", C.c_int32), ("Hz", C.c_int32),
        ("zc", C.POINTER(C.c_double)),
        ("zf", C.POINTER(C.c_double)),
        ("dz", C.POINTER(C.c_double)),
    ]


class FieldSource(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("pad0", C.c_int32),
        ("value", C.c_double),
        ("profile", C.POINTER(C.c_double)),
    ]


OBME_MAX_SCALE_GROUPS = 5
OBME_MAX_SCALE_MEMBERS = 12


class NegativeScaling(C.Structure):
    """obme_negative_scaling (ScaleNegativeTracers groups)."""
    _fields_ = [
        ("n_groups", C.c_int32),
        ("n_members", C.c_int32 * OBME_MAX_SCALE_GROUPS),
        ("member", (C.c_int32 * OBME_MAX_SCALE_MEMBERS) * OBME_MAX_SCALE_GROUPS),
        ("pad0", C.c_int32),
        ("scalefactor", (C.c_double * OBME_MAX_SCALE_MEMBERS) * OBME_MAX_SCALE_GROUPS),
        ("invalid_fill_value", C.c_double * OBME_MAX_SCALE_GROUPS),
    ]


def negative_scaling(groups, names):
    """NegativeScaling from [(member names, scalefactors, invalid_fill_value), ...]
    with member names resolved against the tracer-name tuple `names`."""
    s = NegativeScaling()
    if len(groups) > OBME_MAX_SCALE_GROUPS:
        raise ValueError("too many conserved groups")
    s.n_groups = len(groups)
    for g, (members, factors, fill) in enumerate(groups):
        if len(members) > OBME_MAX_SCALE_MEMBERS or len(members) != len(factors):
            raise ValueError("bad conserved group")
        s.n_members[g] = len(members)
        for m, (name, f) in enumerate(zip(members, factors)):
            s.member[g][m] = names.index(name)
            s.scalefactor[g][m] = float(f)
        s.invalid_fill_value[g] = float(fill)
    return s


OBME_SEDIMENT_NONE, OBME_SEDIMENT_INSTANT_REMINERALISATION, OBME_SEDIMENT_SIMPLE_MULTI_G = 0, 1, 2
OBME_SEDIMENT_AB2 = 0
OBME_MAX_SEDIMENT_TRACERS = 6


class Sediment(C.Structure):
    """obme_sediment (InstantRemineralisation or SimpleMultiG through
    BiogeochemicalSediment)."""
    _fields_ = [
        ("model", C.c_int32),
        ("receiver", C.c_int32),
        ("n_sinking", C.c_int32),
        ("sinking", C.c_int32 * OBME_MAX_SEDIMENT_TRACERS),
        ("timestepper", C.c_int32),
        ("burial_efficiency_constant1", C.c_double),
        ("burial_efficiency_constant2", C.c_double),
        ("burial_efficiency_half_saturation", C.c_double),
        ("ab2_chi", C.c_double),
        ("ammonia", C.c_int32),
        ("nitrate", C.c_int32),
        ("oxygen", C.c_int32),
        ("fast_decay_rate", C.c_double),
        ("slow_decay_rate", C.c_double),
        ("fast_fraction", C.c_double),
        ("slow_fraction", C.c_double),
        ("refactory_fraction", C.c_double),
        ("sedimentation_rate", C.c_double),
        ("anoxia_half_saturation", C.c_double),
        ("sinking_redfield", C.c_double),
        ("nitrate_oxidation_params", C.c_double * 6),
        ("denitrification_params", C.c_double * 6),
        ("anoxic_params", C.c_double * 6),
    ]


class O2Exchange(C.Structure):
    """obme_o2_exchange."""
    _fields_ = [
        ("enabled", C.c_int32),
        ("pad0", C.c_int32),
        ("air_concentration", C.c_double),
        ("wind_speed", C.c_double),
        ("schmidt", C.c_double * 5),
        ("transfer_velocity", C.c_double * 3),
        ("solubility", C.c_double * 6),
    ]


class Oxygen(C.Structure):
    """obme_oxygen (the Oxygen() component of LOBSTER)."""
    _fields_ = [
        ("enabled", C.c_int32),
        ("pad0", C.c_int32),
        ("respiration_oxygen_nitrogen_ratio", C.c_double),
        ("nitrification_oxygen_nitrogen_ratio", C.c_double),
    ]


class Config(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("precision", C.c_int32),
        ("carbonate_system", C.c_int32),
        ("pad0", C.c_int32),
        ("grid", Grid),
        ("npd", NPDParams),
        ("par", TwoBandPAR),
        ("sinking", Sinking),
        ("solver", CarbonSolver),
        ("co2", CO2Exchange),
        ("temperature", FieldSource),
        ("salinity", FieldSource),
        ("scaling", NegativeScaling),
        ("sediment", Sediment),
        ("oxygen", Oxygen),
        ("o2", O2Exchange),
    ]


class BatchParams(C.Structure):
    """obme_batch_params: device pointers to per-column parameter arrays."""
    _fields_ = [("fields", C.c_void_p * 40)]


OBME_PP_COUNT = 159   # obme_pisces_config.params length (2*31 + 2*18 + 61)


class PiscesConfig(C.Structure):
    """obme_pisces_config."""
    _fields_ = [
        ("precision", C.c_int32),
        ("open_bottom", C.c_int32),
        ("grid", Grid),
        ("params", C.c_double * OBME_PP_COUNT),
        ("band_kw", C.c_double * 3),
        ("band_e", C.c_double * 3),
        ("band_chi", C.c_double * 3),
        ("band_division", C.c_double * 3),
        ("surface_PAR", C.c_double),
        ("surface_mode", C.c_int32),
        ("surface_array", C.c_void_p),
        ("solver", CarbonSolver),
        ("scaling", NegativeScaling),
        ("co2", CO2Exchange),
        ("o2", O2Exchange),
        ("sediment", Sediment),
    ]


# per-column parameter rows (OBME_BP_*)
(OBME_BP_NITRATE_HALF_SATURATION, OBME_BP_LIGHT_HALF_SATURATION, OBME_BP_MAX_GROWTH_RATE,
 OBME_BP_EXUDATION_FRACTION, OBME_BP_AMMONIA_FRACTION_OF_EXUDATE, OBME_BP_TEMPERATURE_COEFFICIENT,
 OBME_BP_PHYTO_MORTALITY_RATE, OBME_BP_ZOO_MORTALITY_RATE, OBME_BP_ZOO_EXCRETION_RATE,
 OBME_BP_PHYTO_SOLID_WASTE_FRACTION, OBME_BP_EXCRETION_INORGANIC_FRACTION, OBME_BP_PREFERENCE_FOR_PHYTO,
 OBME_BP_MAX_GRAZING_RATE, OBME_BP_GRAZING_HALF_SATURATION, OBME_BP_ZOO_ASSIMILATION_FRACTION,
 OBME_BP_REMINERALISATION_RATE, OBME_BP_SMALL_PARTICLE_FRACTION, OBME_BP_P_SINKING_SPEED,
 OBME_BP_D_SINKING_SPEED, OBME_BP_SURFACE_PAR, OBME_BP_COUNT) = range(21)


# obme_abi_sizeof(which) indices (see the header)
ABI_STRUCTS = (
    ("obme_config", Config),
    ("obme_grid", Grid),
    ("obme_npd_params", NPDParams),
    ("obme_two_band_par", TwoBandPAR),
    ("obme_sinking", Sinking),
    ("obme_carbon_solver", CarbonSolver),
    ("obme_co2_exchange", CO2Exchange),
    ("obme_field_source", FieldSource),
    ("obme_batch_params", BatchParams),
    ("obme_pisces_config", PiscesConfig),
    ("obme_negative_scaling", NegativeScaling),
    ("obme_o2_exchange", O2Exchange),
    ("obme_sediment", Sediment),
    ("obme_oxygen", Oxygen),
)
