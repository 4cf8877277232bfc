"""PISCES host description: parameters, tracer list, MultiBand light
coefficients and the C-ABI configuration (obme_pisces_config).

Mirrors `PISCES(; grid, ...)` of OceanBioME.jl
(src/Models/AdvectedPopulations/PISCES/PISCES.jl:268-330) with its default
component constructors:
  MixedMondoNanoAndDiatoms   (phytoplankton/mixed_mondo_nano_diatoms.jl:1-28,
                              mixed_mondo.jl:44-76, growth_rate.jl:76-102,
                              nutrient_limitation.jl:9-18)
  MicroAndMesoZooplankton    (zooplankton/defaults.jl:1-22,
                              food_quality_dependant.jl:77-118, micro_and_meso.jl:3-15)
  DissolvedOrganicCarbon     (dissolved_organic_matter/dissolved_organic_carbon.jl:18-36)
  TwoCompartmentCarbonIronParticles (particulate_organic_matter/two_size_class.jl:46-90)
  NitrateAmmonia, SimpleIron, Oxygen (nitrogen/nitrate_ammonia.jl:9-15,
                              iron/simple_iron.jl:9-13, oxygen.jl:15-18)
  MultiBandPhotosyntheticallyActiveRadiation (Light/multi_band.jl:82-140) with
  the Morel (1988) coefficient table (Light/morel_coefficients.jl).
"""
import ctypes as C

import numpy as np

from . import abi
from .units import day

TRACERS = ("P", "PChl", "PFe", "D", "DChl", "DFe", "DSi", "Z", "M", "DOC", "POC", "GOC", "SFe", "BFe", "PSi",
           "CaCO₃", "NO₃", "NH₄", "PO₄", "Fe", "Si", "DIC", "Alk", "O₂", "T", "S")
N_WRITTEN = 24

# Morel (1988) spectral coefficients at 5 nm spacing, 350-700 nm: water
# attenuation, chlorophyll exponent, chlorophyll attenuation
MOREL_LAMBDA = np.arange(350, 701, 5, dtype=np.float64)
MOREL_KW = np.array([
    0.0271, 0.0238, 0.0216, 0.0188, 0.0177, 0.01595, 0.0151, 0.01376, 0.01271, 0.01208,
    0.01042, 0.0089, 0.00812, 0.00765, 0.00758, 0.00768, 0.0077, 0.00792, 0.00885, 0.0099,
    0.01148, 0.01182, 0.01188, 0.01211, 0.01251, 0.0132, 0.01444, 0.01526, 0.0166, 0.01885,
    0.02188, 0.02701, 0.03385, 0.0409, 0.04214, 0.04287, 0.04454, 0.0463, 0.04846, 0.05212,
    0.05746, 0.06053, 0.0628, 0.06507, 0.07034, 0.07801, 0.09038, 0.11076, 0.13584, 0.16792,
    0.2231, 0.25838, 0.26506, 0.26843, 0.27612, 0.284, 0.29218, 0.30176, 0.31134, 0.32553,
    0.34052, 0.3715, 0.41048, 0.42947, 0.43946, 0.44844, 0.46543, 0.48642, 0.5164, 0.55939, 0.62438])
MOREL_E = np.array([
    0.778, 0.767, 0.756, 0.737, 0.72, 0.7, 0.685, 0.673, 0.67, 0.66,
    0.64358, 0.64776, 0.65175, 0.65555, 0.65917, 0.66259, 0.66583, 0.66889, 0.67175, 0.67443,
    0.67692, 0.67923, 0.68134, 0.68327, 0.68501, 0.68657, 0.68794, 0.68903, 0.68955, 0.68947,
    0.6888, 0.68753, 0.68567, 0.6832, 0.68015, 0.67649, 0.67224, 0.66739, 0.66195, 0.65591,
    0.64927, 0.64204, 0.64, 0.63, 0.623, 0.615, 0.61, 0.614, 0.618, 0.622,
    0.626, 0.63, 0.634, 0.638, 0.642, 0.647, 0.653, 0.658, 0.663, 0.667,
    0.672, 0.677, 0.682, 0.687, 0.695, 0.697, 0.693, 0.665, 0.64, 0.62, 0.6])
MOREL_CHI = np.array([
    0.153, 0.149, 0.144, 0.14, 0.136, 0.131, 0.127, 0.123, 0.119, 0.118,
    0.11748, 0.12066, 0.12259, 0.12326, 0.12269, 0.12086, 0.11779, 0.11372, 0.10963, 0.1056,
    0.10165, 0.09776, 0.09393, 0.09018, 0.08649, 0.08287, 0.07932, 0.07584, 0.07242, 0.06907,
    0.06579, 0.06257, 0.05943, 0.05635, 0.05341, 0.05072, 0.04829, 0.04611, 0.04419, 0.04253,
    0.04111, 0.03996, 0.039, 0.0375, 0.036, 0.034, 0.033, 0.0328, 0.0325, 0.033,
    0.034, 0.035, 0.036, 0.0375, 0.0385, 0.04, 0.042, 0.043, 0.044, 0.0445,
    0.045, 0.046, 0.0475, 0.049, 0.0515, 0.052, 0.0505, 0.044, 0.039, 0.034, 0.03])


def _band_mean(lam, c, i1, i2):
    """Trapezoid mean of c over lam[i1..i2] (multi_band.jl:136-140)."""
    s = sum((c[n] + c[n - 1]) * (lam[n] - lam[n - 1]) / 2 for n in range(i1 + 1, i2 + 1))
    return s / (lam[i2] - lam[i1])


def morel_band_coefficients(bands=((400, 500), (500, 600), (600, 700)), lam=MOREL_LAMBDA, kw=MOREL_KW,
                            e=MOREL_E, chi=MOREL_CHI):
    """(kʷ, e, χ) per band (multi_band.jl:89-103): means between the last
    table wavelengths not above each band edge."""
    out_kw, out_e, out_chi = [], [], []
    for b0, b1 in bands:
        i1 = int(np.nonzero(lam <= b0)[0][-1])
        i2 = int(np.nonzero(lam <= b1)[0][-1])
        out_kw.append(_band_mean(lam, kw, i1, i2))
        out_e.append(_band_mean(lam, e, i1, i2))
        out_chi.append(_band_mean(lam, chi, i1, i2))
    return out_kw, out_e, out_chi


# Parameter layout of obme_pisces_config.params (the OBME_PP_* indices of
# obme.h): two phytoplankton blocks, two zooplankton blocks, then scalars.
PHYTO_FIELDS = ("mu0", "bt", "dark_tolerance", "alpha0", "beta", "basal_respiration", "reference_growth",
                "kNH4", "kNO3", "kPO4", "theta_Fe_opt", "si_limited", "kSi_min", "pk", "delta", "beta1",
                "beta2", "beta3", "mort_K", "mort_lin", "mort_w0", "mort_w1", "chl_min", "chl_max", "fe_max",
                "si_K1", "si_K2", "si_theta0", "kFe_uptake", "size_threshold", "size_ratio")
ZOO_FIELDS = ("b", "g0", "pref_P", "pref_D", "pref_POC", "pref_Z", "food_threshold", "J", "K", "ff", "iron_ratio",
              "e0", "sigma", "Km", "m0", "r", "gamma", "eta")
SCALAR_FIELDS = ("rain_ratio", "bZ", "bM", "bact_exp", "bact_max", "K_DOC", "K_NO3_b", "K_NH4_b", "K_PO4_b",
                 "K_Fe_b", "doc_lambda", "doc_bact_ref", "doc_b", "doc_a1", "doc_a2", "doc_a3", "doc_a4", "doc_a5",
                 "pom_b", "pom_lambda", "pom_a1", "pom_a2", "pom_a3", "pom_a4", "fe_scav_min", "fe_scav_load",
                 "bact_fe_eff", "small_bact_fe", "large_bact_fe", "chi0", "si_fast", "si_slow", "caco3_lambda",
                 "caco3_exp", "bact_fe_max", "bact_fe_K", "bact_mu0", "nitrif", "fix_max", "fix_KFe", "fix_KPO4",
                 "fix_E", "fe_excess", "fe_ligand_max", "fe_ligand_ratio", "O2_resp", "O2_nitrif", "anoxia1",
                 "anoxia2", "thetaN", "thetaP", "mixed_layer_shear", "background_shear", "latitude",
                 "silicate_climatology", "mixed_layer_depth", "kappa", "POC_speed", "GOC_min_speed",
                 "GOC_max_speed", "gravity")
PARAM_COUNT = 2 * len(PHYTO_FIELDS) + 2 * len(ZOO_FIELDS) + len(SCALAR_FIELDS)


def _phyto(dark_tolerance, kNH4, kNO3, kPO4, silicate_limited, blue, green, red, max_quadratic, max_chl, kFe):
    return dict(mu0=0.6 / day, bt=1.066, dark_tolerance=dark_tolerance, alpha0=2.0, beta=0.0,
                basal_respiration=0.033 / day, reference_growth=1.0 / day, kNH4=kNH4, kNO3=kNO3, kPO4=kPO4,
                theta_Fe_opt=0.007, si_limited=1.0 if silicate_limited else 0.0, kSi_min=1.0, pk=16.6,
                delta=0.05, beta1=blue, beta2=green, beta3=red, mort_K=0.2, mort_lin=0.01 / day,
                mort_w0=0.01 / day, mort_w1=max_quadratic, chl_min=0.0033, chl_max=max_chl, fe_max=0.06,
                si_K1=2.0, si_K2=20.9, si_theta0=0.159, kFe_uptake=kFe, size_threshold=1.0, size_ratio=3.0)


def _zoo(g0, prefs, quadratic, linear, e0, flux_feeding, calcite, iron):
    return dict(b=1.079, g0=g0, pref_P=prefs[0], pref_D=prefs[1], pref_POC=prefs[2], pref_Z=prefs[3],
                food_threshold=0.3, J=0.001, K=20.0, ff=flux_feeding, iron_ratio=iron, e0=e0, sigma=0.3, Km=0.2,
                m0=quadratic, r=linear, gamma=0.6, eta=calcite)


class PISCES:
    """PISCES(; grid) with the reference defaults.  mixed_layer_depth (m,
    negative down) and the mean mixed-layer diffusivity kappa are prescribed
    constants (the reference keeps zₘₓₗ as a field the user sets and κ̄ = 1
    without a closure).  surface_PAR: a number (constant, W/m²), "default"
    (default_surface_PAR, Light.jl:47, the reference's PISCES default,
    PISCES.jl:325-331) or a 2-D array PAR⁰[j, i] of shape (Ny, Nx)."""

    def __init__(self, grid, surface_PAR="default", mixed_layer_depth=-50.0, kappa=1.0, latitude=45.0,
                 bands=((400, 500), (500, 600), (600, 700)), surface_PAR_division=None, POC_speed=2 / day,
                 open_bottom=True, scale_negatives=False, invalid_fill_value=float("nan"),
                 co2_exchange=None, o2_exchange=None, sediment=None, **overrides):
        """co2_exchange / o2_exchange: the top boundary conditions of DIC and O2
        (a CarbonDioxideGasExchangeBoundaryCondition / OxygenGasExchangeBoundaryCondition,
        or True for the defaults; validation/PISCES/column.jl:63-64).  sediment:
        an InstantRemineralisationSediment over PISCES's tracers (PISCES.jl:233),
        e.g. sinking_tracers=("POC", "GOC", "CaCO₃"), remineralisation_reciever="DIC"."""
        from .models import CarbonDioxideGasExchangeBoundaryCondition, OxygenGasExchangeBoundaryCondition
        self.grid = grid
        def choose(x, make):
            return make() if x is True else (None if x is False else x)

        self.co2_exchange = choose(co2_exchange, CarbonDioxideGasExchangeBoundaryCondition)
        self.o2_exchange = choose(o2_exchange, OxygenGasExchangeBoundaryCondition)
        self.scale_negatives = bool(scale_negatives)
        self.invalid_fill_value = float(invalid_fill_value)
        if isinstance(surface_PAR, str):
            if surface_PAR != "default":
                raise ValueError(f"surface_PAR: a number, 'default' or a 2-D array, not {surface_PAR!r}")
            self.surface_PAR = "default"
        elif np.isscalar(surface_PAR):
            self.surface_PAR = float(surface_PAR)
        else:
            arr = np.ascontiguousarray(surface_PAR, dtype=np.float64)
            if arr.shape != (grid.Ny, grid.Nx):
                raise ValueError(f"surface_PAR array must have shape (Ny, Nx) = {(grid.Ny, grid.Nx)}")
            self.surface_PAR = arr
        self.nano = _phyto(3 * day, 0.013, 0.13, 0.8, False, 2.1, 0.42, 0.4, 0.0, 0.033, 1.0)
        self.diatoms = _phyto(4 * day, 0.039, 0.39, 2.4, True, 1.6, 0.69, 0.7, 0.03 / day, 0.05, 3.0)
        self.micro = _zoo(3 / day, (1.0, 0.5, 0.1, 0.0), 0.004 / day, 0.03 / day, 0.3, 0.0, 0.5, 0.01)
        self.meso = _zoo(0.75 / day, (0.3, 1.0, 0.3, 1.0), 0.03 / day, 0.005 / day, 0.35, 2e3 / 1e6, 0.75, 0.015)
        agg_doc = [x * (1e-6 / day) for x in (0.37, 102, 3530, 5095, 114)]
        agg_pom = [x * (1e-6 / day) for x in (25.9, 4452, 3.3, 47.1)]
        self.scalars = dict(
            rain_ratio=0.3, bZ=0.7, bM=1.4, bact_exp=0.684, bact_max=4.0, K_DOC=417.0, K_NO3_b=0.03, K_NH4_b=0.003,
            K_PO4_b=0.003, K_Fe_b=0.01, doc_lambda=0.3 / day, doc_bact_ref=1.0, doc_b=1.066,
            doc_a1=agg_doc[0], doc_a2=agg_doc[1], doc_a3=agg_doc[2], doc_a4=agg_doc[3], doc_a5=agg_doc[4],
            pom_b=1.066, pom_lambda=0.025 / day, pom_a1=agg_pom[0], pom_a2=agg_pom[1], pom_a3=agg_pom[2],
            pom_a4=agg_pom[3], fe_scav_min=3e-5 / day, fe_scav_load=0.005 / day, bact_fe_eff=0.16,
            small_bact_fe=0.12 / 0.16, large_bact_fe=0.04 / 0.16, chi0=0.5, si_fast=0.025 / day,
            si_slow=0.003 / day, caco3_lambda=0.197 / day, caco3_exp=1.0, bact_fe_max=0.06, bact_fe_K=0.3,
            bact_mu0=0.6 / day, nitrif=0.05 / day, fix_max=0.013 / day, fix_KFe=0.1, fix_KPO4=0.8, fix_E=50.0,
            fe_excess=1000.0, fe_ligand_max=0.6, fe_ligand_ratio=0.09, O2_resp=133 / 122, O2_nitrif=32 / 122,
            anoxia1=6.0, anoxia2=1.0, thetaN=16 / 122, thetaP=1 / 122, mixed_layer_shear=1.0,
            background_shear=0.01, latitude=float(latitude), silicate_climatology=7.5,
            mixed_layer_depth=float(mixed_layer_depth), kappa=float(kappa), POC_speed=float(POC_speed),
            GOC_min_speed=30 / day, GOC_max_speed=200 / day, gravity=9.80665)
        for k, v in overrides.items():
            if k in self.scalars:
                self.scalars[k] = float(v)
            else:
                raise KeyError(k)
        self.kw, self.e, self.chi = morel_band_coefficients(bands)
        nb = len(bands)
        if nb != 3:
            raise NotImplementedError("the device path has three PAR bands")
        self.division = [1 / nb] * nb if surface_PAR_division is None else list(surface_PAR_division)
        self.open_bottom = bool(open_bottom)
        self.sediment = sediment
        from .models import CarbonChemistry
        self.carbon_chemistry = CarbonChemistry()
        self._keep = []

    # Oceananigans.Biogeochemistry API -------------------------------------------
    def required_biogeochemical_tracers(self):
        return TRACERS

    read_only_3d = 2   # T and S are 3-D tracers the step reads

    def written_tracers(self):
        return TRACERS[:N_WRITTEN]

    def required_biogeochemical_auxiliary_fields(self):
        return ("zₘₓₗ", "zₑᵤ", "Si′", "Ω", "κ", "mixed_layer_PAR", "wPOC", "wGOC", "PAR", "PAR₁", "PAR₂", "PAR₃")

    def conserved_tracers(self):
        """conserved_tracers(bgc::PISCES) (PISCES/coupling_utils.jl:11-35):
        carbon, iron, phosphate, silicon, nitrogen — [(names, scalefactors)]."""
        thP, thN = self.scalars["thetaP"], self.scalars["thetaN"]
        carbon = ["P", "D", "Z", "M", "DOC", "POC", "GOC", "DIC", "CaCO₃"]
        iron = (["PFe", "DFe", "Z", "M", "SFe", "BFe", "Fe"],
                [1.0, 1.0, self.micro["iron_ratio"], self.meso["iron_ratio"], 1.0, 1.0, 1.0])
        phosphate = (["P", "D", "Z", "M", "DOC", "POC", "GOC", "PO₄"], [thP] * 7 + [1.0])
        silicon = ["DSi", "Si", "PSi"]
        nitrogen = (["NH₄", "NO₃", "P", "D", "Z", "M", "DOC", "POC", "GOC"], [1.0, 1.0] + [thN] * 7)
        return [(carbon, [1.0] * len(carbon)), iron, phosphate, (silicon, [1.0] * 3), nitrogen]

    def scaling_struct(self):
        """obme_negative_scaling of PISCES(; scale_negatives = true) (PISCES.jl:392-401)."""
        if not self.scale_negatives:
            return abi.NegativeScaling()
        groups = [(n, f, self.invalid_fill_value) for n, f in self.conserved_tracers()]
        return abi.negative_scaling(groups, TRACERS)

    def param_vector(self):
        v = [self.nano[f] for f in PHYTO_FIELDS] + [self.diatoms[f] for f in PHYTO_FIELDS]
        v += [self.micro[f] for f in ZOO_FIELDS] + [self.meso[f] for f in ZOO_FIELDS]
        v += [self.scalars[f] for f in SCALAR_FIELDS]
        return np.array(v, dtype=np.float64)

    def config(self, precision=abi.OBME_FP64):
        g = self.grid
        cfg = abi.PiscesConfig()
        cfg.precision = precision
        zc = np.ascontiguousarray(g.zc, dtype=np.float64)
        zf = np.ascontiguousarray(g.zf, dtype=np.float64)
        dz = np.ascontiguousarray(g.dz, dtype=np.float64)
        self._keep.extend([zc, zf, dz])
        cfg.grid = abi.Grid(Nx=g.Nx, Ny=g.Ny, Nz=g.Nz, Hx=g.Hx, Hy=g.Hy, Hz=g.Hz,
                            zc=zc.ctypes.data_as(C.POINTER(C.c_double)),
                            zf=zf.ctypes.data_as(C.POINTER(C.c_double)),
                            dz=dz.ctypes.data_as(C.POINTER(C.c_double)))
        pv = self.param_vector()
        for n, x in enumerate(pv):
            cfg.params[n] = float(x)
        for n in range(3):
            cfg.band_kw[n] = self.kw[n]
            cfg.band_e[n] = self.e[n]
            cfg.band_chi[n] = self.chi[n]
            cfg.band_division[n] = self.division[n]
        if isinstance(self.surface_PAR, str):
            cfg.surface_mode = abi.OBME_SURFACE_DIURNAL
        elif isinstance(self.surface_PAR, np.ndarray):
            cfg.surface_mode = abi.OBME_SURFACE_ARRAY   # the device array comes from the runtime
        else:
            cfg.surface_mode = abi.OBME_SURFACE_CONSTANT
            cfg.surface_PAR = self.surface_PAR
        cfg.open_bottom = 1 if self.open_bottom else 0
        cc = self.carbon_chemistry
        cfg.solver = cc.solver.struct(cc.initial_pH_guess)
        cfg.scaling = self.scaling_struct()
        if self.co2_exchange is not None:
            cfg.co2 = self.co2_exchange.struct()
        if self.o2_exchange is not None:
            cfg.o2 = self.o2_exchange.struct()
        if self.sediment is not None:
            cfg.sediment = self.sediment.struct(list(self.required_biogeochemical_tracers()))
        return cfg

    def __repr__(self):
        g = self.grid
        return f"PISCES on {g.Nx}x{g.Ny}x{g.Nz}, 24 tracers + T, S"
