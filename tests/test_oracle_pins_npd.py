"""Special-state pins of the FP64 oracle's NPZD and LOBSTER (+ CarbonateSystem)
transitions, term by term.

Each pin evaluates orc_npd_tendencies at a state where the reference's
expressions collapse to a short closed form: Z = 0 switches grazing off,
P = sPOM makes the preference-weighted food equal to P and the preference
the split of the grazing, PAR = kPAR puts the Mondo light limitation at 1/2
and the analytical one at 1/sqrt(2), N = kN the nutrient limitation at 1/2,
T = 10 the Q10 factor at Q10, NH4 = 0 or NO3 = 0 leaves one branch of
LOBSTER's nitrogen limitation, NH4 < 0 takes the max(0, .) branch.  The
expected values are written from the reference's definitions (cited per
pin, paths under src/Models/AdvectedPopulations/NutrientsPlanktonDetritus),
with generic parameters: every parameter distinct and no fraction at 0 or
1, so no term vanishes or coincides with another by accident.

test_planted_mistakes_are_caught compiles the oracle with one plausible
mistake at a time (a dropped term, a flipped sign, a wrong parameter,
exponent or branch) and checks that at least one pin fails for each, so
the pins fix every term of the transitions, which the reference's own
invariant tests (zero state, nitrogen and carbon budgets) cannot.  CPU only."""
import ctypes as C
import math
import shutil

import numpy as np
import pytest

import mutants
import oracle as O
from oceanbiome import abi

NPZD, LOB = abi.OBME_NPZD, abi.OBME_LOBSTER

# generic parameters (PhytoZoo plankton.jl:19-58, NitrateAmmonia nutrients.jl:17-19,
# TwoParticleAndDissolved detritus.jl:25-37, Detritus detritus.jl:278-283)
P_ = dict(
    nitrate_half_saturation=0.7, ammonia_half_saturation=0.013, nitrate_ammonia_inhibition=2.9,
    light_half_saturation=31.0, phytoplankton_maximum_growth_rate=2.3e-5, phytoplankton_exudation_fraction=0.07,
    ammonia_fraction_of_exudate=0.71, temperature_coefficient=1.83, phytoplankton_mortality_rate=6.1e-7,
    zooplankton_mortality_rate=2.2e-6, zooplankton_excretion_rate=5.3e-7, phytoplankton_solid_waste_fraction=0.83,
    excretion_inorganic_fraction=0.47, preference_for_phytoplankton=0.61, maximum_grazing_rate=9.1e-6,
    grazing_half_saturation=0.93, zooplankton_assimilation_fraction=0.68, redfield_ratio=6.56,
    carbon_calcite_ratio=0.11, zooplankton_gut_calcite_dissolution=0.3, phytoplankton_chlorophyll_ratio=1.31,
    nitrification_rate=5.6e-7, remineralisation_inorganic_fraction=0.23, small_remineralisation_rate=5.7e-7,
    large_remineralisation_rate=6.1e-7, dissolved_remineralisation_rate=3.7e-7, small_solid_waste_fraction=0.44,
    detritus_redfield_ratio=6.9, remineralisation_rate=1.3e-6, small_particle_fraction=0.45)
kN, kNH4, psi, kPAR = (P_[k] for k in ("nitrate_half_saturation", "ammonia_half_saturation",
                                        "nitrate_ammonia_inhibition", "light_half_saturation"))
mu0, gam, ax, Q10 = (P_[k] for k in ("phytoplankton_maximum_growth_rate", "phytoplankton_exudation_fraction",
                                      "ammonia_fraction_of_exudate", "temperature_coefficient"))
mP, mZ, muZ, aPs, aZe = (P_[k] for k in ("phytoplankton_mortality_rate", "zooplankton_mortality_rate",
                                          "zooplankton_excretion_rate", "phytoplankton_solid_waste_fraction",
                                          "excretion_inorganic_fraction"))
pref, g, kG, aZa = (P_[k] for k in ("preference_for_phytoplankton", "maximum_grazing_rate",
                                     "grazing_half_saturation", "zooplankton_assimilation_fraction"))
R, rho, nr = P_["redfield_ratio"], P_["carbon_calcite_ratio"], P_["nitrification_rate"]
a_, smu, bmu, dmu, ss, Rd = (P_[k] for k in ("remineralisation_inorganic_fraction", "small_remineralisation_rate",
                                              "large_remineralisation_rate", "dissolved_remineralisation_rate",
                                              "small_solid_waste_fraction", "detritus_redfield_ratio"))
rem, fs = P_["remineralisation_rate"], P_["small_particle_fraction"]


def params(light=abi.OBME_LIGHT_MONDO, q10=0, mort=abi.OBME_QUADRATIC, graz=abi.OBME_QUADRATIC):
    p = abi.NPDParams()
    for k, v in P_.items():
        setattr(p, k, v)
    p.light_limitation = light
    p.use_temperature_coefficient = q10
    p.phytoplankton_mortality_formulation = mort
    p.grazing_concentration_formulation = graz
    return p


def tend(L, model, p, vals, PAR):
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    out = np.zeros(9)
    L.orc_npd_tendencies(model, 1 if model == LOB else 0, C.byref(p), O._dp(vals), float(PAR), O._dp(out))
    return out[:5] if model == NPZD else out


def lobster_expected(mu, up_NO3, up_NH4, nu, NH4, Z, total, onP, onS, sPOM, bPOM, DOM):
    """The nine LOBSTER transitions from their ingredients (nutrients.jl:52-67,
    plankton.jl:92-116,280-346,429-450, detritus.jl:229-287,
    carbonate_system.jl:49-62); used where the pin fixes the ingredients."""
    nit = nr * NH4
    pinw = aZe * muZ * Z + (1 - aPs) * nu
    dinw = a_ * (smu * sPOM + bmu * bPOM) + dmu * DOM
    ponw = (1 - ax) * gam * mu + (1 - aZe) * muZ * Z
    donw = (1 - a_) * (smu * sPOM + bmu * bPOM)
    solid = (1 - aZa) * total + aPs * nu + mZ * Z * Z
    dNO3 = nit - up_NO3
    dNH4 = pinw + dinw - nit - up_NH4
    cdis = (onP + nu) * rho * R
    dDIC = -(1 + rho * (1 - gam) - ax * gam) * mu * R + R * pinw + Rd * dinw + cdis
    dAlk = dNH4 * 15 / 16 - dNO3 * 17 / 16 - 2 * (2 * rho * mu * R) + 2 * cdis
    return np.array([dNO3, dNH4, (1 - gam) * mu - onP - nu, aZa * total - mZ * Z * Z - muZ * Z,
                     ss * solid - onS - smu * sPOM, (1 - ss) * solid - bmu * bPOM, ponw + donw - dmu * DOM,
                     dDIC, dAlk])


# ---- the pins: name -> function(L) returning (got, expected) ----------------

def pin_npzd_growth_analytical_light(L):
    """Z = D = 0: growth and mortality only.  Analytical light at PAR = kPAR
    is 1/sqrt(2) (plankton.jl:220), N = kN gives 1/2 (:180-188), T = 10 gives
    Q10 (:210-214); linear mortality (:86); N, D from nutrients.jl:86-90,
    detritus.jl:373-379."""
    P = 0.3
    mu = mu0 * (1 / math.sqrt(2)) * 0.5 * Q10 * P
    nu = mP * P
    exp = [(1 - aPs) * nu - mu * (1 - ax * gam), (1 - gam) * mu - nu, 0.0, 0.0, (1 - ax) * gam * mu + aPs * nu]
    got = tend(L, NPZD, params(abi.OBME_LIGHT_ANALYTICAL, 1, abi.OBME_LINEAR), [kN, P, 0.0, 10.0, 0.0], kPAR)
    return got, exp


def pin_npzd_growth_mondo_quadratic(L):
    """PAR = 3 kPAR: Mondo light 3/4 (plankton.jl:219); N = 3 kN: 3/4;
    T = 20: Q10^2; quadratic mortality (:87)."""
    P = 0.2
    mu = mu0 * 0.75 * 0.75 * Q10 ** 2 * P
    nu = mP * P * P
    exp = [(1 - aPs) * nu - mu * (1 - ax * gam), (1 - gam) * mu - nu, 0.0, 0.0, (1 - ax) * gam * mu + aPs * nu]
    got = tend(L, NPZD, params(abi.OBME_LIGHT_MONDO, 1, abi.OBME_QUADRATIC), [3 * kN, P, 0.0, 20.0, 0.0],
               3 * kPAR)
    return got, exp


def pin_npzd_without_q10(L):
    """temperature_coefficient = nothing: the factor is one at any T (plankton.jl:208)."""
    P = 0.2
    mu = mu0 * 0.5 * 0.5 * P
    nu = mP * P * P
    exp = [(1 - aPs) * nu - mu * (1 - ax * gam), (1 - gam) * mu - nu, 0.0, 0.0, (1 - ax) * gam * mu + aPs * nu]
    got = tend(L, NPZD, params(abi.OBME_LIGHT_MONDO, 0, abi.OBME_QUADRATIC), [kN, P, 0.0, 17.0, 0.0], kPAR)
    return got, exp


def pin_npzd_grazing_on_phytoplankton(L):
    """D = 0 and PAR = 0: no growth, the preference is 1 (plankton.jl:393-397),
    the food is P and the quadratic concentration limit P^2/(P^2 + kG^2)
    (:89-90); grazing g L Z (:118-135, :352-369); the Z, N and D budgets of
    excretion, mortality and solid waste (:107-116, :292-346)."""
    P, Z = 0.25, 0.4
    Lg = P * P / (P * P + kG * kG)
    G = g * Lg * Z
    nu = mP * P
    exp = [aZe * muZ * Z + (1 - aPs) * nu, -G - nu, aZa * G - mZ * Z * Z - muZ * Z, 0.0,
           (1 - aZe) * muZ * Z + (1 - aZa) * G + aPs * nu + mZ * Z * Z]
    got = tend(L, NPZD, params(abi.OBME_LIGHT_ANALYTICAL, 1, abi.OBME_LINEAR), [1.0, P, Z, 10.0, 0.0], 0.0)
    return got, exp


def pin_npzd_grazing_on_detritus_linear(L):
    """P = 0: the zooplankton graze the small fraction of D alone
    (detritus.jl:386-387, plankton.jl:371-389), linear concentration limit
    sPOM/(sPOM + kG); D remineralises at its rate into N (detritus.jl:389-390)."""
    D, Z = 0.8, 0.3
    sPOM = fs * D
    Lg = sPOM / (sPOM + kG)
    G = g * Lg * Z
    exp = [aZe * muZ * Z + rem * D, 0.0, aZa * G - mZ * Z * Z - muZ * Z, 0.0,
           (1 - aZe) * muZ * Z + (1 - aZa) * G + mZ * Z * Z - G - rem * D]
    got = tend(L, NPZD, params(abi.OBME_LIGHT_ANALYTICAL, 1, abi.OBME_LINEAR, abi.OBME_LINEAR),
               [1.0, 0.0, Z, 10.0, D], 50.0)
    return got, exp


def pin_npzd_grazing_split(L):
    """P = sPOM = c: the preference-weighted food is c and the grazing splits
    pref : (1 - pref) between P and D (plankton.jl:352-397)."""
    c, Z = 0.3, 0.5
    D = c / fs
    Lg = c * c / (c * c + kG * kG)
    tot = g * Lg * Z
    onP, onS = g * pref * Lg * Z, g * (1 - pref) * Lg * Z
    nu = mP * c
    exp = [aZe * muZ * Z + (1 - aPs) * nu + rem * D, -onP - nu, aZa * tot - mZ * Z * Z - muZ * Z, 0.0,
           (1 - aZe) * muZ * Z + (1 - aZa) * tot + aPs * nu + mZ * Z * Z - onS - rem * D]
    got = tend(L, NPZD, params(abi.OBME_LIGHT_ANALYTICAL, 1, abi.OBME_LINEAR), [1.0, c, Z, 10.0, D], 0.0)
    return got, exp


def pin_lobster_nitrate_only(L):
    """NH4 = 0: nitrate limitation NO3/(NO3 + kNO3) = 1/2 at NO3 = kNO3, the
    ammonia term 0, their mean 1/4 (plankton.jl:150-166); Mondo light 1/2;
    all uptake is nitrate (:223-239) and the ammonia uptake is the exudate
    alone, -alpha gamma mu (:241-261); quadratic mortality."""
    P = 0.2
    mu = mu0 * 0.5 * 0.25 * P
    nu = mP * P * P
    exp = lobster_expected(mu, mu, -ax * gam * mu, nu, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    got = tend(L, LOB, params(), [kN, 0.0, P, 0.0, 0.0, 0.0, 0.0, 2100.0, 2300.0], kPAR)
    return got, exp


def pin_lobster_ammonia_only(L):
    """NO3 = 0, NH4 = kNH4: ammonia limitation 1/2, mean 1/4; Mondo light
    3/4 at PAR = 3 kPAR; all uptake is ammonia; nitrification nr NH4
    (nutrients.jl:50)."""
    P = 0.2
    mu = mu0 * 0.75 * 0.25 * P
    nu = mP * P * P
    exp = lobster_expected(mu, 0.0, mu - ax * gam * mu, nu, kNH4, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    got = tend(L, LOB, params(), [0.0, kNH4, P, 0.0, 0.0, 0.0, 0.0, 2100.0, 2300.0], 3 * kPAR)
    return got, exp


def pin_lobster_inhibition(L):
    """NO3 = kNO3, NH4 = kNH4: nitrate limitation exp(-psi kNH4)/2 and
    ammonia limitation 1/2; the uptakes split in their ratio."""
    P = 0.15
    nl, al = 0.5 * math.exp(-psi * kNH4), 0.5
    mu = mu0 * 0.5 * ((nl + al) / 2) * P
    nu = mP * P * P
    exp = lobster_expected(mu, mu * nl / (nl + al), mu * al / (nl + al) - ax * gam * mu, nu, kNH4,
                           0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    got = tend(L, LOB, params(), [kN, kNH4, P, 0.0, 0.0, 0.0, 0.0, 2100.0, 2300.0], kPAR)
    return got, exp


def pin_lobster_negative_ammonia(L):
    """NH4 = -kNH4/2: NH4/(kNH4 + NH4) = -1 and max(0, -1) = 0 (plankton.jl:161);
    the inhibition exp(+psi kNH4/2); NO3 = 2 kNO3."""
    P = 0.15
    NH4 = -kNH4 / 2
    nl = (2.0 / 3.0) * math.exp(psi * kNH4 / 2)
    mu = mu0 * 0.5 * (nl / 2) * P
    nu = mP * P * P
    exp = lobster_expected(mu, mu, -ax * gam * mu, nu, NH4, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    got = tend(L, LOB, params(), [2 * kN, NH4, P, 0.0, 0.0, 0.0, 0.0, 2100.0, 2300.0], kPAR)
    return got, exp


def pin_lobster_detritus_only(L):
    """P = Z = 0: remineralisation alone (detritus.jl:229-287): the inorganic
    fraction of the particles and all of DOM to NH4, the rest of the
    particles to DOM; DIC at the detritus Redfield ratio."""
    sP, bP, DOM = 0.3, 0.2, 0.4
    dinw = a_ * (smu * sP + bmu * bP) + dmu * DOM
    exp = [0.0, dinw, 0.0, 0.0, -smu * sP, -bmu * bP, (1 - a_) * (smu * sP + bmu * bP) - dmu * DOM,
           Rd * dinw, dinw * 15 / 16]
    got = tend(L, LOB, params(), [0.0, 0.0, 0.0, 0.0, sP, bP, DOM, 2100.0, 2300.0], 80.0)
    return got, exp


def pin_lobster_zooplankton_only(L):
    """Only Z: no food, so no grazing; excretion split into NH4 and DOM, the
    quadratic mortality into the particles (plankton.jl:107-116,292-346)."""
    Z = 0.35
    exc = muZ * Z
    exp = [0.0, aZe * exc, 0.0, -mZ * Z * Z - exc, ss * mZ * Z * Z, (1 - ss) * mZ * Z * Z, (1 - aZe) * exc,
           R * aZe * exc, aZe * exc * 15 / 16]
    got = tend(L, LOB, params(), [0.0, 0.0, 0.0, Z, 0.0, 0.0, 0.0, 2100.0, 2300.0], 80.0)
    return got, exp


def _grazing_split(graz):
    c, Z = 0.3, 0.4
    Lg = c * c / (c * c + kG * kG) if graz == abi.OBME_QUADRATIC else c / (c + kG)
    tot = g * Lg * Z
    return c, Z, tot, g * pref * Lg * Z, g * (1 - pref) * Lg * Z


def pin_lobster_grazing_split(L):
    """P = sPOM = c, PAR = 0: grazing pref : (1 - pref) of g L Z with the
    quadratic limit; the calcite dissolution (G_P + nu) rho R of the fixed
    Redfield plankton (plankton.jl:429-441) in DIC and Alk."""
    c, Z, tot, onP, onS = _grazing_split(abi.OBME_QUADRATIC)
    exp = lobster_expected(0.0, 0.0, 0.0, mP * c * c, 0.0, Z, tot, onP, onS, c, 0.0, 0.0)
    got = tend(L, LOB, params(), [0.0, 0.0, c, Z, c, 0.0, 0.0, 2100.0, 2300.0], 0.0)
    return got, exp


def pin_lobster_grazing_linear_mortality_linear(L):
    """As above with the linear concentration limit c/(c + kG) and linear
    phytoplankton mortality."""
    c, Z, tot, onP, onS = _grazing_split(abi.OBME_LINEAR)
    exp = lobster_expected(0.0, 0.0, 0.0, mP * c, 0.0, Z, tot, onP, onS, c, 0.0, 0.0)
    got = tend(L, LOB, params(mort=abi.OBME_LINEAR, graz=abi.OBME_LINEAR),
               [0.0, 0.0, c, Z, c, 0.0, 0.0, 2100.0, 2300.0], 0.0)
    return got, exp


PINS = {n[4:]: f for n, f in sorted(globals().items()) if n.startswith("pin_")}
RTOL = 1e-12


def failing_pins(L):
    bad = {}
    for name, f in PINS.items():
        got, exp = f(L)
        got, exp = np.asarray(got, float), np.asarray(exp, float)
        # per component, relative to the component (floored at 1e-9 of the
        # largest, so an exact zero is compared against the vector's scale)
        scale = np.maximum(np.abs(exp), 1e-9 * np.max(np.abs(exp)) + 1e-300)
        err = np.max(np.abs(got - exp) / scale)
        if not err <= RTOL:
            bad[name] = float(err)
    return bad


def test_pins_hold_for_the_oracle():
    assert failing_pins(O.lib()) == {}


def test_pins_are_not_vacuous():
    """Every pinned tendency vector is non-zero, and the states make the
    grazing, growth and remineralisation terms of comparable size to the rest."""
    for name, f in PINS.items():
        got, exp = f(O.lib())
        assert np.max(np.abs(exp)) > 0, name


# plausible mistakes in oracle/oracle_npd.c: (text, replacement)
MUTATIONS = {
    "mondo_light": ("return PAR / (kPAR + PAR);", "return kPAR / (kPAR + PAR);"),
    "analytical_light": ("PAR / sqrt((PAR * PAR) + (kPAR * kPAR))", "PAR / sqrt((PAR * PAR) + kPAR)"),
    "q10_exponent": ("pow(q->temperature_coefficient, T / 10)", "pow(q->temperature_coefficient, T / 100)"),
    "nutrient_half_saturation": ("double Ln = N / (N + q->nitrate_half_saturation);",
                                 "double Ln = N / (N + q->ammonia_half_saturation);"),
    "mortality_forms_swapped": ("return form == OBME_LINEAR ? m * X : m * (X * X);",
                                "return form == OBME_LINEAR ? m * (X * X) : m * X;"),
    "concentration_limit_quadratic": ("(X * X) / ((X * X) + (k * k))", "(X * X) / ((X * X) + k)"),
    "concentration_limit_linear": ("X / (X + k)", "X / (X + k * k)"),
    "preference_weights": ("pt * P / (pt * P + (1 - pt) * sPOM + EPS0)", "pt * P / (P + sPOM + EPS0)"),
    "grazing_on_small_fraction": ("out.onSmall = g * (1 - p) * L * sPOM", "out.onSmall = g * p * L * sPOM"),
    "exudation_dropped": ("out[1] = (1 - q->phytoplankton_exudation_fraction) * muP - G.onP - nuP;",
                          "out[1] = muP - G.onP - nuP;"),
    "mortality_sign": ("* muP - G.onP - nuP;", "* muP - G.onP + nuP;"),
    "zoo_excretion_dropped": ("- q->zooplankton_excretion_rate * Z;\n    /* N", ";\n    /* N"),
    "pinw_fraction": ("+ (1 - q->phytoplankton_solid_waste_fraction) * nuP;",
                      "+ q->phytoplankton_solid_waste_fraction * nuP;"),
    "npzd_uptake_exudate": ("double uptake = muP * (1 - q->ammonia_fraction_of_exudate * q->phytoplankton_exudation_fraction);",
                            "double uptake = muP;"),
    "npzd_remineralisation_dropped": ("out[4] = ponw + solid - G.onSmall - q->remineralisation_rate * D;",
                                      "out[4] = ponw + solid - G.onSmall;"),
    "small_particle_fraction": ("double sPOM = D * q->small_particle_fraction;", "double sPOM = D;"),
    "solid_waste_zoo_linear": ("+ q->phytoplankton_solid_waste_fraction * nuP + q->zooplankton_mortality_rate * (Z * Z);",
                               "+ q->phytoplankton_solid_waste_fraction * nuP + q->zooplankton_mortality_rate * Z;"),
    "inhibition_by_nitrate": ("exp(-q->nitrate_ammonia_inhibition * NH4)", "exp(-q->nitrate_ammonia_inhibition * NO3)"),
    "ammonia_max_dropped": ("orc_julia_max(0.0, NH4 / (q->ammonia_half_saturation + NH4))",
                            "(NH4 / (q->ammonia_half_saturation + NH4))"),
    "nitrogen_limitation_not_halved": ("double Ln = (nl + al) / 2;", "double Ln = (nl + al);"),
    "uptake_split": ("double up_NO3 = muP * nl / (nl + al + EPS0);", "double up_NO3 = muP * nl / (nl + EPS0);"),
    "nitrification_dropped": ("double dNO3 = nitrification - up_NO3;", "double dNO3 = -up_NO3;"),
    "ammonia_waste_sign": ("double up_NH4 = muP * al / (nl + al + EPS0) - waste;",
                           "double up_NH4 = muP * al / (nl + al + EPS0) + waste;"),
    "dom_rate": ("out[6] = ponw + donw - dmu * DOM;", "out[6] = ponw + donw - smu * DOM;"),
    "small_solid_fraction": ("out[4] = q->small_solid_waste_fraction * solid",
                             "out[4] = (1 - q->small_solid_waste_fraction) * solid"),
    "dinw_dom_dropped": ("double dinw = (a * (smu * sPOM + bmu * bPOM) + dmu * DOM);",
                         "double dinw = (a * (smu * sPOM + bmu * bPOM));"),
    "donw_fraction": ("double donw = (1 - a) * (smu * sPOM + bmu * bPOM);",
                      "double donw = a * (smu * sPOM + bmu * bPOM);"),
    "primary_production": ("double pp = (1 + rho * (1 - gam) - alpha_x * gam) * muP * R;",
                           "double pp = (1 + rho - alpha_x * gam) * muP * R;"),
    "calcite_dissolution_mortality": ("double cdis = (G.onP + nuP) * rho * R;", "double cdis = G.onP * rho * R;"),
    "calcite_uptake": ("double cup = 2 * rho * muP * R;", "double cup = rho * muP * R;"),
    "alkalinity_weights": ("out[8] = dNH4 * 0.9375 - dNO3 * 1.0625", "out[8] = dNH4 * 1.0625 - dNO3 * 0.9375"),
    "detritus_redfield": ("double sPOC = sPOM * Rd, bPOC = bPOM * Rd, DOC = DOM * Rd;",
                          "double sPOC = sPOM * R, bPOC = bPOM * R, DOC = DOM * R;"),
    "inorganic_carbon_waste": ("double picw = R * pinw;", "double picw = pinw;"),
}


@pytest.fixture(scope="module")
def mutant_libs():
    libs, workdir = mutants.build_many(MUTATIONS, "oracle_npd.c")
    yield libs
    shutil.rmtree(workdir, ignore_errors=True)


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistakes_are_caught(name, mutant_libs):
    L = O.load_library(mutant_libs[name])
    assert failing_pins(L), f"the planted mistake {name!r} passes every pin"
