"""Special-state pins of the numpy PISCES oracle (oracle/pisces.py), helper
by helper and for single processes of the assembled transitions.

Each pin evaluates an oracle function where the reference's expression
collapses to a short closed form written from the reference's definitions
(paths under src/Models/AdvectedPopulations/PISCES unless stated): a
Monod term at its half saturation is 1/2, a min/max at a state where one
branch wins, a size factor below its threshold is 1, b^T at T = 0 is 1, a
process alone in an otherwise empty cell (DOC aggregation, POC breakdown,
nitrification, calcite and silica dissolution, the rain ratio).  The free
iron is pinned by the equilibrium its quadratic root must satisfy, the
Morel band coefficients by the table fixture read from the reference
(tests/golden/morel_coefficients.json) and a trapezoid mean of it.

test_planted_mistakes_are_caught loads oracle/pisces.py with one
plausible mistake at a time (a wrong exponent, half saturation, branch,
sign, factor or argument order) and checks that at least one pin fails;
the element balances of the transitions at random states count as one
more pin.  CPU only."""
import importlib.util
import json
import math
import os
import shutil
import tempfile

import numpy as np
import pytest

import oracle as O  # noqa: F401  (puts oracle/ on the path)
import pisces as OP

# the reference's guards divide by eps(0.0) where nothing is grazed: inf, then min(1, inf)
pytestmark = pytest.mark.filterwarnings("ignore:overflow encountered:RuntimeWarning")

HERE = os.path.dirname(os.path.abspath(__file__))
DAY = 86400.0
MOREL = json.load(open(os.path.join(HERE, "golden", "morel_coefficients.json")))


def a1(x):
    return np.array([float(x)])


def empty_state(**kw):
    """All tracers zero except those given; O2 high (no anoxia), T = 0."""
    F = {k: a1(0.0) for k in OP.TRACERS}
    F["O2"] = a1(200.0)
    F["S"] = a1(35.0)
    for k, v in kw.items():
        F[k] = a1(v)
    return F


def aux(**kw):
    A = dict(z=-5.0, zmxl=-10.0, zeu=-10.0, kappa=1.0, mlPAR=40.0, Omega=2.0, PAR=30.0, PAR1=10.0, PAR2=10.0,
             PAR3=10.0, wPOCc=0.0, wGOCc=-40.0 / DAY)
    A.update(kw)
    return {k: a1(v) for k, v in A.items()}


def params(M):
    q = M.default_parameters()
    q["fe_excess"] = 0.0
    return q


# ---- helpers ----------------------------------------------------------------

def pin_anoxia(M):
    """anoxia_factor (common.jl:61-66): min(1, max(0, 0.4 (a1 - O2)/(a2 + O2)))
    with a1 = 6, a2 = 1: 1 at O2 = 0 (clamped from 2.4), 0.16 at O2 = 4, 0 at
    and above a1."""
    q = params(M)
    got = [M.anoxia(q, a1(o))[0] for o in (0.0, 4.0, 6.0, 50.0)]
    return got, [1.0, 0.16, 0.0, 0.0]


def pin_size_factor(M):
    """size_factor (phytoplankton/mixed_mondo.jl:185-193), threshold 1, ratio 3:
    1 up to the threshold, (1 + 3)/2 at 2, (1 + 3*3)/4 at 4."""
    ph = params(M)["nano"]
    got = [M.size_factor(ph, a1(I))[0] for I in (0.5, 1.0, 2.0, 4.0)]
    return got, [1.0, 1.0, 2.0, 2.5]


def pin_nitrogen_limitation(M):
    """nitrogen_limitation (phytoplankton/nutrient_limitation.jl:73):
    K2 N1 / (K1 K2 + K1 N2 + K2 N1): 1/2 at N1 = K1, N2 = 0; 1/3 at N1 = K1,
    N2 = K2."""
    K1, K2 = 0.13, 0.013
    f = M.nitrogen_limitation
    return [f(K1, 0.0, K1, K2), f(K1, K2, K1, K2), f(K2, K1, K2, K1)], [0.5, 1 / 3, 1 / 3]


def _nl(M, which, I, IChl, IFe, **F):
    q = params(M)
    ph = q[which]
    st = empty_state(**F)
    return M.nutrient_limitation(ph, a1(I), a1(IChl), a1(IFe), st, q)


def pin_nutrient_limitation_branches(M):
    """NitrogenIronPhosphateSilicateLimitation (nutrient_limitation.jl:20-71):
    each member of min(LN, LPO4, LFe, LSi) wins in turn.  Below the size
    threshold the half saturations are the minima (nano: NO3 0.13, PO4 0.8;
    diatoms: Si 1 + 7 Si'^2/(pk^2 + Si'^2) with Si' = 2 pk, pk = 16.6)."""
    got, exp = [], []
    # nitrate alone at its half saturation: LNO3 = 1/2, LNH4 = 0
    L, LFe, LPO4, LN, LNO3, LNH4 = _nl(M, "nano", 0.5, 0.0, 1.0, NO3=0.13, NH4=0.0, PO4=100.0, Si=0.0)
    got += [L[0], LN[0], LNO3[0], LNH4[0], LFe[0]]
    exp += [0.5, 0.5, 0.5, 0.0, 1.0]
    # phosphate at its half saturation, nitrogen abundant
    L, LFe, LPO4, *_ = _nl(M, "nano", 0.5, 0.0, 1.0, NO3=1e6, NH4=1e6, PO4=0.8, Si=0.0)
    got += [L[0], LPO4[0]]
    exp += [0.5, 0.5]
    # no iron quota: LFe = max(0, negative) = 0
    L, LFe, *_ = _nl(M, "nano", 0.5, 0.0, 0.0, NO3=1e6, NH4=1e6, PO4=1e6, Si=0.0)
    got += [L[0], LFe[0]]
    exp += [0.0, 0.0]
    # diatoms: silicate at its half saturation 1 + 7 x 4/5 (Si' = 2 pk)
    q = params(M)
    q["silicate_climatology"] = 2 * q["diatoms"]["pk"]
    st = empty_state(NO3=1e6, NH4=1e6, PO4=1e6, Si=1.0 + 7 * 4 / 5)
    L, *_ = M.nutrient_limitation(q["diatoms"], a1(0.5), a1(0.0), a1(1.0), st, q)
    got += [L[0]]
    exp += [0.5]
    # above the size threshold the half saturations scale: I = 2 gives 2x
    L, LFe, LPO4, *_ = _nl(M, "nano", 2.0, 0.0, 4.0, NO3=1e6, NH4=1e6, PO4=1.6, Si=0.0)
    got += [LPO4[0]]
    exp += [0.5]
    return got, exp


def pin_iron_limitation_interior(M):
    """LFe = (θFe - θm)/θopt between its clamps: with no chlorophyll, NO3 = 0
    and NH4 at its half saturation (LN = 1/2, LNO3 = 0), θm = 10^3 x 1.5 x
    1.21e-5 x 14/(55.85 x 7.625) x 1/2 (nutrient_limitation.jl:62), and an
    iron quota θm + θopt/2 gives 1/2."""
    thm = 1e3 * 1.5 * 1.21e-5 * 14 / (55.85 * 7.625) * 0.5
    I = 0.5
    IFe = (thm + 0.007 / 2) * I
    L, LFe, LPO4, LN, LNO3, LNH4 = _nl(M, "nano", I, 0.0, IFe, NO3=0.0, NH4=0.013, PO4=1e6, Si=0.0)
    got, exp = [LFe[0], LN[0], L[0]], [0.5, 0.5, 0.5]
    # nitrate at its half saturation instead (LN = LNO3 = 1/2): the nitrate
    # reduction term 1.15e-4 x 14/(55.85 x 7.625) LNO3 joins θm
    thm = 1e3 * (1.5 * 1.21e-5 * 14 / (55.85 * 7.625) * 0.5 + 1.15e-4 * 14 / (55.85 * 7.625) * 0.5)
    IFe = (thm + 0.007 / 4) * I
    L, LFe, *_ = _nl(M, "nano", I, 0.0, IFe, NO3=0.13, NH4=0.0, PO4=1e6, Si=0.0)
    return got + [LFe[0]], exp + [0.25]


def _growth(M, T=0.0, d=0.0, PARx=10.0, beta=0.3, t=0.0):
    q = params(M)
    ph = q["nano"]
    ph["beta"] = beta
    I, IChl = 0.5, 0.5 * 12 * 0.02
    # the mixed layer reaching d below the euphotic depth: dark residence d^2/κ
    A = aux(zeu=-10.0, zmxl=-10.0 - d, kappa=2.0, PAR1=PARx, PAR2=PARx, PAR3=PARx)
    mu = M.growth_rate(ph, a1(I), a1(IChl), a1(T), A, 0.6, q, t)[0]
    return q, ph, mu


def pin_growth_rate(M):
    """GrowthRespirationLimitedProduction (phytoplankton/growth_rate.jl:3-47,
    :104-111): mu0 bt^T f1 f2 (1 - exp(-α θ PAR / (τ (br + μr)))) L with
    f1 = 1.5 τ / (τ + day/2), f2 = 1 - t_d / (t_d + dark tolerance),
    t_d = max(0, zeu - zmxl)^2 / κ, α = α0 (1 + β e^-PAR), θ = IChl / (12 I)
    = 0.02.  The day length is the reference's day_length(φ, t) call (:30),
    CBMDayLength(t, φ) with its arguments in that order: "time" φ = 45 s,
    "latitude" t."""
    got, exp = [], []
    for T, d, PARx, t in ((0.0, 0.0, 10.0, 0.0), (10.0, 0.0, 10.0, 20.0), (0.0, math.sqrt(2.0 * 3 * DAY), 5.0, 0.0)):
        q, ph, mu = _growth(M, T, d, PARx, 0.3, t)
        tau = OP.cbm_day_length(np.float64(45.0), np.float64(t))
        PAR = (2.1 + 0.42 + 0.4) * PARx
        alpha = 2.0 * (1 + 0.3 * math.exp(-PAR))
        f1 = 1.5 * tau / (tau + 0.5 * DAY)
        f2 = 1.0 if d == 0 else 0.5      # t_d = d^2 / 2 = dark tolerance (3 days)
        fl = 1 - math.exp(-alpha * 0.02 * PAR / (tau * (0.033 / DAY + 1.0 / DAY)))
        got.append(mu)
        exp.append(0.6 / DAY * 1.066 ** T * f1 * f2 * fl * 0.6)
    return got, exp


def pin_phytoplankton_mortality(M):
    """mortality (phytoplankton/mixed_mondo.jl:115-133): linear m I/(I + K) I
    = m K/2 at I = K = 0.2; quadratic shear w I^2 with
    w = w0 + w1 0.25 (1 - L^2)/(0.25 + L^2): w0 at L = 1, w0 + 0.375 w1 at
    L = 1/2; background shear (0.01) below the mixed layer, 1 within."""
    q = params(M)
    got, exp = [], []
    ph = q["diatoms"]
    for L, z, shear in ((1.0, -20.0, 0.01), (0.5, -5.0, 1.0)):
        s = dict(ph=ph, I=a1(0.2), L=a1(L))
        lin, quad = M.phyto_mortality(q, s, aux(z=z, zmxl=-10.0))
        w = ph["mort_w0"] + (0.375 * ph["mort_w1"] if L == 0.5 else 0.0)
        got += [lin[0], quad[0]]
        exp += [ph["mort_lin"] * 0.1, shear * w * 0.04]
    return got, exp


def pin_chlorophyll_growth(M):
    """chlorophyll_growth (mixed_mondo.jl:84-97) with the production to
    energy ratio of growth_rate.jl:113-148: (1 - δ) 12 (θ0 + (θ1 - θ0) ρ) μ I,
    ρ = 12 μ̌ I L / (α IChl PAR), μ̌ = μ τ / f1, where τ and f1 come from the
    day_length(t, φ) call (:143) — the argument order opposite to the
    growth rate's."""
    t = 5 * DAY
    q, ph, _ = _growth(M)
    I, IChl = 0.5, 0.12
    A = aux(PAR1=10.0, PAR2=10.0, PAR3=10.0)
    mu = 3.0e-6
    s = dict(ph=ph, I=a1(I), IChl=a1(IChl), L=a1(0.6), mu=a1(mu))
    got = M.chlorophyll_growth(q, s, A, t)[0]
    tau = OP.cbm_day_length(np.float64(t), np.float64(45.0))
    f1 = 1.5 * tau / (tau + 0.5 * DAY)
    PAR = (2.1 + 0.42 + 0.4) * 10.0
    alpha = 2.0 * (1 + 0.3 * math.exp(-PAR))
    rho = 12 * (mu / f1 * tau) * I / (alpha * IChl * PAR) * 0.6
    exp = (1 - 0.05) * 12 * (0.0033 + (0.033 - 0.0033) * rho) * mu * I
    return [got], [exp]


def pin_iron_uptake(M):
    """iron_uptake (mixed_mondo.jl:161-181): (1 - δ) θmax L1 L2
    max(0, (1 - r)/(1.05 - r)) mu0 bt^T I, L1 = Fe/(Fe + K) = 1/2 at Fe = K
    (kFe below the size threshold), L2 = 4 - 4.5 LFe/(LFe + 1): 4 at LFe = 0,
    1.75 at LFe = 1; r = θFe/θmax: 0 with no iron quota, 1 gives no uptake."""
    q = params(M)
    ph = q["nano"]
    F = empty_state(Fe=ph["kFe_uptake"], T=0.0)
    got, exp = [], []
    for LFe, IFe, fac in ((0.0, 0.0, 4.0 / 1.05), (1.0, 0.0, 1.75 / 1.05), (0.0, 0.06 * 0.5, 0.0)):
        s = dict(ph=ph, I=a1(0.5), IFe=a1(IFe), LFe=a1(LFe))
        got.append(M.iron_uptake(q, s, F)[0])
        exp.append((1 - 0.05) * 0.06 * 0.5 * fac * (0.6 / DAY) * 0.5)
    return got, exp


def pin_silicate_uptake(M):
    """silicate_uptake (mixed_mondo.jl:195-226), diatoms: (1 - δ) θ1 μ I with
    θ1 = θ0 L1 min(5.4, (4.4 e^{-4.23 F1} F2 + 1)(1 + 2 L2)), L1 = Si/(Si + K1),
    F2 = min(1, 2.2 max(0, L1 - 1/2)), F1 = min(μ/(μi L), LFe, LPO4, LN),
    L2 = Si^3/(Si^3 + K2^3) south of the equator, else 0.  At Si = K1: L1 =
    1/2, F2 = 0; at Si = 3 K1: L1 = 3/4, F2 = 0.55, with F1 = 1/2."""
    q = params(M)
    ph = q["diatoms"]
    mui = 0.6 / DAY
    s = dict(ph=ph, I=a1(0.5), mu=a1(mui * 0.6 * 0.5), L=a1(0.6), LFe=a1(0.9), LPO4=a1(0.9), LN=a1(0.9))
    mu = mui * 0.6 * 0.5
    got, exp = [], []
    got.append(M.silicate_uptake(q, s, empty_state(Si=2.0, T=0.0))[0])
    exp.append(0.95 * 0.159 * 0.5 * 1.0 * mu * 0.5)
    got.append(M.silicate_uptake(q, s, empty_state(Si=6.0, T=0.0))[0])
    exp.append(0.95 * 0.159 * 0.75 * min(5.4, 4.4 * math.exp(-4.23 * 0.5) * 0.55 + 1) * mu * 0.5)
    q["latitude"] = -30.0
    L2 = 6.0 ** 3 / (6.0 ** 3 + 20.9 ** 3)
    got.append(M.silicate_uptake(q, s, empty_state(Si=6.0, T=0.0))[0])
    exp.append(0.95 * 0.159 * 0.75 * min(5.4, (4.4 * math.exp(-4.23 * 0.5) * 0.55 + 1) * (1 + 2 * L2)) * mu * 0.5)
    return got, exp


def pin_zooplankton_grazing(M):
    """QualityDependantZooplankton grazing (zooplankton/food_quality_dependant.jl:
    138-183), microzooplankton on nanophytoplankton alone, T = 0:
    available food (P - J) pref, concentration-limited grazing
    max(0, A - min(A/2, threshold)) — A/2 for little food, A - 0.3 for much —
    specific grazing g0 x that / (K + P); the iron grazing ratio
    (PFe/P) / (θFe g) and the efficiency min(1, ratio) min(e0, (1 - σ) ratio),
    each branch of both minima in turn (ratios 50, 0.5 and 0.2)."""
    q = params(M)
    z = q["micro"]
    got, exp = [], []
    for P, ratio in ((0.2, 50.0), (3.0, 0.5), (1.0, 0.2)):
        A = P - z["J"]
        clg = A / 2 if A / 2 < z["food_threshold"] else A - z["food_threshold"]
        tsg = z["g0"] * clg / (z["K"] + P)
        PFe = ratio * z["iron_ratio"] * tsg * P     # (PFe/P) / (θFe tsg) = ratio
        g = M.zoo_grazing(q, "Z", empty_state(P=P, PFe=PFe, Z=0.7, T=0.0))
        got += [g["gI"][0], g["ge"][0], g["gIFe"][0]]
        exp += [tsg * 0.7, min(1.0, ratio) * min(z["e0"], (1 - z["sigma"]) * ratio), PFe / P * tsg * 0.7]
    # below the specific threshold J nothing is grazed
    g = M.zoo_grazing(q, "Z", empty_state(P=0.0005, PFe=1e-6, Z=0.7, T=0.0))
    got.append(g["gI"][0])
    exp.append(0.0)
    return got, exp


def pin_grazing_split(M):
    """grazing(zoo, prey) (food_quality_dependant.jl:215-243): mesozooplankton
    on P and D split as pref (X - J): 0.3 (P - J) : 1.0 (D - J) of the total."""
    q = params(M)
    F = empty_state(P=1.0, D=2.0, M=0.5, T=0.0)
    g = M.zoo_grazing(q, "M", F)
    gP, gD = M.zoo_grazing_on(g, "P", F)[0], M.zoo_grazing_on(g, "D", F)[0]
    J = q["meso"]["J"]
    aP, aD = 0.3 * (1.0 - J), 1.0 * (2.0 - J)
    tot = g["gI"][0]
    return [gP, gD], [tot * aP / (aP + aD), tot * aD / (aP + aD)]


def pin_zooplankton_mortality(M):
    """mortality / linear_mortality (food_quality_dependant.jl:201-227), T = 0:
    I (m0 I + r (I/(I + Km) + 3 Δ)) and r (I/(I + Km) + 3 Δ) I; at I = Km the
    concentration factor is 1/2 and at I = 2 Km it is 2/3; Δ = 0 with O2
    high and 1 without O2."""
    q = params(M)
    z = q["micro"]
    got, exp = [], []
    for I, cf in ((0.2, 0.5), (0.4, 2 / 3)):
        for O2, an in ((200.0, 0.0), (0.0, 1.0)):
            g = dict(z=z, I=a1(I))
            mort, lin = M.zoo_mortality(q, g, empty_state(O2=O2, T=0.0))
            got += [mort[0], lin[0]]
            exp += [I * (z["m0"] * I + z["r"] * (cf + 3 * an)), z["r"] * (cf + 3 * an) * I]
    return got, exp


def pin_upper_trophic(M):
    """Mesozooplankton upper-trophic losses (zooplankton/mortality_waste.jl):
    w = m0 b^T M^2 / (1 - e0), respiration (1 - e0 - σ) w, fecal σ w (T = 0)."""
    q = params(M)
    z = q["meso"]
    R, fecal = M.upper_trophic(q, empty_state(M=2.0, T=0.0))
    w = z["m0"] * 4.0 / (1 - z["e0"])
    return [R[0], fecal[0]], [(1 - z["e0"] - z["sigma"]) * w, z["sigma"] * w]


def pin_bacteria(M):
    """bacteria_concentration / bacteria_activity (zooplankton/micro_and_meso.jl:
    85-132): min(4, bZ Z + bM M) above zm = min(zmxl, zeu), scaled by
    (zm/z)^0.684 below; activity min(LN, LPO4, LFe) DOC/(DOC + K_DOC), with
    LN = NO3/(NO3 + K_NO3) when NH4 = 0."""
    q = params(M)
    got, exp = [], []
    F = empty_state(Z=1.0, M=0.5, NO3=0.03, NH4=0.0, PO4=1e6, Fe=1e6, DOC=417.0)
    B, LB = M.bacteria(q, F, aux(z=-5.0, zmxl=-10.0, zeu=-20.0))
    got += [B[0], LB[0]]
    exp += [0.7 + 0.7, 0.25]
    B, _ = M.bacteria(q, empty_state(Z=3.0, M=2.0, DOC=1.0), aux(z=-5.0, zmxl=-10.0, zeu=-20.0))
    got.append(B[0])
    exp.append(4.0)
    B, _ = M.bacteria(q, F, aux(z=-40.0, zmxl=-10.0, zeu=-20.0))
    got.append(B[0])
    exp.append(0.5 ** 0.684 * 1.4)
    return got, exp


def pin_free_iron(M):
    """free_iron (iron/iron.jl:25-37) is the root of K Fe'^2 + (1 + K L - K Fe) Fe'
    - Fe = 0 in [0, Fe], with L = max(0.6, 0.09 (DOC + 40) - 3) and
    K = exp(16.27 - 1565.7 / (T + 273.15)): the quadratic's residual vanishes
    (to 1e-8: the reference's form of the root cancels digits when K L is
    large, K being about 5e4; a wrong coefficient leaves a residual of order
    one)."""
    q = params(M)
    got, exp = [], []
    for DOC, Fe, T in ((0.0, 0.7, 10.0), (40.0, 0.3, 20.0), (100.0, 2.0, 2.0)):
        Ff = M.free_iron(q, empty_state(DOC=DOC, Fe=Fe, T=T))[0]
        L = max(0.6, 0.09 * (DOC + 40) - 3)
        K = math.exp(16.27 - 1565.7 / (T + 273.15))
        # residual relative to the largest term, and the root inside [0, Fe]
        res = (K * Ff * Ff + (1 + K * L - K * Fe) * Ff - Fe) / max(K * Ff * Ff, Fe)
        got += [float(abs(res) <= 1e-8), float(0 <= Ff <= Fe)]
        exp += [1.0, 1.0]
    return got, exp


def pin_morel_band_means(M):
    """The MultiBand band coefficients (src/Light/multi_band.jl:136-140): the
    trapezoid means of the Morel table over 400-500, 500-600, 600-700 nm,
    from the reference's table (fixture) — the oracle's table and averaging
    must reproduce them."""
    lam = np.array(MOREL["lambda"])
    out = []
    for lo, hi in ((400, 500), (500, 600), (600, 700)):
        i1, i2 = int(np.where(lam == lo)[0][0]), int(np.where(lam == hi)[0][0])
        for key in ("kw", "e", "chi"):
            c = np.array(MOREL[key])
            s = sum((c[n] + c[n - 1]) * (lam[n] - lam[n - 1]) / 2 for n in range(i1 + 1, i2 + 1))
            out.append(s / (hi - lo))
    kw, e, chi = M.morel_band_coefficients()
    got = [x for b in range(3) for x in (kw[b], e[b], chi[b])]
    return got, out


# ---- single processes of the assembled transitions -------------------------

def _G(M, F, A, q=None, t=0.0):
    q = params(M) if q is None else q
    G = M.tendencies(q, F, A, t)
    return {k: float(v[0]) for k, v in G.items()}


def pin_doc_aggregation(M):
    """DOC alone (no bacteria without zooplankton, so no degradation):
    aggregation Φ1 = shear a1 DOC^2 and Φ3 = a5 DOC^2
    (dissolved_organic_carbon.jl:76-93) move DOC into POC
    (particulate_organic_matter/carbon.jl:3-27)."""
    q = params(M)
    a = q["doc_agg"]
    d = 50.0
    got, exp = [], []
    for z, shear in ((-5.0, 1.0), (-20.0, 0.01)):
        G = _G(M, empty_state(DOC=d), aux(z=z, zmxl=-10.0), q)
        agg = shear * a[0] * d * d + a[4] * d * d
        got += [G["DOC"], G["POC"], G["GOC"], G["DIC"]]
        exp += [-agg, agg, 0.0, 0.0]
    return got, exp


def pin_poc_breakdown_and_aggregation(M):
    """POC alone: breakdown λ0 b^T (1 - 0.45 Δ) POC to DOC
    (two_size_class.jl:125-133) and aggregation shear b1 POC^2 + b4 POC^2
    into GOC (:111-123); Δ = 0.16 at O2 = 4 (common.jl:61-66)."""
    q = params(M)
    b = q["pom_agg"]
    p = 8.0
    got, exp = [], []
    for O2, an in ((200.0, 0.0), (4.0, 0.16)):
        G = _G(M, empty_state(POC=p, O2=O2), aux(z=-5.0, zmxl=-10.0), q)
        lam = q["pom_lambda"] * (1 - 0.45 * an)
        agg = 1.0 * b[0] * p * p + b[3] * p * p
        got += [G["POC"], G["GOC"], G["DOC"]]
        exp += [-agg - lam * p, agg, lam * p]
    return got, exp


def pin_nitrification(M):
    """NH4 alone: nitrification λ NH4 / (1 + mean mixed-layer PAR) (1 - Δ)
    (nitrogen/nitrate_ammonia.jl) into NO3; Alk loses twice that, O2 the
    nitrification share θn / θN of it (oxygen.jl); none without O2."""
    q = params(M)
    got, exp = [], []
    for O2, an in ((200.0, 0.0), (4.0, 0.16)):
        G = _G(M, empty_state(NH4=2.0, O2=O2), aux(mlPAR=40.0, PAR=7.0), q)
        nit = q["nitrif"] * 2.0 / (1 + 40.0) * (1 - an)
        got += [G["NO3"], G["NH4"], G["Alk"], G["O2"]]
        exp += [nit, -nit, -2 * nit, -q["O2_nitrif"] * nit / q["thetaN"]]
    return got, exp


def pin_calcite_dissolution(M):
    """CaCO3 alone: dissolution λ max(0, 1 - Ω)^n CaCO3 (particulate_organic_matter/
    calcite.jl) into DIC and twice into Alk (inorganic_carbon.jl); none when
    Ω >= 1."""
    q = params(M)
    got, exp = [], []
    for Om in (0.4, 1.5):
        G = _G(M, empty_state(CaCO3=3.0), aux(Omega=Om), q)
        d = q["caco3_lambda"] * max(0.0, 1 - Om) * 3.0
        got += [G["CaCO3"], G["DIC"], G["Alk"]]
        exp += [-d, d, 2 * d]
    return got, exp


def pin_silica_dissolution(M):
    """PSi alone, T = 0, above zm: dissolution λ PSi with λ = λ0 (0.225 (1 + T/15)
    s + 0.775 ((1 + T/400)^4 s)^9), s = (eq - Si)/eq, eq = 10^(6.44 - 968/T_K),
    λ0 = χ0 λfast + (1 - χ0) λslow (particulate_organic_matter/silicate.jl)."""
    q = params(M)
    eq = 10 ** (6.44 - 968 / 273.15)
    lam0 = q["chi0"] * q["si_fast"] + (1 - q["chi0"]) * q["si_slow"]
    got, exp = [], []
    for Si in (0.0, eq / 2):
        s = (eq - Si) / eq
        lam = lam0 * (0.225 * s + 0.775 * s ** 9)
        G = _G(M, empty_state(PSi=2.0, Si=Si), aux(z=-5.0, zmxl=-10.0, zeu=-10.0), q)
        got += [G["PSi"], G["Si"]]
        exp += [-lam * 2.0, lam * 2.0]
    return got, exp


def pin_rain_ratio(M):
    """Nanophytoplankton mortality feeds calcite at the rain ratio
    (nano_diatom_coupling.jl): R = r min(LN, Fe/(Fe + 0.05), LPO4) max(1, P/2)
    max(0, PAR - 1)/(4 + PAR) 30/(30 + PAR) max(0, T/(T + 0.1))
    (1 + exp(-(T - 10)^2/25)) min(1, -50/zmxl), production R (lin + quad)/2
    with no grazers."""
    q = params(M)
    P, Fe, PAR, T = 4.0, 0.05, 26.0, 10.0
    F = empty_state(P=P, PChl=0.0, PFe=0.0, NO3=1e6, NH4=1e6, PO4=1e6, Fe=Fe, T=T)
    A = aux(z=-5.0, zmxl=-100.0, zeu=-100.0, PAR=PAR, PAR1=0.0, PAR2=0.0, PAR3=0.0, Omega=2.0)
    G = _G(M, F, A, q)
    ph = q["nano"]
    LN = None
    st = empty_state(NO3=1e6, NH4=1e6, PO4=1e6)
    _, _, LPO4, LN, _, _ = M.nutrient_limitation(ph, a1(P), a1(0.0), a1(0.0), st, q)
    Lc = min(LN[0], Fe / (Fe + 0.05), LPO4[0])
    R = (0.3 * Lc * max(1.0, P / 2) * (max(0.0, PAR - 1) / (4 + PAR)) * (30 / (30 + PAR)) * (T / (T + 0.1))
         * (1 + math.exp(0.0)) * min(1.0, 0.5))
    lin = ph["mort_lin"] * P / (P + ph["mort_K"]) * P
    quad = 1.0 * ph["mort_w0"] * P * P
    return [G["CaCO3"]], [R * (lin + quad) / 2]


PINS = {n[4:]: f for n, f in sorted(globals().items()) if n.startswith("pin_")}
RTOL = 1e-12


def balance_errors(M):
    """The carbon, silicon and nitrogen balances of the transitions at random
    states (test_PISCES.jl:68-93 at other states), relative to their terms."""
    rng = np.random.default_rng(5)
    q = params(M)
    q["fix_max"] = 0.0
    base = dict(P=7.0, PChl=1.5, PFe=35e-3, D=7.0, DChl=1.5, DFe=35e-3, DSi=1.0, Z=0.5, M=0.1, DOC=16.0, POC=0.5,
                GOC=0.5, SFe=2e-3, BFe=2e-3, PSi=0.5, CaCO3=0.1, NO3=6.0, NH4=0.26, PO4=0.4, Fe=0.5, Si=7.0,
                DIC=2205.0, Alk=2566.0, O2=317.0, T=12.0, S=35.0)
    n = 32
    F = {k: base[k] * np.exp(0.5 * rng.standard_normal(n)) for k in OP.TRACERS}
    F["O2"] = rng.uniform(0.0, 20.0, n)       # anoxia in play
    one = np.ones(n)
    A = dict(z=rng.uniform(-200, -1, n), zmxl=-30.0 * one, zeu=-40.0 * one, kappa=one, mlPAR=rng.uniform(1, 300, n),
             Omega=rng.uniform(0.5, 4.0, n), PAR=rng.uniform(0, 300, n), PAR1=rng.uniform(0, 100, n),
             PAR2=rng.uniform(0, 100, n), PAR3=rng.uniform(0, 100, n), wPOCc=-2 / DAY * one,
             wGOCc=-rng.uniform(30, 60, n) / DAY)
    G = M.tendencies(q, F, A, 3.0e6)
    groups = {"carbon": {k: 1.0 for k in ("P", "D", "Z", "M", "DOC", "POC", "GOC", "DIC", "CaCO3")},
              "silicon": {"DSi": 1.0, "Si": 1.0, "PSi": 1.0},
              "nitrogen": dict({"NH4": 1.0, "NO3": 1.0},
                               **{k: q["thetaN"] for k in ("P", "D", "Z", "M", "DOC", "POC", "GOC")})}
    out = {}
    for name, w in groups.items():
        s = sum(w[k] * G[k] for k in w)
        scale = sum(abs(w[k] * G[k]) for k in w)
        out[name] = float(np.max(np.abs(s) / scale))
    return out


def failing_pins(M):
    bad = {}
    for name, f in PINS.items():
        got, exp = f(M)
        got, exp = np.asarray(got, float), np.asarray(exp, float)
        scale = np.maximum(np.abs(exp), 1e-9 * np.max(np.abs(exp)) + 1e-300)
        err = np.max(np.abs(got - exp) / scale) if np.any(exp != 0) else np.max(np.abs(got - exp))
        if not err <= RTOL:
            bad[name] = float(err)
    for k, e in balance_errors(M).items():
        if not e <= 1e-12:
            bad["balance_" + k] = e
    return bad


def test_pins_hold_for_the_oracle():
    assert failing_pins(OP) == {}


def test_morel_table_matches_the_reference_fixture():
    """The host package's and the oracle's Morel table are the reference's
    (src/Light/morel_coefficients.jl, read into the fixture)."""
    from oceanbiome import pisces as HP
    assert np.array_equal(HP.MOREL_LAMBDA, np.array(MOREL["lambda"]))
    assert np.array_equal(HP.MOREL_KW, np.array(MOREL["kw"]))
    assert np.array_equal(HP.MOREL_E, np.array(MOREL["e"]))
    assert np.array_equal(HP.MOREL_CHI, np.array(MOREL["chi"]))


# plausible mistakes in oracle/pisces.py: (text, replacement)
MUTATIONS = {
    "anoxia_sign": ('0.4 * (q["anoxia1"] - O2)', '0.4 * (q["anoxia1"] + O2)'),
    "size_ratio": ('(I1 + ph["size_ratio"] * I2)', "(I1 + I2)"),
    "nitrogen_limitation_K": ("K1 * K2 + K1 * N2 + K2 * N1", "K1 * K2 + K2 * N2 + K2 * N1"),
    "phosphate_size_factor": ('F["PO4"] / (F["PO4"] + Kp + EPS0)', 'F["PO4"] / (F["PO4"] + ph["kPO4"] + EPS0)'),
    "iron_limitation_clamp": ('jmin(1.0, jmax(0.0, (thFe - thm) / ph["theta_Fe_opt"]))',
                              'jmax(0.0, (thFe - thm) / ph["theta_Fe_opt"])'),
    "iron_limitation_nitrate_term": ("+ 1.15e-4 * 14 / (55.85 * 7.625) * LNO3)", ")"),
    "iron_limitation_ammonium_term": ("1.5 * 1.21e-5 * 14 / (55.85 * 7.625) * LN", "1.5 * 1.21e-5 * 14 / (55.85 * 7.625)"),
    "silicate_climatology": ('7 * (Sc * Sc) / (ph["pk"] * ph["pk"] + Sc * Sc)', '7 * Sc / (ph["pk"] + Sc)'),
    "silicate_limited_flag": ('if not ph["si_limited"]:', 'if ph["si_limited"]:'),
    "day_length_fraction": ("f1 = 1.5 * dl / (dl + 0.5 * DAY)", "f1 = 1.5 * dl / (dl + DAY)"),
    "growth_day_length_order": ('dl = cbm_day_length(q["latitude"], t)', 'dl = cbm_day_length(t, q["latitude"])'),
    "chlorophyll_day_length_order": ('dl = cbm_day_length(t, q["latitude"])', 'dl = cbm_day_length(q["latitude"], t)'),
    "dark_residence": ('drt = d * d / A["kappa"]', 'drt = d / A["kappa"]'),
    "growth_respiration": ('(dl * (ph["basal_respiration"] + ph["reference_growth"]))', '(dl * ph["reference_growth"])'),
    "low_light_adaptation": ('ph["alpha0"] * (1 + ph["beta"] * np.exp(-PAR))',
                             'ph["alpha0"] * (1 + ph["beta"] * np.exp(PAR))'),
    "growth_temperature": ('mui = ph["mu0"] * np.power(ph["bt"], T)', 'mui = ph["mu0"] * ph["bt"] * T'),
    "mortality_linear_power": ('ph["mort_lin"] * I / (I + ph["mort_K"]) * I', 'ph["mort_lin"] * I / (I + ph["mort_K"])'),
    "mortality_quadratic_weight": ("0.25 * (1 - L2) / (0.25 + L2)", "0.25 * (1 - L2) / (0.5 + L2)"),
    "mortality_shear_swapped": ('shear = np.where(A["z"] < A["zmxl"], q["background_shear"], q["mixed_layer_shear"])',
                                'shear = np.where(A["z"] < A["zmxl"], q["mixed_layer_shear"], q["background_shear"])'),
    "chlorophyll_ratio": ("(th0 + (th1 - th0) * rho)", "(th1 * rho)"),
    "chlorophyll_exudation": ('return (1 - ph["delta"]) * 12 * (th0', "return 12 * (th0"),
    "iron_uptake_L2": ('L2 = 4 - 4.5 * s["LFe"] / (s["LFe"] + 1)', 'L2 = 4 - 4 * s["LFe"] / (s["LFe"] + 1)'),
    "iron_uptake_quota": ("jmax(0.0, (1 - r) / (1.05 - r))", "jmax(0.0, (1 - r))"),
    "silicate_F2": ("2.2 * jmax(0.0, L1 - 0.5)", "2.0 * jmax(0.0, L1 - 0.5)"),
    "silicate_hemisphere": ('if q["latitude"] < 0 else 0.0', 'if q["latitude"] > 0 else 0.0'),
    "silicate_F1_exponent": ("np.exp(-4.23 * F1)", "np.exp(-4.23 / F1)"),
    "grazing_threshold_half": ('jmin(avail / 2, z["food_threshold"])', 'jmin(avail, z["food_threshold"])'),
    "grazing_denominator": ('tsg = base * clg / (z["K"] + total_food)', 'tsg = base * clg / (z["K"] + avail)'),
    "growth_efficiency": ('ge = fq * jmin(z["e0"], (1 - z["sigma"]) * igr)', 'ge = fq * jmin(z["e0"], igr)'),
    "grazing_on_threshold": ('z["pref"][prey] * jmax(0.0, F[prey] - z["J"])', 'z["pref"][prey] * F[prey]'),
    "zoo_mortality_anoxia": ("(cf + 3 * an)", "(cf + an)"),
    "zoo_mortality_concentration": ('cf = I / (I + z["Km"])', 'cf = z["Km"] / (I + z["Km"])'),
    "upper_trophic_efficiency": ('w = 1 / (1 - z["e0"]) * z["m0"] * tf * (I * I)', 'w = z["m0"] * tf * (I * I)'),
    "bacteria_mesozooplankton": ('q["bZ"] * F["Z"] + q["bM"] * F["M"]', 'q["bZ"] * F["Z"] + q["bM"] * F["Z"]'),
    "bacteria_depth_branch": ('np.where(A["z"] >= zm, 1.0, depth_factor)', 'np.where(A["z"] < zm, 1.0, depth_factor)'),
    "bacteria_nitrogen": ('LN = (KNO * F["NH4"] + KNH * F["NO3"])', 'LN = (KNO * F["NO3"] + KNH * F["NH4"])'),
    "free_iron_discriminant": ('4 * K * F["Fe"]', '2 * K * F["Fe"]'),
    "free_iron_ligands": ('jmax(0.6, 0.09 * (F["DOC"] + 40) - 3)', 'jmax(0.6, 0.09 * F["DOC"] - 3)'),
    "doc_aggregation_phi3": ('Phi3 = (a4 * F["POC"] + a5 * DOC) * DOC', 'Phi3 = (a4 * F["POC"]) * DOC'),
    "pom_aggregation_square": ('b4 * (POC * POC)', "b4 * POC"),
    "pom_breakdown_anoxia": ("(1 - 0.45 * an)", "(1 - an)"),
    "nitrification_light": ('F["NH4"] / (1 + A["mlPAR"]) * (1 - an)', 'F["NH4"] / (1 + A["PAR"]) * (1 - an)'),
    "calcite_undersaturation": ('jmax(0.0, 1 - A["Omega"])', 'jmax(0.0, A["Omega"] - 1)'),
    "alkalinity_calcite": ('G["Alk"] = G["NH4"] - G["NO3"] - 2 * G["CaCO3"]', 'G["Alk"] = G["NH4"] - G["NO3"] - G["CaCO3"]'),
    "silica_saturation_linear": ('0.225 * (1 + T / 15) * sat', "0.225 * (1 + T / 15)"),
    "rain_ratio_light": ("(30 / (30 + PAR))", "(30 / (30 - PAR))"),
    "rain_ratio_mixed_layer": ("jmin(1.0, -50 / zmxl)", "jmin(1.0, 50 / zmxl)"),
    "oxygen_nitrification": ('- th_n * nitrif / thN)', "- th_n * nitrif)"),
    "exudate_to_doc": ('exudate = sP["ph"]["delta"] * sP["muI"] + sD["ph"]["delta"] * sD["muI"]',
                       'exudate = sP["ph"]["delta"] * sP["muI"]'),
}

PISCES_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "pisces.py")


def load_mutant(workdir, name, old, new):
    text = open(PISCES_SRC, encoding="utf-8").read()
    if old not in text:
        raise ValueError(f"mutation {name}: {old!r} not found")
    path = os.path.join(workdir, f"pisces_{name}.py")
    open(path, "w", encoding="utf-8").write(text.replace(old, new))
    spec = importlib.util.spec_from_file_location(f"pisces_mutant_{name}", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def workdir():
    d = tempfile.mkdtemp(prefix="pisces_mutants_")
    yield d
    shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistakes_are_caught(name, workdir):
    M = load_mutant(workdir, name, *MUTATIONS[name])
    assert failing_pins(M), f"the planted mistake {name!r} passes every pin"
