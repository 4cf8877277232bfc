"""FP64 numpy restatement of the PISCES biogeochemistry of OceanBioME.jl
v0.17.6 (src/Models/AdvectedPopulations/PISCES/**, src/Light/multi_band.jl,
src/Light/compute_euphotic_depth.jl, src/Utils/Utils.jl CBMDayLength).

TEST INFRASTRUCTURE ONLY: the oracle of the HIP library's PISCES path.
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.

Every function is vectorised over cells: arguments are numpy arrays of any
common shape.  Julia semantics kept: NaN-propagating min/max (np.minimum /
np.maximum), eps(0.0) = 5e-324 as the division guard, literal x^2 and x^3 as
products, float^float through pow, left-to-right evaluation of products.
The calcite saturation per cell uses the C oracle's carbonate chemistry
(pinned separately by the reference's doctests and known answers).

Pins: the element balances and known answers of the reference's PISCES
tests, closed forms of every helper and of single processes of the
transitions at special states (each shown to catch a planted mistake,
tests/test_oracle_pisces_pins.py), and the Morel coefficient table against
a fixture parsed from the reference's morel_coefficients.jl
(tests/golden/make_morel_fixture.py).  Parity unpinned: the carbonate
chemistry's TEOS-10 density away from the doctest point (oracle.h).
"""
import math

import numpy as np

DAY = 86400.0
HOUR = 3600.0
EPS0 = 5e-324          # Julia eps(0.0)
G_EARTH = 9.80665      # Oceananigans defaults.gravitational_acceleration

# tracer order of required_biogeochemical_tracers(PISCES) (PISCES.jl:91-103)
TRACERS = ("P", "PChl", "PFe", "D", "DChl", "DFe", "DSi", "Z", "M", "DOC", "POC", "GOC", "SFe", "BFe", "PSi",
           "CaCO3", "NO3", "NH4", "PO4", "Fe", "Si", "DIC", "Alk", "O2", "T", "S")
WRITTEN = TRACERS[:24]


def jmin(*a):
    r = a[0]
    for x in a[1:]:
        r = np.minimum(r, x)
    return r


def jmax(*a):
    r = a[0]
    for x in a[1:]:
        r = np.maximum(r, x)
    return r


def sind(x):
    return np.sin(np.radians(np.fmod(x, 360.0)))


def cosd(x):
    return np.cos(np.radians(np.fmod(x, 360.0)))


def cbm_day_length(t, phi, p=0.833):
    """CBMDayLength()(t, φ) (src/Utils/Utils.jl:17-33), seconds."""
    J = np.floor(np.mod(t, 365 * DAY) / DAY)
    theta = 0.216310 + 2 * np.arctan(0.9671396 * np.tan(0.00860 * (J - 186)))
    dec = np.degrees(np.arcsin(0.39795 * np.cos(theta)))
    L = jmax(-1.0, jmin(1.0, (sind(p) + sind(phi) * sind(dec)) / (cosd(phi) * cosd(dec))))
    return (24 - 24 / 180 * np.degrees(np.arccos(L))) * HOUR


# ---------------------------------------------------------------- parameters

def default_parameters():
    """PISCES(; grid) defaults (PISCES.jl:268-330 and the component constructors)."""
    def phyto(dark, kNH4, kNO3, kPO4, si_lim, b1, b2, b3, wmax, chl_max, kFe):
        return dict(mu0=0.6 / DAY, bt=1.066, dark_tolerance=dark, alpha0=2.0, beta=0.0,
                    basal_respiration=0.033 / DAY, reference_growth=1.0 / DAY,
                    kNH4=kNH4, kNO3=kNO3, kPO4=kPO4, theta_Fe_opt=0.007, si_limited=si_lim, kSi_min=1.0,
                    pk=16.6, delta=0.05, beta1=b1, beta2=b2, beta3=b3, mort_K=0.2, mort_lin=0.01 / DAY,
                    mort_w0=0.01 / DAY, mort_w1=wmax, chl_min=0.0033, chl_max=chl_max, fe_max=0.06,
                    si_K1=2.0, si_K2=20.9, si_theta0=0.159, kFe_uptake=kFe, size_threshold=1.0, size_ratio=3.0)

    def zoo(g0, prefs, m0, r, e0, ff, eta, fe):
        return dict(b=1.079, g0=g0, pref=prefs, food_threshold=0.3, J=0.001, K=20.0, ff=ff, iron_ratio=fe,
                    e0=e0, sigma=0.3, Km=0.2, m0=m0, r=r, gamma=0.6, eta=eta)

    return dict(
        nano=phyto(3 * DAY, 0.013, 0.13, 0.8, False, 2.1, 0.42, 0.4, 0.0, 0.033, 1.0),
        diatoms=phyto(4 * DAY, 0.039, 0.39, 2.4, True, 1.6, 0.69, 0.7, 0.03 / DAY, 0.05, 3.0),
        rain_ratio=0.3,
        micro=zoo(3 / DAY, dict(P=1.0, D=0.5, POC=0.1, Z=0.0), 0.004 / DAY, 0.03 / DAY, 0.3, 0.0, 0.5, 0.01),
        meso=zoo(0.75 / DAY, dict(P=0.3, D=1.0, POC=0.3, Z=1.0), 0.03 / DAY, 0.005 / DAY, 0.35, 2e3 / 1e6, 0.75,
                 0.015),
        bZ=0.7, bM=1.4, bact_exp=0.684, K_DOC=417.0, K_NO3_b=0.03, K_NH4_b=0.003, K_PO4_b=0.003, K_Fe_b=0.01,
        doc_lambda=0.3 / DAY, doc_bact_ref=1.0, doc_b=1.066,
        doc_agg=tuple(x * (1e-6 / DAY) for x in (0.37, 102, 3530, 5095, 114)),
        pom_b=1.066, pom_lambda=0.025 / DAY, pom_agg=tuple(x * (1e-6 / DAY) for x in (25.9, 4452, 3.3, 47.1)),
        fe_scav_min=3e-5 / DAY, fe_scav_load=0.005 / DAY, bact_fe_eff=0.16, small_bact_fe=0.12 / 0.16,
        large_bact_fe=0.04 / 0.16, chi0=0.5, si_fast=0.025 / DAY, si_slow=0.003 / DAY,
        caco3_lambda=0.197 / DAY, caco3_exp=1.0, bact_fe_max=0.06, bact_fe_K=0.3, bact_mu0=0.6 / DAY,
        nitrif=0.05 / DAY, fix_max=0.013 / DAY, fix_KFe=0.1, fix_KPO4=0.8, fix_E=50.0,
        fe_excess=1000.0, fe_ligand_max=0.6, fe_ligand_ratio=0.09,
        O2_resp=133 / 122, O2_nitrif=32 / 122,
        anoxia1=6.0, anoxia2=1.0, thetaN=16 / 122, thetaP=1 / 122,
        mixed_layer_shear=1.0, background_shear=0.01, latitude=45.0, silicate_climatology=7.5,
    )


# ---------------------------------------------------------------- helpers

def anoxia(q, O2):
    """anoxia_factor (PISCES/common.jl:61-66)."""
    return jmin(1.0, jmax(0.0, 0.4 * (q["anoxia1"] - O2) / (q["anoxia2"] + O2)))


def size_factor(ph, I):
    """mixed_mondo.jl:185-193."""
    I1 = jmin(I, ph["size_threshold"])
    I2 = jmax(0.0, I - ph["size_threshold"])
    return (I1 + ph["size_ratio"] * I2) / (I1 + I2 + EPS0)


def nitrogen_limitation(N1, N2, K1, K2):
    """nutrient_limitation.jl:73."""
    return (K2 * N1) / (K1 * K2 + K1 * N2 + K2 * N1 + EPS0)


def nutrient_limitation(ph, I, IChl, IFe, F, q):
    """NitrogenIronPhosphateSilicateLimitation (nutrient_limitation.jl:20-71):
    (L, LFe, LPO4, LN, LNO3, LNH4)."""
    thFe = np.where(I == 0, 0.0, IFe / (I + EPS0))
    thChl = np.where(I == 0, 0.0, IChl / (12 * I + EPS0))
    Kb = size_factor(ph, I)
    Kno = ph["kNO3"] * Kb
    Knh = ph["kNH4"] * Kb
    Kp = ph["kPO4"] * Kb
    Ksi = ph["kSi_min"] * Kb
    LNO3 = nitrogen_limitation(F["NO3"], F["NH4"], Kno, Knh)
    LNH4 = nitrogen_limitation(F["NH4"], F["NO3"], Knh, Kno)
    LN = LNO3 + LNH4
    LPO4 = F["PO4"] / (F["PO4"] + Kp + EPS0)
    thm = 1000 * (0.0016 / 55.85 * 12 * thChl + 1.5 * 1.21e-5 * 14 / (55.85 * 7.625) * LN
                  + 1.15e-4 * 14 / (55.85 * 7.625) * LNO3)
    LFe = jmin(1.0, jmax(0.0, (thFe - thm) / ph["theta_Fe_opt"]))
    Sc = q["silicate_climatology"]
    KSi = Ksi + 7 * (Sc * Sc) / (ph["pk"] * ph["pk"] + Sc * Sc)
    LSi = F["Si"] / (F["Si"] + KSi)
    if not ph["si_limited"]:
        LSi = np.full_like(LSi, np.inf)
    return jmin(LN, LPO4, LFe, LSi), LFe, LPO4, LN, LNO3, LNH4


def base_production_rate(ph, T):
    """growth_rate.jl:159-164."""
    return ph["mu0"] * np.power(ph["bt"], T)


def growth_rate(ph, I, IChl, T, A, L, q, t):
    """BaseProduction functor with GrowthRespirationLimitedProduction light
    limitation (growth_rate.jl:3-47, :104-111).  The day length is called
    as day_length(φ, t) here, in that argument order (:30)."""
    PAR = ph["beta1"] * A["PAR1"] + ph["beta2"] * A["PAR2"] + ph["beta3"] * A["PAR3"]
    dl = cbm_day_length(q["latitude"], t)
    d = jmax(0.0, A["zeu"] - A["zmxl"])
    drt = d * d / A["kappa"]
    mui = ph["mu0"] * np.power(ph["bt"], T)
    f1 = 1.5 * dl / (dl + 0.5 * DAY)
    f2 = 1 - drt / (drt + ph["dark_tolerance"])
    alpha = ph["alpha0"] * (1 + ph["beta"] * np.exp(-PAR))
    theta = IChl / (12 * I + EPS0)
    fl = 1 - np.exp(-alpha * theta * PAR / (dl * (ph["basal_respiration"] + ph["reference_growth"])))
    return mui * f1 * f2 * fl * L


def phyto_state(q, F, A, t, which):
    """Everything one phytoplankton class's functions share."""
    ph = q["nano"] if which == "P" else q["diatoms"]
    I, IChl, IFe = (F["P"], F["PChl"], F["PFe"]) if which == "P" else (F["D"], F["DChl"], F["DFe"])
    L, LFe, LPO4, LN, LNO3, LNH4 = nutrient_limitation(ph, I, IChl, IFe, F, q)
    mu = growth_rate(ph, I, IChl, F["T"], A, L, q, t)
    return dict(ph=ph, I=I, IChl=IChl, IFe=IFe, L=L, LFe=LFe, LPO4=LPO4, LN=LN, LNO3=LNO3, LNH4=LNH4, mu=mu,
                muI=mu * I)


def phyto_mortality(q, s, A):
    """mixed_mondo.jl:115-133: (linear, quadratic)."""
    ph, I, L = s["ph"], s["I"], s["L"]
    lin = ph["mort_lin"] * I / (I + ph["mort_K"]) * I
    L2 = L * L
    w = ph["mort_w0"] + ph["mort_w1"] * 0.25 * (1 - L2) / (0.25 + L2)
    shear = np.where(A["z"] < A["zmxl"], q["background_shear"], q["mixed_layer_shear"])
    return lin, shear * w * (I * I)


def chlorophyll_growth(q, s, A, t):
    """mixed_mondo.jl:84-97 with production_and_energy_assimilation_absorption_ratio
    (growth_rate.jl:113-148; day_length(t, φ) there)."""
    ph, I, IChl = s["ph"], s["I"], s["IChl"]
    PAR = ph["beta1"] * A["PAR1"] + ph["beta2"] * A["PAR2"] + ph["beta3"] * A["PAR3"]
    dl = cbm_day_length(t, q["latitude"])
    f1 = 1.5 * dl / (dl + 0.5 * DAY)
    mu = s["mu"]
    mucheck = mu / f1 * dl
    alpha = ph["alpha0"] * (1 + ph["beta"] * np.exp(-PAR))
    rho = 12 * mucheck * I / (alpha * IChl * PAR + EPS0) * s["L"]
    th0, th1 = ph["chl_min"], ph["chl_max"]
    return (1 - ph["delta"]) * 12 * (th0 + (th1 - th0) * rho) * mu * I


def iron_uptake(q, s, F):
    """mixed_mondo.jl:161-181."""
    ph, I, IFe = s["ph"], s["I"], s["IFe"]
    thFe = IFe / (I + EPS0)
    mui = base_production_rate(ph, F["T"])
    K = ph["kFe_uptake"] * size_factor(ph, I)
    L1 = F["Fe"] / (F["Fe"] + K + EPS0)
    L2 = 4 - 4.5 * s["LFe"] / (s["LFe"] + 1)
    r = thFe / ph["fe_max"]
    return (1 - ph["delta"]) * ph["fe_max"] * L1 * L2 * jmax(0.0, (1 - r) / (1.05 - r)) * mui * I


def silicate_uptake(q, s, F):
    """mixed_mondo.jl:195-226 (diatoms)."""
    ph, I = s["ph"], s["I"]
    Si = F["Si"]
    mui = base_production_rate(ph, F["T"])
    L1 = Si / (Si + ph["si_K1"] + EPS0)
    K2 = ph["si_K2"]
    L2 = (Si * Si * Si) / (Si * Si * Si + K2 * K2 * K2) if q["latitude"] < 0 else 0.0
    F1 = jmin(s["mu"] / (mui * s["L"] + EPS0), s["LFe"], s["LPO4"], s["LN"])
    F2 = jmin(1.0, 2.2 * jmax(0.0, L1 - 0.5))
    th1 = ph["si_theta0"] * L1 * jmin(5.4, (4.4 * np.exp(-4.23 * F1) * F2 + 1) * (1 + 2 * L2))
    return (1 - ph["delta"]) * th1 * s["mu"] * I


# ---------------------------------------------------------------- zooplankton

FOOD_ORDER = ("P", "D", "POC", "Z")


def zoo_food(zname, F, q):
    """extract_food_availability / extract_iron_availability (defaults.jl:41-59),
    in the (P, D, POC, Z) order; Z grazes the first three, M all four."""
    n = 3 if zname == "Z" else 4
    food = [F["P"], F["D"], F["POC"], F["Z"]][:n]
    iron = [F["PFe"] / (F["P"] + EPS0), F["DFe"] / (F["D"] + EPS0), F["SFe"] / (F["POC"] + EPS0),
            q["micro"]["iron_ratio"]][:n]
    return food, iron


def zoo_grazing(q, zname, F):
    """food_quality_dependant.jl:138-183: (total_specific_grazing, available_total_food,
    grazing g*I, growth efficiency), plus the shared base rate."""
    z = q["micro"] if zname == "Z" else q["meso"]
    I = F[zname]
    base = z["g0"] * np.power(z["b"], F["T"])
    food, iron = zoo_food(zname, F, q)
    p = [z["pref"][k] for k in FOOD_ORDER[:len(food)]]
    total_food = food[0] * p[0]
    avail = jmax(0.0, food[0] - z["J"]) * p[0]
    total_iron = iron[0] * p[0]
    for n in range(1, len(food)):
        total_food = total_food + food[n] * p[n]
        avail = avail + jmax(0.0, food[n] - z["J"]) * p[n]
        total_iron = total_iron + iron[n] * p[n]
    clg = jmax(0.0, avail - jmin(avail / 2, z["food_threshold"]))
    tsg = base * clg / (z["K"] + total_food)
    igr = total_iron / (z["iron_ratio"] * tsg + EPS0)
    fq = jmin(1.0, igr)
    ge = fq * jmin(z["e0"], (1 - z["sigma"]) * igr)
    # iron grazing (iron_grazing.jl:1-36)
    s = jmax(0.0, food[0] - z["J"]) * p[0] * iron[0]
    for n in range(1, len(food)):
        s = s + jmax(0.0, food[n] - z["J"]) * p[n] * iron[n]
    tsig = s * tsg / (avail + EPS0)
    return dict(z=z, I=I, tsg=tsg, avail=avail, gI=tsg * I, ge=ge, gIFe=tsig * I)


def zoo_grazing_on(g, prey, F):
    """grazing(zoo, val_name, val_prey_name, ...) (food_quality_dependant.jl:215-243)."""
    z = g["z"]
    return z["pref"][prey] * jmax(0.0, F[prey] - z["J"]) * g["tsg"] / (g["avail"] + EPS0) * g["I"]


def flux_rates(F, A):
    return dict(POC=F["POC"] * A["wPOCc"], GOC=F["GOC"] * A["wGOCc"], SFe=F["SFe"] * A["wPOCc"],
                BFe=F["BFe"] * A["wGOCc"])


def zoo_flux_feeding(g, F, fr, prey=None):
    """flux_feeding (food_quality_dependant.jl:185-199 and :245-259) and
    iron_flux_feeding (iron_grazing.jl:38-53)."""
    z = g["z"]
    base = z["ff"] * np.power(z["b"], F["T"])
    if prey is None:
        sf = fr["POC"] + fr["GOC"]
    elif prey == "iron":
        sf = fr["SFe"] + fr["BFe"]
    else:
        sf = fr[prey]
    return base * sf * g["I"]


def zoo_mortality(q, g, F):
    """food_quality_dependant.jl:201-213 and :215-227 (linear only)."""
    z, I = g["z"], g["I"]
    tf = np.power(z["b"], F["T"])
    cf = I / (I + z["Km"])
    an = anoxia(q, F["O2"])
    mort = tf * I * (z["m0"] * I + z["r"] * (cf + 3 * an))
    lin = tf * z["r"] * (cf + 3 * an) * I
    return mort, lin


def upper_trophic(q, F):
    """mortality_waste.jl (meso): (respiration product R, fecal production)."""
    z = q["meso"]
    I = F["M"]
    tf = np.power(z["b"], F["T"])
    w = 1 / (1 - z["e0"]) * z["m0"] * tf * (I * I)
    return (1 - z["e0"] - z["sigma"]) * w, z["sigma"] * w


# ---------------------------------------------------------------- DOM / POM / iron

def bacteria(q, F, A):
    """bacteria_concentration and bacteria_activity (micro_and_meso.jl:85-132)."""
    zm = jmin(A["zmxl"], A["zeu"])
    surface = jmin(4.0, q["bZ"] * F["Z"] + q["bM"] * F["M"])
    depth_factor = np.power(zm / A["z"], q["bact_exp"])
    Bact = np.where(A["z"] >= zm, 1.0, depth_factor) * surface
    DOC_limit = F["DOC"] / (F["DOC"] + q["K_DOC"])
    KNO, KNH = q["K_NO3_b"], q["K_NH4_b"]
    LN = (KNO * F["NH4"] + KNH * F["NO3"]) / (KNO * KNH + KNO * F["NH4"] + KNH * F["NO3"])
    LPO4 = F["PO4"] / (F["PO4"] + q["K_PO4_b"])
    LFe = F["Fe"] / (F["Fe"] + q["K_Fe_b"])
    return Bact, jmin(LN, LPO4, LFe) * DOC_limit


def free_iron(q, F):
    """iron.jl:25-37."""
    ligands = jmax(0.6, 0.09 * (F["DOC"] + 40) - 3)
    K = np.exp(16.27 - 1565.7 / jmax(F["T"] + 273.15, 5.0))
    Dl = 1 + K * ligands - K * F["Fe"]
    return (-Dl + np.sqrt(Dl * Dl + 4 * K * F["Fe"])) / (2 * K)


def tendencies(q, F, A, t):
    """The 24 PISCES transitions (the `(bgc::PISCES)(i, j, k, grid, Val(name), ...)`
    methods) at every cell; F: tracers by name, A: auxiliary fields
    (z, zmxl, zeu, kappa, mlPAR, Omega, PAR, PAR1..3, wPOCc, wGOCc)."""
    G = {}
    T = F["T"]
    sP = phyto_state(q, F, A, t, "P")
    sD = phyto_state(q, F, A, t, "D")
    gZ = zoo_grazing(q, "Z", F)
    gM = zoo_grazing(q, "M", F)
    fr = flux_rates(F, A)
    ffZ = zoo_flux_feeding(gZ, F, fr)
    ffM = zoo_flux_feeding(gM, F, fr)
    mZ, linZ = zoo_mortality(q, gZ, F)
    mM, linM = zoo_mortality(q, gM, F)
    an = anoxia(q, F["O2"])
    zmxl, z = A["zmxl"], A["z"]
    shear_lt = np.where(z < zmxl, q["background_shear"], q["mixed_layer_shear"])

    def graz(prey):  # grazing(zoo::MicroAndMeso, prey) = micro + meso
        return zoo_grazing_on(gZ, prey, F) + zoo_grazing_on(gM, prey, F)

    def flux_feed(prey):
        return zoo_flux_feeding(gZ, F, fr, prey) + zoo_flux_feeding(gM, F, fr, prey)

    # phytoplankton
    Plin, Pquad = phyto_mortality(q, sP, A)
    Dlin, Dquad = phyto_mortality(q, sD, A)
    grazP, grazD = graz("P"), graz("D")
    for name, s, lin, quad, gr in (("P", sP, Plin, Pquad, grazP), ("D", sD, Dlin, Dquad, grazD)):
        ph = s["ph"]
        death = lin + quad
        G[name] = (1 - ph["delta"]) * s["muI"] - death - gr
        thChl = s["IChl"] / (12 * s["I"] + EPS0)
        G[name + "Chl"] = chlorophyll_growth(q, s, A, t) - (death + gr) * thChl * 12
        thFe = s["IFe"] / (s["I"] + EPS0)
        G[name + "Fe"] = iron_uptake(q, s, F) - (death + gr) * thFe
    si_uptake = silicate_uptake(q, sD, F)
    thSi = F["DSi"] / (F["D"] + EPS0)
    G["DSi"] = si_uptake - (Dlin + Dquad + grazD) * thSi

    # zooplankton
    G["Z"] = gZ["ge"] * (gZ["gI"] + ffZ) - mZ - zoo_grazing_on(gM, "Z", F)
    G["M"] = gM["ge"] * (gM["gI"] + ffM) - mM

    sigZ, sigM = q["micro"]["sigma"], q["meso"]["sigma"]
    excZ = (1 - sigZ - gZ["ge"]) * (gZ["gI"] + ffZ)
    excM = (1 - sigM - gM["ge"]) * (gM["gI"] + ffM)
    inorg_exc = q["micro"]["gamma"] * excZ + q["meso"]["gamma"] * excM
    org_exc = (1 - q["micro"]["gamma"]) * excZ + (1 - q["meso"]["gamma"]) * excM
    R_ut, fecal = upper_trophic(q, F)
    ut_excretion = (1 - q["meso"]["gamma"]) * R_ut
    ut_respiration = q["meso"]["gamma"] * R_ut

    # DOC
    Bact, LBact = bacteria(q, F, A)
    doc_deg = q["doc_lambda"] * np.power(q["doc_b"], T) * LBact * Bact / q["doc_bact_ref"] * F["DOC"]
    a1, a2, a3, a4, a5 = q["doc_agg"]
    DOC = F["DOC"]
    Phi1 = shear_lt * (a1 * DOC + a2 * F["POC"]) * DOC
    Phi2 = shear_lt * (a3 * F["GOC"]) * DOC
    Phi3 = (a4 * F["POC"] + a5 * DOC) * DOC
    doc_agg = Phi1 + Phi2 + Phi3
    lam_pom = q["pom_lambda"] * np.power(q["pom_b"], T) * (1 - 0.45 * an)
    pom_deg_POC = lam_pom * F["POC"]
    exudate = sP["ph"]["delta"] * sP["muI"] + sD["ph"]["delta"] * sD["muI"]
    G["DOC"] = exudate + ut_excretion + org_exc + pom_deg_POC - doc_deg - doc_agg
    oxic = (1 - an) * doc_deg
    anoxic = an * doc_deg

    # POM aggregation, rain ratio, mortality partitions
    b1, b2, b3, b4 = q["pom_agg"]
    POC, GOC = F["POC"], F["GOC"]
    pom_agg = shear_lt * (b1 * (POC * POC) + b2 * POC * GOC) + b3 * POC * GOC + b4 * (POC * POC)
    PAR = A["PAR"]
    L_CaCO3 = jmin(sP["LN"], F["Fe"] / (F["Fe"] + 0.05), sP["LPO4"])
    R = (q["rain_ratio"] * L_CaCO3 * jmax(1.0, F["P"] / 2) * (jmax(0.0, PAR - 1) / (4 + PAR))
         * (30 / (30 + PAR)) * jmax(0.0, T / (T + 0.1)) * (1 + np.exp(-((T - 10) * (T - 10)) / 25))
         * jmin(1.0, -50 / zmxl))
    small_mort = (1 - R / 2) * (Plin + Pquad) + Dlin / 2
    large_mort = R / 2 * (Plin + Pquad) + Dlin / 2 + Dquad
    thP = F["PFe"] / (F["P"] + EPS0)
    thD = F["DFe"] / (F["D"] + EPS0)
    small_mort_fe = (1 - R / 2) * (Plin + Pquad) * thP + Dlin * thD / 2
    large_mort_fe = R / 2 * (Plin + Pquad) * thP + (Dlin / 2 + Dquad) * thD
    tot_graz_POC = graz("POC") + flux_feed("POC")
    tot_graz_GOC = flux_feed("GOC")

    G["POC"] = (sigZ * (gZ["gI"] + ffZ) + small_mort + mZ + (Phi1 + Phi3) + lam_pom * GOC
                - tot_graz_POC - pom_agg - pom_deg_POC)
    G["GOC"] = (sigM * (gM["gI"] + ffM) + large_mort + linM + fecal + pom_agg + Phi2
                - tot_graz_GOC - lam_pom * GOC)

    # iron in particles
    Fef = free_iron(q, F)
    colloidal = 0.5 * (F["Fe"] - Fef)
    CgFe1 = (Phi1 + Phi3) * colloidal / (DOC + EPS0)
    CgFe2 = Phi2 * colloidal / (DOC + EPS0)
    lamFe = q["fe_scav_min"] + q["fe_scav_load"] * (POC + GOC + F["CaCO3"] + F["PSi"])
    BactFe = (q["bact_mu0"] * np.power(q["pom_b"], T) * LBact * q["bact_fe_max"] * F["Fe"] / (F["Fe"] + q["bact_fe_K"])
              * Bact * q["bact_fe_eff"])
    thS = F["SFe"] / (POC + EPS0)
    thB = F["BFe"] / (GOC + EPS0)
    G["SFe"] = (sigZ * (gZ["gIFe"] + zoo_flux_feeding(gZ, F, fr, "iron")) + small_mort_fe
                + mZ * q["micro"]["iron_ratio"] + lam_pom * F["BFe"] + lamFe * POC * Fef
                + q["small_bact_fe"] * BactFe + CgFe1 - tot_graz_POC * thS - pom_agg * thS - lam_pom * F["SFe"])
    G["BFe"] = (sigM * (gM["gIFe"] + zoo_flux_feeding(gM, F, fr, "iron")) + large_mort_fe
                + linM * q["meso"]["iron_ratio"] + fecal * q["meso"]["iron_ratio"] + lamFe * GOC * Fef
                + q["large_bact_fe"] * BactFe + CgFe2 + pom_agg * thS - tot_graz_GOC * thB - lam_pom * F["BFe"])

    # particulate silicate
    zm = jmin(zmxl, A["zeu"])
    chi = q["chi0"] * np.where(z >= zm, 1.0, np.exp((q["si_fast"] - q["si_slow"]) * (zm - z) / A["wGOCc"]))
    lam0 = chi * q["si_fast"] + (1 - chi) * q["si_slow"]
    eq = np.power(10.0, 6.44 - 968 / (T + 273.15))
    sat = (eq - F["Si"]) / eq
    lam_si = lam0 * (0.225 * (1 + T / 15) * sat + 0.775 * np.power(np.power(1 + T / 400, 4.0) * sat, 9.0))
    si_diss = lam_si * F["PSi"]
    G["PSi"] = (grazD + Dlin + Dquad) * thSi - si_diss

    # calcite
    calc_loss = q["micro"]["eta"] * zoo_grazing_on(gZ, "P", F) + q["meso"]["eta"] * zoo_grazing_on(gM, "P", F)
    calc_prod = R * (calc_loss + (Plin + Pquad) / 2)
    calc_diss = q["caco3_lambda"] * np.power(jmax(0.0, 1 - A["Omega"]), q["caco3_exp"]) * F["CaCO3"]
    G["CaCO3"] = calc_prod - calc_diss

    # nitrogen
    thN = q["thetaN"]
    nitrif = q["nitrif"] * F["NH4"] / (1 + A["mlPAR"]) * (1 - an)
    avail = sP["LN"]
    fix_limit = np.where(avail >= 0.8, 0.01, 1 - avail)
    mu_nano = base_production_rate(q["nano"], T)
    growth_req = jmax(0.0, mu_nano - 2.15)
    nut_lim = jmin(F["Fe"] / (F["Fe"] + q["fix_KFe"]), F["PO4"] / (F["PO4"] + q["fix_KPO4"]))
    light_lim = 1 - np.exp(-PAR / q["fix_E"])
    fixation = q["fix_max"] * growth_req * fix_limit * nut_lim * light_lim
    up_NO3 = sP["muI"] * sP["LNO3"] / (sP["LN"] + EPS0) + sD["muI"] * sD["LNO3"] / (sD["LN"] + EPS0)
    up_NH4 = sP["muI"] * sP["LNH4"] / (sP["LN"] + EPS0) + sD["muI"] * sD["LNH4"] / (sD["LN"] + EPS0)
    G["NO3"] = nitrif + thN * (oxic - up_NO3)
    G["NH4"] = fixation + thN * (anoxic + inorg_exc + ut_respiration - up_NH4) - nitrif

    # phosphate
    total_prod = sP["muI"] + sD["muI"]
    G["PO4"] = q["thetaP"] * (inorg_exc + ut_respiration + doc_deg - total_prod)

    # iron
    ligand = jmax(q["fe_ligand_max"], q["fe_ligand_ratio"] * DOC - q["fe_ligand_max"])
    lig_agg = q["fe_excess"] * lamFe * jmax(0.0, F["Fe"] - ligand) * Fef
    scav = lamFe * (POC + GOC) * Fef
    non_ass_fe = 0.0
    for g, ff, zz in ((gZ, ffZ, q["micro"]), (gM, ffM, q["meso"])):
        gIFe = g["gIFe"]
        gfIFe = zoo_flux_feeding(g, F, fr, "iron")
        non_ass_fe = non_ass_fe + ((gIFe + gfIFe) - zz["sigma"] * (gIFe + gfIFe)
                                   - zz["iron_ratio"] * g["ge"] * (g["gI"] + ff))
    up_Fe = iron_uptake(q, sP, F) + iron_uptake(q, sD, F)
    G["Fe"] = (lam_pom * F["SFe"] + non_ass_fe + q["meso"]["iron_ratio"] * R_ut
               - up_Fe - lig_agg - (CgFe1 + CgFe2) - scav - BactFe)

    # silicate
    G["Si"] = si_diss - si_uptake

    # carbon
    G["DIC"] = inorg_exc + ut_respiration + doc_deg + calc_diss - calc_prod - total_prod
    G["Alk"] = G["NH4"] - G["NO3"] - 2 * G["CaCO3"]

    # oxygen
    th_r, th_n = q["O2_resp"], q["O2_nitrif"]
    G["O2"] = (th_r * up_NH4 + (th_r + th_n) * up_NO3 + th_n * fixation / thN
               - ((th_r + th_n) * oxic + th_r * anoxic) - th_r * inorg_exc - th_r * ut_respiration
               - th_n * nitrif / thN)
    return G


# ---------------------------------------------------------------- light and update state

def morel_band_coefficients(bands=((400, 500), (500, 600), (600, 700))):
    """Band means of the Morel coefficients (multi_band.jl:89-103, :136-140)
    from the table in the host package."""
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "oceanbiome.jl_amd"))
    from oceanbiome.pisces import morel_band_coefficients as m
    return m(bands)


def surface_PAR_at(spec, t):
    """PAR⁰ of the MultiBand light at time t (the surface boundary value,
    multi_band.jl:151): a constant, "default" (default_surface_PAR,
    Light.jl:47, the PISCES default of PISCES.jl:325-331) or a 2-D array
    PAR⁰[j, i], returned per column."""
    import oracle as O
    if isinstance(spec, str):
        if spec != "default":
            raise ValueError(spec)
        return O.default_surface_PAR(t)
    if np.isscalar(spec):
        return float(spec)
    return np.asarray(spec, dtype=np.float64).reshape(-1)


def multi_band_PAR(PAR0, zc, Chl, kw, e, chi, division):
    """update_MultiBandPhotosyntheticallyActiveRadiation! (multi_band.jl:147-165)
    for columns: Chl [Nz, ncol] bottom to top; returns [nbands, Nz, ncol]."""
    nb = len(kw)
    Nz = Chl.shape[0]
    out = np.zeros((nb,) + Chl.shape)
    for n in range(nb):
        out[n, Nz - 1] = PAR0 * division[n] * np.exp(zc[Nz - 1] * (kw[n] + chi[n] * np.power(Chl[Nz - 1], e[n])))
        for k in range(Nz - 2, -1, -1):
            dz = zc[k] - zc[k + 1]
            out[n, k] = out[n, k + 1] * np.exp(dz * (kw[n] + chi[n] * np.power(Chl[k], e[n])))
    return out


def euphotic_depth(PAR, zc, surface_PAR, cutoff=1e-3):
    """_compute_euphotic_depth! (compute_euphotic_depth.jl:3-33), with the
    surface value taken as the boundary value of PAR at z = 0."""
    Nz, ncol = PAR.shape
    zeu = np.full(ncol, -np.inf)
    thr = surface_PAR * cutoff
    for k in range(Nz - 2, -1, -1):
        hit = (PAR[k] <= thr) & np.isinf(zeu)
        val = zc[k] + (np.log(thr) - np.log(PAR[k])) * (zc[k] - zc[k + 1]) / (np.log(PAR[k]) - np.log(PAR[k + 1]))
        zeu = np.where(hit, val, zeu)
    z0 = zc[0] - (zc[1] - zc[0]) if Nz > 1 else zc[0] - 1.0
    return np.where(np.isfinite(zeu), zeu, z0)


def mixed_layer_mean(C, zf, zmxl):
    """_compute_mixed_layer_mean! (mean_mixed_layer_properties.jl:25-47)."""
    Nz = C.shape[0]
    acc = np.zeros(C.shape[1:])
    depth = np.zeros(C.shape[1:])
    for k in range(Nz - 1, -1, -1):
        zk, zk1 = zf[k], zf[k + 1]
        dzk = zk1 - zk
        dzk1 = np.where(zk1 > zmxl, zk1 - zmxl, 0.0)
        dz = np.where(zk >= zmxl, dzk, dzk1)
        acc = acc + C[k] * dz
        depth = depth + dz
    return acc / depth


def goc_face_speed(zc, minimum=30 / DAY, maximum=200 / DAY):
    """DepthDependantSinkingSpeed (PISCES/common.jl:39-55) evaluated once at
    construction, when the mixed-layer and euphotic depths are still 0: face
    k (1..Nz) takes the centre depth of cell k; face Nz+1 is 0."""
    w = np.zeros(len(zc) + 1)
    w[:-1] = -minimum + (maximum - minimum) * np.minimum(0.0, zc - np.minimum(0.0, 0.0)) / 5000
    return w


# ---------------------------------------------------------------- grid drivers

def params_from_vector(pv):
    """The parameter dict of this module from the host package's OBME_PP_*
    vector (the same numbers the device receives)."""
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "oceanbiome.jl_amd"))
    from oceanbiome.pisces import PHYTO_FIELDS, ZOO_FIELDS, SCALAR_FIELDS
    pv = list(pv)
    n = 0

    def take(fields):
        nonlocal n
        d = {}
        for f in fields:
            d[f] = pv[n]
            n += 1
        return d

    nano = take(PHYTO_FIELDS)
    diatoms = take(PHYTO_FIELDS)
    micro = take(ZOO_FIELDS)
    meso = take(ZOO_FIELDS)
    sc = take(SCALAR_FIELDS)
    q = default_parameters()
    for name, src in (("nano", nano), ("diatoms", diatoms)):
        d = dict(src)
        d["si_limited"] = bool(src["si_limited"])
        q[name] = d
    for name, src in (("micro", micro), ("meso", meso)):
        d = {k: v for k, v in src.items() if not k.startswith("pref_")}
        d["pref"] = dict(P=src["pref_P"], D=src["pref_D"], POC=src["pref_POC"], Z=src["pref_Z"])
        q[name] = d
    for k in ("rain_ratio", "bZ", "bM", "bact_exp", "K_DOC", "K_NO3_b", "K_NH4_b", "K_PO4_b", "K_Fe_b",
              "doc_lambda", "doc_bact_ref", "doc_b", "pom_b", "pom_lambda", "fe_scav_min", "fe_scav_load",
              "bact_fe_eff", "small_bact_fe", "large_bact_fe", "chi0", "si_fast", "si_slow", "caco3_lambda",
              "caco3_exp", "bact_fe_max", "bact_fe_K", "bact_mu0", "nitrif", "fix_max", "fix_KFe", "fix_KPO4",
              "fix_E", "fe_excess", "fe_ligand_max", "fe_ligand_ratio", "O2_resp", "O2_nitrif", "anoxia1",
              "anoxia2", "thetaN", "thetaP", "mixed_layer_shear", "background_shear", "latitude",
              "silicate_climatology"):
        q[k] = sc[k]
    q["doc_agg"] = tuple(sc[f"doc_a{i}"] for i in range(1, 6))
    q["pom_agg"] = tuple(sc[f"pom_a{i}"] for i in range(1, 5))
    q["zmxl"] = sc["mixed_layer_depth"]
    q["kappa"] = sc["kappa"]
    q["POC_speed"] = sc["POC_speed"]
    q["GOC_min"] = sc["GOC_min_speed"]
    q["GOC_max"] = sc["GOC_max_speed"]
    q["gravity"] = sc["gravity"]
    return q


def update_state(q, light, F, zc, zf, t, solver, nthreads=8):
    """The PISCES update-state step (update_state.jl:1-17 with the MultiBand
    light before it) on interior arrays [Nz, Ny, Nx]: returns the auxiliary
    fields the transitions read."""
    import oracle as O
    Nz = zc.shape[0]
    shp = F["T"].shape
    Chl = (F["PChl"] + F["DChl"]).reshape(Nz, -1)
    PAR0 = surface_PAR_at(light["surface_PAR"], t)
    bands = multi_band_PAR(PAR0, zc, Chl, light["kw"], light["e"], light["chi"], light["division"])
    PAR = (bands[0] + bands[1]) + bands[2]
    with np.errstate(divide="ignore", invalid="ignore"):
        zeu = euphotic_depth(PAR, zc, PAR0)
    mlPAR = mixed_layer_mean(PAR, zf, q["zmxl"])
    z3 = np.broadcast_to(zc[:, None, None], shp)
    P = np.abs(z3) * q["gravity"] * 1026 / 100000
    Omega = O.calcite_saturation_batch(solver, F["DIC"], F["Alk"], F["T"], F["S"], P, F["Si"], nthreads=nthreads)
    A = dict(z=z3, zmxl=q["zmxl"], zeu=np.broadcast_to(zeu.reshape(shp[1:]), shp), kappa=q["kappa"],
             mlPAR=np.broadcast_to(mlPAR.reshape(shp[1:]), shp), Omega=Omega, PAR=PAR.reshape(shp),
             PAR1=bands[0].reshape(shp), PAR2=bands[1].reshape(shp), PAR3=bands[2].reshape(shp))
    return A


def face_velocities(q, zc, open_bottom=True):
    Nz = len(zc)
    k = np.arange(1, Nz + 1)
    wP = np.zeros(Nz + 1)
    wP[:Nz] = -q["POC_speed"] * (1.0 if open_bottom else (1 - np.exp((1 - k) / 2.0)))
    wG = np.zeros(Nz + 1)
    wG[:Nz] = -q["GOC_min"] + (q["GOC_max"] - q["GOC_min"]) * np.minimum(0.0, zc - 0.0) / 5000
    return wP, wG


SINKING = (("POC", 0), ("SFe", 0), ("GOC", 1), ("BFe", 1), ("PSi", 1), ("CaCO3", 1))


def surface_fluxes(F, gas, solver):
    """Top-cell gas exchange fluxes (GasExchange, gas_exchange.jl:26-38) per
    column: CO2 (CarbonDioxideGasExchangeBoundaryCondition, no silicate or
    phosphate) and O2 (OxygenGasExchangeBoundaryCondition); 0 where off."""
    import oracle as O
    top = lambda name: F[name][-1]
    shp = top("DIC").shape
    Fc = np.zeros(shp)
    Fo = np.zeros(shp)
    co2, o2 = gas.get("co2"), gas.get("o2")
    for idx in np.ndindex(shp):
        T, S = float(top("T")[idx]), float(top("S")[idx])
        if co2 is not None and co2.enabled:
            Fc[idx] = O.co2_flux(co2, solver, float(top("DIC")[idx]), float(top("Alk")[idx]), T, S,
                                 co2.wind_speed, co2.air_concentration)[0]
        if o2 is not None and o2.enabled:
            Fo[idx] = O.o2_flux(o2, float(top("O2")[idx]), T, S)
    return Fc, Fo


def full_tendencies(q, light, F, zc, zf, dz, t, solver, open_bottom=True, nthreads=8, gas=None):
    """Transitions minus the upwind sinking divergence at every cell, with
    the update state at time t; `gas` ({"co2": ..., "o2": ...}) adds the top
    boundary fluxes of DIC and O2 (−F/Δz in the top cell)."""
    A = update_state(q, light, F, zc, zf, t, solver, nthreads)
    wP, wG = face_velocities(q, zc, open_bottom)
    A["wPOCc"] = ((wP[:-1] + wP[1:]) / 2)[:, None, None]
    A["wGOCc"] = ((wG[:-1] + wG[1:]) / 2)[:, None, None]
    G = tendencies(q, F, A, t)
    for name, which in SINKING:
        w = (wP, wG)[which]
        c = F[name]
        cu = np.zeros_like(c)
        cu[:-1] = c[1:]
        div = (w[1:, None, None] * cu - w[:-1, None, None] * c) / dz[:, None, None]
        G[name] = G[name] - div
    if gas:
        Fc, Fo = surface_fluxes(F, gas, solver)
        G["DIC"][-1] = G["DIC"][-1] - Fc / dz[-1]
        G["O2"][-1] = G["O2"][-1] - Fo / dz[-1]
    return G, A


def conserved_groups(q, fill=float("nan")):
    """conserved_tracers(bgc::PISCES) (PISCES/coupling_utils.jl:11-35) as
    (member names, scalefactors, invalid_fill_value), in the reference's
    group and member order."""
    thP, thN = q["thetaP"], q["thetaN"]
    return [
        (("P", "D", "Z", "M", "DOC", "POC", "GOC", "DIC", "CaCO3"), (1.0,) * 9, fill),
        (("PFe", "DFe", "Z", "M", "SFe", "BFe", "Fe"),
         (1.0, 1.0, q["micro"]["iron_ratio"], q["meso"]["iron_ratio"], 1.0, 1.0, 1.0), fill),
        (("P", "D", "Z", "M", "DOC", "POC", "GOC", "PO4"), (thP,) * 7 + (1.0,), fill),
        (("DSi", "Si", "PSi"), (1.0,) * 3, fill),
        (("NH4", "NO3", "P", "D", "Z", "M", "DOC", "POC", "GOC"), (1.0, 1.0) + (thN,) * 7, fill),
    ]


def scale_negatives(F, groups):
    """ScaleNegativeTracers in place on the arrays F (src/Utils/negative_tracers.jl:
    255-276): per cell and group t = Σ s·c, p = Σ s·c over c > 0 (member
    order), t := fill where t < 0, c := c·t/p where c > 0 or c is not finite,
    else 0."""
    for names, factors, fill in groups:
        t = np.zeros_like(F[names[0]])
        p = np.zeros_like(t)
        for n, s in zip(names, factors):
            vs = F[n] * s
            t = t + vs
            p = p + np.where(F[n] > 0, vs, 0.0)
        t = np.where(t < 0, fill, t)
        with np.errstate(divide="ignore", invalid="ignore"):
            for n in names:
                v = F[n]
                F[n] = np.where(~np.isfinite(v) | (v > 0), v * t / p, 0.0)
    return F


def sediment_flux(q, zc, F, sed, open_bottom=True, scaling=None):
    """The sinking flux through the bottom face of every column (Sediments/
    tracked_fields.jl:53-70): -w c of the bottom cell over the sediment's
    sinking tracers, in their listed order, with their bottom-face velocity;
    the bottom cell after the modifiers when `scaling` is given."""
    wP, wG = face_velocities(q, zc, open_bottom)
    bottom = {name: F[name][0].copy() for name in TRACERS}
    if scaling:
        scale_negatives(bottom, scaling)
    which = dict(SINKING)
    f = None
    for qi in range(sed.n_sinking):
        name = TRACERS[sed.sinking[qi]]
        w0 = (wP, wG)[which[name]][0] if name in which else 0.0
        fq = -(w0 * bottom[name])
        f = fq if f is None else f + fq
    return f


def burial_efficiency(sed, f):
    """e = a + b (f / (k + f))^2 (instant_remineralisation.jl:101-123)."""
    qf = f / (sed.burial_efficiency_half_saturation + f)
    return sed.burial_efficiency_constant1 + sed.burial_efficiency_constant2 * (qf * qf)


def sediment_time_step(sed, state, f, dt):
    """One QuasiAdamsBashforth2 step of the storage (timesteppers.jl:16-43 in
    Oceananigans' AB2 time_step!; the oracle's orc_sediment_time_step over
    the columns): the tendency e f evaluated before the first step from this
    flux, Euler where dt differs from the previous step's."""
    if not math.isfinite(dt):
        return
    if state.iteration == 0:
        state.Gn = burial_efficiency(sed, f) * f
    euler = not (dt == state.last_dt)
    chi = -0.5 if euler else sed.ab2_chi
    Gu = (1.5 + chi) * state.Gn - (0.0 if euler else (0.5 + chi) * state.Gm)
    state.storage = state.storage + dt * Gu
    state.Gm = state.Gn
    state.Gn = burial_efficiency(sed, f) * f
    state.last_dt = dt
    state.iteration += 1


def step(q, light, F, zc, zf, dz, nsteps, dt, t0, solver, open_bottom=True, nthreads=8, scaling=None, gas=None,
         sediment=None):
    """nsteps RK3 steps (Oceananigans γ, ζ; the substep form of
    BoxModel/timesteppers.jl:79-92) of the 24 written tracers, in place;
    `scaling` (conserved_groups) runs at the start of every stage.
    `sediment` = (obme_sediment, SedimentState (Ny, Nx)): an
    InstantRemineralisation sediment under every column: the receiver's
    bottom cell gains (1 - e) f / Δz at every stage (tracer_coupling.jl:30-38)
    and after every substep the storage takes its step of the stage Δt
    (Sediments/update_state.jl:6-17) from the flux of the new state."""
    g = (8 / 15, 5 / 12, 3 / 4)
    z = (None, -17 / 60, -5 / 12)
    cst = (0.0, 8 / 15, 2 / 3)
    sed, state = sediment if sediment is not None else (None, None)
    for s in range(nsteps):
        Gm = None
        for st in range(3):
            t = t0 + s * dt + cst[st] * dt
            if scaling:
                scale_negatives(F, scaling)
            G, _ = full_tendencies(q, light, F, zc, zf, dz, t, solver, open_bottom, nthreads, gas)
            if sed is not None:
                f = sediment_flux(q, zc, F, sed, open_bottom)
                recv = TRACERS[sed.receiver]
                G[recv][0] = G[recv][0] + (1 - burial_efficiency(sed, f)) * f / dz[0]
            for name in WRITTEN:
                if st == 0:
                    F[name] = F[name] + (dt * g[0]) * G[name]
                else:
                    F[name] = F[name] + dt * (g[st] * G[name] + z[st] * Gm[name])
            Gm = G
            if sed is not None:
                sdt = (g[0], g[1] + z[1], g[2] + z[2])[st] * dt
                sediment_time_step(sed, state, sediment_flux(q, zc, F, sed, open_bottom, scaling), sdt)
    return F
