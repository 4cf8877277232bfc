// obme_pisces.h — PISCES (OceanBioME.jl src/Models/AdvectedPopulations/PISCES)
// on gfx950: the 24 per-cell transitions, the MultiBand light, euphotic
// depth and mixed-layer means of the update-state step, the per-cell calcite
// saturation (FP64 carbonate solve with silicate and pressure), and the
// sinking of the particle tracers.
//
// Kernels (one RK3 stage each):
//   k_pisces_columns  thread per water column: MultiBand PAR from the surface
//                     down (multi_band.jl:147-165), euphotic depth
//                     (compute_euphotic_depth.jl:3-33), mixed-layer mean PAR
//                     (mean_mixed_layer_properties.jl:25-47);
//   k_pisces_omega    thread per cell: calcite saturation (FP64,
//                     compute_calcite_saturation.jl:22-36);
//   k_pisces_cell     thread per cell: transitions, sinking divergence and
//                     the RK3 substep (or the tendency output).
// FT is the arithmetic type of the transitions (double: the reference's
// expressions in order; float: the FP32-mixed path).  Julia semantics:
// NaN-propagating min/max (jmin/jmax), eps(0.0) guards, x^2 as x*x.
#pragma once

#include "obme_device.h"

namespace obme {
namespace pisces {

constexpr double DAYS = 86400.0;
constexpr double EPS0 = 4.9406564584124654e-324;  // Julia eps(0.0)

template <typename FT>
__device__ __forceinline__ FT eps0()
{
    if constexpr (sizeof(FT) == 8) return EPS0;
    else return 1.17549435e-38f;  // the smallest normal float keeps the FP32 guards finite
}

// The quotients of the transitions.  FP64: the IEEE quotient (the
// reference's expressions, unchanged).  FP32-mixed: the dividend times the
// hardware reciprocal (within 2 ulp; two instructions, where the compiler's
// quotient scales both operands for denormal and huge divisors, which these
// divisors are not: tracer sums guarded by the smallest normal, parameters).
template <typename FT>
__device__ __forceinline__ FT qdiv(FT a, FT b)
{
    if constexpr (sizeof(FT) == 8) return a / b;
    else return a * __builtin_amdgcn_rcpf(b);
}

// a / k for a constant k (FP32-mixed: times 1/k, folded at compile time).
template <typename FT>
__device__ __forceinline__ FT qdivk(FT a, double k)
{
    if constexpr (sizeof(FT) == 8) return a / k;
    else return a * (float)(1.0 / k);
}

template <typename FT>
__device__ __forceinline__ FT jmin3(FT a, FT b, FT c) { return jmin(jmin(a, b), c); }
template <typename FT>
__device__ __forceinline__ FT jmin4(FT a, FT b, FT c, FT d) { return jmin(jmin(jmin(a, b), c), d); }

template <typename FT>
__device__ __forceinline__ FT m_log(FT x)
{
    if constexpr (sizeof(FT) == 8) return log(x);
    else return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
}

// sind / cosd with the argument reduced modulo 360 first
__device__ __forceinline__ double sind(double x) { return sin(fmod(x, 360.0) * (3.141592653589793 / 180.0)); }
__device__ __forceinline__ double cosd(double x) { return cos(fmod(x, 360.0) * (3.141592653589793 / 180.0)); }

// CBMDayLength()(t, φ) (src/Utils/Utils.jl:17-33) in seconds, FP64.
__device__ __forceinline__ double day_length(double t, double phi)
{
    const double p = 0.833;
    double J = floor(fmod(t, 365 * DAYS) / DAYS);
    if (fmod(t, 365 * DAYS) < 0) J = floor((fmod(t, 365 * DAYS) + 365 * DAYS) / DAYS);
    double theta = 0.216310 + 2 * atan(0.9671396 * tan(0.00860 * (J - 186)));
    double dec = asin(0.39795 * cos(theta)) * (180.0 / 3.141592653589793);
    double L = jmax(-1.0, jmin(1.0, (sind(p) + sind(phi) * sind(dec)) / (cosd(phi) * cosd(dec))));
    return (24 - 24.0 / 180 * (acos(L) * (180.0 / 3.141592653589793))) * 3600.0;
}

// Per-stage uniform inputs of the transitions.
template <typename FT>
struct StageConst {
    FT dl_growth;   // day_length(φ, t): the argument order of growth_rate.jl:30
    FT dl_chl;      // day_length(t, φ): growth_rate.jl:128
};

// Auxiliary values of one cell.
template <typename FT>
struct Aux {
    FT z, zmxl, zeu, kappa, mlPAR, Omega, PAR, PAR1, PAR2, PAR3, wPOCc, wGOCc;
};

// The factors of the transitions that depend on the temperature alone: the
// growth rates' bᵀ of the two phytoplankton and two zooplankton classes
// (growth_rate.jl:25, food_quality_dependant.jl:157), of DOC and POM
// degradation (dissolved_organic_carbon.jl:62, two_size_class.jl:154), the
// rain ratio's 1 + exp(-(T-10)²/25) (nano_diatom_coupling.jl), the ligand
// constant of the free iron (iron/iron.jl:29), the silicate equilibrium
// 10^(6.44 - 968/T_K) and (1 + T/400)^4 (silicate.jl:30-34).  The same
// expressions whether evaluated per stage or once per stepper call.
template <typename FT>
struct TFac {
    FT btN, btD, bZ, bM, doc, pom, rain, kfe, eq, t4;
};
constexpr int TFAC_FIELDS = 10;

template <typename FT>
__device__ __forceinline__ void temperature_factors(const FT* pp, FT T, TFac<FT>& f)
{
    const FT* sc = pp + OBME_PP_SCALARS;
    f.btN = m_pow(pp[OBME_PP_NANO + OBME_PH_BT], T);
    f.btD = m_pow(pp[OBME_PP_DIATOMS + OBME_PH_BT], T);
    f.bZ = m_pow(pp[OBME_PP_MICRO + OBME_PZ_B], T);
    f.bM = m_pow(pp[OBME_PP_MESO + OBME_PZ_B], T);
    f.doc = m_pow(sc[OBME_PS_DOC_B], T);
    f.pom = m_pow(sc[OBME_PS_POM_B], T);
    const FT Tm = T - 10;
    f.rain = 1 + m_exp(qdivk(-(Tm * Tm), 25));
    f.kfe = m_exp(FT(16.27) - qdiv(FT(1565.7), jmax(T + FT(273.15), FT(5))));
    f.eq = m_pow(FT(10), FT(6.44) - qdiv(FT(968), T + FT(273.15)));
    f.t4 = m_pow(1 + qdivk(T, 400), FT(4));
}

// One phytoplankton class (parameters at block b of pp).
template <typename FT>
struct Phyto {
    FT I, IChl, IFe;
    FT L, LFe, LPO4, LN, LNO3, LNH4;
    FT mu, muI;
};

template <typename FT>
__device__ __forceinline__ FT size_factor(const FT* ph, FT I)
{
    const FT Im = ph[OBME_PH_SIZE_THRESHOLD];
    FT I1 = jmin(I, Im);
    FT I2 = jmax(FT(0), I - Im);
    return qdiv(I1 + ph[OBME_PH_SIZE_RATIO] * I2, I1 + I2 + eps0<FT>());
}

template <typename FT>
__device__ __forceinline__ FT nitrogen_lim(FT N1, FT N2, FT K1, FT K2)
{
    return qdiv(K2 * N1, K1 * K2 + K1 * N2 + K2 * N1 + eps0<FT>());
}

// NitrogenIronPhosphateSilicateLimitation (nutrient_limitation.jl:20-71) and
// the growth rate (growth_rate.jl:3-47, GrowthRespirationLimitedProduction :104-111).
template <typename FT>
__device__ __forceinline__ void phyto_state(const FT* ph, const FT* sc, const FT* v, const Aux<FT>& A,
                                            const StageConst<FT>& sk, FT btT, FT I, FT IChl, FT IFe, Phyto<FT>& s)
{
    const FT NO3 = v[OBME_PI_NO3], NH4 = v[OBME_PI_NH4], PO4 = v[OBME_PI_PO4], Si = v[OBME_PI_SI];
    s.I = I;
    s.IChl = IChl;
    s.IFe = IFe;
    const FT thFe = (I == FT(0)) ? FT(0) : qdiv(IFe, I + eps0<FT>());
    const FT thChl = (I == FT(0)) ? FT(0) : qdiv(IChl, 12 * I + eps0<FT>());
    const FT Kb = size_factor(ph, I);
    const FT Kno = ph[OBME_PH_KNO3] * Kb, Knh = ph[OBME_PH_KNH4] * Kb, Kp = ph[OBME_PH_KPO4] * Kb,
             Ksi = ph[OBME_PH_KSI_MIN] * Kb;
    s.LNO3 = nitrogen_lim(NO3, NH4, Kno, Knh);
    s.LNH4 = nitrogen_lim(NH4, NO3, Knh, Kno);
    s.LN = s.LNO3 + s.LNH4;
    s.LPO4 = qdiv(PO4, PO4 + Kp + eps0<FT>());
    const FT thm = FT(1000) * (FT(0.0016) / FT(55.85) * FT(12) * thChl
                               + FT(1.5) * FT(1.21e-5) * FT(14) / (FT(55.85) * FT(7.625)) * s.LN
                               + FT(1.15e-4) * FT(14) / (FT(55.85) * FT(7.625)) * s.LNO3);
    s.LFe = jmin(FT(1), jmax(FT(0), qdiv(thFe - thm, ph[OBME_PH_THETA_FE_OPT])));
    const FT Sc = sc[OBME_PS_SILICATE_CLIMATOLOGY];
    const FT pk = ph[OBME_PH_PK];
    const FT KSi = Ksi + qdiv(7 * (Sc * Sc), pk * pk + Sc * Sc);
    FT LSi = qdiv(Si, Si + KSi);
    if (ph[OBME_PH_SI_LIMITED] == FT(0)) LSi = FT(INFINITY);
    s.L = jmin4(s.LN, s.LPO4, s.LFe, LSi);
    // growth rate μ(L)
    const FT PAR = ph[OBME_PH_BETA1] * A.PAR1 + ph[OBME_PH_BETA2] * A.PAR2 + ph[OBME_PH_BETA3] * A.PAR3;
    const FT dl = sk.dl_growth;
    const FT d = jmax(FT(0), A.zeu - A.zmxl);
    const FT drt = qdiv(d * d, A.kappa);
    const FT mui = ph[OBME_PH_MU0] * btT;
    const FT f1 = qdiv(FT(1.5) * dl, dl + FT(0.5 * DAYS));
    const FT f2 = 1 - qdiv(drt, drt + ph[OBME_PH_DARK_TOLERANCE]);
    const FT alpha = ph[OBME_PH_ALPHA0] * (1 + ph[OBME_PH_BETA] * m_exp(-PAR));
    const FT theta = qdiv(IChl, 12 * I + eps0<FT>());
    const FT fl = 1 - m_exp(qdiv(-alpha * theta * PAR, dl * (ph[OBME_PH_BASAL_RESPIRATION] + ph[OBME_PH_REFERENCE_GROWTH])));
    s.mu = mui * f1 * f2 * fl * s.L;
    s.muI = s.mu * I;
}

// mortality (mixed_mondo.jl:115-133): linear, quadratic
template <typename FT>
__device__ __forceinline__ void phyto_mortality(const FT* ph, const FT* sc, const Phyto<FT>& s, const Aux<FT>& A,
                                                FT& lin, FT& quad)
{
    const FT I = s.I;
    lin = qdiv(ph[OBME_PH_MORT_LIN] * I, I + ph[OBME_PH_MORT_K]) * I;
    const FT L2 = s.L * s.L;
    const FT w = ph[OBME_PH_MORT_W0] + qdiv(ph[OBME_PH_MORT_W1] * FT(0.25) * (1 - L2), FT(0.25) + L2);
    const FT shear = (A.z < A.zmxl) ? sc[OBME_PS_BACKGROUND_SHEAR] : sc[OBME_PS_MIXED_LAYER_SHEAR];
    quad = shear * w * (I * I);
}

// chlorophyll growth (mixed_mondo.jl:84-97; growth_rate.jl:113-148)
template <typename FT>
__device__ __forceinline__ FT chlorophyll_growth(const FT* ph, const Phyto<FT>& s, const Aux<FT>& A,
                                                 const StageConst<FT>& sk)
{
    const FT PAR = ph[OBME_PH_BETA1] * A.PAR1 + ph[OBME_PH_BETA2] * A.PAR2 + ph[OBME_PH_BETA3] * A.PAR3;
    const FT dl = sk.dl_chl;
    const FT f1 = qdiv(FT(1.5) * dl, dl + FT(0.5 * DAYS));
    const FT mucheck = qdiv(s.mu, f1) * dl;
    const FT alpha = ph[OBME_PH_ALPHA0] * (1 + ph[OBME_PH_BETA] * m_exp(-PAR));
    const FT rho = qdiv(12 * mucheck * s.I, alpha * s.IChl * PAR + eps0<FT>()) * s.L;
    const FT th0 = ph[OBME_PH_CHL_MIN], th1 = ph[OBME_PH_CHL_MAX];
    return (1 - ph[OBME_PH_DELTA]) * 12 * (th0 + (th1 - th0) * rho) * s.mu * s.I;
}

// iron uptake (mixed_mondo.jl:161-181)
template <typename FT>
__device__ __forceinline__ FT iron_uptake(const FT* ph, const Phyto<FT>& s, FT btT, FT Fe)
{
    const FT fe_max = ph[OBME_PH_FE_MAX];
    const FT mui = ph[OBME_PH_MU0] * btT;
    const FT K = ph[OBME_PH_KFE_UPTAKE] * size_factor(ph, s.I);
    const FT L1 = qdiv(Fe, Fe + K + eps0<FT>());
    const FT L2 = 4 - qdiv(FT(4.5) * s.LFe, s.LFe + 1);
    // (1 - r) / (1.05 - r), r = θFe / θFe_max, θFe = IFe / (I + ε)
    FT pole;
    if constexpr (sizeof(FT) == 8) {
        const FT thFe = s.IFe / (s.I + eps0<FT>());
        const FT r = thFe / fe_max;
        pole = (1 - r) / (FT(1.05) - r);
    } else {
        // the same quotient times F / F, F = θFe_max (I + ε):
        // (F - IFe) / (1.05 F - IFe), with F's product exact inside the FMAs
        // and 1.05 split into two floats.  Near the pole (r -> 1.05) the
        // denominator is a small difference, and rounding r first would lose
        // its leading digits.
        const FT Ie = s.I + eps0<FT>();
        const FT Fh = fe_max * Ie, Fl = fmaf(fe_max, Ie, -Fh);
        const FT c = 1.05f, cl = (float)(1.05 - (double)1.05f);
        const FT num = fmaf(fe_max, Ie, -s.IFe);
        const FT den = fmaf(c, Fh, -s.IFe) + fmaf(c, Fl, cl * Fh);
        pole = qdiv(num, den);
    }
    return (1 - ph[OBME_PH_DELTA]) * fe_max * L1 * L2 * jmax(FT(0), pole) * mui * s.I;
}

// silicate uptake (mixed_mondo.jl:195-226)
template <typename FT>
__device__ __forceinline__ FT silicate_uptake(const FT* ph, const FT* sc, const Phyto<FT>& s, FT btT, FT Si)
{
    const FT mui = ph[OBME_PH_MU0] * btT;
    const FT L1 = qdiv(Si, Si + ph[OBME_PH_SI_K1] + eps0<FT>());
    const FT K2 = ph[OBME_PH_SI_K2];
    const FT L2 = (sc[OBME_PS_LATITUDE] < FT(0)) ? qdiv(Si * Si * Si, Si * Si * Si + K2 * K2 * K2) : FT(0);
    const FT F1 = jmin4(qdiv(s.mu, mui * s.L + eps0<FT>()), s.LFe, s.LPO4, s.LN);
    const FT F2 = jmin(FT(1), FT(2.2) * jmax(FT(0), L1 - FT(0.5)));
    const FT th1 = ph[OBME_PH_SI_THETA0] * L1 * jmin(FT(5.4), (FT(4.4) * m_exp(FT(-4.23) * F1) * F2 + 1) * (1 + 2 * L2));
    return (1 - ph[OBME_PH_DELTA]) * th1 * s.mu * s.I;
}

// QualityDependantZooplankton grazing of one class (food_quality_dependant.jl:138-183,
// iron_grazing.jl:1-36); prey order P, D, POC (, Z for the meso class).
template <typename FT>
struct Zoo {
    FT I, tsg, avail, gI, ge, gIFe, ff, ffFe, tf;
};

template <typename FT, int NPREY>
__device__ __forceinline__ void zoo_state(const FT* zp, const FT* v, FT bT, FT micro_iron, FT wPOCc, FT wGOCc, FT I,
                                          Zoo<FT>& g)
{
    g.I = I;
    g.tf = bT;
    const FT base = zp[OBME_PZ_G0] * g.tf;
    const FT food[4] = {v[OBME_PI_P], v[OBME_PI_D], v[OBME_PI_POC], v[OBME_PI_Z]};
    const FT iron[4] = {qdiv(v[OBME_PI_PFE], v[OBME_PI_P] + eps0<FT>()), qdiv(v[OBME_PI_DFE], v[OBME_PI_D] + eps0<FT>()),
                        qdiv(v[OBME_PI_SFE], v[OBME_PI_POC] + eps0<FT>()), micro_iron};
    const FT pref[4] = {zp[OBME_PZ_PREF_P], zp[OBME_PZ_PREF_D], zp[OBME_PZ_PREF_POC], zp[OBME_PZ_PREF_Z]};
    const FT J = zp[OBME_PZ_J];
    FT total_food = food[0] * pref[0];
    FT avail = jmax(FT(0), food[0] - J) * pref[0];
    FT total_iron = iron[0] * pref[0];
    FT s = jmax(FT(0), food[0] - J) * pref[0] * iron[0];
#pragma unroll
    for (int n = 1; n < NPREY; n++) {
        total_food = total_food + food[n] * pref[n];
        avail = avail + jmax(FT(0), food[n] - J) * pref[n];
        total_iron = total_iron + iron[n] * pref[n];
        s = s + jmax(FT(0), food[n] - J) * pref[n] * iron[n];
    }
    const FT clg = jmax(FT(0), avail - jmin(avail / 2, zp[OBME_PZ_FOOD_THRESHOLD]));
    g.tsg = qdiv(base * clg, zp[OBME_PZ_K] + total_food);
    g.avail = avail;
    const FT igr = qdiv(total_iron, zp[OBME_PZ_IRON_RATIO] * g.tsg + eps0<FT>());
    const FT fq = jmin(FT(1), igr);
    g.ge = fq * jmin(zp[OBME_PZ_E0], (1 - zp[OBME_PZ_SIGMA]) * igr);
    g.gI = g.tsg * I;
    g.gIFe = qdiv(s * g.tsg, avail + eps0<FT>()) * I;
    // flux feeding on the sinking particles (food_quality_dependant.jl:185-199, iron_grazing.jl:38-53)
    const FT ffb = zp[OBME_PZ_FF] * g.tf;
    g.ff = ffb * (v[OBME_PI_POC] * wPOCc + v[OBME_PI_GOC] * wGOCc) * I;
    g.ffFe = ffb * (v[OBME_PI_SFE] * wPOCc + v[OBME_PI_BFE] * wGOCc) * I;
}

// grazing of one class on one prey (food_quality_dependant.jl:215-243)
template <typename FT>
__device__ __forceinline__ FT zoo_on(const FT* zp, const Zoo<FT>& g, FT pref, FT prey)
{
    return qdiv(pref * jmax(FT(0), prey - zp[OBME_PZ_J]) * g.tsg, g.avail + eps0<FT>()) * g.I;
}

template <typename FT>
__device__ __forceinline__ FT anoxia(const FT* sc, FT O2)
{
    return jmin(FT(1), jmax(FT(0), qdiv(FT(0.4) * (sc[OBME_PS_ANOXIA1] - O2), sc[OBME_PS_ANOXIA2] + O2)));
}

// The 24 transitions of one cell (v: tracers in OBME_PI order).
template <typename FT>
__device__ __forceinline__ void pisces_tendencies(const FT* pp, const FT* v, const Aux<FT>& A,
                                                  const StageConst<FT>& sk, const TFac<FT>& tf, FT* G)
{
    const FT* phN = pp + OBME_PP_NANO;
    const FT* phD = pp + OBME_PP_DIATOMS;
    const FT* zpZ = pp + OBME_PP_MICRO;
    const FT* zpM = pp + OBME_PP_MESO;
    const FT* sc = pp + OBME_PP_SCALARS;
    const FT T = v[OBME_PI_T];
    const FT e0 = eps0<FT>();

    Phyto<FT> sP, sD;
    phyto_state(phN, sc, v, A, sk, tf.btN, v[OBME_PI_P], v[OBME_PI_PCHL], v[OBME_PI_PFE], sP);
    phyto_state(phD, sc, v, A, sk, tf.btD, v[OBME_PI_D], v[OBME_PI_DCHL], v[OBME_PI_DFE], sD);
    Zoo<FT> gZ, gM;
    const FT micro_iron = zpZ[OBME_PZ_IRON_RATIO];
    zoo_state<FT, 3>(zpZ, v, tf.bZ, micro_iron, A.wPOCc, A.wGOCc, v[OBME_PI_Z], gZ);
    zoo_state<FT, 4>(zpM, v, tf.bM, micro_iron, A.wPOCc, A.wGOCc, v[OBME_PI_M], gM);
    const FT O2 = v[OBME_PI_O2];
    const FT an = anoxia(sc, O2);
    // zooplankton mortality (food_quality_dependant.jl:201-227)
    auto zmort = [&](const FT* zp, const Zoo<FT>& g, FT& mort, FT& lin) {
        const FT cf = qdiv(g.I, g.I + zp[OBME_PZ_KM]);
        mort = g.tf * g.I * (zp[OBME_PZ_M0] * g.I + zp[OBME_PZ_R] * (cf + 3 * an));
        lin = g.tf * zp[OBME_PZ_R] * (cf + 3 * an) * g.I;
    };
    FT mZ, linZ, mM, linM;
    zmort(zpZ, gZ, mZ, linZ);
    zmort(zpM, gM, mM, linM);
    (void)linZ;
    auto graz = [&](int prey_pref, FT prey) {
        return zoo_on(zpZ, gZ, zpZ[prey_pref], prey) + zoo_on(zpM, gM, zpM[prey_pref], prey);
    };
    const FT shear_lt = (A.z < A.zmxl) ? sc[OBME_PS_BACKGROUND_SHEAR] : sc[OBME_PS_MIXED_LAYER_SHEAR];

    // phytoplankton (mixed_mondo_nano_diatoms.jl:47-111)
    FT Plin, Pquad, Dlin, Dquad;
    phyto_mortality(phN, sc, sP, A, Plin, Pquad);
    phyto_mortality(phD, sc, sD, A, Dlin, Dquad);
    const FT grazP = graz(OBME_PZ_PREF_P, v[OBME_PI_P]);
    const FT grazD = graz(OBME_PZ_PREF_D, v[OBME_PI_D]);
    {
        const FT death = Plin + Pquad;
        G[OBME_PI_P] = (1 - phN[OBME_PH_DELTA]) * sP.muI - death - grazP;
        const FT thChl = qdiv(sP.IChl, 12 * sP.I + e0);
        G[OBME_PI_PCHL] = chlorophyll_growth(phN, sP, A, sk) - (death + grazP) * thChl * 12;
        const FT thFe = qdiv(sP.IFe, sP.I + e0);
        G[OBME_PI_PFE] = iron_uptake(phN, sP, tf.btN, v[OBME_PI_FE]) - (death + grazP) * thFe;
    }
    {
        const FT death = Dlin + Dquad;
        G[OBME_PI_D] = (1 - phD[OBME_PH_DELTA]) * sD.muI - death - grazD;
        const FT thChl = qdiv(sD.IChl, 12 * sD.I + e0);
        G[OBME_PI_DCHL] = chlorophyll_growth(phD, sD, A, sk) - (death + grazD) * thChl * 12;
        const FT thFe = qdiv(sD.IFe, sD.I + e0);
        G[OBME_PI_DFE] = iron_uptake(phD, sD, tf.btD, v[OBME_PI_FE]) - (death + grazD) * thFe;
    }
    const FT si_uptake = silicate_uptake(phD, sc, sD, tf.btD, v[OBME_PI_SI]);
    const FT thSi = qdiv(v[OBME_PI_DSI], v[OBME_PI_D] + e0);
    G[OBME_PI_DSI] = si_uptake - (Dlin + Dquad + grazD) * thSi;

    // zooplankton (micro_and_meso.jl:36-48)
    G[OBME_PI_Z] = gZ.ge * (gZ.gI + gZ.ff) - mZ - zoo_on(zpM, gM, zpM[OBME_PZ_PREF_Z], v[OBME_PI_Z]);
    G[OBME_PI_M] = gM.ge * (gM.gI + gM.ff) - mM;

    const FT sigZ = zpZ[OBME_PZ_SIGMA], sigM = zpM[OBME_PZ_SIGMA];
    const FT excZ = (1 - sigZ - gZ.ge) * (gZ.gI + gZ.ff);
    const FT excM = (1 - sigM - gM.ge) * (gM.gI + gM.ff);
    const FT inorg_exc = zpZ[OBME_PZ_GAMMA] * excZ + zpM[OBME_PZ_GAMMA] * excM;
    const FT org_exc = (1 - zpZ[OBME_PZ_GAMMA]) * excZ + (1 - zpM[OBME_PZ_GAMMA]) * excM;
    // upper trophic waste of the meso class (mortality_waste.jl)
    const FT utw = qdiv(FT(1), 1 - zpM[OBME_PZ_E0]) * zpM[OBME_PZ_M0] * gM.tf * (gM.I * gM.I);
    const FT R_ut = (1 - zpM[OBME_PZ_E0] - zpM[OBME_PZ_SIGMA]) * utw;
    const FT fecal = zpM[OBME_PZ_SIGMA] * utw;
    const FT ut_excretion = (1 - zpM[OBME_PZ_GAMMA]) * R_ut;
    const FT ut_respiration = zpM[OBME_PZ_GAMMA] * R_ut;

    // bacteria (micro_and_meso.jl:85-132)
    const FT zm = jmin(A.zmxl, A.zeu);
    const FT surface_bact = jmin(sc[OBME_PS_BACT_MAX], sc[OBME_PS_BZ] * v[OBME_PI_Z] + sc[OBME_PS_BM] * v[OBME_PI_M]);
    const FT Bact = ((A.z >= zm) ? FT(1) : m_pow(qdiv(zm, A.z), sc[OBME_PS_BACT_EXP])) * surface_bact;
    const FT DOC = v[OBME_PI_DOC], NO3 = v[OBME_PI_NO3], NH4 = v[OBME_PI_NH4], PO4 = v[OBME_PI_PO4],
             Fe = v[OBME_PI_FE];
    const FT KNO = sc[OBME_PS_K_NO3_B], KNH = sc[OBME_PS_K_NH4_B];
    const FT LBact = jmin3(qdiv(KNO * NH4 + KNH * NO3, KNO * KNH + KNO * NH4 + KNH * NO3),
                           qdiv(PO4, PO4 + sc[OBME_PS_K_PO4_B]), qdiv(Fe, Fe + sc[OBME_PS_K_FE_B]))
                     * qdiv(DOC, DOC + sc[OBME_PS_K_DOC]);

    // DOC (dissolved_organic_carbon.jl:39-106)
    const FT doc_deg = qdiv(sc[OBME_PS_DOC_LAMBDA] * tf.doc * LBact * Bact, sc[OBME_PS_DOC_BACT_REF]) * DOC;
    const FT POC = v[OBME_PI_POC], GOC = v[OBME_PI_GOC];
    const FT Phi1 = shear_lt * (sc[OBME_PS_DOC_A1] * DOC + sc[OBME_PS_DOC_A2] * POC) * DOC;
    const FT Phi2 = shear_lt * (sc[OBME_PS_DOC_A3] * GOC) * DOC;
    const FT Phi3 = (sc[OBME_PS_DOC_A4] * POC + sc[OBME_PS_DOC_A5] * DOC) * DOC;
    const FT doc_agg = Phi1 + Phi2 + Phi3;
    const FT lam_pom = sc[OBME_PS_POM_LAMBDA] * tf.pom * (1 - FT(0.45) * an);
    const FT pom_deg_POC = lam_pom * POC;
    const FT exudate = phN[OBME_PH_DELTA] * sP.muI + phD[OBME_PH_DELTA] * sD.muI;
    G[OBME_PI_DOC] = exudate + ut_excretion + org_exc + pom_deg_POC - doc_deg - doc_agg;
    const FT oxic = (1 - an) * doc_deg;
    const FT anoxic = an * doc_deg;

    // particles: aggregation, rain ratio and the mortality partitions
    // (two_size_class.jl:109-143, nano_diatom_coupling.jl)
    const FT pom_agg = shear_lt * (sc[OBME_PS_POM_A1] * (POC * POC) + sc[OBME_PS_POM_A2] * POC * GOC)
                       + sc[OBME_PS_POM_A3] * POC * GOC + sc[OBME_PS_POM_A4] * (POC * POC);
    const FT PAR = A.PAR;
    const FT L_CaCO3 = jmin3(sP.LN, qdiv(Fe, Fe + FT(0.05)), sP.LPO4);
    const FT R = sc[OBME_PS_RAIN_RATIO] * L_CaCO3 * jmax(FT(1), v[OBME_PI_P] / 2)
                 * qdiv(jmax(FT(0), PAR - 1), 4 + PAR) * qdiv(FT(30), 30 + PAR) * jmax(FT(0), qdiv(T, T + FT(0.1)))
                 * tf.rain * jmin(FT(1), qdiv(FT(-50), A.zmxl));
    const FT small_mort = (1 - R / 2) * (Plin + Pquad) + Dlin / 2;
    const FT large_mort = R / 2 * (Plin + Pquad) + Dlin / 2 + Dquad;
    const FT thP = qdiv(v[OBME_PI_PFE], v[OBME_PI_P] + e0);
    const FT thD = qdiv(v[OBME_PI_DFE], v[OBME_PI_D] + e0);
    const FT small_mort_fe = (1 - R / 2) * (Plin + Pquad) * thP + Dlin * thD / 2;
    const FT large_mort_fe = R / 2 * (Plin + Pquad) * thP + (Dlin / 2 + Dquad) * thD;
    const FT ffbZ = zpZ[OBME_PZ_FF] * gZ.tf, ffbM = zpM[OBME_PZ_FF] * gM.tf;
    const FT tot_graz_POC = graz(OBME_PZ_PREF_POC, POC) + (ffbZ * (POC * A.wPOCc) * gZ.I + ffbM * (POC * A.wPOCc) * gM.I);
    const FT tot_graz_GOC = ffbZ * (GOC * A.wGOCc) * gZ.I + ffbM * (GOC * A.wGOCc) * gM.I;

    // carbon.jl:3-50
    G[OBME_PI_POC] = (sigZ * (gZ.gI + gZ.ff) + small_mort + mZ + (Phi1 + Phi3) + lam_pom * GOC
                      - tot_graz_POC - pom_agg - pom_deg_POC);
    G[OBME_PI_GOC] = (sigM * (gM.gI + gM.ff) + large_mort + linM + fecal + pom_agg + Phi2
                      - tot_graz_GOC - lam_pom * GOC);

    // iron in particles (iron.jl:2-89) with the free iron of SimpleIron (iron/iron.jl:25-37)
    const FT ligands = jmax(FT(0.6), FT(0.09) * (DOC + 40) - 3);
    const FT Kfe = tf.kfe;
    const FT Dl = 1 + Kfe * ligands - Kfe * Fe;
    const FT Fef = qdiv(-Dl + m_sqrt(Dl * Dl + 4 * Kfe * Fe), 2 * Kfe);
    const FT colloidal = FT(0.5) * (Fe - Fef);
    const FT CgFe1 = qdiv((Phi1 + Phi3) * colloidal, DOC + e0);
    const FT CgFe2 = qdiv(Phi2 * colloidal, DOC + e0);
    const FT lamFe = sc[OBME_PS_FE_SCAV_MIN] + sc[OBME_PS_FE_SCAV_LOAD] * (POC + GOC + v[OBME_PI_CACO3] + v[OBME_PI_PSI]);
    const FT BactFe = qdiv(sc[OBME_PS_BACT_MU0] * tf.pom * LBact * sc[OBME_PS_BACT_FE_MAX] * Fe,
                           Fe + sc[OBME_PS_BACT_FE_K]) * Bact * sc[OBME_PS_BACT_FE_EFF];
    const FT thS = qdiv(v[OBME_PI_SFE], POC + e0);
    const FT thB = qdiv(v[OBME_PI_BFE], GOC + e0);
    G[OBME_PI_SFE] = (sigZ * (gZ.gIFe + gZ.ffFe) + small_mort_fe + mZ * zpZ[OBME_PZ_IRON_RATIO]
                      + lam_pom * v[OBME_PI_BFE] + lamFe * POC * Fef + sc[OBME_PS_SMALL_BACT_FE] * BactFe + CgFe1
                      - tot_graz_POC * thS - pom_agg * thS - lam_pom * v[OBME_PI_SFE]);
    G[OBME_PI_BFE] = (sigM * (gM.gIFe + gM.ffFe) + large_mort_fe + linM * zpM[OBME_PZ_IRON_RATIO]
                      + fecal * zpM[OBME_PZ_IRON_RATIO] + lamFe * GOC * Fef + sc[OBME_PS_LARGE_BACT_FE] * BactFe
                      + CgFe2 + pom_agg * thS - tot_graz_GOC * thB - lam_pom * v[OBME_PI_BFE]);

    // particulate silicate (silicate.jl:1-48)
    const FT chi = sc[OBME_PS_CHI0] * ((A.z >= zm) ? FT(1)
                                                   : m_exp(qdiv((sc[OBME_PS_SI_FAST] - sc[OBME_PS_SI_SLOW]) * (zm - A.z), A.wGOCc)));
    const FT lam0 = chi * sc[OBME_PS_SI_FAST] + (1 - chi) * sc[OBME_PS_SI_SLOW];
    const FT eq = tf.eq;
    const FT sat = qdiv(eq - v[OBME_PI_SI], eq);
    const FT t4 = tf.t4;
    const FT lam_si = lam0 * (FT(0.225) * (1 + qdivk(T, 15)) * sat + FT(0.775) * m_pow(t4 * sat, FT(9)));
    const FT si_diss = lam_si * v[OBME_PI_PSI];
    G[OBME_PI_PSI] = (grazD + Dlin + Dquad) * thSi - si_diss;

    // calcite (calcite.jl:1-19, nano_diatom_coupling.jl:96-104)
    const FT calc_loss = zpZ[OBME_PZ_ETA] * zoo_on(zpZ, gZ, zpZ[OBME_PZ_PREF_P], v[OBME_PI_P])
                         + zpM[OBME_PZ_ETA] * zoo_on(zpM, gM, zpM[OBME_PZ_PREF_P], v[OBME_PI_P]);
    const FT calc_prod = R * (calc_loss + (Plin + Pquad) / 2);
    const FT calc_diss = sc[OBME_PS_CACO3_LAMBDA] * m_pow(jmax(FT(0), 1 - A.Omega), sc[OBME_PS_CACO3_EXP])
                         * v[OBME_PI_CACO3];
    G[OBME_PI_CACO3] = calc_prod - calc_diss;

    // nitrogen (nitrogen/nitrate_ammonia.jl:22-88)
    const FT thN = sc[OBME_PS_THETAN];
    const FT nitrif = qdiv(sc[OBME_PS_NITRIF] * NH4, 1 + A.mlPAR) * (1 - an);
    const FT fix_limit = (sP.LN >= FT(0.8)) ? FT(0.01) : 1 - sP.LN;
    const FT mu_nano = phN[OBME_PH_MU0] * tf.btN;
    const FT growth_req = jmax(FT(0), mu_nano - FT(2.15));
    const FT nut_lim = jmin(qdiv(Fe, Fe + sc[OBME_PS_FIX_KFE]), qdiv(PO4, PO4 + sc[OBME_PS_FIX_KPO4]));
    const FT light_lim = 1 - m_exp(qdiv(-PAR, sc[OBME_PS_FIX_E]));
    const FT fixation = sc[OBME_PS_FIX_MAX] * growth_req * fix_limit * nut_lim * light_lim;
    const FT up_NO3 = qdiv(sP.muI * sP.LNO3, sP.LN + e0) + qdiv(sD.muI * sD.LNO3, sD.LN + e0);
    const FT up_NH4 = qdiv(sP.muI * sP.LNH4, sP.LN + e0) + qdiv(sD.muI * sD.LNH4, sD.LN + e0);
    G[OBME_PI_NO3] = nitrif + thN * (oxic - up_NO3);
    G[OBME_PI_NH4] = fixation + thN * (anoxic + inorg_exc + ut_respiration - up_NH4) - nitrif;

    // phosphate (phosphate.jl:21-33)
    const FT total_prod = sP.muI + sD.muI;
    G[OBME_PI_PO4] = sc[OBME_PS_THETAP] * (inorg_exc + ut_respiration + doc_deg - total_prod);

    // iron (iron/simple_iron.jl:19-53)
    const FT ligand = jmax(sc[OBME_PS_FE_LIGAND_MAX], sc[OBME_PS_FE_LIGAND_RATIO] * DOC - sc[OBME_PS_FE_LIGAND_MAX]);
    const FT lig_agg = sc[OBME_PS_FE_EXCESS] * lamFe * jmax(FT(0), Fe - ligand) * Fef;
    const FT scav = lamFe * (POC + GOC) * Fef;
    const FT naZ = (gZ.gIFe + gZ.ffFe) - sigZ * (gZ.gIFe + gZ.ffFe) - zpZ[OBME_PZ_IRON_RATIO] * gZ.ge * (gZ.gI + gZ.ff);
    const FT naM = (gM.gIFe + gM.ffFe) - sigM * (gM.gIFe + gM.ffFe) - zpM[OBME_PZ_IRON_RATIO] * gM.ge * (gM.gI + gM.ff);
    const FT up_Fe = iron_uptake(phN, sP, tf.btN, Fe) + iron_uptake(phD, sD, tf.btD, Fe);
    G[OBME_PI_FE] = (lam_pom * v[OBME_PI_SFE] + (naZ + naM) + zpM[OBME_PZ_IRON_RATIO] * R_ut
                     - up_Fe - lig_agg - (CgFe1 + CgFe2) - scav - BactFe);

    // silicate (silicate.jl:20-26)
    G[OBME_PI_SI] = si_diss - si_uptake;

    // inorganic carbon (inorganic_carbon.jl:32-57)
    G[OBME_PI_DIC] = inorg_exc + ut_respiration + doc_deg + calc_diss - calc_prod - total_prod;
    G[OBME_PI_ALK] = G[OBME_PI_NH4] - G[OBME_PI_NO3] - 2 * G[OBME_PI_CACO3];

    // oxygen (oxygen.jl:30-50)
    const FT th_r = sc[OBME_PS_O2_RESP], th_n = sc[OBME_PS_O2_NITRIF];
    G[OBME_PI_O2] = (th_r * up_NH4 + (th_r + th_n) * up_NO3 + qdiv(th_n * fixation, thN)
                     - ((th_r + th_n) * oxic + th_r * anoxic) - th_r * inorg_exc - th_r * ut_respiration
                     - qdiv(th_n * nitrif, thN));
}

// ---------------------------------------------------------------- calcite saturation (FP64)

// Millero (2007) pressure correction (equilibrium_constants.jl:29-38), a: a0 a1 a2 b0 b1
__device__ __forceinline__ double pressure_correction(const double* a, double Tk, double P)
{
    const double Tc = Tk - 273.15;
    const double dV = a[0] + a[1] * Tc + a[2] * (Tc * Tc);
    const double dk = a[3] + a[4] * Tc;
    const double RT = 83.14472 * Tk;
    return exp((-dV + 0.5 * dk * P) * P / RT);
}

// Constants of the calcite saturation that depend only on T, S and the
// pressure (fixed through a step of frozen physics): the unit conversion
// 1e-3/ρ(T, S, P), the pressure-corrected equilibrium constants that enter
// the residual (phosphate is 0, so KP1-3 do not), the salinity totals, the
// calcium concentration and KSP.
struct OmegaConst {
    double conv, K1, K2, KB, KW, KS, KF, KSi, boron, sulfate, fluoride, ca, KSP;
};
constexpr int OMEGA_CONST_FIELDS = 13;

// The salinity totals and the calcium concentration (functions of S alone).
__device__ __forceinline__ void omega_salinity_terms(double S, OmegaConst& c)
{
    c.boron = 0.000232 / 10.811 * S / 1.80655;
    c.sulfate = 0.14 / 96.06 * S / 1.80655;
    c.fluoride = 0.000067 / 18.9984 * S / 1.80655;
    c.ca = 0.0103 * S / 35;
}

// calcite_saturation (CarbonChemistry/calcite_concentration.jl:1-79) with
// silicate and the pressure P in bar: constants pressure-corrected, density
// at P.  First half: the constants.
__device__ __forceinline__ void omega_constants(double T, double S, double P, OmegaConst& c)
{
    static constexpr double PC_K1[5] = {-25.50, 0.1271, 0.0, -0.00308, 0.0000877};
    static constexpr double PC_K2[5] = {-15.82, -0.0219, 0.0, 0.00113, -0.0001475};
    static constexpr double PC_KB[5] = {-29.48, 0.1622, -0.0026080, -0.00284, 0.0};
    static constexpr double PC_KW[5] = {-20.02, 0.1119, -0.001409, -0.00513, 0.0000794};
    static constexpr double PC_KS[5] = {-18.03, 0.0466, 0.000316, -0.00453, 0.00009};
    static constexpr double PC_KF[5] = {-9.78, -0.0090, -0.000942, -0.00391, 0.000054};
    static constexpr double PC_CAL[5] = {-48.76, 0.5304, -0.0, -0.01176, 0.0003692};
    CarbonParams p;
    omega_salinity_terms(S, c);
    const double rho = teos10_density(T, S, P);
    const double Tk = T + 273.15;
    c.conv = 1e-3 / rho;
    carbon_constants(Tk, S, p);
    c.K1 = p.K1 * pressure_correction(PC_K1, Tk, P);
    c.K2 = p.K2 * pressure_correction(PC_K2, Tk, P);
    c.KB = p.KB * pressure_correction(PC_KB, Tk, P);
    c.KW = p.KW * pressure_correction(PC_KW, Tk, P);
    c.KS = p.KS * pressure_correction(PC_KS, Tk, P);
    c.KF = p.KF * pressure_correction(PC_KF, Tk, P);
    c.KSi = p.KSi;
    const double lnK_therm = -171.9065 + -0.077993 * Tk + 2839.319 / Tk + 71.595 * log10(Tk);
    const double lnK_sea = ((-0.77712 + 0.0028426 * Tk + 178.34 / Tk) * sqrt(S) + -0.07711 * S + 0.0041249 * pow(S, 1.5));
    c.KSP = pressure_correction(PC_CAL, Tk, P) * pow(10.0, lnK_therm + lnK_sea);
}

__device__ __forceinline__ CarbonParams omega_params(const OmegaConst& c, double DIC, double Alk, double Si)
{
    CarbonParams p;
    p.boron = c.boron;
    p.sulfate = c.sulfate;
    p.fluoride = c.fluoride;
    p.Alk = Alk * c.conv;
    p.DIC = DIC * c.conv;
    p.phosphate = 0.0;
    p.silicate = Si * c.conv;
    p.K1 = c.K1;
    p.K2 = c.K2;
    p.KB = c.KB;
    p.KW = c.KW;
    p.KS = c.KS;
    p.KF = c.KF;
    p.KP1 = p.KP2 = p.KP3 = 0.0;
    p.KSi = c.KSi;
    p.K0 = 0.0;
    p.Tk = 0.0;
    return p;
}

// Second half: Ω from [H+] (calcite_concentration.jl: CO3 from H, then
// Ca·CO3/KSP).
__device__ __forceinline__ double omega_from_H(const OmegaConst& c, const CarbonParams& p, double H)
{
    const double denom1 = (H * (H + p.K1));
    const double denom2 = (1.0 + p.K1 * p.K2 / denom1);
    const double co3 = p.DIC * p.K1 * p.K2 / denom1 / denom2;
    return c.ca * co3 / c.KSP;
}

// The reference solve: damped Newton from 10^-pH₀.
__device__ __forceinline__ double calcite_saturation(double DIC, double Alk, double T, double S, double Si, double P,
                                                     const obme_carbon_solver& s, int& evals)
{
    OmegaConst c;
    omega_constants(T, S, P, c);
    const CarbonParams p = omega_params(c, DIC, Alk, Si);
    const double H = solve_for_H<true>(p, s, pow(10.0, -s.initial_pH_guess), evals);
    return omega_from_H(c, p, H);
}

}  // namespace pisces
}  // namespace obme
