// obme_device.h — device-side building blocks of the gfx950 biogeochemistry
// step: Julia scalar semantics, precision-dispatched math, the
// NutrientsPlanktonDetritus tendencies (NPZD, LOBSTER + CarbonateSystem) and
// the FP64 carbonate chemistry (constants, alkalinity residual, damped
// Newton, pCO2, CO2 gas-exchange flux).
//
// FT = double is the plain FP64 path (the reference's arithmetic, operand
// order kept; contraction into FMA is the only change).  FT = float is the
// FP32-mixed path: native v_exp/v_log/v_rcp based transcendentals and
// divisions on the vector ALUs; the carbonate chemistry is always FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "../../include/obme.h"

namespace obme {

// ---------------------------------------------------------------- scalars

// Julia Base min/max for IEEE floats (NaN propagates; ±0 ordered).
template <typename FT>
__device__ __forceinline__ FT jmin(FT x, FT y)
{
    FT d = x - y;
    FT a = __builtin_signbit(d) ? x : y;
    return (__builtin_isnan(x) || __builtin_isnan(y)) ? d : a;
}
template <typename FT>
__device__ __forceinline__ FT jmax(FT x, FT y)
{
    FT d = x - y;
    FT a = __builtin_signbit(d) ? y : x;
    return (__builtin_isnan(x) || __builtin_isnan(y)) ? d : a;
}

// Julia eps(x): spacing of floats at |x|; eps(0) is the smallest subnormal.
__device__ __forceinline__ double jeps(double x)
{
    uint64_t b = __double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff);
    uint64_t r = (e > 52) ? ((uint64_t)(e - 52) << 52) : ((uint64_t)1 << ((e > 1 ? e : 1) - 1));
    double out = __longlong_as_double(r);
    return (e == 0x7ff) ? __longlong_as_double(0x7ff8000000000000ull) : out;
}
__device__ __forceinline__ float jeps(float x)
{
    uint32_t b = __float_as_uint(x);
    int e = (int)((b >> 23) & 0xff);
    uint32_t r = (e > 23) ? ((uint32_t)(e - 23) << 23) : ((uint32_t)1 << ((e > 1 ? e : 1) - 1));
    float out = __uint_as_float(r);
    return (e == 0xff) ? __uint_as_float(0x7fc00000u) : out;
}

template <typename FT> struct Tiny;
template <> struct Tiny<double> { static constexpr double eps0 = 4.9406564584124654e-324; };
template <> struct Tiny<float> { static constexpr float eps0 = 1.40129846e-45f; };

// Precision-dispatched math.  FP64: the accurate device library.  FP32:
// the hardware instructions v_exp_f32 (2^x), v_log_f32 (log2 x),
// v_rcp_f32 and v_sqrt_f32 (about 1 ulp each).
__device__ __forceinline__ double m_div(double a, double b) { return a / b; }
__device__ __forceinline__ float m_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double m_exp2(double x) { return exp2(x); }
__device__ __forceinline__ float m_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ double m_log2(double x) { return log2(x); }
__device__ __forceinline__ float m_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ double m_exp(double x) { return exp(x); }
__device__ __forceinline__ float m_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ double m_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float m_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double m_pow(double x, double y) { return pow(x, y); }
// x^e1 and x^e2 of one base in FP64 as exp(e log x) with the logarithm
// shared: within a few ulp of pow (the rounding of e log x, |e log x| of
// order 10, is what exp amplifies), at a third of two pow calls.  x <= 0,
// NaN and a zero exponent take pow's own special cases.
__device__ __forceinline__ void pow_pair(double x, double e1, double e2, double& r1, double& r2)
{
    if (x > 0.0 && e1 != 0.0 && e2 != 0.0) {
        const double L = log(x);
        r1 = exp(e1 * L);
        r2 = exp(e2 * L);
    } else {
        r1 = pow(x, e1);
        r2 = pow(x, e2);
    }
}
// x^y for x >= 0 as 2^(y log2 x); a negative x gives NaN, as on the GPU
// reference path.
__device__ __forceinline__ float m_pow(float x, float y) { return m_exp2(y * m_log2(x)); }

// ---------------------------------------------------------------- buffer I/O

// Raw buffer access: a 128-bit descriptor in SGPRs per array plus a 32-bit
// byte offset per lane, shared by every array of the same layout.  Offsets
// at or beyond `bytes` read 0 and drop stores.
using rsrc_t = __amdgpu_buffer_rsrc_t;

__device__ __forceinline__ rsrc_t make_rsrc(const void* p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ void buf_load(rsrc_t r, unsigned off, float& v)
{
    v = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void buf_load(rsrc_t r, unsigned off, double& v)
{
    v = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ void buf_store(rsrc_t r, unsigned off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}
__device__ __forceinline__ void buf_store(rsrc_t r, unsigned off, double v)
{
    using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, off, 0, 0);
}

// Plain global access at a 32-bit byte offset from an array's base: the
// base stays in scalar registers and one per-lane offset serves every array
// of the same layout (global_load/store with a scalar base address).
template <typename FT>
__device__ __forceinline__ FT gload(const FT* p, unsigned boff)
{
    return *reinterpret_cast<const FT*>(reinterpret_cast<const char*>(p) + boff);
}
template <typename FT>
__device__ __forceinline__ void gstore(FT* p, unsigned boff, FT v)
{
    *reinterpret_cast<FT*>(reinterpret_cast<char*>(p) + boff) = v;
}

// ---------------------------------------------------------------- parameters

// obme_npd_params converted to the arithmetic type.
template <typename FT>
struct NPDParams {
    FT kNO3, kNH4, psi, kPAR, mu0, gamma, alpha_x, Q10;
    FT mP, mZ, muZ, aP_solid, aZ_exc, pref, g, kG, aZ_ass, R, rho_c, R_chl;
    FT nitrif, a_rem, s_rem, b_rem, d_rem, small_solid, Rd, rem, small_frac;
    FT Rp, Rn;    // Oxygen(): respiration and nitrification O:N ratios (set from obme_oxygen)
    int light_form, use_Q10, mort_form, graz_form;
};

template <typename FT>
inline NPDParams<FT> convert_params(const obme_npd_params& p)
{
    NPDParams<FT> q;
    q.Rp = q.Rn = FT(0);
    q.kNO3 = (FT)p.nitrate_half_saturation;
    q.kNH4 = (FT)p.ammonia_half_saturation;
    q.psi = (FT)p.nitrate_ammonia_inhibition;
    q.kPAR = (FT)p.light_half_saturation;
    q.mu0 = (FT)p.phytoplankton_maximum_growth_rate;
    q.gamma = (FT)p.phytoplankton_exudation_fraction;
    q.alpha_x = (FT)p.ammonia_fraction_of_exudate;
    q.Q10 = (FT)p.temperature_coefficient;
    q.mP = (FT)p.phytoplankton_mortality_rate;
    q.mZ = (FT)p.zooplankton_mortality_rate;
    q.muZ = (FT)p.zooplankton_excretion_rate;
    q.aP_solid = (FT)p.phytoplankton_solid_waste_fraction;
    q.aZ_exc = (FT)p.excretion_inorganic_fraction;
    q.pref = (FT)p.preference_for_phytoplankton;
    q.g = (FT)p.maximum_grazing_rate;
    q.kG = (FT)p.grazing_half_saturation;
    q.aZ_ass = (FT)p.zooplankton_assimilation_fraction;
    q.R = (FT)p.redfield_ratio;
    q.rho_c = (FT)p.carbon_calcite_ratio;
    q.R_chl = (FT)p.phytoplankton_chlorophyll_ratio;
    q.nitrif = (FT)p.nitrification_rate;
    q.a_rem = (FT)p.remineralisation_inorganic_fraction;
    q.s_rem = (FT)p.small_remineralisation_rate;
    q.b_rem = (FT)p.large_remineralisation_rate;
    q.d_rem = (FT)p.dissolved_remineralisation_rate;
    q.small_solid = (FT)p.small_solid_waste_fraction;
    q.Rd = (FT)p.detritus_redfield_ratio;
    q.rem = (FT)p.remineralisation_rate;
    q.small_frac = (FT)p.small_particle_fraction;
    q.light_form = p.light_limitation;
    q.use_Q10 = p.use_temperature_coefficient;
    q.mort_form = p.phytoplankton_mortality_formulation;
    q.graz_form = p.grazing_concentration_formulation;
    return q;
}

// ---------------------------------------------------------------- tendencies

// Shared plankton terms (NutrientsPlanktonDetritus/plankton.jl).
template <typename FT>
struct Grazing {
    FT total, onP, onSmall;
};

template <typename FT>
__device__ __forceinline__ FT light_limitation(const NPDParams<FT>& q, FT PAR)
{   // plankton.jl:219-220
    if (q.light_form == OBME_LIGHT_MONDO) return m_div(PAR, q.kPAR + PAR);
    return m_div(PAR, m_sqrt(PAR * PAR + q.kPAR * q.kPAR));
}

template <typename FT>
__device__ __forceinline__ Grazing<FT> grazing(const NPDParams<FT>& q, FT P, FT Z, FT sPOM)
{   // total_grazing :118-135, grazing(Val{:P}) :352-369, grazing(Val{:sPOM}/Val{:D}) :371-389,
    // weighted_phytoplankton_preference :393-397, concentration_limit :89-90
    FT p = m_div(q.pref * P, q.pref * P + (FT(1) - q.pref) * sPOM + Tiny<FT>::eps0);
    FT food = p * P + (FT(1) - p) * sPOM;
    FT L = (q.graz_form == OBME_LINEAR) ? m_div(food, food + q.kG)
                                        : m_div(food * food, food * food + q.kG * q.kG);
    FT den = food + jeps(food);
    Grazing<FT> G;
    G.total = q.g * L * Z;
    G.onP = m_div(q.g * p * L * P, den) * Z;
    G.onSmall = m_div(q.g * (FT(1) - p) * L * sPOM, den) * Z;
    return G;
}

template <typename FT>
__device__ __forceinline__ FT mortality(const NPDParams<FT>& q, FT P)
{   // plankton.jl:86-87
    return q.mort_form == OBME_LINEAR ? q.mP * P : q.mP * (P * P);
}

// NPZD transitions (N, P, Z, D); T has zero tendency.  With dO2 the O2
// transition of the Oxygen component's Nutrient method (oxygen.jl:35-46):
// Rp μP - (Rp - Rn) respiration, where respiration is the plankton's
// inorganic nitrogen waste alone (the line that adds the detritus waste,
// oxygen.jl:42, begins with a unary + and is a separate, discarded
// expression in Julia).
template <typename FT>
__device__ __forceinline__ void npzd_tendencies(const NPDParams<FT>& q, FT N, FT P, FT Z, FT T, FT D, FT PAR,
                                                FT Lt, FT& dN, FT& dP, FT& dZ, FT& dD, FT* dO2 = nullptr)
{
    // phytoplankton_growth (plankton.jl:191-206), Nutrient limitation (:180-188)
    FT Ln = m_div(N, N + q.kNO3);
    FT Ll = light_limitation(q, PAR);
    FT muP = q.mu0 * Ll * Ln * Lt * P;
    FT sPOM = D * q.small_frac;  // detritus.jl:386-387
    Grazing<FT> G = grazing(q, P, Z, sPOM);
    FT nuP = mortality(q, P);
    dP = (FT(1) - q.gamma) * muP - G.onP - nuP;                     // plankton.jl:92-105
    dZ = q.aZ_ass * G.total - q.mZ * (Z * Z) - q.muZ * Z;          // :107-116
    FT pinw = q.aZ_exc * q.muZ * Z + (FT(1) - q.aP_solid) * nuP;  // :292-306
    FT dinw = D * q.rem;                                           // detritus.jl:389-390
    FT uptake = muP * (FT(1) - q.alpha_x * q.gamma);               // plankton.jl:271-278
    dN = pinw + dinw - uptake;                                     // nutrients.jl:86-90
    FT ponw = (FT(1) - q.alpha_x) * q.gamma * muP + (FT(1) - q.aZ_exc) * q.muZ * Z;  // :312-325
    FT solid = (FT(1) - q.aZ_ass) * G.total + q.aP_solid * nuP + q.mZ * (Z * Z);    // :330-346
    dD = ponw + solid - G.onSmall - q.rem * D;                     // detritus.jl:373-379
    if (dO2) *dO2 = q.Rp * muP - (q.Rp - q.Rn) * pinw;
}

// LOBSTER transitions; v/out in tracer order NO3, NH4, P, Z, sPOM, bPOM, DOM[, DIC, Alk][, O2].
// CARB: bit 1 CarbonateSystem, bit 2 Oxygen.
template <typename FT, int CARB>
__device__ __forceinline__ void lobster_tendencies(const NPDParams<FT>& q, const FT* v, FT PAR, FT* out)
{
    FT NO3 = v[0], NH4 = v[1], P = v[2], Z = v[3], sPOM = v[4], bPOM = v[5], DOM = v[6];
    // nitrogen_limitation (plankton.jl:150-166)
    FT nl = m_div(NO3 * m_exp(-q.psi * NH4), NO3 + q.kNO3);
    FT al = jmax(FT(0), m_div(NH4, q.kNH4 + NH4));
    FT Ln = (nl + al) / FT(2);
    FT Ll = light_limitation(q, PAR);
    FT muP = q.mu0 * Ll * Ln * FT(1) * P;
    Grazing<FT> G = grazing(q, P, Z, sPOM);
    FT nuP = mortality(q, P);
    FT nitrification = q.nitrif * NH4;                       // nutrients.jl:55
    FT share = nl + al + Tiny<FT>::eps0;
    FT up_NO3 = m_div(muP * nl, share);                     // plankton.jl:223-239
    FT waste = q.alpha_x * q.gamma * muP;
    FT up_NH4 = m_div(muP * al, share) - waste;             // :241-261
    FT pinw = q.aZ_exc * q.muZ * Z + (FT(1) - q.aP_solid) * nuP;
    FT dinw = (q.a_rem * (q.s_rem * sPOM + q.b_rem * bPOM) + q.d_rem * DOM);     // detritus.jl:252-263
    FT ponw = (FT(1) - q.alpha_x) * q.gamma * muP + (FT(1) - q.aZ_exc) * q.muZ * Z;
    FT donw = (FT(1) - q.a_rem) * (q.s_rem * sPOM + q.b_rem * bPOM);             // :265-274
    FT solid = (FT(1) - q.aZ_ass) * G.total + q.aP_solid * nuP + q.mZ * (Z * Z);
    FT dNO3 = nitrification - up_NO3;                        // nutrients.jl:57-60
    FT dNH4 = pinw + dinw - nitrification - up_NH4;          // :62-67
    out[0] = dNO3;
    out[1] = dNH4;
    out[2] = (FT(1) - q.gamma) * muP - G.onP - nuP;
    out[3] = q.aZ_ass * G.total - q.mZ * (Z * Z) - q.muZ * Z;
    out[4] = q.small_solid * solid - G.onSmall - q.s_rem * sPOM;       // detritus.jl:240-244
    out[5] = (FT(1) - q.small_solid) * solid - q.b_rem * bPOM;        // :246-249
    out[6] = ponw + donw - q.d_rem * DOM;                              // :234-238
    if constexpr ((CARB & 1) != 0) {
        FT pp = (FT(1) + q.rho_c * (FT(1) - q.gamma) - q.alpha_x * q.gamma) * muP * q.R;  // plankton.jl:280-290
        FT picw = q.R * pinw;                                                           // :308-309
        FT dicw = q.a_rem * (q.s_rem * (sPOM * q.Rd) + q.b_rem * (bPOM * q.Rd)) + q.d_rem * (DOM * q.Rd);
        FT cdis = (G.onP + nuP) * q.rho_c * q.R;                                        // :429-441
        FT cup = FT(2) * q.rho_c * muP * q.R;                                           // :443-450
        out[7] = -pp + picw + dicw + cdis;                                              // carbonate_system.jl:50-55
        out[8] = dNH4 * FT(0.9375) - dNO3 * FT(1.0625) - FT(2) * cup + FT(2) * cdis;    // :57-62
    }
    if constexpr ((CARB & 2) != 0) {
        // oxygen.jl:21-31: the NH4 tendency stands for "nitrate_production"
        out[(CARB & 1) ? 9 : 7] = q.Rp * muP - (q.Rp - q.Rn) * dNH4 - q.Rp * nitrification;
    }
}

// ---------------------------------------------------------------- FP32 fast tendencies

// Derived constants of the FP32-mixed path, formed once on the host in
// FP64 and rounded: the vector ALUs of gfx950 do all float arithmetic
// (there is no scalar float ALU), so uniform products are precomputed.
struct FastParams {
    float nPSI2;       // -psi * log2(e): exp(-psi NH4) = 2^(nPSI2 NH4)
    float kNO3, kNH4, kPAR, kPAR2;
    float mu0, mu0h;   // mu0, mu0/2 (LOBSTER's (L_NO3 + L_NH4)/2)
    float pref, prefc; // pref, 1 - pref
    float kG, kG2, g;
    float mP, mZ, muZ;
    float one_m_gamma, axg;
    float nitrif;
    float pinw_z;      // aZ_exc * muZ
    float one_m_aPs, aPs;
    float s_rem, b_rem, d_rem, a_rem, one_m_arem;
    float ponw_mu;     // (1 - alpha_x) gamma
    float ponw_z;      // (1 - aZ_exc) muZ
    float one_m_aZass, aZass;
    float ss, one_m_ss;
    float ppc;         // (1 + rho (1-gamma) - alpha_x gamma) R
    float R, Rd, rhoR2, rhoR4;  // R, Rd, 2 rho R, 4 rho R
    float rem, small_frac, uptake_N;  // Detritus rate, fraction, (1 - alpha_x gamma)
    float log2Q10_10;  // log2(Q10)/10: Q10^(T/10) = 2^(T log2Q10_10)
    float oRp, oRpn;   // Oxygen(): Rp, Rp - Rn (set from obme_oxygen)
    int light_form, use_Q10, mort_form, graz_form;
};

inline FastParams make_fast_params(const obme_npd_params& p)
{
    FastParams f;
    f.oRp = f.oRpn = 0.0f;
    const double LOG2E = 1.4426950408889634;
    f.nPSI2 = (float)(-p.nitrate_ammonia_inhibition * LOG2E);
    f.kNO3 = (float)p.nitrate_half_saturation;
    f.kNH4 = (float)p.ammonia_half_saturation;
    f.kPAR = (float)p.light_half_saturation;
    f.kPAR2 = (float)(p.light_half_saturation * p.light_half_saturation);
    f.mu0 = (float)p.phytoplankton_maximum_growth_rate;
    f.mu0h = (float)(p.phytoplankton_maximum_growth_rate / 2);
    f.pref = (float)p.preference_for_phytoplankton;
    f.prefc = (float)(1 - p.preference_for_phytoplankton);
    f.kG = (float)p.grazing_half_saturation;
    f.kG2 = (float)(p.grazing_half_saturation * p.grazing_half_saturation);
    f.g = (float)p.maximum_grazing_rate;
    f.mP = (float)p.phytoplankton_mortality_rate;
    f.mZ = (float)p.zooplankton_mortality_rate;
    f.muZ = (float)p.zooplankton_excretion_rate;
    f.one_m_gamma = (float)(1 - p.phytoplankton_exudation_fraction);
    f.axg = (float)(p.ammonia_fraction_of_exudate * p.phytoplankton_exudation_fraction);
    f.nitrif = (float)p.nitrification_rate;
    f.pinw_z = (float)(p.excretion_inorganic_fraction * p.zooplankton_excretion_rate);
    f.one_m_aPs = (float)(1 - p.phytoplankton_solid_waste_fraction);
    f.aPs = (float)p.phytoplankton_solid_waste_fraction;
    f.s_rem = (float)p.small_remineralisation_rate;
    f.b_rem = (float)p.large_remineralisation_rate;
    f.d_rem = (float)p.dissolved_remineralisation_rate;
    f.a_rem = (float)p.remineralisation_inorganic_fraction;
    f.one_m_arem = (float)(1 - p.remineralisation_inorganic_fraction);
    f.ponw_mu = (float)((1 - p.ammonia_fraction_of_exudate) * p.phytoplankton_exudation_fraction);
    f.ponw_z = (float)((1 - p.excretion_inorganic_fraction) * p.zooplankton_excretion_rate);
    f.one_m_aZass = (float)(1 - p.zooplankton_assimilation_fraction);
    f.aZass = (float)p.zooplankton_assimilation_fraction;
    f.ss = (float)p.small_solid_waste_fraction;
    f.one_m_ss = (float)(1 - p.small_solid_waste_fraction);
    f.ppc = (float)((1 + p.carbon_calcite_ratio * (1 - p.phytoplankton_exudation_fraction)
                     - p.ammonia_fraction_of_exudate * p.phytoplankton_exudation_fraction) * p.redfield_ratio);
    f.R = (float)p.redfield_ratio;
    f.Rd = (float)p.detritus_redfield_ratio;
    f.rhoR2 = (float)(2 * p.carbon_calcite_ratio * p.redfield_ratio);
    f.rhoR4 = (float)(4 * p.carbon_calcite_ratio * p.redfield_ratio);
    f.rem = (float)p.remineralisation_rate;
    f.small_frac = (float)p.small_particle_fraction;
    f.uptake_N = (float)(1 - p.ammonia_fraction_of_exudate * p.phytoplankton_exudation_fraction);
    f.log2Q10_10 = (float)(std::log2(p.temperature_coefficient > 0 ? p.temperature_coefficient : 1.0) / 10);
    f.light_form = p.light_limitation;
    f.use_Q10 = p.use_temperature_coefficient;
    f.mort_form = p.phytoplankton_mortality_formulation;
    f.graz_form = p.grazing_concentration_formulation;
    return f;
}

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
// max(0, x) that keeps a NaN x (Julia's max propagates NaN)
__device__ __forceinline__ float fmax0(float x) { return (x < 0.0f) ? 0.0f : x; }
constexpr float F_TINY = 1.17549435e-38f;  // smallest normal float

// Grazing of the PhytoZoo plankton (plankton.jl:118-135, :352-397) in FP32,
// reformulated over n = pref P^2 + (1-pref) sPOM^2 and d = pref P + (1-pref) sPOM:
// the preference-weighted food is n/d, so the total grazing is
// g Z n^2 / (n^2 + kG^2 d^2) (quadratic) or g Z n / (n + kG d) (linear), and
// the grazed fractions of the total are pref P^2 / n and (1-pref) sPOM^2 / n.
// With fr = total / n = g Z n / (n^2 + kG^2 d^2) (or g Z / (n + kG d)) the
// three terms are fr n, fr pref P^2 and fr (1-pref) sPOM^2: one reciprocal
// instead of the reference's three divisions; F_TINY keeps 0/0 away as the
// reference's eps(food) does.
struct FGrazing {
    float total, onP, onSmall;
};
// Formulation choices of the FP32 path: FM < 0 reads them from FastParams
// at run time; otherwise they are compiled in, FM = light (bit 0: 0 Mondo,
// 1 analytical) | phytoplankton mortality << 1 (0 linear, 1 quadratic) |
// grazing concentration << 2 (0 linear, 1 quadratic) | Q10 in use << 3.
template <int FM> __device__ __forceinline__ int fm_light(const FastParams& f) { if constexpr (FM < 0) return f.light_form; else return FM & 1; }
template <int FM> __device__ __forceinline__ int fm_mort(const FastParams& f) { if constexpr (FM < 0) return f.mort_form; else return (FM >> 1) & 1; }
template <int FM> __device__ __forceinline__ int fm_graz(const FastParams& f) { if constexpr (FM < 0) return f.graz_form; else return (FM >> 2) & 1; }
template <int FM> __device__ __forceinline__ int fm_q10(const FastParams& f) { if constexpr (FM < 0) return f.use_Q10; else return (FM >> 3) & 1; }
inline int fm_code(const FastParams& f)
{
    return (f.light_form & 1) | ((f.mort_form & 1) << 1) | ((f.graz_form & 1) << 2) | ((f.use_Q10 ? 1 : 0) << 3);
}

template <int FM = -1>
__device__ __forceinline__ FGrazing fast_grazing(const FastParams& f, float P, float Z, float sPOM)
{
    const float pP = f.pref * P;
    const float cS = f.prefc * sPOM;
    const float sS = cS * sPOM;
    const float n = fmaf(pP, P, sS);
    const float d = pP + cS;
    const float gZ = f.g * Z;
    float fr;
    if (fm_graz<FM>(f) == OBME_LINEAR) fr = gZ * frcp(fmaf(f.kG, d, n) + F_TINY);
    else {
        const float kd = f.kG * d;
        fr = (gZ * n) * frcp(fmaf(kd, kd, n * n) + F_TINY);
    }
    FGrazing G;
    G.total = fr * n;
    G.onP = fr * (pP * P);
    G.onSmall = fr * sS;
    return G;
}

template <int FM = -1>
__device__ __forceinline__ float fast_light(const FastParams& f, float PAR)
{
    if (fm_light<FM>(f) == OBME_LIGHT_MONDO) return PAR * frcp(f.kPAR + PAR);
    return PAR * __builtin_amdgcn_rsqf(fmaf(PAR, PAR, f.kPAR2));
}

// LOBSTER (+ CarbonateSystem) transitions in FP32; v/out in tracer order.
// The nitrate and ammonia uptakes are mu0/2 L_PAR P times their own
// limitation (the reference forms them as fractions of the total
// production, plankton.jl:223-261, which is the same product).
template <int CARB, int FM = -1>
__device__ __forceinline__ void lobster_fast(const FastParams& f, const float* v, float PAR, float* out)
{
    const float NO3 = v[0], NH4 = v[1], P = v[2], Z = v[3], sPOM = v[4], bPOM = v[5], DOM = v[6];
    const float a = NO3 + f.kNO3, b = f.kNH4 + NH4;
    float up_NO3, muA;
    if (fm_light<FM>(f) == OBME_LIGHT_MONDO) {
        // one reciprocal for both Monod terms and the light limitation
        // PAR / (kPAR + PAR): with c = kPAR + PAR > 0, 1/(a c) = b r and
        // 1/(b c) = a r for r = 1/(a b c), and max(0, x / c) = max(0, x) / c
        const float r = frcp((a * b) * (f.kPAR + PAR));
        const float nlc = NO3 * __builtin_amdgcn_exp2f(f.nPSI2 * NH4) * (b * r);
        const float alc = fmax0(NH4 * (a * r));
        const float muLPc = f.mu0h * PAR * P;
        up_NO3 = muLPc * nlc;
        muA = muLPc * alc;
    } else {
        // one reciprocal for both Monod terms
        const float rab = frcp(a * b);
        const float nl = NO3 * __builtin_amdgcn_exp2f(f.nPSI2 * NH4) * (b * rab);
        const float al = fmax0(NH4 * (a * rab));
        const float muLP = f.mu0h * fast_light<FM>(f, PAR) * P;
        up_NO3 = muLP * nl;
        muA = muLP * al;
    }
    const float muP = up_NO3 + muA;
    const FGrazing G = fast_grazing<FM>(f, P, Z, sPOM);
    const float nuP = (fm_mort<FM>(f) == OBME_LINEAR) ? f.mP * P : f.mP * P * P;
    const float nit = f.nitrif * NH4;
    const float up_NH4 = fmaf(-f.axg, muP, muA);
    const float pinw = fmaf(f.pinw_z, Z, f.one_m_aPs * nuP);
    const float sb = fmaf(f.s_rem, sPOM, f.b_rem * bPOM);
    const float dinw = fmaf(f.a_rem, sb, f.d_rem * DOM);
    const float ponw = fmaf(f.ponw_mu, muP, f.ponw_z * Z);
    const float Z2 = Z * Z;
    const float solid = fmaf(f.one_m_aZass, G.total, fmaf(f.aPs, nuP, f.mZ * Z2));
    const float dNO3 = nit - up_NO3;
    const float dNH4 = pinw + dinw - nit - up_NH4;
    out[0] = dNO3;
    out[1] = dNH4;
    out[2] = fmaf(f.one_m_gamma, muP, -G.onP) - nuP;
    out[3] = fmaf(f.aZass, G.total, -f.mZ * Z2) - f.muZ * Z;
    out[4] = fmaf(f.ss, solid, -G.onSmall) - f.s_rem * sPOM;
    out[5] = fmaf(f.one_m_ss, solid, -f.b_rem * bPOM);
    out[6] = fmaf(f.one_m_arem, sb, ponw) - f.d_rem * DOM;
    if constexpr ((CARB & 1) != 0) {
        // calcite dissolution (G_P + ν) ρ R (plankton.jl:429-441), uptake 2 ρ μP R (:443-450)
        const float cdis2 = (G.onP + nuP) * f.rhoR2;            // 2 cdis
        out[7] = fmaf(-f.ppc, muP, fmaf(f.R, pinw, fmaf(f.Rd, dinw, 0.5f * cdis2)));
        out[8] = fmaf(0.9375f, dNH4, -1.0625f * dNO3) - f.rhoR4 * muP + cdis2;
    }
    if constexpr ((CARB & 2) != 0) {
        // Rp (μP - nitrification) - (Rp - Rn) G_NH4 (oxygen.jl:21-31)
        out[(CARB & 1) ? 9 : 7] = fmaf(f.oRp, muP - nit, -f.oRpn * dNH4);
    }
}

// NPZD transitions in FP32 (N, P, Z, D); Lt = Q10^(T/10); dO2 as in npzd_tendencies.
template <int FM = -1>
__device__ __forceinline__ void npzd_fast(const FastParams& f, float N, float P, float Z, float D, float PAR,
                                          float Lt, float& dN, float& dP, float& dZ, float& dD,
                                          float* dO2 = nullptr)
{
    const float muP = f.mu0 * fast_light<FM>(f, PAR) * (N * frcp(N + f.kNO3)) * Lt * P;
    const float sPOM = D * f.small_frac;
    const FGrazing G = fast_grazing<FM>(f, P, Z, sPOM);
    const float nuP = (fm_mort<FM>(f) == OBME_LINEAR) ? f.mP * P : f.mP * P * P;
    const float Z2 = Z * Z;
    dP = fmaf(f.one_m_gamma, muP, -G.onP) - nuP;
    dZ = fmaf(f.aZass, G.total, -f.mZ * Z2) - f.muZ * Z;
    const float pinw = fmaf(f.pinw_z, Z, f.one_m_aPs * nuP);
    dN = fmaf(D, f.rem, pinw) - muP * f.uptake_N;
    const float ponw = fmaf(f.ponw_mu, muP, f.ponw_z * Z);
    const float solid = fmaf(f.one_m_aZass, G.total, fmaf(f.aPs, nuP, f.mZ * Z2));
    dD = ponw + solid - G.onSmall - f.rem * D;
    if (dO2) *dO2 = fmaf(f.oRp, muP, -f.oRpn * pinw);
}

// ScaleNegativeTracers groups for the kernels: members as bits of the
// tracer index and a scalefactor per tracer index (obme_negative_scaling
// converted on the host).
template <typename FT, int NTMAX>
struct NegScale {
    int n_groups;
    unsigned mask[OBME_MAX_SCALE_GROUPS];
    FT sf[OBME_MAX_SCALE_GROUPS][NTMAX];
    FT fill[OBME_MAX_SCALE_GROUPS];
};

// The scaling of one cell's NT tracer values x[] (src/Utils/negative_tracers.jl:
// 198-251): per group t = Σ s·c, p = Σ s·c over c > 0 (summed in tracer
// order: a reassociation of the reference's member order), t := fill where
// t < 0, then c := c·t/p where c > 0 or c is not finite, else 0.
template <int NT, typename FT, int NTMAX>
__device__ __forceinline__ void scale_negatives_cell(const NegScale<FT, NTMAX>& s, FT* x)
{
#pragma unroll 1
    for (int g = 0; g < s.n_groups; g++) {
        const unsigned mask = s.mask[g];
        FT t = FT(0), p = FT(0);
#pragma unroll
        for (int n = 0; n < NT; n++) {
            if (mask & (1u << n)) {
                const FT vs = x[n] * s.sf[g][n];
                t = t + vs;
                if (x[n] > FT(0)) p = p + vs;
            }
        }
        t = (t < FT(0)) ? s.fill[g] : t;
#pragma unroll
        for (int n = 0; n < NT; n++) {
            if (mask & (1u << n)) {
                const FT v = x[n];
                x[n] = (!isfinite(v) || v > FT(0)) ? v * t / p : FT(0);
            }
        }
    }
}

// ---------------------------------------------------------------- carbonate chemistry (FP64)

// teos10_polynomial_approximation (Models/seawater_density.jl:33-39) over
// SeawaterPolynomials' TEOS10 _ρ (Roquet et al. 2015 polynomial).
__device__ __forceinline__ double teos10_density(double T, double S, double Pbar)
{
    const double Sau = 40.0 * 35.16504 / 35.0;
    double t = T / 40.0;
    double s = sqrt((S + 32.0) / Sau);
    double z = -(10.0 * Pbar) / 1e4;
    double r0 = (((((-1.7243708991e-03 * z + 1.5616995503e-02) * z + 6.4326772569e-02) * z
                   + 2.2601900708e-01) * z - 5.2099962525e+00) * z + 4.6494977072e+01) * z;
    double n3 = 3.7969820455e-01 * t + -1.8507636718e-02 * s + -2.3342758797e-02;
    double n2 = (-1.2419983026e+00 * t + -2.1311365518e-01 * s + 2.0564311499e+00) * t
                + (2.5019633244e+00 * s + -4.9527603989e+00) * s + 2.0660924175e+00;
    double n1 = (((5.5927935970e-01 * t + -5.5077101279e-01 * s + -2.4649669534e+00) * t
                  + (-1.8795372996e+00 * s + 3.5063081279e+00) * s + 6.7080479603e+00) * t
                 + ((-6.5399043664e-01 * s + 5.0042598061e+00) * s + -4.4870114575e+00) * s + -1.3336301113e+01) * t
                + (((6.6051753097e+00 * s + -3.0938076334e+01) * s + 5.0774768218e+01) * s + -4.2549998214e+01) * s
                + 1.9681925209e+01;
    double n0 = (((((-1.9083568888e-01 * t + 4.8169980163e-01 * s + 5.4048723791e-01) * t
                    + (-5.3563304045e+00 * s + 1.1311538584e+01) * s + -8.3627885467e+00) * t
                   + ((-3.1742946532e+00 * s + 1.9717078466e+01) * s + -3.3449108469e+01) * s + 2.1661789529e+01) * t
                  + (((-5.4723692739e+00 * s + 2.9130021253e+01) * s + -6.0362551501e+01) * s + 6.1548258127e+01) * s
                  + -3.7074170417e+01) * t
                 + ((((-1.9193502195e+00 * s + 1.7681814114e+01) * s + -5.6888046321e+01) * s + 8.1770425108e+01) * s
                    + -6.5281885265e+01) * s + 2.6010145068e+01) * t
                + (((((-6.0579916612e+01 * s + 4.3227585684e+02) * s + -1.2849161071e+03) * s + 2.0375295546e+03) * s
                    + -1.7864682637e+03) * s + 8.6672408165e+02) * s + 8.0189615746e+02;
    return r0 + (((n3 * z + n2) * z + n1) * z + n0);
}

struct CarbonParams {
    double DIC, Alk, boron, sulfate, fluoride, silicate, phosphate;
    double K1, K2, KB, KW, KS, KF, KP1, KP2, KP3, KSi, K0, Tk;
};

// Equilibrium constants at the surface (P = nothing): CarbonChemistry/
// equilibrium_constants.jl K0 :65-80, K1 :103-126, K2 :149-172, KB :202-250,
// KW :279-313, IonicStrength :335-341, KS :366-419, KF :451-488,
// KP1-3 :511-651, KSi :674-713.  log(T), √S, √Is are computed once.
__device__ __forceinline__ double k0_constant(double T, double S, double lT)
{
    return exp(-60.2409 + (93.4517 * 100) / T + 23.3585 * (lT - log(100.0)) + 0.0 * (T * T)
               + (0.023517 + (-0.023656 / 100) * T + (0.0047036 / 10000.0) * (T * T)) * S);
}

__device__ __forceinline__ void carbon_constants(double Tk, double S, CarbonParams& p)
{
    double lT = log(Tk);
    double sS = sqrt(S);
    double Is = 19.924 * S / (1000.0 + -1.005 * S);
    double sIs = sqrt(Is);
    p.K0 = k0_constant(Tk, S, lT);
    p.K1 = pow(10.0, 61.2172 + -3633.86 / Tk + -9.67770 * lT + 0.011555 * S + -0.0001152 * (S * S));
    p.K2 = pow(10.0, -25.9290 + -471.78 / Tk + 0.01781 * S + -0.0001122 * (S * S) + 3.16967 * lT);
    p.KB = exp(148.0248 + (-8966.90 + -2890.53 * sS + -77.942 * S + 1.728 * pow(S, 1.5) + -0.0996 * (S * S)) / Tk
               + 137.1942 * sS + 1.62142 * S + (-24.4344 + -25.085 * sS + -0.2474 * S) * lT + 0.053105 * sS * Tk);
    p.KW = exp(148.9652 + -13847.26 / Tk + -23.6521 * lT + (-5.977 + 118.67 / Tk + 1.0495 * lT) * sS + -0.01615 * S);
    double l1S = log(1 + -0.001005 * S);
    p.KS = exp(141.328 + -4276.1 / Tk + -23.093 * lT + (324.57 + -13856.0 / Tk + -47.986 * lT) * sIs
               + (-771.54 + 35474.0 / Tk + 114.723 * lT) * Is + -2698.0 * pow(Is, 1.5) / Tk
               + 1776.0 * (Is * Is) / Tk + l1S);
    p.KF = exp(-9.68 + 874.0 / Tk + 0.111 * sS + log(1 + 0.0 * S) + log(1 + 0.0 * S / p.KS));
    p.KP1 = exp(115.525 + -4576.752 / Tk + -18.453 * lT + (0.69171 + -106.736 / Tk) * sS + (-0.01844 + -0.65643 / Tk) * S);
    p.KP2 = exp(172.0883 + -8814.715 / Tk + -27.927 * lT + (1.3566 + -160.340 / Tk) * sS + (-0.05778 + 0.37335 / Tk) * S);
    p.KP3 = exp(-18.141 + -3070.75 / Tk + 0.0 * lT + (2.81197 + 17.27039 / Tk) * sS + (-0.09984 + -44.99486 / Tk) * S);
    p.KSi = exp(117.385 + -8904.2 / Tk + -19.334 * lT + (3.5913 + -458.79 / Tk) * sIs + (-1.5998 + 188.74 / Tk) * Is
                + (0.07871 + -12.1652 / Tk) * (Is * Is) + l1S);
}

// Conversion of the functor's inputs (carbon_chemistry.jl:116-135): ρ at
// Pbar = 1 (:123), kelvin, mmol/m³ → mol/kg.
__device__ __forceinline__ void carbon_setup(double DIC, double Alk, double T, double S, CarbonParams& p)
{
    p.boron = 0.000232 / 10.811 * S / 1.80655;
    p.sulfate = 0.14 / 96.06 * S / 1.80655;
    p.fluoride = 0.000067 / 18.9984 * S / 1.80655;
    double rho = teos10_density(T, S, 1.0);
    p.Tk = T + 273.15;
    p.Alk = Alk * (1e-3 / rho);
    p.DIC = DIC * (1e-3 / rho);
    p.phosphate = 0.0;
    p.silicate = 0.0;
    carbon_constants(p.Tk, S, p);
}

// alkalinity_residual / ∂ₕ (CarbonChemistry/alkalinity_residual.jl:18-75)
// with phosphate = 0 and, unless SI, silicate = 0 (the defaults).  Those
// terms are then ±0 and adding them to the nonzero partial sums is exact, so
// they are left out; with SI the silicate term enters at its place in the sum.
template <bool SI = false>
__device__ __forceinline__ double alk_residual(double H, const CarbonParams& p)
{
    double cd = H * H + p.K1 * H + p.K1 * p.K2;
    double sd = 1 + p.sulfate / p.KS;
    double bicarbonate = p.K1 * H * p.DIC / cd;
    double carbonate = 2 * p.DIC * p.K1 * p.K2 / cd;
    double borate = p.boron / (1 + H / p.KB);
    double hydroxide = p.KW / H;
    double free_hydrogen = -H / sd;
    double hydrogen_sulfate = -p.sulfate / (1 + p.KS / H * sd);
    double hydrogen_fluoride = -p.fluoride / (1 + p.KF / H);
    double acc = bicarbonate + carbonate + borate + hydroxide;
    if constexpr (SI) acc = acc + p.silicate / (1 + H / p.KSi);
    return acc + free_hydrogen + hydrogen_sulfate + hydrogen_fluoride - p.Alk;
}

template <bool SI = false>
__device__ __forceinline__ double alk_residual_derivative(double H, const CarbonParams& p)
{
    double cd = H * H + p.K1 * H + p.K1 * p.K2;
    double sd = 1 + p.sulfate / p.KS;
    double dcd = 2 * H + p.K1;
    double d_bicarbonate = p.K1 * p.DIC * (p.K1 * p.K2 - H * H) / (cd * cd);
    double d_carbonate = -dcd * 2 * p.DIC * p.K1 * p.K2 / (cd * cd);
    double bd = 1 + H / p.KB;
    double d_borate = -p.boron / (bd * bd) / p.KB;
    double d_hydroxide = -p.KW / (H * H);
    double d_free_hydrogen = -1 / sd;
    double hsd = 1 + p.KS / H * sd;
    double d_hydrogen_sulfate = p.sulfate / (hsd * hsd) * (p.KS / H * 0.0 - p.KS / (H * H) * sd);
    double fd = 1 + p.KF / H;
    double d_hydrogen_fluoride = -p.fluoride / (fd * fd) * p.KF / (H * H);
    double acc = d_bicarbonate + d_carbonate + d_borate + d_hydroxide;
    if constexpr (SI) {
        const double sdn = 1 + H / p.KSi;
        acc = acc + -p.silicate / (sdn * sdn) / p.KSi;
    }
    return acc + d_free_hydrogen + d_hydrogen_sulfate + d_hydrogen_fluoride;
}

__device__ __forceinline__ double jsign(double x)
{
    return (x > 0) ? 1.0 : ((x < 0) ? -1.0 : x);
}

// DampedNewtonRaphsonSolver (Utils/solvers.jl:90-121) with bounded_λ for
// the lower bound 0 (:126-131).  An iteration that leaves x unchanged is an
// exact fixed point of the loop (same x, same f(x), same δ, same λ
// sequence), so the remaining iterations are skipped with the identical
// result.  evals counts residual evaluations.
template <bool SI = false>
__device__ __forceinline__ double solve_for_H(const CarbonParams& p, const obme_carbon_solver& s, double x0,
                                              int& evals)
{
    double x = x0;
    int N = 0;
    double fx = alk_residual<SI>(x, p);
    int ne = 1;
    while ((fabs(fx) > s.atol) && (N < s.max_iters)) {
        double delta = fx / alk_residual_derivative<SI>(x, p);
        double lb = (x - 0.0) / delta;
        lb = (lb < 0) ? 1.0 : lb;
        double lambda = jmin(1.0, lb);
        double fnew = alk_residual<SI>(x - lambda * delta, p);
        ne++;
        while ((fabs(fnew) >= fabs(fx) - s.armijo_constant * lambda * fx * jsign(fx)) && (lambda > s.min_damping)) {
            lambda *= s.damping;
            fnew = alk_residual<SI>(x - lambda * delta, p);
            ne++;
        }
        double xn = x - lambda * delta;
        N += 1;
        bool stuck = (xn == x);
        x = xn;
        fx = fnew;
        if (stuck) N = s.max_iters;
    }
    evals += ne;
    return x;
}

// FP32 first guess for [H+]: Newton on the residual in FP32 from pH 8, each
// step limited to half the distance to the bound H = 0, at most 16 steps.
__device__ __forceinline__ float guess_H_fp32(const CarbonParams& p)
{
    float K1 = (float)p.K1, K2 = (float)p.K2, KB = (float)p.KB, KW = (float)p.KW, KS = (float)p.KS,
          KF = (float)p.KF, DIC = (float)p.DIC, Alk = (float)p.Alk, boron = (float)p.boron,
          sulfate = (float)p.sulfate, fluoride = (float)p.fluoride;
    float sd = 1.0f + sulfate / KS;
    float rsd = 1.0f / sd;
    float K12 = K1 * K2;
    float x = 1e-8f;
#pragma unroll 1
    for (int it = 0; it < 16; it++) {
        float rx = 1.0f / x;
        float cd = x * x + K1 * x + K12;
        float rcd = 1.0f / cd;
        float rbd = 1.0f / (1.0f + x / KB);
        float rhsd = 1.0f / (1.0f + KS * rx * sd);
        float rfd = 1.0f / (1.0f + KF * rx);
        float f = (K1 * x + 2.0f * K12) * DIC * rcd + boron * rbd + KW * rx - x * rsd - sulfate * rhsd
                  - fluoride * rfd - Alk;
        float df = K1 * DIC * (K12 - x * x) * rcd * rcd - (2.0f * x + K1) * 2.0f * DIC * K12 * rcd * rcd
                   - boron * rbd * rbd / KB - KW * rx * rx - rsd - sulfate * rhsd * rhsd * KS * rx * rx * sd
                   - fluoride * rfd * rfd * KF * rx * rx;
        float d = f / df;
        float lam = (d > 0.0f) ? fminf(1.0f, 0.5f * x / d) : 1.0f;
        x = x - lam * d;
        if (fabsf(d) <= 2e-7f * x) break;
    }
    return x;
}

// The residual and its derivative together for the refinement of the
// mixed path: the same function as alk_residual / alk_residual_derivative,
// each fraction rewritten over one reciprocal of its denominator
// (bicarbonate + carbonate = DIC K1 (H + 2 K2) / cd, borate = boron KB /
// (KB + H), ...), so 8 FP64 divisions instead of 24.  The FP64 path keeps the
// reference's expressions.
template <bool SI>
__device__ __forceinline__ void alk_residual_and_derivative_fast(double H, const CarbonParams& p, double& f,
                                                                 double& df)
{
    const double cd = H * H + p.K1 * H + p.K1 * p.K2;
    const double rcd = 1.0 / cd;
    const double sd = 1.0 + p.sulfate / p.KS;
    const double rsd = 1.0 / sd;
    const double rH = 1.0 / H;
    const double rB = 1.0 / (p.KB + H);
    const double KSsd = p.KS * sd;
    const double rS = 1.0 / (H + KSsd);
    const double rF = 1.0 / (H + p.KF);
    const double DK1 = p.DIC * p.K1;
    const double carb = DK1 * (H + 2.0 * p.K2) * rcd;
    const double dcarb = DK1 * (cd - (H + 2.0 * p.K2) * (2.0 * H + p.K1)) * (rcd * rcd);
    const double bor = p.boron * p.KB * rB;
    const double dbor = -bor * rB;
    const double hyd = p.KW * rH;
    const double dhyd = -hyd * rH;
    const double hso4 = p.sulfate * H * rS;
    const double dhso4 = p.sulfate * KSsd * (rS * rS);
    const double hf = p.fluoride * H * rF;
    const double dhf = p.fluoride * p.KF * (rF * rF);
    double acc = carb + bor + hyd;
    double dacc = dcarb + dbor + dhyd;
    if constexpr (SI) {
        const double rSi = 1.0 / (p.KSi + H);
        const double sil = p.silicate * p.KSi * rSi;
        acc = acc + sil;
        dacc = dacc - sil * rSi;
    }
    f = acc - H * rsd - hso4 - hf - p.Alk;
    df = dacc - rsd - dhso4 - dhf;
}

// Refinement of a close first guess by plain FP64 Newton on the same
// residual; returns false when it has not converged after 6 steps.  It
// stops after the step that is at most tol relative to the iterate: Newton
// converges quadratically here, so the error left after a step of relative
// size tol is of order tol^2 (tol = 1e-8 leaves ~1e-16, the FP64 rounding
// of H; the default 1e-14 takes one confirming step more).
// The stop of the FP32-mixed paths, whose [H+] feeds FP32 arithmetic: after
// a step below 1e-4 relative, H is within ~1e-8 of the root (measured on
// the CPU oracle from warm starts off by up to 1e-2: at most 9e-9).
constexpr double FP32_NEWTON_TOL = 1e-4;
// The stop of the FP64 warm-started refinements: a step below 1e-14
// relative leaves H at the root to rounding.
constexpr double FP64_NEWTON_TOL = 1e-14;
template <bool SI = false>
__device__ __forceinline__ bool refine_H(const CarbonParams& p, double& x, int& evals, double tol = 1e-14)
{
#pragma unroll 1
    for (int it = 0; it < 6; it++) {
        double f, df;
        alk_residual_and_derivative_fast<SI>(x, p, f, df);
        evals++;
        double d = f / df;
        double xn = x - d;
        if (!(xn > 0.0)) return false;
        x = xn;
        if (fabs(d) <= tol * x) return true;
    }
    return false;
}

// pCO2 from H (carbon_chemistry.jl:160-162 and selected_output(::Val{:pCO₂}) :170-193).
__device__ __forceinline__ double pco2_from_H(const CarbonParams& p, double H)
{
    double CO2 = p.DIC * (H * H) / (H * H + p.K1 * H + p.K1 * p.K2);
    double fCO2 = (CO2 / p.K0) * 1e6;
    double P = 101325.0;
    double Tk = p.Tk;
    double B = (-1636.75 + 12.0408 * Tk + (-3.27957 * 1e-2) * (Tk * Tk) + (3.16528 * 1e-5) * (Tk * Tk * Tk)) * 1e-6;
    double dl = (57.7 + -0.118 * Tk) * 1e-6;
    fCO2 *= 0.09807;
    double phi = 1.0;
    double x = fCO2 / (phi * P);
    for (int n = 0; n < 3; n++) {
        double om = 1.0 - x;
        phi = exp((B + 2.0 * (om * om) * dl) * P / (8.31446261815324 * Tk));
        x = fCO2 / (phi * P);
    }
    double pCO2 = fCO2 / phi;
    return pCO2 / 0.09807;
}

// Julia's literal x^4 (compensated power by squaring).
__device__ __forceinline__ double pow4_compensated(double x)
{
    double hi2 = x * x;
    double lo2 = fma(x, x, -hi2);
    double err = hi2 * 2.0 * lo2;
    double hi4 = hi2 * hi2;
    double lo4 = fma(hi2, hi2, -hi4);
    double lo = lo4 + err;
    return (isfinite(hi4) && isfinite(lo)) ? fma(hi4, 1.0, lo) : hi4;
}

// Gas transfer velocity (gas_transfer_velocity.jl:33-34, :61; schmidt
// number gas_solubility.jl:211-212; solubility gas_solubility.jl:204).
__device__ __forceinline__ double transfer_velocity(const obme_co2_exchange& g, double u10, double T, double S)
{
    double k660 = g.transfer_velocity[0] + g.transfer_velocity[1] * u10 + g.transfer_velocity[2] * (u10 * u10);
    double sc = g.schmidt[0] + g.schmidt[1] * T + g.schmidt[2] * (T * T) + g.schmidt[3] * (T * T * T)
                + g.schmidt[4] * pow4_compensated(T);
    double Tk = T + 273.15;
    double K0 = k0_constant(Tk, S, log(Tk));
    double sol = K0 * teos10_density(T, S, 0.0) / 1000.0;
    return k660 / sqrt(sc / 660.0) * sol;
}

// Everything of the CO2 flux that depends on T and S only: the carbonate
// parameters without DIC/Alk, the mmol/m³ → mol/kg factor 1e-3/ρ and the
// transfer velocity k (a function of u10, T, S).
struct SurfaceCarbon {
    CarbonParams p;   // DIC and Alk unset
    double conv;      // 1e-3 / ρ(T, S, Pbar = 1)
    double k;         // transfer velocity
};

__device__ __forceinline__ void surface_carbon_setup(const obme_co2_exchange& g, double T, double S, double u10,
                                                     SurfaceCarbon& sc)
{
    CarbonParams& p = sc.p;
    p.boron = 0.000232 / 10.811 * S / 1.80655;
    p.sulfate = 0.14 / 96.06 * S / 1.80655;
    p.fluoride = 0.000067 / 18.9984 * S / 1.80655;
    double rho = teos10_density(T, S, 1.0);
    p.Tk = T + 273.15;
    sc.conv = 1e-3 / rho;
    p.phosphate = 0.0;
    p.silicate = 0.0;
    carbon_constants(p.Tk, S, p);
    sc.k = transfer_velocity(g, u10, T, S);
}

// Solve for H and return the flux F = k (pCO2_w - pCO2_air)
// (GasExchange/gas_exchange.jl:26-38).  mode 0: the reference damped Newton
// from 10^-initial_pH_guess; mode 1: FP32 first guess refined by FP64
// Newton; with FALLBACK, a refinement that has not converged is redone
// with the mode-0 solve (the fused step kernels omit the fallback to keep
// their register budget; a non-converged refinement keeps its last iterate).
template <bool FALLBACK = true>
__device__ __forceinline__ double surface_flux(const SurfaceCarbon& sc, const obme_carbon_solver& s, double DIC,
                                               double Alk, double pCO2_air, int mode, double& pH, int& evals)
{
    CarbonParams p = sc.p;
    p.Alk = Alk * sc.conv;
    p.DIC = DIC * sc.conv;
    double H = 0.0;
    if (mode == 1) {
        H = (double)guess_H_fp32(p);
        bool done = refine_H(p, H, evals);
        if (FALLBACK && !done) H = solve_for_H(p, s, pow(10.0, -s.initial_pH_guess), evals);
    } else {
        H = solve_for_H(p, s, pow(10.0, -s.initial_pH_guess), evals);
    }
    pH = -log10(H);
    double pCO2 = pco2_from_H(p, H);
    return sc.k * (pCO2 - pCO2_air);
}

__device__ __forceinline__ double co2_flux(const obme_co2_exchange& g, const obme_carbon_solver& s, double DIC,
                                           double Alk, double T, double S, double u10, double pCO2_air, int mode,
                                           double& pH, int& evals)
{
    SurfaceCarbon sc;
    surface_carbon_setup(g, T, S, u10, sc);
    return surface_flux(sc, s, DIC, Alk, pCO2_air, mode, pH, evals);
}

// Flux with the mixed path's warm start: H from the previous stage (or an
// FP32 guess) refined by FP64 Newton (stopping after a step below tol
// relative, see refine_H), with the reference solve as fallback.
__device__ __forceinline__ double surface_flux_warm(const SurfaceCarbon& sc, const obme_carbon_solver& s,
                                                    double DIC, double Alk, double pCO2_air, int mode,
                                                    double& H, double& pH, int& evals, double tol = 1e-14)
{
    if (mode == 0) return surface_flux<true>(sc, s, DIC, Alk, pCO2_air, 0, pH, evals);
    CarbonParams p = sc.p;
    p.Alk = Alk * sc.conv;
    p.DIC = DIC * sc.conv;
    double x = (H > 0.0) ? H : (double)guess_H_fp32(p);
    bool done = refine_H(p, x, evals, tol);
    if (!done) {
        x = (double)guess_H_fp32(p);
        done = refine_H(p, x, evals, tol);
        if (!done) x = solve_for_H(p, s, pow(10.0, -s.initial_pH_guess), evals);
    }
    H = x;
    pH = -log10(x);
    return sc.k * (pco2_from_H(p, x) - pCO2_air);
}

// OxygenGasExchangeBoundaryCondition (GasExchange/GasExchange.jl:127-146):
// F = k (O2 - C_air α), k = k660(u10) / sqrt(Sc(T) / 660) · 1
// (gas_transfer_velocity.jl:33-34), Sc the quartic of schmidt_number.jl:15-16
// (generic_parameterisations.jl:31-32), α = β / Tk with
// β = exp(A1 + A2/(Tk/100) + A3 log(Tk/100) + S (B1 + B2 Tk/100 + B2 (Tk/100)²))
// (gas_solubility.jl:37-48: B2 in both salinity terms, as there).
// The T- and S-dependent part: the transfer velocity k and the air-side
// concentration C_air α.
__device__ __forceinline__ void o2_exchange_coefficients(const obme_o2_exchange& g, double T, double S, double& k,
                                                         double& air)
{
    const double* c = g.schmidt;
    const double Sc = c[0] + c[1] * T + c[2] * (T * T) + c[3] * (T * T * T) + c[4] * pow4_compensated(T);
    const double* v = g.transfer_velocity;
    const double u = g.wind_speed;
    const double k660 = v[0] + v[1] * u + v[2] * (u * u);
    k = k660 / sqrt(Sc / 660.0) * 1.0;
    const double Tk = T + 273.15;
    const double Tk100 = Tk / 100.0;
    const double* a = g.solubility;
    const double beta = exp(a[0] + a[1] / Tk100 + a[2] * log(Tk100) + S * (a[3] + a[4] * Tk100 + a[4] * (Tk100 * Tk100)));
    air = g.air_concentration * (beta / Tk);
}

__device__ __forceinline__ double o2_flux(const obme_o2_exchange& g, double O2, double T, double S)
{
    double k, air;
    o2_exchange_coefficients(g, T, S, k, air);
    return k * (O2 - air);
}

// SurfaceCarbon (without DIC and Alk) and the O2 coefficients of a column
// as 21 doubles, for a stepper with T and S fixed.
constexpr int SURF_PACK = 21;
__device__ __forceinline__ void surface_pack(const SurfaceCarbon& sc, double kO2, double airO2, double* o)
{
    const CarbonParams& p = sc.p;
    const double f[SURF_PACK] = {p.boron, p.sulfate, p.fluoride, p.silicate, p.phosphate, p.K1, p.K2,
                                 p.KB, p.KW, p.KS, p.KF, p.KP1, p.KP2, p.KP3, p.KSi, p.K0, p.Tk,
                                 sc.conv, sc.k, kO2, airO2};
#pragma unroll
    for (int q = 0; q < SURF_PACK; q++) o[q] = f[q];
}
__device__ __forceinline__ void surface_unpack(const double* o, SurfaceCarbon& sc, double& kO2, double& airO2)
{
    CarbonParams& p = sc.p;
    p.DIC = 0.0;
    p.Alk = 0.0;
    p.boron = o[0];
    p.sulfate = o[1];
    p.fluoride = o[2];
    p.silicate = o[3];
    p.phosphate = o[4];
    p.K1 = o[5];
    p.K2 = o[6];
    p.KB = o[7];
    p.KW = o[8];
    p.KS = o[9];
    p.KF = o[10];
    p.KP1 = o[11];
    p.KP2 = o[12];
    p.KP3 = o[13];
    p.KSi = o[14];
    p.K0 = o[15];
    p.Tk = o[16];
    sc.conv = o[17];
    sc.k = o[18];
    kO2 = o[19];
    airO2 = o[20];
}

// ---- wave-level helpers (the column sweeps)

// Cross-lane data movement with DPP (GFX9 family): row_shr:d shifts within
// rows of 16 lanes, row_bcast:15/31 carries row ends into later rows,
// wave_shr:1 shifts the whole wave by one lane.  bound_ctrl gives 0 for a
// source lane outside the row (or wave), so a shifted-in 0 is the identity
// of the sums below.
template <int CTRL, int ROWMASK, bool BC>
__device__ __forceinline__ float dpp(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, ROWMASK,
                                                                 0xf, BC));
}
template <int CTRL, int ROWMASK, bool BC>
__device__ __forceinline__ double dpp(double x)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffull), CTRL, ROWMASK, 0xf, BC);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROWMASK, 0xf, BC);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// value of lane r-1 (lane 0 of the wave: 0)
template <typename T>
__device__ __forceinline__ T lane_shr1(T x)
{
    return dpp<0x138, 0xf, true>(x);
}

// Inclusive prefix sum over segments of W lanes (W = 16, 32 or 64).
template <int W, typename T>
__device__ __forceinline__ T seg_prefix(T x)
{
    x = x + dpp<0x111, 0xf, true>(x);
    x = x + dpp<0x112, 0xf, true>(x);
    x = x + dpp<0x114, 0xf, true>(x);
    x = x + dpp<0x118, 0xf, true>(x);
    if constexpr (W >= 32) x = x + dpp<0x142, 0xa, false>(x);
    if constexpr (W >= 64) x = x + dpp<0x143, 0xc, false>(x);
    return x;
}

// Tile of workgroup b of an nb-workgroup launch.  The dispatcher hands
// workgroup b to XCD b % 8 (round robin over the 8 XCDs of MI355X, each
// with its own L2).  Each XCD is given a contiguous run of tiles instead,
// so tiles that are neighbours along i, which share the 64-byte segments
// at their edges (a tile row is TXB values from an arbitrary offset), are
// read into and written back from one L2, by workgroups that run at about
// the same time.  A bijection of [0, nb).
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb)
{
    constexpr unsigned X = 8;
    const unsigned x = b % X, s = b / X, q = nb / X, rem = nb % X;
    return x * q + (x < rem ? x : rem) + s;
}

}  // namespace obme
