/*
 * oracle_carbon.c — FP64 restatement of OceanBioME's CarbonChemistry and CO2
 * GasExchange path.  TEST INFRASTRUCTURE ONLY (the oracle; see oracle.h).
 *
 * Third-party arithmetic restated here:
 *  - SeawaterPolynomials.jl 0.3 (Project.toml:28), TEOS10._ρ(τ, s, ζ): the
 *    Roquet et al. (2015, Ocean Modelling 90, 29-43) polynomial for in-situ
 *    density in (Θ, S_A, Z): ρ = r₀(ζ) + r(τ, s, ζ) with τ = Θ/40,
 *    s = √((S_A + 32)/S_au), S_au = 40·35.16504/35, ζ = −Z/10⁴.  The
 *    published coefficients are typed below.  Pinned through the reference's
 *    doctest values (carbon_chemistry.jl:55-62), which depend on ρ(10 °C, 35,
 *    Pbar = 1).
 *  - Julia Base scalar semantics: NaN-propagating min/max, eps(x), and the
 *    compensated integer power used by a literal x^4.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

/* ---------------- Julia scalar semantics ---------------- */

/* Julia Base min/max for IEEE floats: NaN propagates; ±0 ordered. */
double orc_julia_min(double x, double y)
{
    double diff = x - y;
    double argmin = signbit(diff) ? x : y;
    return (isnan(x) || isnan(y)) ? diff : argmin;
}

double orc_julia_max(double x, double y)
{
    double diff = x - y;
    double argmax = signbit(diff) ? y : x;
    return (isnan(x) || isnan(y)) ? diff : argmax;
}

/* Julia eps(x): spacing of floats at |x| (eps(0.0) = 2^-1074). */
double orc_julia_eps(double x)
{
    if (!isfinite(x)) return NAN;
    uint64_t b;
    memcpy(&b, &x, 8);
    int e = (int)((b >> 52) & 0x7ff);
    uint64_t r;
    if (e > 52) r = (uint64_t)(e - 52) << 52;
    else r = (uint64_t)1 << ((e > 1 ? e : 1) - 1);
    double out;
    memcpy(&out, &r, 8);
    return out;
}

/* x^n (n >= 0) as Julia evaluates Float64^Integer: power by squaring with
 * each product carried as a double-double (hi, lo) pair, result hi*y + lo.
 * A literal x^2 or x^3 is plain multiplication in Julia and is written as
 * such at the call sites. */
double orc_pow_int(double x, int n)
{
    if (n == 0) return 1.0;
    if (n == 3) return x * x * x;
    double y = 1.0, xnlo = 0.0, ynlo = 0.0;
    while (n > 1) {
        if (n & 1) {
            double err = fma(y, xnlo, x * ynlo);
            double hi = x * y;
            double lo = fma(x, y, -hi);
            y = hi;
            ynlo = lo + err;
        }
        double err = x * 2 * xnlo;
        double hi = x * x;
        double lo = fma(x, x, -hi);
        x = hi;
        xnlo = lo + err;
        n >>= 1;
    }
    double err = fma(y, xnlo, x * ynlo);
    if (isfinite(x) && isfinite(err)) return fma(x, y, err);
    return x * y;
}

double orc_pow_int4(double x) { return orc_pow_int(x, 4); }

/* ---------------- TEOS-10 polynomial density ---------------- */

/* Roquet et al. (2015) coefficients for ρ(τ, s, ζ) [kg/m^3]. */
static const double EOS000 = 8.0189615746e+02, EOS100 = 8.6672408165e+02,
    EOS200 = -1.7864682637e+03, EOS300 = 2.0375295546e+03, EOS400 = -1.2849161071e+03,
    EOS500 = 4.3227585684e+02, EOS600 = -6.0579916612e+01, EOS010 = 2.6010145068e+01,
    EOS110 = -6.5281885265e+01, EOS210 = 8.1770425108e+01, EOS310 = -5.6888046321e+01,
    EOS410 = 1.7681814114e+01, EOS510 = -1.9193502195e+00, EOS020 = -3.7074170417e+01,
    EOS120 = 6.1548258127e+01, EOS220 = -6.0362551501e+01, EOS320 = 2.9130021253e+01,
    EOS420 = -5.4723692739e+00, EOS030 = 2.1661789529e+01, EOS130 = -3.3449108469e+01,
    EOS230 = 1.9717078466e+01, EOS330 = -3.1742946532e+00, EOS040 = -8.3627885467e+00,
    EOS140 = 1.1311538584e+01, EOS240 = -5.3563304045e+00, EOS050 = 5.4048723791e-01,
    EOS150 = 4.8169980163e-01, EOS060 = -1.9083568888e-01, EOS001 = 1.9681925209e+01,
    EOS101 = -4.2549998214e+01, EOS201 = 5.0774768218e+01, EOS301 = -3.0938076334e+01,
    EOS401 = 6.6051753097e+00, EOS011 = -1.3336301113e+01, EOS111 = -4.4870114575e+00,
    EOS211 = 5.0042598061e+00, EOS311 = -6.5399043664e-01, EOS021 = 6.7080479603e+00,
    EOS121 = 3.5063081279e+00, EOS221 = -1.8795372996e+00, EOS031 = -2.4649669534e+00,
    EOS131 = -5.5077101279e-01, EOS041 = 5.5927935970e-01, EOS002 = 2.0660924175e+00,
    EOS102 = -4.9527603989e+00, EOS202 = 2.5019633244e+00, EOS012 = 2.0564311499e+00,
    EOS112 = -2.1311365518e-01, EOS022 = -1.2419983026e+00, EOS003 = -2.3342758797e-02,
    EOS103 = -1.8507636718e-02, EOS013 = 3.7969820455e-01;
static const double R00 = 4.6494977072e+01, R01 = -5.2099962525e+00, R02 = 2.2601900708e-01,
    R03 = 6.4326772569e-02, R04 = 1.5616995503e-02, R05 = -1.7243708991e-03;

static double teos10_rho(double tau, double s, double zeta)
{
    double r0 = (((((R05 * zeta + R04) * zeta + R03) * zeta + R02) * zeta + R01) * zeta + R00) * zeta;
    double t = tau;
    double n3 = EOS013 * t + EOS103 * s + EOS003;
    double n2 = (EOS022 * t + EOS112 * s + EOS012) * t + (EOS202 * s + EOS102) * s + EOS002;
    double n1 = (((EOS041 * t + EOS131 * s + EOS031) * t + (EOS221 * s + EOS121) * s + EOS021) * t
                 + ((EOS311 * s + EOS211) * s + EOS111) * s + EOS011) * t
                + (((EOS401 * s + EOS301) * s + EOS201) * s + EOS101) * s + EOS001;
    double n0 = (((((EOS060 * t + EOS150 * s + EOS050) * t
                    + (EOS240 * s + EOS140) * s + EOS040) * t
                   + ((EOS330 * s + EOS230) * s + EOS130) * s + EOS030) * t
                  + (((EOS420 * s + EOS320) * s + EOS220) * s + EOS120) * s + EOS020) * t
                 + ((((EOS510 * s + EOS410) * s + EOS310) * s + EOS210) * s + EOS110) * s + EOS010) * t
                + (((((EOS600 * s + EOS500) * s + EOS400) * s + EOS300) * s + EOS200) * s + EOS100) * s + EOS000;
    double r = ((n3 * zeta + n2) * zeta + n1) * zeta + n0;
    return r0 + r;
}

/* teos10_polynomial_approximation (src/Models/seawater_density.jl:33-39):
 * Z = 10·Pbar (a positive "height"), S_A = S_p. */
double orc_teos10_density(double T, double Sp, double Pbar)
{
    double Z = 10.0 * Pbar;
    double Sa = Sp;
    const double Sau = 40.0 * 35.16504 / 35.0;
    double tau = T / 40.0;
    double s = sqrt((Sa + 32.0) / Sau);
    double zeta = -Z / 1e4;
    return teos10_rho(tau, s, zeta);
}

/* ---------------- equilibrium constants ---------------- */

/* PressureCorrection (CarbonChemistry/equilibrium_constants.jl:11-40).
 * which: 1 K1, 2 K2, 3 KB, 4 KW, 5 KS, 6 KF, 7 KP1, 8 KP2, 9 KP3,
 *        10 KSP calcite, 11 KSP aragonite. */
static const double PC[12][5] = {
    {0, 0, 0, 0, 0},
    {-25.50, 0.1271, 0.0, -0.00308, 0.0000877},      /* K1  :118 */
    {-15.82, -0.0219, 0.0, 0.00113, -0.0001475},     /* K2  :164 */
    {-29.48, 0.1622, -0.0026080, -0.00284, 0.0},     /* KB  :231 */
    {-20.02, 0.1119, -0.001409, -0.00513, 0.0000794},/* KW  :298 */
    {-18.03, 0.0466, 0.000316, -0.00453, 0.00009},   /* KS  :399 */
    {-9.78, -0.0090, -0.000942, -0.00391, 0.000054}, /* KF  :474 */
    {-14.51, 0.1211, -0.000321, -0.00267, 0.0000427},/* KP1 :567 */
    {-23.12, 0.1758, -0.002647, -0.00515, 0.00009},  /* KP2 :605 */
    {-26.57, 0.2020, -0.0030420, -0.00408, 0.0000714},/* KP3 :643 */
    {-48.76, 0.5304, -0.0, -0.01176, 0.0003692},     /* KSP calcite :800 */
    {-45.96, 0.5304, -0.0, -0.01176, 0.0003692},     /* KSP aragonite :840 */
};

double orc_pressure_correction(int which, double Tk, double P)
{
    const double* c = PC[which];
    const double R = 83.14472;
    double Tc = Tk - 273.15;
    double dV = c[0] + c[1] * Tc + c[2] * (Tc * Tc);
    double dk = c[3] + c[4] * Tc;
    double RT = R * Tk;
    return exp((-dV + 0.5 * dk * P) * P / RT);
}

/* K0 (equilibrium_constants.jl:65-80) — T in kelvin. */
static double k0(double T, double S)
{
    const double constant = -60.2409, inverse_T = 93.4517 * 100, log_T = 23.3585, T2 = 0.0,
                 Sc = 0.023517, ST = -0.023656 / 100, ST2 = 0.0047036 / 10000.0;
    return exp(constant + inverse_T / T + log_T * (log(T) - log(100.0)) + T2 * (T * T)
               + (Sc + ST * T + ST2 * (T * T)) * S);
}

/* K1 (:103-126) */
static double k1(double T, double S)
{
    return pow(10.0, 61.2172 + -3633.86 / T + -9.67770 * log(T) + 0.011555 * S + -0.0001152 * (S * S));
}

/* K2 (:149-172) */
static double k2(double T, double S)
{
    return pow(10.0, -25.9290 + -471.78 / T + 0.01781 * S + -0.0001122 * (S * S) + 3.16967 * log(T));
}

/* KB (:202-250) */
static double kb(double T, double S)
{
    double sS = sqrt(S);
    return exp(148.0248
               + (-8966.90 + -2890.53 * sS + -77.942 * S + 1.728 * pow(S, 1.5) + -0.0996 * (S * S)) / T
               + 137.1942 * sS
               + 1.62142 * S
               + (-24.4344 + -25.085 * sS + -0.2474 * S) * log(T)
               + 0.053105 * sS * T);
}

/* KW (:279-313) */
static double kw(double T, double S)
{
    return exp(148.9652
               + -13847.26 / T
               + -23.6521 * log(T)
               + (-5.977 + 118.67 / T + 1.0495 * log(T)) * sqrt(S)
               + -0.01615 * S);
}

/* IonicStrength (:335-341) */
static double ionic_strength(double S)
{
    return 19.924 * S / (1000.0 + -1.005 * S);
}

/* KS (:366-419) */
static double ks(double T, double S, double Is)
{
    return exp(141.328
               + -4276.1 / T
               + -23.093 * log(T)
               + (324.57 + -13856.0 / T + -47.986 * log(T)) * sqrt(Is)
               + (-771.54 + 35474.0 / T + 114.723 * log(T)) * Is
               + -2698.0 * pow(Is, 1.5) / T
               + 1776.0 * (Is * Is) / T
               + log(1 + -0.001005 * S));
}

/* KF (:451-488) */
static double kf(double T, double S, double Is, double KS)
{
    (void)Is;
    return exp(-9.68
               + 874.0 / T
               + 0.111 * sqrt(S)
               + log(1 + 0.0 * S)
               + log(1 + 0.0 * S / KS));
}

/* KP (:511-529) with the KP1/KP2/KP3 coefficients (:558-651) */
static double kp(const double* c, double T, double S)
{
    return exp(c[0]
               + c[1] / T
               + c[2] * log(T)
               + (c[3] + c[4] / T) * sqrt(S)
               + (c[5] + c[6] / T) * S);
}
static const double KP1C[7] = {115.525, -4576.752, -18.453, 0.69171, -106.736, -0.01844, -0.65643};
static const double KP2C[7] = {172.0883, -8814.715, -27.927, 1.3566, -160.340, -0.05778, 0.37335};
static const double KP3C[7] = {-18.141, -3070.75, 0.0, 2.81197, 17.27039, -0.09984, -44.99486};

/* KSi (:674-713) */
static double ksi(double T, double S, double Is)
{
    return exp(117.385
               + -8904.2 / T
               + -19.334 * log(T)
               + (3.5913 + -458.79 / T) * sqrt(Is)
               + (-1.5998 + 188.74 / T) * Is
               + (0.07871 + -12.1652 / T) * (Is * Is)
               + log(1 + -0.001005 * S));
}

/* KSP (:754-764) */
static double ksp(const double* c, int which, double T, double S, int has_P, double P)
{
    double pc = has_P ? orc_pressure_correction(which, T, P) : 1.0;
    double lnK_therm = c[0] + c[1] * T + c[2] / T + c[3] * log10(T);
    double lnK_sea = ((c[4] + c[5] * T + c[6] / T) * sqrt(S) + c[7] * S + c[8] * pow(S, 1.5));
    return pc * pow(10.0, lnK_therm + lnK_sea);
}
static const double KSPC[9] = {-171.9065, -0.077993, 2839.319, 71.595, -0.77712, 0.0028426,
                               178.34, -0.07711, 0.0041249};  /* KSP_calcite :789-810 */
static const double KSPA[9] = {-171.945, -0.077993, 2903.293, 71.595, -0.068393, 0.0017276,
                               88.135, -0.10018, 0.0059415};  /* KSP_aragonite :829-850 */

double orc_ksp_calcite(double Tk, double S, int has_P, double P) { return ksp(KSPC, 10, Tk, S, has_P, P); }
double orc_ksp_aragonite(double Tk, double S, int has_P, double P) { return ksp(KSPA, 11, Tk, S, has_P, P); }

void orc_equilibrium_constants(double Tk, double S, int has_P, double P, orc_constants* o)
{
    double Is = ionic_strength(S);
    o->Is = Is;
    o->K0 = k0(Tk, S);
    o->K1 = (has_P ? orc_pressure_correction(1, Tk, P) : 1.0) * k1(Tk, S);
    o->K2 = (has_P ? orc_pressure_correction(2, Tk, P) : 1.0) * k2(Tk, S);
    o->KB = (has_P ? orc_pressure_correction(3, Tk, P) : 1.0) * kb(Tk, S);
    o->KW = (has_P ? orc_pressure_correction(4, Tk, P) : 1.0) * kw(Tk, S);
    o->KS = (has_P ? orc_pressure_correction(5, Tk, P) : 1.0) * ks(Tk, S, Is);
    o->KF = (has_P ? orc_pressure_correction(6, Tk, P) : 1.0) * kf(Tk, S, Is, o->KS);
    o->KP1 = (has_P ? orc_pressure_correction(7, Tk, P) : 1.0) * kp(KP1C, Tk, S);
    o->KP2 = (has_P ? orc_pressure_correction(8, Tk, P) : 1.0) * kp(KP2C, Tk, S);
    o->KP3 = (has_P ? orc_pressure_correction(9, Tk, P) : 1.0) * kp(KP3C, Tk, S);
    o->KSi = ksi(Tk, S, Is);
}

/* Virial coefficients (CarbonChemistry/virial_coefficients.jl:1-17). */
double orc_virial_B(double Tk)
{
    const double a = -1636.75, b = 12.0408, c = -3.27957 * 1e-2, d = 3.16528 * 1e-5;
    return (a + b * Tk + c * (Tk * Tk) + d * (Tk * Tk * Tk)) * 1e-6;
}

double orc_virial_delta(double Tk)
{
    return (57.7 + -0.118 * Tk) * 1e-6;
}

/* ---------------- alkalinity residual ---------------- */

/* alkalinity_residual and ∂ₕ_alkalinity_residual
 * (CarbonChemistry/alkalinity_residual.jl:18-75). */
double orc_alkalinity_residual(double H, const orc_carbon_params* p)
{
    double cd = H * H + p->K1 * H + p->K1 * p->K2;
    double pd = H * H * H + p->KP1 * (H * H) + p->KP1 * p->KP2 * H + p->KP1 * p->KP2 * p->KP3;
    double sd = 1 + p->sulfate / p->KS;
    double bicarbonate = p->K1 * H * p->DIC / cd;
    double carbonate = 2 * p->DIC * p->K1 * p->K2 / cd;
    double borate = p->boron / (1 + H / p->KB);
    double hydroxide = p->KW / H;
    double hydrogen_phosphate = p->phosphate * p->KP1 * p->KP2 * H / pd;
    double phosphate = 2 * p->phosphate * p->KP1 * p->KP2 * p->KP3 / pd;
    double silicate = p->silicate / (1 + H / p->KSi);
    double free_hydrogen = -H / sd;
    double hydrogen_sulfate = -p->sulfate / (1 + p->KS / H * sd);
    double hydrogen_fluoride = -p->fluoride / (1 + p->KF / H);
    double phosphoric_acid = -p->phosphate * (H * H * H) / pd;
    return bicarbonate + carbonate + borate + hydroxide + hydrogen_phosphate + phosphate + silicate
           + free_hydrogen + hydrogen_sulfate + hydrogen_fluoride + phosphoric_acid - p->Alk;
}

double orc_dalkalinity_residual(double H, const orc_carbon_params* p)
{
    double cd = H * H + p->K1 * H + p->K1 * p->K2;
    double pd = H * H * H + p->KP1 * (H * H) + p->KP1 * p->KP2 * H + p->KP1 * p->KP2 * p->KP3;
    double sd = 1 + p->sulfate / p->KS;
    double dcd = 2 * H + p->K1;
    double dpd = 3 * (H * H) + 2 * p->KP1 * H + p->KP1 * p->KP2;
    double dsd = 0.0;
    double d_bicarbonate = p->K1 * p->DIC * (p->K1 * p->K2 - H * H) / (cd * cd);
    double d_carbonate = -dcd * 2 * p->DIC * p->K1 * p->K2 / (cd * cd);
    double bd = 1 + H / p->KB;
    double d_borate = -p->boron / (bd * bd) / p->KB;
    double d_hydroxide = -p->KW / (H * H);
    double d_hydrogen_phosphate = p->phosphate * p->KP1 * p->KP2 / pd
                                  - dpd * p->phosphate * p->KP1 * p->KP2 * H / (pd * pd);
    double d_phosphate = -dpd * 2 * p->phosphate * p->KP1 * p->KP2 * p->KP3 / (pd * pd);
    double sid = 1 + H / p->KSi;
    double d_silicate = -p->silicate / (sid * sid) / p->KSi;
    double d_free_hydrogen = -1 / sd;
    double hsd = 1 + p->KS / H * sd;
    double d_hydrogen_sulfate = p->sulfate / (hsd * hsd) * (p->KS / H * dsd - p->KS / (H * H) * sd);
    double fd = 1 + p->KF / H;
    double d_hydrogen_fluoride = -p->fluoride / (fd * fd) * p->KF / (H * H);
    double d_phosphoric_acid = -3 * p->phosphate * (H * H) / pd - dpd * p->phosphate * (H * H * H) / (pd * pd);
    return d_bicarbonate + d_carbonate + d_borate + d_hydroxide + d_hydrogen_phosphate + d_phosphate
           + d_silicate + d_free_hydrogen + d_hydrogen_sulfate + d_hydrogen_fluoride + d_phosphoric_acid;
}

/* ---------------- damped Newton ---------------- */

static double julia_sign(double x)
{
    if (isnan(x)) return x;
    if (x > 0) return 1.0;
    if (x < 0) return -1.0;
    return x; /* ±0 */
}

typedef double (*orc_fn)(double, const void*);

/* DampedNewtonRaphsonSolver (Utils/solvers.jl:90-121) with bounded_λ
 * (:123-141).  fast_exit: when an iteration leaves x unchanged, every later
 * iteration repeats it exactly (same x, same f(x), same δ, same λ
 * sequence), so the remaining iterations are skipped with an identical
 * result. */
static double damped_newton(orc_fn f, orc_fn fp, const void* params, double x0, int max_iters,
                            double atol, double damping, double c, double lambda_min,
                            int has_lower, double lower, int fast_exit, int* iters, int* evals)
{
    double x = x0;
    int N = 0, ne = 0;
    double fx = f(x, params);
    ne++;
    while ((fabs(fx) > atol) && (N < max_iters)) {
        double delta = fx / fp(x, params);
        double lambda;
        if (has_lower) {
            double lb = (x - lower) / delta;
            lb = (lb < 0) ? 1.0 : lb;
            lambda = orc_julia_min(1.0, lb);
        } else {
            lambda = 1.0;
        }
        double fnew = f(x - lambda * delta, params);
        ne++;
        while ((fabs(fnew) >= fabs(fx) - c * lambda * fx * julia_sign(fx)) && (lambda > lambda_min)) {
            lambda *= damping;
            fnew = f(x - lambda * delta, params);
            ne++;
        }
        double xn = x - lambda * delta;
        N += 1;
        int stuck = (xn == x);
        x = xn;
        fx = fnew;
        if (fast_exit && stuck) N = max_iters > N ? max_iters : N;
    }
    if (iters) *iters = N;
    if (evals) *evals = ne;
    return x;
}

static double resid_cb(double H, const void* p) { return orc_alkalinity_residual(H, (const orc_carbon_params*)p); }
static double dresid_cb(double H, const void* p) { return orc_dalkalinity_residual(H, (const orc_carbon_params*)p); }

/* solve_for_H (carbon_chemistry.jl:217-218): x0 = 10^(-initial_pH_guess),
 * bounds (lower = 0, upper = nothing) (:81). */
double orc_solve_for_H(const orc_carbon_params* p, const obme_carbon_solver* s, double x0,
                       int fast_exit, int* iters, int* evals)
{
    return damped_newton(resid_cb, dresid_cb, p, x0, s->max_iters, s->atol, s->damping,
                         s->armijo_constant, s->min_damping, 1, 0.0, fast_exit, iters, evals);
}

/* test/test_solvers.jl:8-25 analytic functions */
static double tf1(double x, const void* p) { (void)p; return tanh(x); }
static double tf1p(double x, const void* p) { (void)p; double s = 1.0 / cosh(x); return s * s; }
static double tf2(double x, const void* p) { (void)p; return x * x - 2; }
static double tf2p(double x, const void* p) { (void)p; return 2 * x; }
static double tf3(double x, const void* p) { (void)p; return exp(x) - 1000; }
static double tf3p(double x, const void* p) { (void)p; return exp(x); }
static double tf4(double x, const void* p) { (void)p; return 1 / (1 + x * x) - 0.5; }
static double tf4p(double x, const void* p) { (void)p; double d = 1 + x * x; return -2 * x / (d * d); }
static double tf5(double x, const void* p) { (void)p; return 1.0 * (x * x) + -2.0; }
static double tf5p(double x, const void* p) { (void)p; return 2 * 1.0 * x; }

double orc_damped_newton_test(int which, double x0, double damping, int has_lower, double lower,
                              int fast_exit, int* iters)
{
    orc_fn f = tf1, fp = tf1p;
    switch (which) {
    case 2: f = tf2; fp = tf2p; break;
    case 3: f = tf3; fp = tf3p; break;
    case 4: f = tf4; fp = tf4p; break;
    case 5: f = tf5; fp = tf5p; break;
    default: break;
    }
    /* min_damping = damping^10 (solvers.jl:86), Julia's Float64^Int power. */
    double lm = orc_pow_int(damping, 10);
    return damped_newton(f, fp, NULL, x0, 100, 1e-20, damping, 0.5, lm, has_lower, lower,
                         fast_exit, iters, NULL);
}

/* ---------------- CarbonChemistry functor ---------------- */

static void fill_params(double DIC, double Alk, double T, double S, int densP, double Pbar_dens,
                        int has_P, double P, double silicate, double phosphate,
                        orc_carbon_params* p, orc_constants* k, double* Tk_out)
{
    /* carbon_chemistry.jl:116-118 defaults */
    double boron = 0.000232 / 10.811 * S / 1.80655;
    double sulfate = 0.14 / 96.06 * S / 1.80655;
    double fluoride = 0.000067 / 18.9984 * S / 1.80655;
    (void)densP;
    double rho = orc_teos10_density(T, S, Pbar_dens);
    double Tk = T + 273.15;
    Alk *= 1e-3 / rho;
    DIC *= 1e-3 / rho;
    phosphate *= 1e-3 / rho;
    silicate *= 1e-3 / rho;
    orc_equilibrium_constants(Tk, S, has_P, P, k);
    p->DIC = DIC; p->Alk = Alk; p->boron = boron; p->sulfate = sulfate; p->fluoride = fluoride;
    p->silicate = silicate; p->phosphate = phosphate;
    p->K1 = k->K1; p->K2 = k->K2; p->KB = k->KB; p->KW = k->KW; p->KS = k->KS; p->KF = k->KF;
    p->KP1 = k->KP1; p->KP2 = k->KP2; p->KP3 = k->KP3; p->KSi = k->KSi;
    *Tk_out = Tk;
}

/* (p::CarbonChemistry)(; DIC, T, S, Alk, pH, output) (carbon_chemistry.jl:111-165)
 * with selected_output (:167-211); P = nothing (surface). */
double orc_carbon_chemistry(double DIC, double Alk, double T, double S, double pH, int output,
                            const obme_carbon_solver* s, int fast_exit, double silicate,
                            double phosphate, int* evals)
{
    orc_carbon_params p;
    orc_constants k;
    double Tk;
    /* ρₒ = density(T, S, ifelse(isnothing(P), 1, P)) (:123) */
    fill_params(DIC, Alk, T, S, 1, 1.0, 0, 0.0, silicate, phosphate, &p, &k, &Tk);
    double H;
    if (pH > -100) {
        H = pow(10.0, -pH); /* solve_for_H(pH) (:213) */
        if (evals) *evals = 0;
    } else {
        H = orc_solve_for_H(&p, s, pow(10.0, -s->initial_pH_guess), fast_exit, NULL, evals);
    }
    double K0 = k.K0;
    double CO2 = p.DIC * (H * H) / (H * H + k.K1 * H + k.K1 * k.K2);
    double fCO2 = (CO2 / K0) * 1e6;
    switch (output) {
    case 0:
        return fCO2;
    case 2:
        return -log10(H);
    case 3: {
        double HSO4 = p.sulfate / (1 + k.KS / H);
        return -log10(H + HSO4);
    }
    case 4: {
        double HSO4 = p.sulfate / (1 + k.KS / H);
        double HF = p.fluoride / (1 + k.KF / H);
        return -log10(H + HSO4 + HF);
    }
    default: {
        /* selected_output(::Val{:pCO₂}) (:170-193) */
        double P = 1.0;
        P *= 101325.0;
        double B = orc_virial_B(Tk);
        double d = orc_virial_delta(Tk);
        fCO2 *= 0.09807;
        double phi = 1.0;
        double x = fCO2 / (phi * P);
        for (int n = 0; n < 3; n++) {
            double om = 1.0 - x;
            phi = exp((B + 2.0 * (om * om) * d) * P / (8.31446261815324 * Tk));
            x = fCO2 / (phi * P);
        }
        double pCO2 = fCO2 / phi;
        pCO2 /= 0.09807;
        return pCO2;
    }
    }
}

/* carbonate_concentration (CarbonChemistry/calcite_concentration.jl:1-53):
 * density at ifelse(isnothing(P), 0, P). */
double orc_carbonate_concentration(double DIC, double Alk, double T, double S, int has_P, double P,
                                   const obme_carbon_solver* s)
{
    orc_carbon_params p;
    orc_constants k;
    double Tk;
    fill_params(DIC, Alk, T, S, 0, has_P ? P : 0.0, has_P, P, 0.0, 0.0, &p, &k, &Tk);
    double H = orc_solve_for_H(&p, s, pow(10.0, -s->initial_pH_guess), 1, NULL, NULL);
    double denom1 = (H * (H + k.K1));
    double denom2 = (1.0 + k.K1 * k.K2 / denom1);
    return p.DIC * k.K1 * k.K2 / denom1 / denom2;
}

/* calcite_saturation (calcite_concentration.jl:55-79) */
double orc_calcite_saturation(double DIC, double Alk, double T, double S, int has_P, double P,
                              const obme_carbon_solver* s)
{
    double ca = 0.0103 * S / 35;
    double co3 = orc_carbonate_concentration(DIC, Alk, T, S, has_P, P, s);
    double K = orc_ksp_calcite(T + 273.15, S, has_P, P);
    return ca * co3 / K;
}

/* calcite_saturation with a silicate concentration (mmol/m3), as PISCES
 * calls it per cell (PISCES/compute_calcite_saturation.jl:22-36). */
double orc_calcite_saturation_si(double DIC, double Alk, double T, double S, int has_P, double P, double silicate,
                                 const obme_carbon_solver* s)
{
    orc_carbon_params p;
    orc_constants k;
    double Tk;
    fill_params(DIC, Alk, T, S, 0, has_P ? P : 0.0, has_P, P, silicate, 0.0, &p, &k, &Tk);
    double H = orc_solve_for_H(&p, s, pow(10.0, -s->initial_pH_guess), 1, NULL, NULL);
    double denom1 = (H * (H + k.K1));
    double denom2 = (1.0 + k.K1 * k.K2 / denom1);
    double co3 = p.DIC * k.K1 * k.K2 / denom1 / denom2;
    double ca = 0.0103 * S / 35;
    return ca * co3 / orc_ksp_calcite(T + 273.15, S, has_P, P);
}

void orc_calcite_saturation_batch(const double* DIC, const double* Alk, const double* T, const double* S,
                                  const double* P, const double* Si, double* out, long n,
                                  const obme_carbon_solver* s, int nthreads)
{
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (long q = 0; q < n; q++) out[q] = orc_calcite_saturation_si(DIC[q], Alk[q], T[q], S[q], 1, P[q], Si[q], s);
}

/* ---------------- gas exchange ---------------- */

/* CarbonDioxidePolynomialSchmidtNumber (GasExchange/gas_solubility.jl:211-212)
 * evaluated as PolynomialParameterisation{4} (generic_parameterisations.jl:31-32). */
double orc_schmidt_co2(const obme_co2_exchange* g, double T)
{
    const double* c = g->schmidt;
    return c[0] + c[1] * T + c[2] * (T * T) + c[3] * (T * T * T) + c[4] * orc_pow_int4(T);
}

/* SchmidtScaledTransferVelocity (gas_transfer_velocity.jl:33-34) with Ho06
 * (:61) and MolPerKgPerAtmToMMolPerCubicMPerMicroAtm (gas_solubility.jl:204):
 * k = k660(u)/√(Sc/660) · K0(T+273.15, S)·ρ(T, S, 0)/10³. */
double orc_transfer_velocity(const obme_co2_exchange* g, double u10, double T, double S)
{
    const double* c = g->transfer_velocity;
    double k660 = c[0] + c[1] * u10 + c[2] * (u10 * u10);
    double sc = orc_schmidt_co2(g, T);
    double sol = k0(T + 273.15, S) * orc_teos10_density(T, S, 0.0) / 1000.0;
    return k660 / sqrt(sc / 660.0) * sol;
}

/* GasExchange functor (gas_exchange.jl:26-38) for O2
 * (OxygenGasExchangeBoundaryCondition, GasExchange.jl:127-146):
 * F = k (O2 - C_air beta / Tk), k = k660(u) / sqrt(Sc(T) / 660)
 * (gas_transfer_velocity.jl:33-34), Sc the quartic of schmidt_number.jl:15-16,
 * beta the solubility of gas_solubility.jl:37-48 (B2 in both salinity terms,
 * as written there). */
double orc_o2_flux(const obme_o2_exchange* g, double O2, double T, double S)
{
    const double* c = g->schmidt;
    double Sc = c[0] + c[1] * T + c[2] * (T * T) + c[3] * (T * T * T) + c[4] * orc_pow_int4(T);
    const double* v = g->transfer_velocity;
    double u = g->wind_speed;
    double k = (v[0] + v[1] * u + v[2] * (u * u)) / sqrt(Sc / 660.0) * 1.0;
    const double* a = g->solubility;
    double Tk = T + 273.15, Tk100 = Tk / 100.0;
    double beta = exp(a[0] + a[1] / Tk100 + a[2] * log(Tk100) + S * (a[3] + a[4] * Tk100 + a[4] * (Tk100 * Tk100)));
    return k * (O2 - g->air_concentration * (beta / Tk));
}

/* GasExchange functor (gas_exchange.jl:26-38) with CarbonDioxideConcentration
 * surface_value (carbon_dioxide_concentration.jl:48-60). */
double orc_co2_flux(const obme_co2_exchange* g, const obme_carbon_solver* s, double DIC, double Alk,
                    double T, double S, double u10, double pCO2_air, int fast_exit, double* pH_out,
                    int* evals)
{
    double k = orc_transfer_velocity(g, u10, T, S);
    /* pCO₂ and the pH of the same solve */
    orc_carbon_params p;
    orc_constants kc;
    double Tk;
    fill_params(DIC, Alk, T, S, 1, 1.0, 0, 0.0, 0.0, 0.0, &p, &kc, &Tk);
    double H = orc_solve_for_H(&p, s, pow(10.0, -s->initial_pH_guess), fast_exit, NULL, evals);
    if (pH_out) *pH_out = -log10(H);
    double CO2 = p.DIC * (H * H) / (H * H + kc.K1 * H + kc.K1 * kc.K2);
    double fCO2 = (CO2 / kc.K0) * 1e6;
    /* surface_value calls the CarbonChemistry functor without P, so
     * selected_output uses P = 1 atm (carbon_chemistry.jl:171). */
    double P = 1.0;
    P *= 101325.0;
    double B = orc_virial_B(Tk), d = orc_virial_delta(Tk);
    fCO2 *= 0.09807;
    double phi = 1.0;
    double x = fCO2 / (phi * P);
    for (int n = 0; n < 3; n++) {
        double om = 1.0 - x;
        phi = exp((B + 2.0 * (om * om) * d) * P / (8.31446261815324 * Tk));
        x = fCO2 / (phi * P);
    }
    double pCO2 = fCO2 / phi;
    pCO2 /= 0.09807;
    return k * (pCO2 - pCO2_air);
}

void orc_co2_flux_batch(const obme_co2_exchange* g, const obme_carbon_solver* s, const double* DIC,
                        const double* Alk, const double* T, const double* S, const double* u10,
                        double pCO2_air, double* out_pH, double* out_flux, long n, int nthreads)
{
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (long i = 0; i < n; i++) {
        double pH;
        double u = u10 ? u10[i] : g->wind_speed;
        double F = orc_co2_flux(g, s, DIC[i], Alk[i], T[i], S[i], u, pCO2_air, 1, &pH, NULL);
        if (out_pH) out_pH[i] = pH;
        if (out_flux) out_flux[i] = F;
    }
}
