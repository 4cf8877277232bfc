/*
 * oracle_npd.c — FP64 restatement of OceanBioME's NutrientsPlanktonDetritus
 * tendencies (NPZD and LOBSTER + CarbonateSystem), the TwoBand PAR column
 * recurrence, the sinking flux divergence and the RK3 stage sequence.
 * TEST INFRASTRUCTURE ONLY (the oracle; see oracle.h).
 *
 * Reference files (relative to OceanBioME.jl/src):
 *   Models/AdvectedPopulations/NutrientsPlanktonDetritus/{plankton,nutrients,
 *   detritus,carbonate_system}.jl, Light/2band.jl, Light/Light.jl,
 *   Utils/sinking_velocity_fields.jl, BoxModel/timesteppers.jl, OceanBioME.jl.
 *
 * Each reference helper recomputes shared sub-expressions (growth, grazing);
 * in FP64 a recomputation yields the identical value, so each is computed
 * once here and reused, with the reference's operand order kept.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPS0 4.9406564584124654e-324 /* eps(0.0) */

int orc_npd_ntracers(int model, int carbonate)
{
    if (model == OBME_NPZD) return 5;
    if (model == OBME_LOBSTER) return carbonate ? 9 : 7;
    return 0;
}

/* mortality / concentration_limit (plankton.jl:86-90) */
static double mortality(int form, double X, double m) { return form == OBME_LINEAR ? m * X : m * (X * X); }
static double concentration_limit(int form, double X, double k)
{
    return form == OBME_LINEAR ? X / (X + k) : (X * X) / ((X * X) + (k * k));
}

/* light_limitation (plankton.jl:219-220) */
static double light_limitation(int form, double PAR, double kPAR)
{
    if (form == OBME_LIGHT_MONDO) return PAR / (kPAR + PAR);
    return PAR / sqrt((PAR * PAR) + (kPAR * kPAR));
}

/* weighted_phytoplankton_preference (plankton.jl:393-397) */
static double preference(const obme_npd_params* q, double P, double sPOM)
{
    double pt = q->preference_for_phytoplankton;
    return pt * P / (pt * P + (1 - pt) * sPOM + EPS0);
}

typedef struct grazing_terms {
    double total;   /* total_grazing (plankton.jl:118-135)      */
    double onP;     /* grazing(Val{:P}) (:352-369)              */
    double onSmall; /* grazing(Val{:sPOM}/Val{:D}) (:371-389)   */
} grazing_terms;

static grazing_terms grazing(const obme_npd_params* q, double P, double Z, double sPOM)
{
    double kG = q->grazing_half_saturation;
    double g = q->maximum_grazing_rate;
    double p = preference(q, P, sPOM);
    double food = p * P + (1 - p) * sPOM;
    double L = concentration_limit(q->grazing_concentration_formulation, food, kG);
    grazing_terms out;
    out.total = g * L * Z;
    out.onP = g * p * L * P / (food + orc_julia_eps(food)) * Z;
    out.onSmall = g * (1 - p) * L * sPOM / (food + orc_julia_eps(food)) * Z;
    return out;
}

/* NPZD: Nutrient + PhytoZoo + Detritus; tracers (N, P, Z, T, D). */
static void npzd_tendencies(const obme_npd_params* q, const double* v, double PAR, double* out, const obme_oxygen* ox)
{
    double N = v[0], P = v[1], Z = v[2], T = v[3], D = v[4];

    /* phytoplankton_growth (plankton.jl:191-206) with nutrient_limitation
     * (Nutrient, :180-188) and temperature_limitation (:208-214) */
    double Ln = N / (N + q->nitrate_half_saturation);
    double Ll = light_limitation(q->light_limitation, PAR, q->light_half_saturation);
    double Lt = q->use_temperature_coefficient ? pow(q->temperature_coefficient, T / 10) : 1.0;
    double muP = q->phytoplankton_maximum_growth_rate * Ll * Ln * Lt * P;

    /* small_particulate_concentration(::Detritus) (detritus.jl:386-387) */
    double sPOM = D * q->small_particle_fraction;
    grazing_terms G = grazing(q, P, Z, sPOM);
    double nuP = mortality(q->phytoplankton_mortality_formulation, P, q->phytoplankton_mortality_rate);

    /* P (plankton.jl:92-105) */
    out[1] = (1 - q->phytoplankton_exudation_fraction) * muP - G.onP - nuP;
    /* Z (plankton.jl:107-116) */
    out[2] = q->zooplankton_assimilation_fraction * G.total - q->zooplankton_mortality_rate * (Z * Z)
             - q->zooplankton_excretion_rate * Z;
    /* N (nutrients.jl:86-90): plankton_inorganic_nitrogen_waste (plankton.jl:292-306)
     * + detritus_inorganic_nitrogen_waste (detritus.jl:389-390)
     * - nutrient_uptake(Val{:N}) (plankton.jl:271-278) */
    double pinw = q->excretion_inorganic_fraction * q->zooplankton_excretion_rate * Z
                  + (1 - q->phytoplankton_solid_waste_fraction) * nuP;
    double dinw = D * q->remineralisation_rate;
    double uptake = muP * (1 - q->ammonia_fraction_of_exudate * q->phytoplankton_exudation_fraction);
    out[0] = pinw + dinw - uptake;
    /* T: no method -> zero(grid) (NutrientsPlanktonDetritus.jl:88) */
    out[3] = 0.0;
    /* D (detritus.jl:373-379): plankton_organic_nitrogen_waste (plankton.jl:312-325)
     * + solid_waste (:330-346) - grazing(Val{:D}) - remineralisation */
    double ponw = (1 - q->ammonia_fraction_of_exudate) * q->phytoplankton_exudation_fraction * muP
                  + (1 - q->excretion_inorganic_fraction) * q->zooplankton_excretion_rate * Z;
    double solid = (1 - q->zooplankton_assimilation_fraction) * G.total
                   + q->phytoplankton_solid_waste_fraction * nuP + q->zooplankton_mortality_rate * (Z * Z);
    out[4] = ponw + solid - G.onSmall - q->remineralisation_rate * D;
    if (ox) {
        /* Oxygen, Nutrient method (oxygen.jl:35-46). Julia ends the statement
         * `respiration = plankton_inorganic_nitrogen_waste(...)` at the line
         * break; the next line, `+ detritus_inorganic_nitrogen_waste(...)`,
         * is a separate unary-plus expression whose value is discarded. So
         * respiration is the plankton's inorganic nitrogen waste alone. */
        double Rp = ox->respiration_oxygen_nitrogen_ratio;
        double Rn = ox->nitrification_oxygen_nitrogen_ratio;
        out[5] = Rp * muP - (Rp - Rn) * pinw;
    }
}

/* LOBSTER: NitrateAmmonia + PhytoZoo + TwoParticleAndDissolved [+ CarbonateSystem]
 * [+ Oxygen]; tracers (NO3, NH4, P, Z, sPOM, bPOM, DOM [, DIC, Alk] [, O2]).
 * ox is NULL without the Oxygen component. */
static void lobster_tendencies(const obme_npd_params* q, int carbonate, const obme_oxygen* ox,
                               const double* v, double PAR, double* out)
{
    double NO3 = v[0], NH4 = v[1], P = v[2], Z = v[3], sPOM = v[4], bPOM = v[5], DOM = v[6];

    /* nitrogen_limitation (plankton.jl:150-166) */
    double nl = NO3 * exp(-q->nitrate_ammonia_inhibition * NH4) / (NO3 + q->nitrate_half_saturation);
    double al = orc_julia_max(0.0, NH4 / (q->ammonia_half_saturation + NH4));
    double Ln = (nl + al) / 2;
    double Ll = light_limitation(q->light_limitation, PAR, q->light_half_saturation);
    /* temperature_coefficient = nothing -> one(eltype) (plankton.jl:208) */
    double Lt = 1.0;
    double muP = q->phytoplankton_maximum_growth_rate * Ll * Ln * Lt * P;

    grazing_terms G = grazing(q, P, Z, sPOM);
    double nuP = mortality(q->phytoplankton_mortality_formulation, P, q->phytoplankton_mortality_rate);

    double gam = q->phytoplankton_exudation_fraction;
    double alpha_x = q->ammonia_fraction_of_exudate;
    double nitrification = q->nitrification_rate * NH4; /* nutrients.jl:55 */

    /* nutrient_uptake NO3 / NH4 (plankton.jl:223-261) */
    double up_NO3 = muP * nl / (nl + al + EPS0);
    double waste = alpha_x * gam * muP;
    double up_NH4 = muP * al / (nl + al + EPS0) - waste;

    /* plankton_inorganic_nitrogen_waste (plankton.jl:292-306) */
    double pinw = q->excretion_inorganic_fraction * q->zooplankton_excretion_rate * Z
                  + (1 - q->phytoplankton_solid_waste_fraction) * nuP;
    /* detritus_inorganic_nitrogen_waste (detritus.jl:252-263) */
    double a = q->remineralisation_inorganic_fraction, smu = q->small_remineralisation_rate,
           bmu = q->large_remineralisation_rate, dmu = q->dissolved_remineralisation_rate;
    double dinw = (a * (smu * sPOM + bmu * bPOM) + dmu * DOM);
    /* plankton_organic_nitrogen_waste (plankton.jl:312-325) */
    double ponw = (1 - alpha_x) * gam * muP
                  + (1 - q->excretion_inorganic_fraction) * q->zooplankton_excretion_rate * Z;
    /* detritus_organic_nitrogen_waste (detritus.jl:265-274) */
    double donw = (1 - a) * (smu * sPOM + bmu * bPOM);
    /* solid_waste (plankton.jl:330-346) */
    double solid = (1 - q->zooplankton_assimilation_fraction) * G.total
                   + q->phytoplankton_solid_waste_fraction * nuP + q->zooplankton_mortality_rate * (Z * Z);

    /* NO3, NH4 (nutrients.jl:57-67) */
    double dNO3 = nitrification - up_NO3;
    double dNH4 = pinw + dinw - nitrification - up_NH4;
    out[0] = dNO3;
    out[1] = dNH4;
    /* P, Z (plankton.jl:92-116) */
    out[2] = (1 - gam) * muP - G.onP - nuP;
    out[3] = q->zooplankton_assimilation_fraction * G.total - q->zooplankton_mortality_rate * (Z * Z)
             - q->zooplankton_excretion_rate * Z;
    /* sPOM, bPOM, DOM (detritus.jl:234-249) */
    out[4] = q->small_solid_waste_fraction * solid - G.onSmall - smu * sPOM;
    out[5] = (1 - q->small_solid_waste_fraction) * solid - bmu * bPOM;
    out[6] = ponw + donw - dmu * DOM;

    if (carbonate) {
        double R = q->redfield_ratio, rho = q->carbon_calcite_ratio;
        double Rd = q->detritus_redfield_ratio;
        /* phytoplankton_primary_production (plankton.jl:280-290) */
        double pp = (1 + rho * (1 - gam) - alpha_x * gam) * muP * R;
        /* plankton_inorganic_carbon_waste (plankton.jl:308-309) */
        double picw = R * pinw;
        /* detritus_inorganic_carbon_waste (detritus.jl:276-287) */
        double sPOC = sPOM * Rd, bPOC = bPOM * Rd, DOC = DOM * Rd;
        double dicw = a * (smu * sPOC + bmu * bPOC) + dmu * DOC;
        /* calcite_dissolution, fixed-Redfield override (plankton.jl:429-441) */
        double cdis = (G.onP + nuP) * rho * R;
        /* calcite_uptake (plankton.jl:443-450) */
        double cup = 2 * rho * muP * R;
        /* DIC, Alk (carbonate_system.jl:50-62) */
        out[7] = -pp + picw + dicw + cdis;
        out[8] = dNH4 * 0.9375 - dNO3 * 1.0625 - 2.0 * cup + 2.0 * cdis;
    }
    if (ox) {
        /* Oxygen (oxygen.jl:21-31), NitrateAmmonia method: phytoplankton_growth
         * is muP, "nitrate_production" is the NH4 tendency, and the
         * nitrification term is muNH4 = nitrification_rate NH4 */
        double Rp = ox->respiration_oxygen_nitrogen_ratio;
        double Rn = ox->nitrification_oxygen_nitrogen_ratio;
        out[carbonate ? 9 : 7] = Rp * muP - (Rp - Rn) * dNH4 - Rp * nitrification;
    }
}

void orc_npd_tendencies(int model, int carbonate, const obme_npd_params* p, const double* vals,
                        double PAR, double* out)
{
    if (model == OBME_NPZD) npzd_tendencies(p, vals, PAR, out, NULL);
    else lobster_tendencies(p, carbonate, NULL, vals, PAR, out);
}

void orc_npd_tendencies_oxygen(int model, int carbonate, const obme_npd_params* p, const obme_oxygen* ox,
                               const double* vals, double PAR, double* out)
{
    if (model == OBME_NPZD) npzd_tendencies(p, vals, PAR, out, ox);
    else lobster_tendencies(p, carbonate, ox, vals, PAR, out);
}

/* tracer count and tendencies of a configuration (its optional components included) */
static int cfg_ntracers(const obme_config* cfg)
{
    int n = orc_npd_ntracers(cfg->model, cfg->carbonate_system);
    return n + (cfg->oxygen.enabled ? 1 : 0);
}

static void cfg_tendencies(const obme_config* cfg, const double* vals, double PAR, double* out)
{
    if (cfg->model == OBME_NPZD) npzd_tendencies(&cfg->npd, vals, PAR, out, cfg->oxygen.enabled ? &cfg->oxygen : NULL);
    else lobster_tendencies(&cfg->npd, cfg->carbonate_system, cfg->oxygen.enabled ? &cfg->oxygen : NULL,
                            vals, PAR, out);
}

/* ---------------- light ---------------- */

/* default_surface_PAR (Light/Light.jl:217-219); hour = 3600 s. */
double orc_default_surface_PAR(double t)
{
    return 100 * orc_julia_max(0.0, cos(t * M_PI / (12 * 3600.0)));
}

/* update_TwoBandPhotosyntheticallyActiveRadiation! (Light/2band.jl:1-33). */
void orc_two_band_column(const obme_two_band_par* par, int Nz, const double* zc, const double* zf,
                         const double* P, double PAR0, double* PAR)
{
    double kr = par->water_red_attenuation, kb = par->water_blue_attenuation;
    double xr = par->chlorophyll_red_attenuation, xb = par->chlorophyll_blue_attenuation;
    double er = par->chlorophyll_red_exponent, eb = par->chlorophyll_blue_exponent;
    double r = par->pigment_ratio, Rcp = par->phytoplankton_chlorophyll_ratio;
    int t = Nz - 1; /* 0-based top cell */
    double chlr = (zf[Nz] - zc[t]) * pow(P[t] * Rcp / r, er);
    double chlb = (zf[Nz] - zc[t]) * pow(P[t] * Rcp / r, eb);
    PAR[t] = PAR0 * (exp(kr * zc[t] - xr * chlr) + exp(kb * zc[t] - xb * chlb)) / 2;
    for (int k = Nz - 2; k >= 0; k--) {
        chlr += (zc[k + 1] - zf[k + 1]) * pow(P[k + 1] * Rcp / r, er) + (zf[k + 1] - zc[k]) * pow(P[k] * Rcp / r, er);
        chlb += (zc[k + 1] - zf[k + 1]) * pow(P[k + 1] * Rcp / r, eb) + (zf[k + 1] - zc[k]) * pow(P[k] * Rcp / r, eb);
        PAR[k] = PAR0 * (exp(kr * zc[k] - xr * chlr) + exp(kb * zc[k] - xb * chlb)) / 2;
    }
}

/* ---------------- sinking ---------------- */

/* setup_velocity_fields (Utils/sinking_velocity_fields.jl:10-35): w on faces
 * 1..Nz (0-based 0..Nz-1) = -speed * (open ? 1 : 1 - exp((1-k)/d)); top face 0. */
static void sinking_profile(const obme_sinking* s, double speed, int Nz, double* w)
{
    double d = s->smoothing_distance > 0 ? s->smoothing_distance : 2.0;
    for (int k = 1; k <= Nz; k++)
        w[k - 1] = -speed * (s->open_bottom ? 1.0 : (1 - exp((1 - k) / d)));
    w[Nz] = 0.0;
}

static double sq(double x) { return x * x; }

/* WENO-Z weights (Borges et al. 2008, J. Comput. Phys. 227:3191) on the
 * Jiang & Shu (1996) smoothness indicators: alpha_r = d_r (1 + (tau /
 * (beta_r + eps))^2), normalised; eps = 1e-8.  This is the form of
 * Oceananigans' WENO reconstruction; Oceananigans is not part of the
 * reference tree, so parity with it is unpinned (the closed forms below are
 * the textbook scheme). */
#define WENO_EPS 1e-8

/* The value at face k reconstructed from above (w <= 0: the upwind cell is
 * k, the cells above it k+1, k+2 are further upwind): fifth order from the
 * cells k-2 .. k+2, third order from k-1 .. k+1.  The mirror images of the
 * left-biased textbook stencils. */
static double weno5_from_above(double cm2, double cm1, double c0, double cp1, double cp2)
{
    const double q0 = (2 * cp2 - 7 * cp1 + 11 * c0) / 6;
    const double q1 = (-cp1 + 5 * c0 + 2 * cm1) / 6;
    const double q2 = (2 * c0 + 5 * cm1 - cm2) / 6;
    const double b0 = 13.0 / 12 * sq(cp2 - 2 * cp1 + c0) + 0.25 * sq(cp2 - 4 * cp1 + 3 * c0);
    const double b1 = 13.0 / 12 * sq(cp1 - 2 * c0 + cm1) + 0.25 * sq(cp1 - cm1);
    const double b2 = 13.0 / 12 * sq(c0 - 2 * cm1 + cm2) + 0.25 * sq(3 * c0 - 4 * cm1 + cm2);
    const double tau = fabs(b0 - b2);
    const double a0 = 0.1 * (1 + sq(tau / (b0 + WENO_EPS)));
    const double a1 = 0.6 * (1 + sq(tau / (b1 + WENO_EPS)));
    const double a2 = 0.3 * (1 + sq(tau / (b2 + WENO_EPS)));
    return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}

static double weno3_from_above(double cm1, double c0, double cp1)
{
    const double q0 = (-cp1 + 3 * c0) / 2;
    const double q1 = (c0 + cm1) / 2;
    const double b0 = sq(c0 - cp1);
    const double b1 = sq(cm1 - c0);
    const double tau = fabs(b0 - b1);
    const double a0 = (1.0 / 3) * (1 + sq(tau / (b0 + WENO_EPS)));
    const double a1 = (2.0 / 3) * (1 + sq(tau / (b1 + WENO_EPS)));
    return (a0 * q0 + a1 * q1) / (a0 + a1);
}

/* The face value at an inner face k (1 <= k <= Nz-1).  Near the boundaries
 * the order drops as Oceananigans' schemes do in a Bounded direction (the
 * face must lie at least the scheme's halo size from the boundary): WENO5
 * on faces 3 .. Nz-3, WENO3 on faces 2 and Nz-2, first-order upwind on
 * faces 1 and Nz-1. */
static double face_value(int scheme, int Nz, const double* c, int k)
{
    if (scheme == OBME_CENTERED2) return (c[k - 1] + c[k]) / 2;
    if (scheme == OBME_WENO5) {
        if (k >= 3 && k <= Nz - 3) return weno5_from_above(c[k - 2], c[k - 1], c[k], c[k + 1], c[k + 2]);
        if (k >= 2 && k <= Nz - 2) return weno3_from_above(c[k - 1], c[k], c[k + 1]);
    }
    return c[k];
}

/* -∂z(w c) in flux form for one tracer column (Oceananigans div_Uc with the
 * drift velocity; the in-tree flux use is Sediments/tracked_fields.jl:59-60).
 * Face k (0-based, between cells k-1 and k): w <= 0, so the upwind cell is
 * k; at the bottom face the zero-gradient halo gives cell 0 for every
 * scheme; the top face carries w = 0. */
static void sinking_divergence(int scheme, int Nz, const double* w, const double* c, const double* dz,
                               double* div)
{
    for (int k = 0; k < Nz; k++) {
        const double cf_hi = (k + 1 < Nz) ? face_value(scheme, Nz, c, k + 1) : c[k];
        const double cf_lo = (k > 0) ? face_value(scheme, Nz, c, k) : c[k];
        double Fhi = w[k + 1] * cf_hi;
        double Flo = w[k] * cf_lo;
        div[k] = (Fhi - Flo) / dz[k];
    }
}

double orc_sinking_face_value(int scheme, int Nz, const double* c, int k)
{
    if (k <= 0) return c[0];
    if (k >= Nz) return c[Nz - 1];
    return face_value(scheme, Nz, c, k);
}

void orc_sinking_profile(const obme_sinking* s, double speed, int Nz, double* w)
{
    sinking_profile(s, speed, Nz, w);
}

void orc_sinking_divergence(int scheme, int Nz, const double* w, const double* c, const double* dz,
                            double* div)
{
    sinking_divergence(scheme, Nz, w, c, dz, div);
}

/* ---------------- column stepper ---------------- */

/* ScaleNegativeTracers on the tracer values of one cell, group after group
 * (src/Utils/negative_tracers.jl: scale_for_negs_cpu! :255-276, the same
 * arithmetic as the GPU kernel :198-251): t = Σ s·c, p = Σ s·c over c > 0,
 * t := fill if t < 0, c := c·t/p if c > 0 or c is not finite, else 0. */
void orc_scale_negatives(const obme_negative_scaling* s, double* vals)
{
    for (int g = 0; g < s->n_groups; g++) {
        double t = 0.0, p = 0.0;
        for (int m = 0; m < s->n_members[g]; m++) {
            double value = vals[s->member[g][m]];
            double sf = s->scalefactor[g][m];
            t += value * sf;
            if (value > 0) p += value * sf;
        }
        t = (t < 0) ? s->invalid_fill_value[g] : t;
        for (int m = 0; m < s->n_members[g]; m++) {
            double value = vals[s->member[g][m]];
            vals[s->member[g][m]] = (!isfinite(value) || value > 0) ? value * t / p : 0.0;
        }
    }
}

static void scale_column(const obme_config* cfg, int Nz, int nt, double* U)
{
    if (cfg->scaling.n_groups <= 0) return;
    double vals[OBME_MAX_TRACERS];
    for (int k = 0; k < Nz; k++) {
        for (int n = 0; n < nt; n++) vals[n] = U[n * Nz + k];
        orc_scale_negatives(&cfg->scaling, vals);
        for (int n = 0; n < nt; n++) U[n * Nz + k] = vals[n];
    }
}

/* ---------------- sediment ---------------- */

/* The tracked sinking flux (Sediments/tracked_fields.jl:53-70): for each of
 * the sediment's sinking tracers, -w c at the bottom face (k = 1, 1-based)
 * per unit area, w from the drift velocity profile and c the upwind (bottom)
 * cell, summed in the listed order (instant_remineralisation.jl:104).  A
 * tracer without a drift velocity has a zero flux.  (The reference divides
 * the advective flux Az w c by Az: the same value to rounding.) */
double orc_sediment_flux(const obme_config* cfg, const double* U, int Nz)
{
    const obme_sediment* s = &cfg->sediment;
    double f = 0.0;
    double* w = (double*)malloc(sizeof(double) * (Nz + 1));
    for (int q = 0; q < s->n_sinking; q++) {
        int n = s->sinking[q];
        double sp = cfg->sinking.speed[n];
        double fq = 0.0;
        if (sp != 0.0) {
            sinking_profile(&cfg->sinking, sp, Nz, w);
            fq = -(w[0] * U[n * Nz + 0]);
        }
        f = (q == 0) ? fq : f + fq;
    }
    free(w);
    return f;
}

/* e = a + b (f / (k + f))^2 (instant_remineralisation.jl:101-123; the
 * docstring's b f / (k + f)^2 is not what the code computes). */
double orc_burial_efficiency(const obme_sediment* s, double f)
{
    double x = f / (s->burial_efficiency_half_saturation + f);
    return s->burial_efficiency_constant1 + s->burial_efficiency_constant2 * (x * x);
}

int orc_sediment_fields(const obme_sediment* s)
{
    if (s->model == OBME_SEDIMENT_INSTANT_REMINERALISATION) return 1;
    if (s->model == OBME_SEDIMENT_SIMPLE_MULTI_G) return 3;
    return 0;
}

/* The tendencies of the sediment's fields (compute_tendencies.jl:7-31):
 * InstantRemineralisation e f (instant_remineralisation.jl:101-110);
 * SimpleMultiG's nitrogen form (simple_multi_G.jl:165-193):
 * Ns: (1 - fr) fs f - λs Ns, Nf: (1 - fr) ff f - λf Nf, Nr: fr f. */
void orc_sediment_tendencies(const obme_sediment* s, const double* u, double f, double* G)
{
    if (s->model == OBME_SEDIMENT_INSTANT_REMINERALISATION) {
        G[0] = orc_burial_efficiency(s, f) * f;
    } else if (s->model == OBME_SEDIMENT_SIMPLE_MULTI_G) {
        const double fr = s->refactory_fraction;
        G[0] = (1 - fr) * s->slow_fraction * f - s->slow_decay_rate * u[0];
        G[1] = (1 - fr) * s->fast_fraction * f - s->fast_decay_rate * u[1];
        G[2] = fr * f;
    }
}

/* SimpleMultiG's fractions (simple_multi_G.jl:372-427), day = 86400 s;
 * a non-finite fraction is zero, as the reference's ifelse(isfinite(p), p, 0). */
static const double SMG_DAY = 86400.0;

static double smg_finite(double p) { return isfinite(p) ? p : 0.0; }

static double smg_ammonia_oxidation(const obme_sediment* s, double Nr, double Cr, double k, double NH4, double O2)
{
    const double* q = s->nitrate_oxidation_params;
    double lC = log(Cr * SMG_DAY);
    double ln = (q[0] + q[1] * lC * log(O2) + q[2] * (lC * lC) + q[3] * log(k * SMG_DAY) * log(NH4) + q[4] * lC
                 + q[5] * lC * log(NH4));
    double p = exp(ln) / (Nr * SMG_DAY) * O2 / (s->anoxia_half_saturation + O2);
    return smg_finite(p);
}

static double smg_denitrification(const obme_sediment* s, double Cr, double k, double NO3, double O2)
{
    const double* q = s->denitrification_params;
    double lC = log(Cr * SMG_DAY), lN = log(NO3), lk = log(k * SMG_DAY);
    double ln = (q[0] + q[1] * lC + q[2] * (lN * lN) + q[3] * (lC * lC) + q[4] * (lk * lk) + q[5] * log(O2) * log(k));
    double p = exp(ln) / (Cr * SMG_DAY) * O2 / (s->anoxia_half_saturation + O2);
    return smg_finite(p);
}

static double smg_anoxic(const obme_sediment* s, double Cr, double k, double NO3, double O2)
{
    const double* q = s->anoxic_params;
    double lC = log(Cr * SMG_DAY), lN = log(NO3);
    double ln = (q[0] + q[1] * lC + q[2] * (lC * lC) + q[3] * log(k * SMG_DAY) + q[4] * log(O2) * log(k)
                 + q[5] * (lN * lN));
    double p = exp(ln) / (Cr * SMG_DAY);
    return smg_finite(p);
}

void orc_smg_coupled(const obme_sediment* s, const double* u, double NO3, double NH4, double O2,
                     double* dNH4, double* dNO3, double* dO2)
{
    const double R = s->sinking_redfield, ls = s->slow_decay_rate, lf = s->fast_decay_rate;
    const double Ns = u[0], Nf = u[1];
    const double Nr = ls * Ns + lf * Nf;
    const double Cr = Nr * R;
    /* reactivity(s, Ns R, Nf R) (simple_multi_G.jl:372-379), eps(0.0) = 5e-324 */
    const double Cs = Ns * R, Cf = Nf * R;
    const double k = (ls * Cs + lf * Cf) / (Cs + Cf + 4.9406564584124654e-324);
    const double pn = smg_ammonia_oxidation(s, Nr, Cr, k, NH4, O2);
    const double pd = smg_denitrification(s, Cr, k, NO3, O2);
    const double pa = smg_anoxic(s, Cr, k, NO3, O2);
    const double ps = 0.223 * pow(s->sedimentation_rate, 0.336);
    *dNH4 = (1 - pn) * Nr + 0.8 * pd * Cr;                       /* :195-209 */
    *dNO3 = pn * Nr - 0.8 * pd * Cr;                             /* :211-226 */
    *dO2 = -(1 - pa * ps - pd) * O2 / (s->anoxia_half_saturation + O2) * Cr - 2 * pn * Nr;   /* :228-247 */
}

/* time_step!(sediment, Δt) of update_biogeochemical_state!
 * (Sediments/update_state.jl:6-17) with its QuasiAdamsBashforth2 stepper:
 * at iteration 0 the tendencies are computed first (update_state!); the
 * step is forward Euler (χ = -1/2, G⁻ multiplied by a strong zero) when Δt
 * differs from the clock's last Δt, else (3/2 + χ) Gⁿ - (1/2 + χ) G⁻
 * (timesteppers.jl:16-43); then G⁻ = Gⁿ (cache_previous_tendencies!,
 * :81-91) and Gⁿ from the new fields and the tracked flux
 * (compute_tendencies.jl:7-31). */
void orc_sediment_time_step(const obme_sediment* s, orc_sediment_column* c, double f, double dt)
{
    if (!isfinite(dt)) return;
    int nf = orc_sediment_fields(s);
    if (c->iteration == 0) orc_sediment_tendencies(s, c->u, f, c->Gn);
    int euler = !(dt == c->last_dt);
    double chi = euler ? -0.5 : s->ab2_chi;
    for (int q = 0; q < nf; q++) {
        double Gu = (1.5 + chi) * c->Gn[q] - (euler ? 0.0 : (0.5 + chi) * c->Gm[q]);
        c->u[q] = c->u[q] + dt * Gu;
        c->Gm[q] = c->Gn[q];
    }
    orc_sediment_tendencies(s, c->u, f, c->Gn);
    c->last_dt = dt;
    c->iteration += 1;
}

/* The flux of the bottom cell after the modifiers (the scaling runs first
 * in update_biogeochemical_state!, before the sediment's tracked fields). */
static double sediment_flux_scaled(const obme_config* cfg, const double* U, int Nz, int nt)
{
    if (cfg->scaling.n_groups <= 0) return orc_sediment_flux(cfg, U, Nz);
    /* the bottom cell alone, as a one-level column */
    double* B = (double*)malloc(sizeof(double) * nt);
    for (int n = 0; n < nt; n++) B[n] = U[n * Nz + 0];
    orc_scale_negatives(&cfg->scaling, B);
    double f = orc_sediment_flux(cfg, B, 1);
    free(B);
    return f;
}

static double surface_PAR(const orc_column_env* env, double t)
{
    if (env->cfg->par.surface_mode == OBME_SURFACE_DIURNAL) return orc_default_surface_PAR(t);
    return env->PAR0_const;
}

void orc_column_tendencies(const orc_column_env* env, const double* U, double t, double* G, double* PAR)
{
    const obme_config* cfg = env->cfg;
    int Nz = env->Nz;
    int nt = cfg_ntracers(cfg);
    int Pidx = (cfg->model == OBME_NPZD) ? 1 : 2;

    /* update_biogeochemical_state!: light (src/OceanBioME.jl:161-167) */
    orc_two_band_column(&cfg->par, Nz, env->zc, env->zf, U + Pidx * Nz, surface_PAR(env, t), PAR);

    /* biogeochemical transitions cell by cell */
    double vals[OBME_MAX_TRACERS], out[OBME_MAX_TRACERS];
    for (int k = 0; k < Nz; k++) {
        for (int n = 0; n < nt; n++) vals[n] = U[n * Nz + k];
        if (cfg->model == OBME_NPZD && env->Tcol) vals[3] = env->Tcol[k];
        cfg_tendencies(cfg, vals, PAR[k], out);
        for (int n = 0; n < nt; n++) G[n * Nz + k] = out[n];
    }

    /* minus the drift-velocity flux divergence */
    double* w = (double*)malloc(sizeof(double) * (Nz + 1));
    double* div = (double*)malloc(sizeof(double) * Nz);
    for (int n = 0; n < nt; n++) {
        double sp = cfg->sinking.speed[n];
        if (sp == 0.0) continue;
        sinking_profile(&cfg->sinking, sp, Nz, w);
        sinking_divergence(cfg->sinking.scheme, Nz, w, U + n * Nz, env->dz, div);
        for (int k = 0; k < Nz; k++) G[n * Nz + k] = G[n * Nz + k] - div[k];
    }
    free(w);
    free(div);

    /* top DIC flux boundary condition: G[Nz] -= F/Δz (GasExchange) */
    if (cfg->model == OBME_LOBSTER && cfg->carbonate_system && cfg->co2.enabled) {
        int top = Nz - 1;
        double DIC = U[7 * Nz + top], Alk = U[8 * Nz + top];
        double F = orc_co2_flux(&cfg->co2, &cfg->solver, DIC, Alk, env->T_surface, env->S_surface,
                                cfg->co2.wind_speed, cfg->co2.air_concentration, env->fast_exit,
                                NULL, NULL);
        G[7 * Nz + top] = G[7 * Nz + top] - F / env->dz[top];
    }

    /* top O2 flux boundary condition of LOBSTER's O2: G[Nz] -= F/Δz */
    if (cfg->model == OBME_LOBSTER && cfg->oxygen.enabled && cfg->o2.enabled) {
        int top = Nz - 1, n = nt - 1;
        double F = orc_o2_flux(&cfg->o2, U[n * Nz + top], env->T_surface, env->S_surface);
        G[n * Nz + top] = G[n * Nz + top] - F / env->dz[top];
    }

    /* the sediment's coupled tendency (Sediments/tracer_coupling.jl:3-38):
     * the receiver's bottom cell gains (1 - e) f / Δz, Δz at the bottom face
     * taken as the bottom cell's thickness */
    if (cfg->sediment.model == OBME_SEDIMENT_INSTANT_REMINERALISATION) {
        double f = orc_sediment_flux(cfg, U, Nz);
        double e = orc_burial_efficiency(&cfg->sediment, f);
        int r = cfg->sediment.receiver;
        G[r * Nz + 0] = G[r * Nz + 0] + ((1 - e) * f) / env->dz[0];
    } else if (cfg->sediment.model == OBME_SEDIMENT_SIMPLE_MULTI_G) {
        /* from the sediment's fields now (an empty sediment without them)
         * and the bottom cell's tracked NO3, NH4, O2 */
        const obme_sediment* sd = &cfg->sediment;
        const double zero[3] = {0.0, 0.0, 0.0};
        const double* u = env->sed ? env->sed->u : zero;
        double dNH4, dNO3, dO2;
        orc_smg_coupled(sd, u, U[sd->nitrate * Nz + 0], U[sd->ammonia * Nz + 0], U[sd->oxygen * Nz + 0], &dNH4,
                        &dNO3, &dO2);
        G[sd->ammonia * Nz + 0] = G[sd->ammonia * Nz + 0] + dNH4 / env->dz[0];
        G[sd->nitrate * Nz + 0] = G[sd->nitrate * Nz + 0] + dNO3 / env->dz[0];
        G[sd->oxygen * Nz + 0] = G[sd->oxygen * Nz + 0] + dO2 / env->dz[0];
    }
}

/* RungeKutta3 (Oceananigans): γ = 8/15, 5/12, 3/4; ζ = -17/60, -5/12; the
 * substep form of BoxModel/timesteppers.jl:79-92; stage clocks t, t+8/15Δt,
 * t+2/3Δt. */
void orc_column_run(const orc_column_env* env, double* U, int nsteps, double dt, double t0,
                    double* PAR_out)
{
    int Nz = env->Nz;
    int nt = cfg_ntracers(env->cfg);
    int ncell = nt * Nz;
    double* G = (double*)malloc(sizeof(double) * ncell);
    double* Gm = (double*)malloc(sizeof(double) * ncell);
    double* PAR = (double*)malloc(sizeof(double) * Nz);
    const double g1 = 8.0 / 15.0, g2 = 5.0 / 12.0, g3 = 3.0 / 4.0;
    const double z2 = -17.0 / 60.0, z3 = -5.0 / 12.0;
    int readonly = (env->cfg->model == OBME_NPZD) ? 3 : -1;
    const obme_sediment* sd = &env->cfg->sediment;
    int sed = env->sed && orc_sediment_fields(sd) > 0;
    /* the modifiers run first in update_biogeochemical_state!
     * (src/OceanBioME.jl:161-167): scaling before every tendency evaluation;
     * after every substep the update-state hooks end with the sediment's
     * time step of the stage Δt (Sediments/update_state.jl:6-17), from the
     * flux of the substep's (scaled) state */
    for (int s = 0; s < nsteps; s++) {
        double t = t0 + s * dt;
        scale_column(env->cfg, Nz, nt, U);
        orc_column_tendencies(env, U, t, G, PAR);
        for (int i = 0; i < ncell; i++) {
            if (i / Nz == readonly) continue;
            U[i] += dt * g1 * G[i];
        }
        if (sed) orc_sediment_time_step(sd, env->sed, sediment_flux_scaled(env->cfg, U, Nz, nt), g1 * dt);
        memcpy(Gm, G, sizeof(double) * ncell);
        scale_column(env->cfg, Nz, nt, U);
        orc_column_tendencies(env, U, t + g1 * dt, G, PAR);
        for (int i = 0; i < ncell; i++) {
            if (i / Nz == readonly) continue;
            U[i] += dt * (g2 * G[i] + z2 * Gm[i]);
        }
        if (sed) orc_sediment_time_step(sd, env->sed, sediment_flux_scaled(env->cfg, U, Nz, nt), (g2 + z2) * dt);
        memcpy(Gm, G, sizeof(double) * ncell);
        scale_column(env->cfg, Nz, nt, U);
        orc_column_tendencies(env, U, t + (g1 + g2 + z2) * dt, G, PAR);
        for (int i = 0; i < ncell; i++) {
            if (i / Nz == readonly) continue;
            U[i] += dt * (g3 * G[i] + z3 * Gm[i]);
        }
        if (sed) orc_sediment_time_step(sd, env->sed, sediment_flux_scaled(env->cfg, U, Nz, nt), (g3 + z3) * dt);
    }
    if (PAR_out) memcpy(PAR_out, PAR, sizeof(double) * Nz);
    free(G);
    free(Gm);
    free(PAR);
}

/* ---------------- 3-D driver ---------------- */

static long lidx(const obme_grid* g, int i, int j, int k)
{
    long sx = g->Nx + 2 * g->Hx, sy = g->Ny + 2 * g->Hy;
    return (long)(i + g->Hx) + sx * ((long)(j + g->Hy) + sy * (long)(k + g->Hz));
}

static double field_value(const obme_field_source* f, const double* arr3, const obme_grid* g, int i,
                          int j, int k)
{
    if (arr3) return arr3[lidx(g, i, j, k)];
    if (f->mode == OBME_FIELD_PROFILE && f->profile) return f->profile[k];
    return f->value;
}

static void env_for_column(const obme_config* cfg, const double* T3, const double* S3, int i, int j,
                           double* Tcol, orc_column_env* env, int fast_exit)
{
    const obme_grid* g = &cfg->grid;
    env->cfg = cfg;
    env->Nz = g->Nz;
    env->zc = g->zc;
    env->zf = g->zf;
    env->dz = g->dz;
    env->PAR0_const = cfg->par.surface_value;
    env->fast_exit = fast_exit;
    env->T_surface = field_value(&cfg->temperature, T3, g, i, j, g->Nz - 1);
    env->S_surface = field_value(&cfg->salinity, S3, g, i, j, g->Nz - 1);
    env->Tcol = NULL;
    env->sed = NULL;
    (void)Tcol;
}

static int run_one_column(const obme_config* cfg, double* const* tracers, const double* T3,
                          const double* S3, const double* PAR0, int i, int j, int nsteps, double dt,
                          double t0, int fast_exit, orc_sediment_column* sed)
{
    const obme_grid* g = &cfg->grid;
    int Nz = g->Nz;
    int nt = cfg_ntracers(cfg);
    double* U = (double*)malloc(sizeof(double) * nt * Nz);
    double* Tcol = (double*)malloc(sizeof(double) * Nz);
    orc_column_env env;
    env_for_column(cfg, T3, S3, i, j, Tcol, &env, fast_exit);
    if (PAR0) env.PAR0_const = PAR0[i + g->Nx * j];
    env.sed = sed;
    for (int n = 0; n < nt; n++)
        for (int k = 0; k < Nz; k++) U[n * Nz + k] = tracers[n][lidx(g, i, j, k)];
    if (cfg->model == OBME_NPZD) {
        for (int k = 0; k < Nz; k++) Tcol[k] = tracers[3][lidx(g, i, j, k)];
        env.Tcol = Tcol;
    }
    orc_column_run(&env, U, nsteps, dt, t0, NULL);
    for (int n = 0; n < nt; n++) {
        if (cfg->model == OBME_NPZD && n == 3) continue;
        for (int k = 0; k < Nz; k++) tracers[n][lidx(g, i, j, k)] = U[n * Nz + k];
    }
    free(U);
    free(Tcol);
    return 0;
}

int orc_grid_run_columns(const obme_config* cfg, double* const* tracers, const double* T3,
                         const double* S3, const double* PAR0, int nsteps, double dt, double t0,
                         const int* col_i, const int* col_j, int ncols, int nthreads, int fast_exit)
{
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (int c = 0; c < ncols; c++)
        run_one_column(cfg, tracers, T3, S3, PAR0, col_i[c], col_j[c], nsteps, dt, t0, fast_exit, NULL);
    return 0;
}

int orc_grid_run_sediment(const obme_config* cfg, double* const* tracers, const double* T3,
                          const double* S3, const double* PAR0, int nsteps, double dt, double t0,
                          int nthreads, int fast_exit, double* storage, double* Gn, double* Gm,
                          double* last_dt, long* iteration)
{
    const obme_grid* g = &cfg->grid;
    long ncols = (long)g->Nx * g->Ny;
    const int nf = orc_sediment_fields(&cfg->sediment);
    int sed = nf > 0 && storage && Gn && Gm && last_dt && iteration;
    double last_dt_after = sed ? *last_dt : 0.0;
    long iteration_after = sed ? *iteration : 0;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (long c = 0; c < ncols; c++) {
        orc_sediment_column sc;
        if (sed) {
            for (int q = 0; q < nf; q++) {
                sc.u[q] = storage[q * ncols + c];
                sc.Gn[q] = Gn[q * ncols + c];
                sc.Gm[q] = Gm[q * ncols + c];
            }
            sc.last_dt = *last_dt;
            sc.iteration = *iteration;
        }
        run_one_column(cfg, tracers, T3, S3, PAR0, (int)(c % g->Nx), (int)(c / g->Nx), nsteps, dt, t0,
                       fast_exit, sed ? &sc : NULL);
        if (sed) {
            for (int q = 0; q < nf; q++) {
                storage[q * ncols + c] = sc.u[q];
                Gn[q * ncols + c] = sc.Gn[q];
                Gm[q * ncols + c] = sc.Gm[q];
            }
            if (c == 0) {
                last_dt_after = sc.last_dt;
                iteration_after = sc.iteration;
            }
        }
    }
    if (sed) {
        *last_dt = last_dt_after;
        *iteration = iteration_after;
    }
    return 0;
}

int orc_grid_run(const obme_config* cfg, double* const* tracers, const double* T3, const double* S3,
                 const double* PAR0, int nsteps, double dt, double t0, int nthreads, int fast_exit)
{
    return orc_grid_run_sediment(cfg, tracers, T3, S3, PAR0, nsteps, dt, t0, nthreads, fast_exit, NULL, NULL,
                                 NULL, NULL, NULL);
}

int orc_grid_sediment_update_state(const obme_config* cfg, const double* const* tracers,
                                   double* storage, double* Gn, double* Gm, double* last_dt,
                                   long* iteration, double dt)
{
    const obme_grid* g = &cfg->grid;
    int nt = cfg_ntracers(cfg);
    double* B = (double*)malloc(sizeof(double) * nt);
    long ncols = (long)g->Nx * g->Ny;
    orc_sediment_column sc;
    memset(&sc, 0, sizeof(sc));
    const int nf = orc_sediment_fields(&cfg->sediment);
    for (long c = 0; c < ncols; c++) {
        int i = (int)(c % g->Nx), j = (int)(c / g->Nx);
        for (int n = 0; n < nt; n++) B[n] = tracers[n][lidx(g, i, j, 0)];
        for (int q = 0; q < nf; q++) {
            sc.u[q] = storage[q * ncols + c];
            sc.Gn[q] = Gn[q * ncols + c];
            sc.Gm[q] = Gm[q * ncols + c];
        }
        sc.last_dt = *last_dt;
        sc.iteration = *iteration;
        orc_sediment_time_step(&cfg->sediment, &sc, orc_sediment_flux(cfg, B, 1), dt);
        for (int q = 0; q < nf; q++) {
            storage[q * ncols + c] = sc.u[q];
            Gn[q * ncols + c] = sc.Gn[q];
            Gm[q * ncols + c] = sc.Gm[q];
        }
    }
    if (ncols > 0) {
        *last_dt = sc.last_dt;
        *iteration = sc.iteration;
    }
    free(B);
    return 0;
}

/* The scaling modifier in place on every interior cell (the first part of
 * update_biogeochemical_state!). */
int orc_grid_scale_negatives(const obme_config* cfg, double* const* tracers)
{
    const obme_grid* g = &cfg->grid;
    int nt = cfg_ntracers(cfg);
    double vals[OBME_MAX_TRACERS];
    for (int k = 0; k < g->Nz; k++)
        for (int j = 0; j < g->Ny; j++)
            for (int i = 0; i < g->Nx; i++) {
                long q = lidx(g, i, j, k);
                for (int n = 0; n < nt; n++) vals[n] = tracers[n][q];
                orc_scale_negatives(&cfg->scaling, vals);
                for (int n = 0; n < nt; n++) tracers[n][q] = vals[n];
            }
    return 0;
}

/* update_biogeochemical_state! for the light model: PAR interior at time t. */
int orc_grid_update_state(const obme_config* cfg, const double* const* tracers, const double* PAR0,
                          double* PAR, double t)
{
    const obme_grid* g = &cfg->grid;
    int Nz = g->Nz;
    int Pidx = (cfg->model == OBME_NPZD) ? 1 : 2;
    double* P = (double*)malloc(sizeof(double) * Nz);
    double* col = (double*)malloc(sizeof(double) * Nz);
    for (int j = 0; j < g->Ny; j++)
        for (int i = 0; i < g->Nx; i++) {
            for (int k = 0; k < Nz; k++) P[k] = tracers[Pidx][lidx(g, i, j, k)];
            double s0 = PAR0 ? PAR0[i + g->Nx * j]
                             : (cfg->par.surface_mode == OBME_SURFACE_DIURNAL ? orc_default_surface_PAR(t)
                                                                              : cfg->par.surface_value);
            orc_two_band_column(&cfg->par, Nz, g->zc, g->zf, P, s0, col);
            for (int k = 0; k < Nz; k++) PAR[lidx(g, i, j, k)] = col[k];
        }
    free(P);
    free(col);
    return 0;
}

/* Tendencies at time t (update-state, transitions, drift, top flux) ADDED
 * into G (the update_tendencies! hook form). PAR (interior) is written. */
int orc_grid_tendencies(const obme_config* cfg, const double* const* tracers, double* const* G,
                        const double* T3, const double* S3, const double* PAR0, double t, double* PAR)
{
    return orc_grid_tendencies_sediment(cfg, tracers, G, T3, S3, PAR0, t, PAR, NULL);
}

int orc_grid_tendencies_sediment(const obme_config* cfg, const double* const* tracers, double* const* G,
                                 const double* T3, const double* S3, const double* PAR0, double t, double* PAR,
                                 const double* fields)
{
    const int nf = orc_sediment_fields(&cfg->sediment);
    const long ncols = (long)cfg->grid.Nx * cfg->grid.Ny;
    const obme_grid* g = &cfg->grid;
    int Nz = g->Nz;
    int nt = cfg_ntracers(cfg);
    double* U = (double*)malloc(sizeof(double) * nt * Nz);
    double* Gc = (double*)malloc(sizeof(double) * nt * Nz);
    double* Pc = (double*)malloc(sizeof(double) * Nz);
    double* Tcol = (double*)malloc(sizeof(double) * Nz);
    for (int j = 0; j < g->Ny; j++)
        for (int i = 0; i < g->Nx; i++) {
            orc_column_env env;
            env_for_column(cfg, T3, S3, i, j, Tcol, &env, 1);
            if (PAR0) env.PAR0_const = PAR0[i + g->Nx * j];
            orc_sediment_column sc;
    memset(&sc, 0, sizeof(sc));
            if (fields && nf > 0) {
                for (int q = 0; q < nf; q++) sc.u[q] = fields[q * ncols + i + (long)g->Nx * j];
                env.sed = &sc;
            }
            for (int n = 0; n < nt; n++)
                for (int k = 0; k < Nz; k++) U[n * Nz + k] = tracers[n][lidx(g, i, j, k)];
            if (cfg->model == OBME_NPZD) {
                for (int k = 0; k < Nz; k++) Tcol[k] = tracers[3][lidx(g, i, j, k)];
                env.Tcol = Tcol;
            }
            orc_column_tendencies(&env, U, t, Gc, Pc);
            for (int n = 0; n < nt; n++)
                for (int k = 0; k < Nz; k++) G[n][lidx(g, i, j, k)] += Gc[n * Nz + k];
            if (PAR)
                for (int k = 0; k < Nz; k++) PAR[lidx(g, i, j, k)] = Pc[k];
        }
    free(U);
    free(Gc);
    free(Pc);
    free(Tcol);
    return 0;
}

/* ---------------- batched NPZD ensemble ---------------- */

/* Independent NPZD columns with per-column parameters: each default
 * multiplied by a per-column factor, per-column surface PAR, uniform Δz and
 * T.  The semantics are N independent column runs of the model above. */
int orc_batch_run_sediment(const obme_config* cfg, int n_columns, const int* offsets, const double* const* params,
                           double dz, double T_const, double* const* U, int nsteps, double dt, double t0,
                           int nthreads, double* storage, double* Gn, double* Gm, double* last_dt, long* iteration)
{
    const int nf = orc_sediment_fields(&cfg->sediment);
    const int sed = nf == 1 && storage && Gn && Gm && last_dt && iteration;
    double last_dt_after = sed ? *last_dt : 0.0;
    long iteration_after = sed ? *iteration : 0;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (int c = 0; c < n_columns; c++) {
        int Nz = offsets[c + 1] - offsets[c];
        obme_config local = *cfg;
        obme_npd_params* q = &local.npd;
        const double* const* P = params;
        q->nitrate_half_saturation = P[OBME_BP_NITRATE_HALF_SATURATION][c];
        q->light_half_saturation = P[OBME_BP_LIGHT_HALF_SATURATION][c];
        q->phytoplankton_maximum_growth_rate = P[OBME_BP_MAX_GROWTH_RATE][c];
        q->phytoplankton_exudation_fraction = P[OBME_BP_EXUDATION_FRACTION][c];
        q->ammonia_fraction_of_exudate = P[OBME_BP_AMMONIA_FRACTION_OF_EXUDATE][c];
        q->temperature_coefficient = P[OBME_BP_TEMPERATURE_COEFFICIENT][c];
        q->phytoplankton_mortality_rate = P[OBME_BP_PHYTO_MORTALITY_RATE][c];
        q->zooplankton_mortality_rate = P[OBME_BP_ZOO_MORTALITY_RATE][c];
        q->zooplankton_excretion_rate = P[OBME_BP_ZOO_EXCRETION_RATE][c];
        q->phytoplankton_solid_waste_fraction = P[OBME_BP_PHYTO_SOLID_WASTE_FRACTION][c];
        q->excretion_inorganic_fraction = P[OBME_BP_EXCRETION_INORGANIC_FRACTION][c];
        q->preference_for_phytoplankton = P[OBME_BP_PREFERENCE_FOR_PHYTO][c];
        q->maximum_grazing_rate = P[OBME_BP_MAX_GRAZING_RATE][c];
        q->grazing_half_saturation = P[OBME_BP_GRAZING_HALF_SATURATION][c];
        q->zooplankton_assimilation_fraction = P[OBME_BP_ZOO_ASSIMILATION_FRACTION][c];
        q->remineralisation_rate = P[OBME_BP_REMINERALISATION_RATE][c];
        q->small_particle_fraction = P[OBME_BP_SMALL_PARTICLE_FRACTION][c];
        memset(local.sinking.speed, 0, sizeof(local.sinking.speed));
        local.sinking.speed[1] = P[OBME_BP_P_SINKING_SPEED][c];
        local.sinking.speed[4] = P[OBME_BP_D_SINKING_SPEED][c];
        local.par.surface_mode = OBME_SURFACE_CONSTANT;

        double* zc = (double*)malloc(sizeof(double) * Nz);
        double* zf = (double*)malloc(sizeof(double) * (Nz + 1));
        double* dzs = (double*)malloc(sizeof(double) * Nz);
        double* Tcol = (double*)malloc(sizeof(double) * Nz);
        double* col = (double*)malloc(sizeof(double) * 5 * Nz);
        for (int k = 0; k <= Nz; k++) zf[k] = -dz * (Nz - k);
        for (int k = 0; k < Nz; k++) {
            zc[k] = -dz * (Nz - k) + dz / 2;
            dzs[k] = dz;
            Tcol[k] = T_const;
        }
        orc_column_env env;
        env.cfg = &local;
        env.Nz = Nz;
        env.zc = zc;
        env.zf = zf;
        env.dz = dzs;
        env.PAR0_const = P[OBME_BP_SURFACE_PAR][c];
        env.T_surface = T_const;
        env.S_surface = 35.0;
        env.Tcol = Tcol;
        env.fast_exit = 1;
        /* the column's sediment (storage, Gⁿ, G⁻ and the shared clock) */
        orc_sediment_column sc;
        memset(&sc, 0, sizeof(sc));
        if (sed) {
            sc.u[0] = storage[c];
            sc.Gn[0] = Gn[c];
            sc.Gm[0] = Gm[c];
            sc.last_dt = *last_dt;
            sc.iteration = *iteration;
        }
        env.sed = sed ? &sc : NULL;
        /* tracer order N, P, Z, T, D */
        const int map[5] = {0, 1, 2, -1, 3};
        for (int n = 0; n < 5; n++)
            for (int k = 0; k < Nz; k++)
                col[n * Nz + k] = map[n] < 0 ? T_const : U[map[n]][offsets[c] + k];
        orc_column_run(&env, col, nsteps, dt, t0, NULL);
        for (int n = 0; n < 5; n++)
            if (map[n] >= 0)
                for (int k = 0; k < Nz; k++) U[map[n]][offsets[c] + k] = col[n * Nz + k];
        if (sed) {
            storage[c] = sc.u[0];
            Gn[c] = sc.Gn[0];
            Gm[c] = sc.Gm[0];
            if (c == 0) {
                last_dt_after = sc.last_dt;
                iteration_after = sc.iteration;
            }
        }
        free(zc);
        free(zf);
        free(dzs);
        free(Tcol);
        free(col);
    }
    if (sed) {
        *last_dt = last_dt_after;
        *iteration = iteration_after;
    }
    return 0;
}

int orc_batch_run(const obme_config* cfg, int n_columns, const int* offsets, const double* const* params,
                  double dz, double T_const, double* const* U, int nsteps, double dt, double t0,
                  int nthreads)
{
    return orc_batch_run_sediment(cfg, n_columns, offsets, params, dz, T_const, U, nsteps, dt, t0, nthreads,
                                  NULL, NULL, NULL, NULL, NULL);
}
