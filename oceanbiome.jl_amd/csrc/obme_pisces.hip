// obme_pisces.hip — PISCES kernels and their C-ABI entry points
// (obme_pisces_*, include/obme.h).
//
// obme_pisces_update_state / obme_pisces_add_tendencies launch
//   k_pisces_columns  MultiBand PAR, euphotic depth, mixed-layer mean PAR;
//   k_pisces_omega    calcite saturation per cell (FP64);
//   k_pisces_cell     transitions + sinking (tendency output).
// obme_pisces_step holds T and S fixed, so it computes the T-, S- and
// pressure-dependent carbonate constants once per call (k_pisces_prep) and
// each RK3 stage is two launches: k_pisces_columns, then k_pisces_cell with
// the calcite saturation solved inline from those constants, and the
// substep.
// The stage input is read and the stage output written to different
// arrays (the caller's tracers and a library-owned copy, alternately), so
// the sinking stencil's k+1 reads never race with the update.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "obme_internal.h"
#include "obme_pisces.h"

using namespace obme;
using namespace obme::pisces;

namespace {

constexpr int NW = OBME_PISCES_WRITTEN;  // 24 written tracers
constexpr int NTR = OBME_PI_COUNT;       // + T, S

template <typename FT>
struct PiscesArgs {
    const FT* in[NTR];      // stage input (parent arrays)
    FT* out[NW];            // stage output (parent arrays) or tendencies (TILE mode)
    FT* Gm;                 // previous-stage transitions [NW][ncell] (compact)
    FT* PAR;                // [4][ncell] compact: total, band 1, 2, 3
    FT* Omega;              // [ncell] compact
    const FT* oc;           // [OMEGA_CONST_FIELDS][ncell]: per-cell carbonate constants (the
                            // element type: FP32-mixed Ω solves in FP64 on FP32-rounded constants)
    FT* tfac;               // [TFAC_FIELDS][ncell]: temperature factors of the FP64 step, or null
    float* Hs;              // [ncell]: [H+] of the last stage (FP32-mixed warm start)
    int have_H;             // Hs holds the previous stage's [H+]
    double* sflux;          // [2][Nx*Ny]: top CO2 and O2 fluxes of the stage (gas exchange on)
    double* Hsurf;          // [Nx*Ny]: [H+] of the surface pCO2 solve (FP32-mixed warm start)
    obme_co2_exchange co2;
    obme_o2_exchange o2;
    FT* zeu;                // [Nx*Ny]
    FT* mlPAR;              // [Nx*Ny]
    const FT* pp;           // parameters (OBME_PP_*)
    const FT* lev;          // [4][Nz]: zc, zf (bottom face), dz, 1/dz
    const FT* w;            // [2][Nz+1]: POC and GOC face velocities
    double zf_top;          // top face position
    int Nx, Ny, Nz;
    unsigned sx, sxy, base0;
    FT surface_PAR;         // PAR⁰ at this call's clock (constant and diurnal modes)
    const FT* surf;         // OBME_SURFACE_ARRAY: PAR⁰[Nx*Ny], else null
    FT kw[3], e[3], chi[3], division[3];
    StageConst<FT> sk;
    FT A, B;                // RK3: Δt γ and Δt ζ of the stage (FP32 path)
    double dt;
    int stage;
    obme_carbon_solver solver;
    // InstantRemineralisation sediment (sed_on): the flux through the bottom
    // face f = Σ sedw[n] c[n] of the bottom cell over the sinking tracers
    // (sedw = -w of the tracer's bottom face), burial e = a + b (f/(k + f))^2,
    // the receiver's coupled tendency (1 - e) f / Δz; in the step mode the
    // column's storage (when given) takes the sediment step that follows
    // this stage's substep, with the tendency e f of the stage's input
    int sed_on, sed_recv, sed_init, sed_ab2;
    FT sedw[NW];
    FT sed_a, sed_b, sed_k;
    double sed_dt[3], sed_cn, sed_cm;
    double* storage;        // [Nx*Ny] or null
    double* sedGm;          // [Nx*Ny] G⁻: the tendency of the sediment's last step
    double* sedGn;          // [Nx*Ny] Gⁿ: e f after the sediment's last step
};

template <typename FT>
__device__ __forceinline__ unsigned pidx(const PiscesArgs<FT>& a, int i, int j, int k)
{
    return a.base0 + (unsigned)i + a.sx * (unsigned)j + a.sxy * (unsigned)k;
}

// Thread per column: MultiBand PAR (multi_band.jl:147-165) top-down, the
// euphotic depth (compute_euphotic_depth.jl:3-33, surface value = the
// boundary value of PAR at z = 0) and the mixed-layer mean of the total PAR
// (mean_mixed_layer_properties.jl:25-47) in the same sweep.
template <typename FT>
__global__ void __launch_bounds__(256) k_pisces_columns(PiscesArgs<FT> a)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= a.Nx * a.Ny) return;
    const int i = col % a.Nx, j = col / a.Nx;
    const int Nz = a.Nz;
    const long ncell = (long)a.Nx * a.Ny * Nz;
    const FT* zc = a.lev;
    const FT* zf = a.lev + Nz;
    const FT* sc = a.pp + OBME_PP_SCALARS;
    const FT zmxl = sc[OBME_PS_MIXED_LAYER_DEPTH];
    const FT PAR0 = a.surf ? a.surf[col] : a.surface_PAR;
    const FT thr = PAR0 * FT(1e-3);
    const FT lthr = m_log(thr);
    FT band[3] = {FT(0), FT(0), FT(0)};
    FT acc = 0, depth = 0;
    FT PARup = 0;
    // the euphotic crossing: the first level below the top with PAR <= thr,
    // its PAR and the PAR of the level above (the logarithms after the walk)
    int kc = -1;
    FT Pc = FT(0), Pu = FT(0);
    // Chl^e of the three bands from one logarithm (FP64: exp(e ln Chl),
    // within a few ulp of pow; pow for Chl <= 0 or a zero exponent)
    const bool pw = !(a.e[0] != FT(0) && a.e[1] != FT(0) && a.e[2] != FT(0));
    const long plane = (long)a.Nx * a.Ny;
    // One thread per column, one wave per SIMD: the walk goes in chunks of
    // four levels, whose loads and logarithms and exponentials are
    // independent of each other; only the bands' running products, the
    // stores and the sums follow level by level.
    constexpr int CH = 4;
    for (int k0 = Nz - 1; k0 >= 0; k0 -= CH) {
        FT E[CH][3];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int k = k0 - u;
            const int kk = k >= 0 ? k : 0;
            const unsigned p = pidx(a, i, j, kk);
            const FT Chl = a.in[OBME_PI_PCHL][p] + a.in[OBME_PI_DCHL][p];
            const bool via_log = Chl > FT(0) && !pw;
            const FT lChl = via_log ? m_log(Chl) : FT(0);
            // the exponent increment from the level above (the top level:
            // from the surface)
            const FT dzl = (kk == Nz - 1) ? zc[kk] : zc[kk] - zc[kk + 1 < Nz ? kk + 1 : kk];
#pragma unroll
            for (int n = 0; n < 3; n++) {
                const FT ce = via_log ? m_exp(a.e[n] * lChl) : m_pow(Chl, a.e[n]);
                const FT att = a.kw[n] + a.chi[n] * ce;
                E[u][n] = m_exp(dzl * att);
            }
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int k = k0 - u;
            if (k < 0) break;
#pragma unroll
            for (int n = 0; n < 3; n++)
                band[n] = (k == Nz - 1) ? PAR0 * a.division[n] * E[u][n] : band[n] * E[u][n];
            const FT PAR = band[0] + band[1] + band[2];
            const long c = (long)col + plane * k;
            a.PAR[c] = PAR;
            a.PAR[ncell + c] = band[0];
            a.PAR[2 * ncell + c] = band[1];
            a.PAR[3 * ncell + c] = band[2];
            const bool cross = kc < 0 && k < Nz - 1 && PAR <= thr;
            kc = cross ? k : kc;
            Pc = cross ? PAR : Pc;
            Pu = cross ? PARup : Pu;
            PARup = PAR;
            // mixed-layer integral, top cell first
            const FT zk = zf[k];
            const FT zk1 = (k + 1 < Nz) ? zf[k + 1] : FT(a.zf_top);
            const FT dzk = zk1 - zk;
            const FT dzk1 = (zk1 > zmxl) ? zk1 - zmxl : FT(0);
            const FT dz = (zk >= zmxl) ? dzk : dzk1;
            acc = acc + PAR * dz;
            depth = depth + dz;
        }
    }
    // the euphotic depth, interpolated in log PAR with the level above
    FT zeu = FT(-INFINITY);
    if (kc >= 0) {
        const FT lPAR = m_log(Pc), lPARup = m_log(Pu);
        zeu = zc[kc] + (lthr - lPAR) * (zc[kc] - zc[kc + 1]) / (lPAR - lPARup);
    }
    if (!(zeu > FT(-INFINITY) && zeu < FT(INFINITY))) zeu = zc[0] - (Nz > 1 ? zc[1] - zc[0] : FT(1));
    a.zeu[col] = zeu;
    a.mlPAR[col] = acc / depth;
}

// Thread per cell: calcite saturation at the hydrostatic pressure
// |z| g 1026 / 1e5 bar (compute_calcite_saturation.jl:22-36), FP64.
template <typename FT>
__global__ void __launch_bounds__(256) k_pisces_omega(PiscesArgs<FT> a)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ncol = (long)a.Nx * a.Ny;
    if (c >= ncol * a.Nz) return;
    const int k = (int)(c / ncol);
    const int col = (int)(c % ncol);
    const int i = col % a.Nx, j = col / a.Nx;
    const unsigned p = pidx(a, i, j, k);
    const double z = (double)a.lev[k];
    const double g = (double)a.pp[OBME_PP_SCALARS + OBME_PS_GRAVITY];
    const double P = fabs(z) * g * 1026 / 100000;
    int ev = 0;
    a.Omega[c] = (FT)calcite_saturation((double)a.in[OBME_PI_DIC][p], (double)a.in[OBME_PI_ALK][p],
                                        (double)a.in[OBME_PI_T][p], (double)a.in[OBME_PI_S][p],
                                        (double)a.in[OBME_PI_SI][p], P, a.solver, ev);
}

// Thread per cell: ScaleNegativeTracers in place on the stage input (the
// modifier runs first in update_biogeochemical_state!, src/OceanBioME.jl:
// 161-167; src/Utils/negative_tracers.jl:198-251), the groups of
// PISCES/coupling_utils.jl:11-35.
template <typename FT>
__global__ void __launch_bounds__(256) k_pisces_scale(PiscesArgs<FT> a, NegScale<FT, 32> ns)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ncol = (long)a.Nx * a.Ny;
    if (c >= ncol * a.Nz) return;
    const int k = (int)(c / ncol);
    const int col = (int)(c % ncol);
    const unsigned p = pidx(a, col % a.Nx, col / a.Nx, k);
    unsigned members = 0;
    for (int g = 0; g < ns.n_groups; g++) members |= ns.mask[g];
    FT x[NW];
#pragma unroll
    for (int n = 0; n < NW; n++) x[n] = (members & (1u << n)) ? a.in[n][p] : FT(0);
    scale_negatives_cell<NW>(ns, x);
#pragma unroll
    for (int n = 0; n < NW; n++)
        if (members & (1u << n)) const_cast<FT*>(a.in[n])[p] = x[n];
}

// Thread per column: the top-cell CO2 and O2 gas exchange fluxes of the
// stage input (GasExchange, gas_exchange.jl:26-38: F = k (c_water - c_air),
// positive out of the ocean), for the cell kernel to take out of the top
// cell's tendencies.  CO2: CarbonDioxideGasExchangeBoundaryCondition with no
// silicate or phosphate (validation/PISCES/column.jl:63), the surface solve
// of the LOBSTER path; O2: OxygenGasExchangeBoundaryCondition (:64).
template <typename FT>
__global__ void __launch_bounds__(256) k_pisces_surface(PiscesArgs<FT> a)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = a.Nx * a.Ny;
    if (col >= ncol) return;
    const unsigned p = pidx(a, col % a.Nx, col / a.Nx, a.Nz - 1);
    const double T = (double)a.in[OBME_PI_T][p], S = (double)a.in[OBME_PI_S][p];
    double Fc = 0.0, Fo = 0.0;
    if (a.co2.enabled) {
        SurfaceCarbon sc;
        surface_carbon_setup(a.co2, T, S, a.co2.wind_speed, sc);
        double H = (sizeof(FT) == 4) ? a.Hsurf[col] : 0.0;
        double pH = 0.0;
        int ev = 0;
        Fc = surface_flux_warm(sc, a.solver, (double)a.in[OBME_PI_DIC][p], (double)a.in[OBME_PI_ALK][p],
                               a.co2.air_concentration, sizeof(FT) == 4 ? 1 : 0, H, pH, ev,
                               sizeof(FT) == 4 ? FP32_NEWTON_TOL : 1e-14);
        if (sizeof(FT) == 4) a.Hsurf[col] = H;
    }
    if (a.o2.enabled) Fo = o2_flux(a.o2, (double)a.in[OBME_PI_O2][p], T, S);
    a.sflux[col] = Fc;
    a.sflux[ncol + col] = Fo;
}

// Thread per cell: the T-, S- and pressure-dependent constants of the
// calcite saturation (the first half of calcite_saturation).
template <typename FT>
__global__ void __launch_bounds__(256) k_pisces_prep(PiscesArgs<FT> a)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ncol = (long)a.Nx * a.Ny;
    const long ncell = ncol * a.Nz;
    if (c >= ncell) return;
    const int k = (int)(c / ncol);
    const int col = (int)(c % ncol);
    const int i = col % a.Nx, j = col / a.Nx;
    const unsigned p = pidx(a, i, j, k);
    const double z = (double)a.lev[k];
    const double g = (double)a.pp[OBME_PP_SCALARS + OBME_PS_GRAVITY];
    const double P = fabs(z) * g * 1026 / 100000;
    OmegaConst o;
    omega_constants((double)a.in[OBME_PI_T][p], (double)a.in[OBME_PI_S][p], P, o);
    FT* out = const_cast<FT*>(a.oc);
    const double f[OMEGA_CONST_FIELDS] = {o.conv, o.K1, o.K2, o.KB, o.KW, o.KS, o.KF, o.KSi,
                                          o.boron, o.sulfate, o.fluoride, o.ca, o.KSP};
#pragma unroll
    for (int n = 0; n < OMEGA_CONST_FIELDS; n++) out[(long)n * ncell + c] = (FT)f[n];
    if (a.tfac) {
        // the temperature factors of the transitions, with T fixed
        TFac<FT> t;
        temperature_factors(a.pp, a.in[OBME_PI_T][p], t);
        const FT tv[TFAC_FIELDS] = {t.btN, t.btD, t.bZ, t.bM, t.doc, t.pom, t.rain, t.kfe, t.eq, t.t4};
#pragma unroll
        for (int n = 0; n < TFAC_FIELDS; n++) a.tfac[(long)n * ncell + c] = tv[n];
    }
}

// Thread per cell: transitions, minus the upwind sinking divergence of
// POC, SFe (wPOC) and GOC, BFe, PSi, CaCO3 (wGOC), then the RK3 substep
// (STEP) or the addition into the tendency arrays (TENDENCY).  OM = 1:
// the calcite saturation is solved here from the k_pisces_prep constants
// (FP64: the reference damped Newton from 10^-pH₀, as k_pisces_omega;
// FP32-mixed: FP64 Newton warm-started from the previous stage's [H+], or
// from an FP32 first guess, falling back to the reference solve);
// OM = 0: it is read from the Omega array.
// Waves per SIMD the registers are held to: FP32 tendencies (the hook) 4;
// the FP32 step 2 (at 4 its Ω Newton and substep spilled 416 B per lane and
// the step took 4.46 ms at 256x256x64, at 2 none and 2.45 ms); FP64 2.
#ifndef OBME_PISCES_F32_WAVES
#define OBME_PISCES_F32_WAVES 4
#endif
template <typename FT, int MODE, int OM, bool SED = false>
__global__ void __launch_bounds__(256, (sizeof(FT) == 4 && MODE == 1) ? OBME_PISCES_F32_WAVES : 2)
k_pisces_cell(PiscesArgs<FT> a)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ncol = (long)a.Nx * a.Ny;
    const long ncell = ncol * a.Nz;
    if (c >= ncell) return;
    const int k = (int)(c / ncol);
    const int col = (int)(c % ncol);
    const int i = col % a.Nx, j = col / a.Nx;
    const int Nz = a.Nz;
    const unsigned p = pidx(a, i, j, k);
    FT v[NTR];
#pragma unroll
    for (int n = 0; n < NTR; n++) v[n] = a.in[n][p];
    // the sediment under the column (SED; Sediments/tracked_fields.jl:53-70,
    // instant_remineralisation.jl:101-123): the bottom cell's flux f of the
    // listed tracers only and the burial e, formed before the tendencies
    FT sed_f = FT(0), sed_e = FT(0);
    const bool sed_cell = SED && k == 0;
    if (sed_cell) {
        bool first = true;
#pragma unroll
        for (int n = 0; n < NW; n++)
            if (a.sedw[n] != FT(0)) {
                sed_f = first ? a.sedw[n] * v[n] : sed_f + a.sedw[n] * v[n];
                first = false;
            }
        const FT qf = sed_f / (a.sed_k + sed_f);
        sed_e = a.sed_a + a.sed_b * (qf * qf);
    }
    const FT* sc = a.pp + OBME_PP_SCALARS;
    const FT* wP = a.w;
    const FT* wG = a.w + (Nz + 1);
    Aux<FT> A;
    A.z = a.lev[k];
    A.zmxl = sc[OBME_PS_MIXED_LAYER_DEPTH];
    A.zeu = a.zeu[col];
    A.kappa = sc[OBME_PS_KAPPA];
    A.mlPAR = a.mlPAR[col];
    if constexpr (OM == 1) {
        OmegaConst o;
        const FT* oc = a.oc + c;
        o.conv = oc[0];
        o.K1 = oc[ncell];
        o.K2 = oc[2 * ncell];
        o.KB = oc[3 * ncell];
        o.KW = oc[4 * ncell];
        o.KS = oc[5 * ncell];
        o.KF = oc[6 * ncell];
        o.KSi = oc[7 * ncell];
        o.boron = oc[8 * ncell];
        o.sulfate = oc[9 * ncell];
        o.fluoride = oc[10 * ncell];
        o.ca = oc[11 * ncell];
        o.KSP = oc[12 * ncell];
        const CarbonParams cp = omega_params(o, (double)v[OBME_PI_DIC], (double)v[OBME_PI_ALK], (double)v[OBME_PI_SI]);
        // Newton from the previous stage's [H+] (stored in FP32; the first
        // stage of a call from the FP32 guess), with the reference solve
        // for a cell whose refinement fails.  FP32-mixed stops after a step
        // below 1e-4 relative: quadratic convergence leaves H within ~1e-8 of
        // the root, below the FP32 rounding of Ω.  FP64 stops after a step
        // below 1e-14 relative: H at the root to rounding, where the
        // reference solve from pH 8 also ends, after many more iterations
        // (and a wave waits for its slowest lane)
        int ev = 0;
        double H = a.have_H ? (double)a.Hs[c] : (double)guess_H_fp32(cp);
        if (!refine_H<true>(cp, H, ev, sizeof(FT) == 8 ? FP64_NEWTON_TOL : FP32_NEWTON_TOL))
            H = solve_for_H<true>(cp, a.solver, pow(10.0, -a.solver.initial_pH_guess), ev);
        a.Hs[c] = (float)H;
        A.Omega = (FT)omega_from_H(o, cp, H);
    } else {
        A.Omega = a.Omega[c];
    }
    A.PAR = a.PAR[c];
    A.PAR1 = a.PAR[ncell + c];
    A.PAR2 = a.PAR[2 * ncell + c];
    A.PAR3 = a.PAR[3 * ncell + c];
    A.wPOCc = (wP[k] + wP[k + 1]) / 2;
    A.wGOCc = (wG[k] + wG[k + 1]) / 2;
    FT G[NW];
    // the FP64 step reads the temperature factors k_pisces_prep stored; the
    // other modes evaluate them here
    TFac<FT> tf;
    if constexpr (sizeof(FT) == 8 && MODE == 0) {
        const FT* t = a.tfac + c;
        tf.btN = t[0];
        tf.btD = t[ncell];
        tf.bZ = t[2 * ncell];
        tf.bM = t[3 * ncell];
        tf.doc = t[4 * ncell];
        tf.pom = t[5 * ncell];
        tf.rain = t[6 * ncell];
        tf.kfe = t[7 * ncell];
        tf.eq = t[8 * ncell];
        tf.t4 = t[9 * ncell];
    } else {
        temperature_factors(a.pp, v[OBME_PI_T], tf);
    }
    pisces_tendencies(a.pp, v, A, a.sk, tf, G);
    // flux-form upwind sinking: (w(k+1) c(k+1) - w(k) c(k)) / dz, w = 0 on the top face
    const FT idz = a.lev[3 * Nz + k];
    const unsigned pu = (k + 1 < Nz) ? pidx(a, i, j, k + 1) : p;
    auto sink = [&](int n, const FT* w) {
        const FT cu = (k + 1 < Nz) ? a.in[n][pu] : FT(0);
        G[n] = G[n] - (w[k + 1] * cu - w[k] * v[n]) * idz;
    };
    sink(OBME_PI_POC, wP);
    sink(OBME_PI_SFE, wP);
    sink(OBME_PI_GOC, wG);
    sink(OBME_PI_BFE, wG);
    sink(OBME_PI_PSI, wG);
    sink(OBME_PI_CACO3, wG);
    if (k == Nz - 1 && a.sflux) {
        // the top boundary fluxes of DIC and O2
        G[OBME_PI_DIC] = G[OBME_PI_DIC] - (FT)a.sflux[col] * idz;
        G[OBME_PI_O2] = G[OBME_PI_O2] - (FT)a.sflux[ncol + col] * idz;
    }
    if (sed_cell) {
        // the receiver's coupled tendency (tracer_coupling.jl:30-38) and, in
        // the step mode, the storage's step that follows this stage's substep
        const FT f = sed_f, e = sed_e;
        const FT cpl = ((FT(1) - e) * f) / a.lev[2 * Nz];
#pragma unroll
        for (int n = 0; n < NW; n++)
            if (n == a.sed_recv) G[n] = G[n] + cpl;
        if (MODE == 0 && a.storage) {
            const double ge = (double)(e * f);
            double su = a.storage[col];
            if (a.stage == 0) {
                if (!a.sed_init) su += a.sed_dt[0] * (a.sed_ab2 ? a.sed_cn * ge - a.sed_cm * a.sedGm[col] : ge);
            } else {
                if (a.stage == 1 && a.sed_init) su += a.sed_dt[0] * ge;
                su += a.sed_dt[a.stage] * ge;
            }
            a.storage[col] = su;
            a.sedGm[col] = ge;
        }
    }
    if constexpr (MODE == 1) {
#pragma unroll
        for (int n = 0; n < NW; n++) a.out[n][p] = a.out[n][p] + G[n];
    } else {
        const double g1 = 8.0 / 15.0, g2 = 5.0 / 12.0, g3 = 3.0 / 4.0, z2 = -17.0 / 60.0, z3 = -5.0 / 12.0;
        FT fnew = FT(0);
        bool fnew_first = true;
#pragma unroll
        for (int n = 0; n < NW; n++) {
            FT u = v[n];
            const FT gm = (a.stage == 0) ? FT(0) : a.Gm[(long)n * ncell + c];
            if constexpr (sizeof(FT) == 8) {
                // the substep form of BoxModel/timesteppers.jl:79-92
                if (a.stage == 0) u += (a.dt * g1) * G[n];
                else if (a.stage == 1) u += a.dt * (g2 * G[n] + z2 * gm);
                else u += a.dt * (g3 * G[n] + z3 * gm);
            } else {
                u = u + fmaf(a.A, G[n], a.B * gm);
            }
            a.out[n][p] = u;
            a.Gm[(long)n * ncell + c] = G[n];
            if (SED && sed_cell && a.stage == 2 && a.sedw[n] != FT(0)) {
                fnew = fnew_first ? a.sedw[n] * u : fnew + a.sedw[n] * u;
                fnew_first = false;
            }
        }
        if (SED && sed_cell && a.stage == 2 && a.storage) {
            // Gⁿ = e f of the step's final state, for a sediment step by the hook
            const FT qf = fnew / (a.sed_k + fnew);
            a.sedGn[col] = (double)((a.sed_a + a.sed_b * (qf * qf)) * fnew);
        }
    }
}

// One time step of the sediment's storage under PISCES (Sediments/
// update_state.jl:6-17; QuasiAdamsBashforth2, timesteppers.jl:16-43): the
// tendency Gⁿ (at the first step, e f of the tracked flux), a forward Euler
// step or the χ-weighted step with G⁻, then G⁻ = Gⁿ and Gⁿ = e f.  One
// thread per column.
template <typename FT>
__global__ void __launch_bounds__(256) k_pisces_sediment_state(PiscesArgs<FT> a, double* storage, double dt,
                                                               int euler, int init)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= a.Nx * a.Ny) return;
    const unsigned p = pidx(a, col % a.Nx, col / a.Nx, 0);
    FT f = FT(0);
    bool first = true;
#pragma unroll
    for (int n = 0; n < NW; n++)
        if (a.sedw[n] != FT(0)) {
            f = first ? a.sedw[n] * a.in[n][p] : f + a.sedw[n] * a.in[n][p];
            first = false;
        }
    const FT qf = f / (a.sed_k + f);
    const double ge = (double)((a.sed_a + a.sed_b * (qf * qf)) * f);
    const double Gn = init ? ge : a.sedGn[col];
    const double Gu = euler ? Gn : a.sed_cn * Gn - a.sed_cm * a.sedGm[col];
    if (storage) storage[col] = storage[col] + dt * Gu;
    a.sedGm[col] = Gn;
    a.sedGn[col] = ge;
}

// Thread per column: the surface constants of the CO2 and O2 exchange
// (T and S fixed within a stepper call), SURF_PACK doubles per column.
template <typename FT>
__global__ void __launch_bounds__(256) k_pisces_surface_prep(PiscesArgs<FT> a, double* out)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = a.Nx * a.Ny;
    if (col >= ncol) return;
    const unsigned p = pidx(a, col % a.Nx, col / a.Nx, a.Nz - 1);
    const double T = (double)a.in[OBME_PI_T][p], S = (double)a.in[OBME_PI_S][p];
    SurfaceCarbon sc;
    std::memset(&sc, 0, sizeof(sc));
    if (a.co2.enabled) surface_carbon_setup(a.co2, T, S, a.co2.wind_speed, sc);
    double kO2 = 0.0, airO2 = 0.0;
    if (a.o2.enabled) o2_exchange_coefficients(a.o2, T, S, kO2, airO2);
    double o[SURF_PACK];
    surface_pack(sc, kO2, airO2, o);
#pragma unroll
    for (int q = 0; q < SURF_PACK; q++) out[(long)q * ncol + col] = o[q];
}

// Per-call constants of the fused step.
// Steps per launch of the fused step at most: columns are independent (T
// and S frozen, sinking within the column), so a wave carries its column
// through several steps with the state in registers.
constexpr int PF_MAX_STEPS = 16;

struct PiscesFused {
    NegScale<float, 32> ns;         // ScaleNegativeTracers groups (SCALE)
    float A[3], B[3];               // Δt γ, Δt ζ per stage
    float div[3];                   // the share of each band
    int have_H;                     // Hs holds the [H+] of the previous step
    int nsteps;                     // steps of this launch (1 .. PF_MAX_STEPS)
    const double* surf;             // [SURF_PACK][Nx*Ny] from k_pisces_surface_prep
    StageConst<float> sk[PF_MAX_STEPS][3];   // day lengths at the stage clocks of each step
    float PAR0[PF_MAX_STEPS][3];             // surface PAR at the stage clocks (constant, diurnal)
};

// The fused FP32-mixed PISCES step: one wave per water column, lane r
// holding level k = Nz - 1 - r (r = 0 at the top; Nz <= 64, lanes r >= Nz
// idle), the three RK3 stages in registers, one read and one write of each
// tracer per step, in place.  Per stage, as the per-stage kernels do:
// the scaling modifier (SCALE); the MultiBand PAR as an inclusive prefix
// sum from the top of the per-level log attenuation (multi_band.jl:147-165);
// the euphotic depth (the first level from the top where PAR falls to 1e-3
// of the surface value, by ballot; compute_euphotic_depth.jl:3-33) and the
// mixed-layer mean PAR (a wave sum; mean_mixed_layer_properties.jl:25-47);
// Ω from the prepared constants (FP64 Newton warm-started from the
// previous stage's [H+]); the top fluxes of CO2 and O2 by the top lane; the
// 24 transitions; the upwind sinking with the value of the level above
// from the neighbouring lane; the substep.  The workgroup's PF_WAVES
// columns are neighbours along i, in the XCD-aware tile order.
#ifndef OBME_PISCES_PF_WAVES
#define OBME_PISCES_PF_WAVES 4
#endif
constexpr int PF_WAVES = OBME_PISCES_PF_WAVES;
// waves per SIMD the register allocation is held to
#ifndef OBME_PISCES_PF_OCC
#define OBME_PISCES_PF_OCC 2
#endif

// the carbonate constants the fused step stages in (the indices of
// conv, K1, K2, KB, KW, KS, KF, KSi, KSP in k_pisces_prep's order); the
// salinity terms are formed from S in the kernel (16 B less to read per
// cell and step)
constexpr int FUSED_OMEGA_FIELDS = 9;
__device__ __forceinline__ constexpr int fused_omega_field(int q) { return q < 8 ? q : 12; }

template <bool SCALE>
__global__ void __launch_bounds__(PF_WAVES * 64, OBME_PISCES_PF_OCC) k_pisces_fused(PiscesArgs<float> a, PiscesFused f)
{
    using FT = float;
    const int r = threadIdx.x & 63;
    // the wave's column, as a scalar: the values derived from it (the column
    // index, its surface PAR and gas constants) stay out of the vector registers
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned nbx = ((unsigned)a.Nx + PF_WAVES - 1) / PF_WAVES;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int j = (int)(tile / nbx);
    const int ib = (int)(tile % nbx) * PF_WAVES;
    const int i = ib + wv;
    // a wave past the last column computes on zeros and stores nothing
    // (it stays for the workgroup barriers)
    const bool colok = i < a.Nx;
    const int Nz = a.Nz;
    const bool act = colok && r < Nz;
    const int k = (r < Nz) ? Nz - 1 - r : 0;
    const unsigned ncol = (unsigned)a.Nx * (unsigned)a.Ny;
    const unsigned col = (unsigned)(colok ? i : a.Nx - 1) + (unsigned)a.Nx * (unsigned)j;
    const unsigned long ncell = (unsigned long)ncol * (unsigned)Nz;
    const int tid = threadIdx.x;

    // the lane's carbonate constants (T, S and pressure fixed) and the
    // previous stage's transitions live in LDS, each lane in its own slots,
    // to leave the registers to the state and the transitions
    // the constants that depend on T and P (the salinity terms come from S)
    __shared__ float s_oc[FUSED_OMEGA_FIELDS][PF_WAVES * 64];
    __shared__ FT s_gm[NW][PF_WAVES * 64];
    __shared__ FT s_ts[2][PF_WAVES * 64];
    __shared__ float s_h[PF_WAVES * 64];
    __shared__ double s_surf[PF_WAVES][SURF_PACK];   // the top lane's surface constants
    __shared__ double s_hsurf[PF_WAVES];             // and its [H+] warm start (not held in registers)
    __shared__ float s_tf[TFAC_FIELDS][PF_WAVES * 64];   // the lane's temperature factors (T is fixed)

    // stage in: thread t takes level kt = t / PF_WAVES of column ct = t %
    // PF_WAVES (PF_WAVES neighbouring values along i per level: coalesced)
    // and puts it into the LDS slot of the lane that owns that cell
    const int kt = tid / PF_WAVES, ct = tid % PF_WAVES;
    const bool tok = kt < Nz && ib + ct < a.Nx;
    const int slot = ct * 64 + (Nz - 1 - (kt < Nz ? kt : 0));
    const unsigned pt = pidx(a, tok ? ib + ct : 0, j, tok ? kt : 0);
    // (32-bit: a field holds fewer than 2^31 cells, checked at create)
    const unsigned ct_cell = (unsigned)(tok ? ib + ct : 0) + (unsigned)a.Nx * (unsigned)j +
                             ncol * (unsigned)(tok ? kt : 0);
    {
        FT stg[NTR];
#pragma unroll
        for (int n = 0; n < NTR; n++) stg[n] = tok ? a.in[n][pt] : FT(0);
        float so[FUSED_OMEGA_FIELDS];
#pragma unroll
        for (int q = 0; q < FUSED_OMEGA_FIELDS; q++)
            so[q] = tok ? (a.oc + (unsigned long)fused_omega_field(q) * ncell)[ct_cell] : 1.0f;
        const float sh = (tok && f.have_H) ? a.Hs[ct_cell] : -1.0f;
        if (tok) {
#pragma unroll
            for (int n = 0; n < NW; n++) s_gm[n][slot] = stg[n];
            s_ts[0][slot] = stg[OBME_PI_T];
            s_ts[1][slot] = stg[OBME_PI_S];
#pragma unroll
            for (int q = 0; q < FUSED_OMEGA_FIELDS; q++) s_oc[q][slot] = so[q];
            s_h[slot] = sh;
        }
    }
    __syncthreads();
    FT v[NTR];
#pragma unroll
    for (int n = 0; n < NW; n++) v[n] = act ? s_gm[n][tid] : FT(0);
    v[OBME_PI_T] = act ? s_ts[0][tid] : FT(0);
    v[OBME_PI_S] = act ? s_ts[1][tid] : FT(0);
    double H = act ? (double)s_h[tid] : -1.0;      // < 0: the FP32 first guess
    __syncthreads();
    {
        // the temperature factors, once per step
        TFac<FT> t;
        temperature_factors(a.pp, v[OBME_PI_T], t);
        const FT tv[TFAC_FIELDS] = {t.btN, t.btD, t.bZ, t.bM, t.doc, t.pom, t.rain, t.kfe, t.eq, t.t4};
#pragma unroll
        for (int q = 0; q < TFAC_FIELDS; q++) s_tf[q][tid] = tv[q];
    }
    // the level's geometry
    const FT* zcv = a.lev;
    const FT* zfv = a.lev + Nz;
    const FT zc = zcv[k];
    const FT idz = a.lev[3 * Nz + k];
    const FT* wP = a.w;
    const FT* wG = a.w + (Nz + 1);
    const FT wPlo = wP[k], wPhi = wP[k + 1], wGlo = wG[k], wGhi = wG[k + 1];
    // the exponent increment of the light from the level above (the top
    // level: from the surface, zc of the top cell)
    const FT dzl = act ? (r == 0 ? zc : zc - zcv[k + 1 < Nz ? k + 1 : k]) : FT(0);
    const FT* sc = a.pp + OBME_PP_SCALARS;
    const FT zmxl = sc[OBME_PS_MIXED_LAYER_DEPTH];
    FT wml;
    {
        const FT zk = zfv[k];
        const FT zk1 = (k + 1 < Nz) ? zfv[k + 1] : FT(a.zf_top);
        const FT dzk = zk1 - zk;
        const FT dzk1 = (zk1 > zmxl) ? zk1 - zmxl : FT(0);
        wml = act ? ((zk >= zmxl) ? dzk : dzk1) : FT(0);
    }
    const FT depth = __builtin_bit_cast(FT, __builtin_amdgcn_readlane(__builtin_bit_cast(int, seg_prefix<64>(wml)), 63));
    // the column's surface PAR: a 2-D array, or per stage from the host
    const FT PAR0col = a.surf ? a.surf[col] : FT(0);
    FT zeu_none;
    {
        const FT z0 = zcv[0];
        zeu_none = z0 - (Nz > 1 ? zcv[1] - z0 : FT(1));
    }

    const bool top = r == 0 && colok;
    const bool gas = a.co2.enabled || a.o2.enabled;
    if (top && gas) {
#pragma unroll
        for (int q = 0; q < SURF_PACK; q++) s_surf[wv][q] = f.surf[(unsigned long)q * ncol + col];
        s_hsurf[wv] = a.co2.enabled ? a.Hsurf[col] : 0.0;
    }

#pragma unroll 1
    for (int sn = 0; sn < f.nsteps; sn++) {
        // G⁻ of a step's first stage is zero (ζ = 0 there; the zero also keeps a
        // non-finite tendency of the previous step out of the product)
#pragma unroll
        for (int n = 0; n < NW; n++) s_gm[n][tid] = FT(0);
#pragma unroll 1
        for (int st = 0; st < 3; st++) {
            const FT* pps = a.pp;
            if constexpr (SCALE) scale_negatives_cell<NW>(f.ns, v);

            // the top boundary fluxes of DIC and O2 (the surface kernel's solve),
            // first, while few values are live
            double Fc = 0.0, Fo = 0.0;
            if (top && gas) {
                SurfaceCarbon scb;
                double kO2, airO2;
                surface_unpack(s_surf[wv], scb, kO2, airO2);
                if (a.co2.enabled) {
                    double pH = 0.0, Hsurf = s_hsurf[wv];
                    int ev = 0;
                    Fc = surface_flux_warm(scb, a.solver, (double)v[OBME_PI_DIC], (double)v[OBME_PI_ALK],
                                           a.co2.air_concentration, 1, Hsurf, pH, ev, FP32_NEWTON_TOL);
                    s_hsurf[wv] = Hsurf;
                }
                if (a.o2.enabled) Fo = kO2 * ((double)v[OBME_PI_O2] - airO2);
            }

            // MultiBand PAR: band n = PAR0 share · exp(Σ from the top of Δz·att)
            const FT PAR0 = a.surf ? PAR0col : f.PAR0[sn][st];
            const FT thr = PAR0 * FT(1e-3);
            const FT lthr = m_log(thr);
            const FT Chl = v[OBME_PI_PCHL] + v[OBME_PI_DCHL];
            const FT lChl = m_log2(Chl);   // Chl^e = 2^(e log2 Chl), one logarithm for the three bands
            FT band[3];
#pragma unroll
            for (int n = 0; n < 3; n++) {
                const FT att = a.kw[n] + a.chi[n] * m_exp2(a.e[n] * lChl);
                const FT S = seg_prefix<64>(act ? dzl * att : FT(0));
                band[n] = (PAR0 * f.div[n]) * m_exp(S);
            }
            const FT PAR = band[0] + band[1] + band[2];

            // euphotic depth: the first level below the top cell, from the top,
            // with PAR <= thr, interpolated in log PAR with the level above
            const FT lPAR = m_log(PAR);
            const unsigned long long cand = __ballot(act && r >= 1 && PAR <= thr);
            FT zeu = zeu_none;
            if (cand) {
                const int rs = __ffsll((long long)cand) - 1;
                const FT zk = __builtin_bit_cast(FT, __builtin_amdgcn_readlane(__builtin_bit_cast(int, zc), rs));
                const FT zu = __builtin_bit_cast(FT, __builtin_amdgcn_readlane(__builtin_bit_cast(int, zc), rs - 1));
                const FT lk = __builtin_bit_cast(FT, __builtin_amdgcn_readlane(__builtin_bit_cast(int, lPAR), rs));
                const FT lu = __builtin_bit_cast(FT, __builtin_amdgcn_readlane(__builtin_bit_cast(int, lPAR), rs - 1));
                zeu = zk + (lthr - lk) * (zk - zu) / (lk - lu);
                if (!(zeu > FT(-INFINITY) && zeu < FT(INFINITY))) zeu = zeu_none;
            }
            // mixed-layer mean of the total PAR
            const FT acc =
                __builtin_bit_cast(FT, __builtin_amdgcn_readlane(__builtin_bit_cast(int, seg_prefix<64>(PAR * wml)), 63));

            Aux<FT> A;
            A.z = zc;
            A.zmxl = zmxl;
            A.zeu = zeu;
            A.kappa = pps[OBME_PP_SCALARS + OBME_PS_KAPPA];
            A.mlPAR = acc / depth;
            A.PAR = PAR;
            A.PAR1 = band[0];
            A.PAR2 = band[1];
            A.PAR3 = band[2];
            A.wPOCc = (wPlo + wPhi) / 2;
            A.wGOCc = (wGlo + wGhi) / 2;
            A.Omega = FT(0);
            if (act) {
                OmegaConst o;
                o.conv = s_oc[0][tid];
                o.K1 = s_oc[1][tid];
                o.K2 = s_oc[2][tid];
                o.KB = s_oc[3][tid];
                o.KW = s_oc[4][tid];
                o.KS = s_oc[5][tid];
                o.KF = s_oc[6][tid];
                o.KSi = s_oc[7][tid];
                o.KSP = s_oc[8][tid];
                omega_salinity_terms((double)v[OBME_PI_S], o);
                const CarbonParams cp =
                    omega_params(o, (double)v[OBME_PI_DIC], (double)v[OBME_PI_ALK], (double)v[OBME_PI_SI]);
                int ev = 0;
                if (H < 0) H = (double)guess_H_fp32(cp);
                if (!refine_H<true>(cp, H, ev, FP32_NEWTON_TOL)) H = solve_for_H<true>(cp, a.solver, pow(10.0, -a.solver.initial_pH_guess), ev);
                A.Omega = (FT)omega_from_H(o, cp, H);
            }
            FT G[NW];
            TFac<FT> tf;
            tf.btN = s_tf[0][tid];
            tf.btD = s_tf[1][tid];
            tf.bZ = s_tf[2][tid];
            tf.bM = s_tf[3][tid];
            tf.doc = s_tf[4][tid];
            tf.pom = s_tf[5][tid];
            tf.rain = s_tf[6][tid];
            tf.kfe = s_tf[7][tid];
            tf.eq = s_tf[8][tid];
            tf.t4 = s_tf[9][tid];
            pisces_tendencies(pps, v, A, f.sk[sn][st], tf, G);
            // flux-form upwind sinking: (w(k+1) c(k+1) - w(k) c(k)) / dz, the
            // level above from lane r - 1, w = 0 on the top face
            auto sink = [&](int n, FT wlo, FT whi) {
                const FT cu = lane_shr1(v[n]);
                G[n] = G[n] - (whi * cu - wlo * v[n]) * idz;
            };
            sink(OBME_PI_POC, wPlo, wPhi);
            sink(OBME_PI_SFE, wPlo, wPhi);
            sink(OBME_PI_GOC, wGlo, wGhi);
            sink(OBME_PI_BFE, wGlo, wGhi);
            sink(OBME_PI_PSI, wGlo, wGhi);
            sink(OBME_PI_CACO3, wGlo, wGhi);
            if (top) {
                G[OBME_PI_DIC] = G[OBME_PI_DIC] - (FT)Fc * idz;
                G[OBME_PI_O2] = G[OBME_PI_O2] - (FT)Fo * idz;
            }
            const FT As = f.A[st], Bs = f.B[st];
#pragma unroll
            for (int n = 0; n < NW; n++) {
                v[n] = v[n] + fmaf(As, G[n], Bs * s_gm[n][tid]);
                s_gm[n][tid] = G[n];
            }
        }
        // between steps [H+] is rounded as the FP32 warm-start store rounds it:
        // a launch of several steps gives the results of one launch per step
        H = (double)(float)H;
    }

    // stage out: each lane into its own slots (G⁻ is no longer needed),
    // then the coalesced stores of the stage-in pattern
#pragma unroll
    for (int n = 0; n < NW; n++) s_gm[n][tid] = v[n];
    s_h[tid] = (float)H;
    __syncthreads();
    {
        // the stage-in indices again, from a copy of the thread index the
        // compiler cannot see through: they are not held in registers through
        // the stage loop
        int t2 = tid;
        asm volatile("" : "+v"(t2));
        const int kt2 = t2 / PF_WAVES, ct2 = t2 % PF_WAVES;
        if (kt2 < Nz && ib + ct2 < a.Nx) {
            const int slot2 = ct2 * 64 + (Nz - 1 - kt2);
            const unsigned pt2 = pidx(a, ib + ct2, j, kt2);
#pragma unroll
            for (int n = 0; n < NW; n++) a.out[n][pt2] = s_gm[n][slot2];
            a.Hs[(unsigned)(ib + ct2) + (unsigned)a.Nx * (unsigned)j + ncol * (unsigned)kt2] = s_h[slot2];
        }
    }
    if (top && a.co2.enabled) a.Hsurf[col] = s_hsurf[wv];
}

}  // namespace

struct obme_pisces {
    obme_pisces_config cfg;
    int prec;
    size_t es;
    long ncell, ncol;
    void* d_pp = nullptr;
    void* d_lev = nullptr;
    void* d_w = nullptr;
    void* d_PAR = nullptr;
    void* d_Omega = nullptr;
    void* d_zeu = nullptr;
    void* d_mlPAR = nullptr;
    void* d_Gm = nullptr;
    void* d_oc = nullptr;
    void* d_H = nullptr;
    void* d_sflux = nullptr;
    void* d_Hsurf = nullptr;
    void* d_surfc = nullptr;   // [SURF_PACK][Nx*Ny] surface constants of the fused step
    void* d_tfac = nullptr;    // [TFAC_FIELDS][ncell] temperature factors of the FP64 step
    void* d_copy[NW] = {};
    size_t field_bytes = 0;
    std::vector<double> zc, zf, dz;
    std::vector<double> w_host;   // [2][Nz+1]: the POC and GOC face velocities (FP64)
    bool scaling = false;
    bool fused = false;        // FP32-mixed with Nz <= 64: k_pisces_fused (OBME_PISCES_FUSED=0 turns it off)
    int fused_steps = PF_MAX_STEPS;   // steps per fused launch (OBME_PISCES_FUSED_STEPS sets 1 .. 16)
    // an InstantRemineralisation sediment under every column (cfg.sediment)
    bool sediment = false;
    double* d_sedGm = nullptr;         // [Nx*Ny] G⁻ of the sediment's last step
    double* d_sedGn = nullptr;         // [Nx*Ny] Gⁿ: e f after the sediment's last step
    long long sed_iter = 0;            // sediment steps taken (its clock's iteration)
    double sed_last_dt = HUGE_VAL;     // Δt of its last step (Inf before the first)
    // the current step call's sediment: storage (or null), and this step's
    // lazy first tendency and AB2 first step
    double* call_storage = nullptr;
    int call_init = 0, call_ab2 = 0;
};

extern "C" obme_status obme_pisces_destroy(obme_pisces* h)
{
    if (!h) return OBME_OK;
    for (void* p : {h->d_pp, h->d_lev, h->d_w, h->d_PAR, h->d_Omega, h->d_zeu, h->d_mlPAR, h->d_Gm, h->d_oc, h->d_H,
                    h->d_sflux, h->d_Hsurf, h->d_surfc, h->d_tfac})
        if (p) (void)hipFree(p);
    for (void* p : h->d_copy)
        if (p) (void)hipFree(p);
    if (h->d_sedGm) (void)hipFree(h->d_sedGm);
    if (h->d_sedGn) (void)hipFree(h->d_sedGn);
    delete h;
    return OBME_OK;
}

template <typename FT>
static obme_status pisces_upload(obme_pisces* h)
{
    const obme_pisces_config& c = h->cfg;
    const int Nz = c.grid.Nz;
    std::vector<FT> pp(OBME_PP_COUNT);
    for (int n = 0; n < OBME_PP_COUNT; n++) pp[n] = (FT)c.params[n];
    std::vector<FT> lev(4 * (size_t)Nz);
    for (int k = 0; k < Nz; k++) {
        lev[k] = (FT)h->zc[k];
        lev[Nz + k] = (FT)h->zf[k];
        lev[2 * Nz + k] = (FT)h->dz[k];
        lev[3 * Nz + k] = (FT)(1.0 / h->dz[k]);
    }
    // face velocities: POC constant speed (Utils/sinking_velocity_fields.jl:10-35);
    // GOC from DepthDependantSinkingSpeed evaluated once at construction with
    // zmxl = zeu = 0 (PISCES/common.jl:39-55): face k takes the centre depth of cell k
    const double* sc = c.params + OBME_PP_SCALARS;
    const double sPOC = sc[OBME_PS_POC_SPEED], wmin = sc[OBME_PS_GOC_MIN_SPEED], wmax = sc[OBME_PS_GOC_MAX_SPEED];
    h->w_host.assign(2 * ((size_t)Nz + 1), 0.0);
    for (int k = 1; k <= Nz; k++) {
        h->w_host[k - 1] = -sPOC * (c.open_bottom ? 1.0 : (1 - std::exp((1 - k) / 2.0)));
        const double z = h->zc[k - 1];
        h->w_host[(Nz + 1) + (k - 1)] = -wmin + (wmax - wmin) * std::min(0.0, z - std::min(0.0, 0.0)) / 5000;
    }
    std::vector<FT> w(h->w_host.size());
    for (size_t q = 0; q < w.size(); q++) w[q] = (FT)h->w_host[q];
    auto up = [&](void** d, const void* src, size_t bytes) -> obme_status {
        OBME_HIP_TRY(hipMalloc(d, bytes));
        OBME_HIP_TRY(hipMemcpy(*d, src, bytes, hipMemcpyHostToDevice));
        return OBME_OK;
    };
    obme_status st = up(&h->d_pp, pp.data(), pp.size() * sizeof(FT));
    if (st == OBME_OK) st = up(&h->d_lev, lev.data(), lev.size() * sizeof(FT));
    if (st == OBME_OK) st = up(&h->d_w, w.data(), w.size() * sizeof(FT));
    return st;
}

extern "C" obme_status obme_pisces_create(const obme_pisces_config* cfg, obme_pisces** out)
{
    if (!cfg || !out) return obme_fail(OBME_ERR_INVALID, "null argument");
    *out = nullptr;
    const obme_grid& g = cfg->grid;
    if (cfg->precision != OBME_FP64 && cfg->precision != OBME_FP32_MIXED)
        return obme_fail(OBME_ERR_INVALID, "unknown precision");
    if (g.Nx < 1 || g.Ny < 1 || g.Nz < 2 || g.Hx < 0 || g.Hy < 0 || g.Hz < 0)
        return obme_fail(OBME_ERR_INVALID, "invalid grid size or halo");
    if (!g.zc || !g.zf || !g.dz) return obme_fail(OBME_ERR_INVALID, "grid coordinates missing");
    if (cfg->surface_mode != OBME_SURFACE_CONSTANT && cfg->surface_mode != OBME_SURFACE_DIURNAL &&
        cfg->surface_mode != OBME_SURFACE_ARRAY)
        return obme_fail(OBME_ERR_INVALID, "unknown surface PAR mode");
    if (cfg->surface_mode == OBME_SURFACE_ARRAY && !cfg->surface_array)
        return obme_fail(OBME_ERR_INVALID, "surface PAR array missing");
    const size_t es = (cfg->precision == OBME_FP64) ? 8 : 4;
    const size_t parent = (size_t)(g.Nx + 2 * g.Hx) * (g.Ny + 2 * g.Hy) * (g.Nz + 2 * g.Hz);
    if (parent >= ((size_t)1 << 31)) return obme_fail(OBME_ERR_UNSUPPORTED, "field too large for 32-bit offsets");
    if (!(cfg->params[OBME_PP_SCALARS + OBME_PS_MIXED_LAYER_DEPTH] < 0))
        return obme_fail(OBME_ERR_INVALID, "mixed_layer_depth must be negative (below the surface)");
    {
        const obme_negative_scaling& s = cfg->scaling;
        if (s.n_groups < 0 || s.n_groups > OBME_MAX_SCALE_GROUPS)
            return obme_fail(OBME_ERR_INVALID, "bad number of scaling groups");
        for (int gi = 0; gi < s.n_groups; gi++) {
            if (s.n_members[gi] < 1 || s.n_members[gi] > OBME_MAX_SCALE_MEMBERS)
                return obme_fail(OBME_ERR_INVALID, "bad number of scaling group members");
            unsigned seen = 0;
            for (int m = 0; m < s.n_members[gi]; m++) {
                const int t = s.member[gi][m];
                if (t < 0 || t >= NW) return obme_fail(OBME_ERR_INVALID, "scaling member is not a written tracer");
                if (seen & (1u << t)) return obme_fail(OBME_ERR_INVALID, "scaling member listed twice in a group");
                seen |= 1u << t;
            }
        }
    }
    {
        const obme_sediment& sd = cfg->sediment;
        if (sd.model != OBME_SEDIMENT_NONE && sd.model != OBME_SEDIMENT_INSTANT_REMINERALISATION)
            return obme_fail(OBME_ERR_UNSUPPORTED, "PISCES takes an InstantRemineralisation sediment");
        if (sd.model == OBME_SEDIMENT_INSTANT_REMINERALISATION) {
            if (sd.receiver < 0 || sd.receiver >= NW)
                return obme_fail(OBME_ERR_INVALID, "the sediment's receiver is not a written tracer");
            if (sd.n_sinking < 1 || sd.n_sinking > OBME_MAX_SEDIMENT_TRACERS)
                return obme_fail(OBME_ERR_INVALID, "the sediment needs sinking tracers");
            for (int q = 0; q < sd.n_sinking; q++) {
                const int t = sd.sinking[q];
                if (t != OBME_PI_POC && t != OBME_PI_SFE && t != OBME_PI_GOC && t != OBME_PI_BFE &&
                    t != OBME_PI_PSI && t != OBME_PI_CACO3)
                    return obme_fail(OBME_ERR_INVALID, "a sediment sinking tracer is not one of PISCES's "
                                                       "sinking tracers (POC, SFe, GOC, BFe, PSi, CaCO3)");
                for (int q2 = 0; q2 < q; q2++)
                    if (sd.sinking[q2] == t)
                        return obme_fail(OBME_ERR_INVALID, "a sinking tracer of the sediment is listed twice");
            }
            if (sd.timestepper != OBME_SEDIMENT_AB2)
                return obme_fail(OBME_ERR_UNSUPPORTED, "only the QuasiAdamsBashforth2 sediment time stepper is built");
        }
    }
    obme_pisces* h = new obme_pisces();
    h->cfg = *cfg;
    h->prec = cfg->precision;
    h->es = es;
    h->ncol = (long)g.Nx * g.Ny;
    h->ncell = h->ncol * g.Nz;
    h->field_bytes = parent * es;
    h->zc.assign(g.zc, g.zc + g.Nz);
    h->zf.assign(g.zf, g.zf + g.Nz + 1);
    h->dz.assign(g.dz, g.dz + g.Nz);
    h->cfg.grid.zc = h->cfg.grid.zf = h->cfg.grid.dz = nullptr;
    h->scaling = cfg->scaling.n_groups > 0;
    h->sediment = cfg->sediment.model == OBME_SEDIMENT_INSTANT_REMINERALISATION;
    // the fused step: FP32-mixed, one wave per column of at most 64 levels
    // (with a sediment, the per-stage kernels)
    h->fused = h->prec == OBME_FP32_MIXED && g.Nz <= 64 && !h->sediment;
    if (const char* e = std::getenv("OBME_PISCES_FUSED")) h->fused = h->fused && std::atoi(e) != 0;
    if (const char* e = std::getenv("OBME_PISCES_FUSED_STEPS")) {
        const int n = std::atoi(e);
        if (n >= 1 && n <= PF_MAX_STEPS) h->fused_steps = n;
    }
    obme_status st = (h->prec == OBME_FP64) ? pisces_upload<double>(h) : pisces_upload<float>(h);
    auto alloc = [&](void** d, size_t bytes) {
        if (st == OBME_OK && hipMalloc(d, bytes) != hipSuccess) st = obme_fail(OBME_ERR_NOMEM, "PISCES buffers");
    };
    alloc(&h->d_PAR, 4 * h->ncell * es);
    alloc(&h->d_Omega, h->ncell * es);
    alloc(&h->d_zeu, h->ncol * es);
    alloc(&h->d_mlPAR, h->ncol * es);
    alloc(&h->d_Gm, (size_t)NW * h->ncell * es);
    if (h->sediment) {
        alloc((void**)&h->d_sedGm, (size_t)h->ncol * sizeof(double));
        alloc((void**)&h->d_sedGn, (size_t)h->ncol * sizeof(double));
        if (st == OBME_OK && (hipMemset(h->d_sedGm, 0, (size_t)h->ncol * sizeof(double)) != hipSuccess ||
                              hipMemset(h->d_sedGn, 0, (size_t)h->ncol * sizeof(double)) != hipSuccess))
            st = obme_fail(OBME_ERR_HIP, "hipMemset");
    }
    if (cfg->co2.enabled || cfg->o2.enabled) {
        alloc(&h->d_sflux, 2 * (size_t)h->ncol * sizeof(double));
        alloc(&h->d_Hsurf, (size_t)h->ncol * sizeof(double));
        if (st == OBME_OK && hipMemset(h->d_Hsurf, 0, (size_t)h->ncol * sizeof(double)) != hipSuccess)
            st = obme_fail(OBME_ERR_HIP, "hipMemset");
    }
    if (st != OBME_OK) {
        obme_pisces_destroy(h);
        return st;
    }
    *out = h;
    return OBME_OK;
}

// PAR⁰ at time t: the constant, or default_surface_PAR (Light.jl:47),
// 100 max(0, cos(t π / 12 h)) with Julia's max (NaN propagates, ±0 ordered)
static double pisces_surface_value(const obme_pisces_config& c, double t)
{
    if (c.surface_mode != OBME_SURFACE_DIURNAL) return c.surface_PAR;
    const double cv = std::cos(t * 3.141592653589793 / (12 * 3600.0));
    const double d = 0.0 - cv;
    const double v = std::isnan(cv) ? d : (std::signbit(d) ? cv : 0.0);
    return 100 * v;
}

template <typename FT>
static PiscesArgs<FT> pisces_args(const obme_pisces* h, double t)
{
    const obme_pisces_config& c = h->cfg;
    const obme_grid& g = c.grid;
    PiscesArgs<FT> a;
    std::memset(&a, 0, sizeof(a));
    a.Gm = (FT*)h->d_Gm;
    a.PAR = (FT*)h->d_PAR;
    a.Omega = (FT*)h->d_Omega;
    a.zeu = (FT*)h->d_zeu;
    a.mlPAR = (FT*)h->d_mlPAR;
    a.oc = (const FT*)h->d_oc;
    a.tfac = (FT*)h->d_tfac;
    a.Hs = (float*)h->d_H;
    a.sflux = (double*)h->d_sflux;
    a.Hsurf = (double*)h->d_Hsurf;
    a.co2 = c.co2;
    a.o2 = c.o2;
    a.pp = (const FT*)h->d_pp;
    a.lev = (const FT*)h->d_lev;
    a.w = (const FT*)h->d_w;
    a.zf_top = h->zf[g.Nz];
    a.Nx = g.Nx;
    a.Ny = g.Ny;
    a.Nz = g.Nz;
    a.sx = (unsigned)(g.Nx + 2 * g.Hx);
    a.sxy = a.sx * (unsigned)(g.Ny + 2 * g.Hy);
    a.base0 = (unsigned)g.Hx + a.sx * (unsigned)g.Hy + a.sxy * (unsigned)g.Hz;
    a.surface_PAR = (FT)pisces_surface_value(c, t);
    a.surf = (c.surface_mode == OBME_SURFACE_ARRAY) ? (const FT*)c.surface_array : nullptr;
    for (int n = 0; n < 3; n++) {
        a.kw[n] = (FT)c.band_kw[n];
        a.e[n] = (FT)c.band_e[n];
        a.chi[n] = (FT)c.band_chi[n];
        a.division[n] = (FT)c.band_division[n];
    }
    const double lat = c.params[OBME_PP_SCALARS + OBME_PS_LATITUDE];
    a.solver = c.solver;
    (void)t;
    (void)lat;
    if (h->sediment) {
        // -w of each listed tracer's bottom face (the face velocities the
        // sinking uses: w[0] of the POC or GOC profile)
        const obme_sediment& sd = c.sediment;
        a.sed_on = 1;
        a.sed_recv = sd.receiver;
        for (int q = 0; q < sd.n_sinking; q++) {
            const int t = sd.sinking[q];
            const bool poc = (t == OBME_PI_POC || t == OBME_PI_SFE);
            a.sedw[t] = -(poc ? h->w_host[0] : h->w_host[h->cfg.grid.Nz + 1]);
        }
        a.sed_a = (FT)sd.burial_efficiency_constant1;
        a.sed_b = (FT)sd.burial_efficiency_constant2;
        a.sed_k = (FT)sd.burial_efficiency_half_saturation;
        a.sed_cn = 1.5 + sd.ab2_chi;
        a.sed_cm = 0.5 + sd.ab2_chi;
        a.storage = h->call_storage;
        a.sedGm = h->d_sedGm;
        a.sedGn = h->d_sedGn;
    }
    return a;
}

// day lengths of the stage clock, on the host in FP64 (the same expressions
// as the device helper)
static double host_sind(double x) { return std::sin(std::fmod(x, 360.0) * (3.141592653589793 / 180.0)); }
static double host_cosd(double x) { return std::cos(std::fmod(x, 360.0) * (3.141592653589793 / 180.0)); }
static double host_day_length(double t, double phi)
{
    const double p = 0.833;
    double m = std::fmod(t, 365 * 86400.0);
    if (m < 0) m += 365 * 86400.0;
    double J = std::floor(m / 86400.0);
    double theta = 0.216310 + 2 * std::atan(0.9671396 * std::tan(0.00860 * (J - 186)));
    double dec = std::asin(0.39795 * std::cos(theta)) * (180.0 / 3.141592653589793);
    double r = (host_sind(p) + host_sind(phi) * host_sind(dec)) / (host_cosd(phi) * host_cosd(dec));
    double L = std::isnan(r) ? r : std::max(-1.0, std::min(1.0, r));
    return (24 - 24.0 / 180 * (std::acos(L) * (180.0 / 3.141592653589793))) * 3600.0;
}

template <typename FT>
static obme_status pisces_scale(obme_pisces* h, const void* const* in, hipStream_t s);

// One stage.  mode 0: the substep of obme_pisces_step (Ω inline from the
// prepared constants; have_H: the [H+] of the previous stage is stored);
// mode 1: tendencies added into out (Ω by its own kernel).
template <typename FT>
static obme_status pisces_stage(obme_pisces* h, const void* const* in, void* const* out, double t, double dt,
                                int stage, int mode, int have_H, hipStream_t s)
{
    PiscesArgs<FT> a = pisces_args<FT>(h, t);
    for (int n = 0; n < NTR; n++) a.in[n] = (const FT*)in[n];
    for (int n = 0; n < NW; n++) a.out[n] = (FT*)out[n];
    const double lat = h->cfg.params[OBME_PP_SCALARS + OBME_PS_LATITUDE];
    a.sk.dl_growth = (FT)host_day_length(lat, t);  // day_length(φ, t), growth_rate.jl:30
    a.sk.dl_chl = (FT)host_day_length(t, lat);     // day_length(t, φ), growth_rate.jl:128
    a.dt = dt;
    a.stage = stage;
    if (h->sediment) {
        a.sed_dt[0] = (8.0 / 15.0) * dt;
        a.sed_dt[1] = (5.0 / 12.0 + -17.0 / 60.0) * dt;
        a.sed_dt[2] = (3.0 / 4.0 + -5.0 / 12.0) * dt;
        a.sed_init = h->call_init;
        a.sed_ab2 = h->call_ab2;
        if (mode != 0) a.storage = nullptr;
    }
    const double gam[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0}, zet[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
    a.A = (FT)(dt * gam[stage]);
    a.B = (FT)(dt * zet[stage]);
    a.have_H = have_H;
    if (mode == 0) {
        // the fused step: the scaling modifier first, on the stage input
        obme_status r = pisces_scale<FT>(h, in, s);
        if (r != OBME_OK) return r;
    }
    const int bs = 256;
    hipLaunchKernelGGL(k_pisces_columns<FT>, dim3((unsigned)((h->ncol + bs - 1) / bs)), dim3(bs), 0, s, a);
    OBME_HIP_TRY(hipGetLastError());
    if (a.sflux) {
        hipLaunchKernelGGL(k_pisces_surface<FT>, dim3((unsigned)((h->ncol + bs - 1) / bs)), dim3(bs), 0, s, a);
        OBME_HIP_TRY(hipGetLastError());
    }
    const unsigned nb = (unsigned)((h->ncell + bs - 1) / bs);
    if (mode == 0) {
        if (h->sediment) hipLaunchKernelGGL((k_pisces_cell<FT, 0, 1, true>), dim3(nb), dim3(bs), 0, s, a);
        else hipLaunchKernelGGL((k_pisces_cell<FT, 0, 1>), dim3(nb), dim3(bs), 0, s, a);
    } else {
        hipLaunchKernelGGL(k_pisces_omega<FT>, dim3(nb), dim3(bs), 0, s, a);
        OBME_HIP_TRY(hipGetLastError());
        if (h->sediment) hipLaunchKernelGGL((k_pisces_cell<FT, 1, 0, true>), dim3(nb), dim3(bs), 0, s, a);
        else hipLaunchKernelGGL((k_pisces_cell<FT, 1, 0>), dim3(nb), dim3(bs), 0, s, a);
    }
    OBME_HIP_TRY(hipGetLastError());
    return OBME_OK;
}

template <typename FT>
static NegScale<FT, 32> pisces_scaling(const obme_negative_scaling& s)
{
    NegScale<FT, 32> d;
    std::memset(&d, 0, sizeof(d));
    d.n_groups = s.n_groups;
    for (int g = 0; g < s.n_groups; g++) {
        for (int m = 0; m < s.n_members[g]; m++) {
            d.mask[g] |= 1u << s.member[g][m];
            d.sf[g][s.member[g][m]] = (FT)s.scalefactor[g][m];
        }
        d.fill[g] = (FT)s.invalid_fill_value[g];
    }
    return d;
}

// the scaling modifier in place on the arrays `in` (no-op without groups)
template <typename FT>
static obme_status pisces_scale(obme_pisces* h, const void* const* in, hipStream_t s)
{
    if (!h->scaling) return OBME_OK;
    PiscesArgs<FT> a = pisces_args<FT>(h, 0.0);
    for (int n = 0; n < NTR; n++) a.in[n] = (const FT*)in[n];
    const int bs = 256;
    hipLaunchKernelGGL(k_pisces_scale<FT>, dim3((unsigned)((h->ncell + bs - 1) / bs)), dim3(bs), 0, s, a,
                       pisces_scaling<FT>(h->cfg.scaling));
    OBME_HIP_TRY(hipGetLastError());
    return OBME_OK;
}

template <typename FT>
static obme_status pisces_prep(obme_pisces* h, const void* const* in, hipStream_t s)
{
    PiscesArgs<FT> a = pisces_args<FT>(h, 0.0);
    for (int n = 0; n < NTR; n++) a.in[n] = (const FT*)in[n];
    const int bs = 256;
    hipLaunchKernelGGL(k_pisces_prep<FT>, dim3((unsigned)((h->ncell + bs - 1) / bs)), dim3(bs), 0, s, a);
    OBME_HIP_TRY(hipGetLastError());
    return OBME_OK;
}

static obme_status pisces_stage_any(obme_pisces* h, const void* const* in, void* const* out, double t, double dt,
                                    int stage, int mode, int have_H, hipStream_t s)
{
    return (h->prec == OBME_FP64) ? pisces_stage<double>(h, in, out, t, dt, stage, mode, have_H, s)
                                  : pisces_stage<float>(h, in, out, t, dt, stage, mode, have_H, s);
}

template <typename FT>
static obme_status pisces_copy_aux(obme_pisces* h, void* PAR, void* const* PARn, void* Omega, void* zeu,
                                   void* mlPAR, hipStream_t s)
{
    // compact interior -> parent layout for the 3-D outputs
    const obme_grid& g = h->cfg.grid;
    const size_t es = sizeof(FT);
    const size_t sx = (size_t)(g.Nx + 2 * g.Hx), sxy = sx * (size_t)(g.Ny + 2 * g.Hy);
    auto put3 = [&](void* dst, const char* src) -> obme_status {
        if (!dst) return OBME_OK;
        for (int k = 0; k < g.Nz; k++) {
            char* d = (char*)dst + ((size_t)g.Hx + sx * (size_t)g.Hy + sxy * (size_t)(k + g.Hz)) * es;
            const char* c = src + (size_t)k * g.Nx * g.Ny * es;
            OBME_HIP_TRY(hipMemcpy2DAsync(d, sx * es, c, (size_t)g.Nx * es, (size_t)g.Nx * es, (size_t)g.Ny,
                                          hipMemcpyDeviceToDevice, s));
        }
        return OBME_OK;
    };
    const char* P = (const char*)h->d_PAR;
    obme_status st = put3(PAR, P);
    for (int n = 0; n < 3 && st == OBME_OK; n++)
        if (PARn) st = put3(PARn[n], P + (size_t)(n + 1) * h->ncell * es);
    if (st == OBME_OK) st = put3(Omega, (const char*)h->d_Omega);
    if (st == OBME_OK && zeu) OBME_HIP_TRY(hipMemcpyAsync(zeu, h->d_zeu, h->ncol * es, hipMemcpyDeviceToDevice, s));
    if (st == OBME_OK && mlPAR)
        OBME_HIP_TRY(hipMemcpyAsync(mlPAR, h->d_mlPAR, h->ncol * es, hipMemcpyDeviceToDevice, s));
    return st;
}

extern "C" obme_status obme_pisces_update_state(obme_pisces* h, const void* const* tracers, void* PAR,
                                                void* const* PARn, void* Omega, void* euphotic_depth,
                                                void* mixed_layer_PAR, double t, void* stream)
{
    if (!h || !tracers) return obme_fail(OBME_ERR_INVALID, "null argument");
    for (int n = 0; n < NTR; n++)
        if (!tracers[n]) return obme_fail(OBME_ERR_INVALID, "null tracer pointer");
    hipStream_t s = (hipStream_t)stream;
    const int bs = 256;
    {
        // the modifiers first: the tracers are rescaled in place
        obme_status r = (h->prec == OBME_FP64) ? pisces_scale<double>(h, tracers, s) : pisces_scale<float>(h, tracers, s);
        if (r != OBME_OK) return r;
    }
    if (h->prec == OBME_FP64) {
        PiscesArgs<double> a = pisces_args<double>(h, t);
        for (int n = 0; n < NTR; n++) a.in[n] = (const double*)tracers[n];
        hipLaunchKernelGGL(k_pisces_columns<double>, dim3((unsigned)((h->ncol + bs - 1) / bs)), dim3(bs), 0, s, a);
        hipLaunchKernelGGL(k_pisces_omega<double>, dim3((unsigned)((h->ncell + bs - 1) / bs)), dim3(bs), 0, s, a);
        OBME_HIP_TRY(hipGetLastError());
        return pisces_copy_aux<double>(h, PAR, PARn, Omega, euphotic_depth, mixed_layer_PAR, s);
    }
    PiscesArgs<float> a = pisces_args<float>(h, t);
    for (int n = 0; n < NTR; n++) a.in[n] = (const float*)tracers[n];
    hipLaunchKernelGGL(k_pisces_columns<float>, dim3((unsigned)((h->ncol + bs - 1) / bs)), dim3(bs), 0, s, a);
    hipLaunchKernelGGL(k_pisces_omega<float>, dim3((unsigned)((h->ncell + bs - 1) / bs)), dim3(bs), 0, s, a);
    OBME_HIP_TRY(hipGetLastError());
    return pisces_copy_aux<float>(h, PAR, PARn, Omega, euphotic_depth, mixed_layer_PAR, s);
}

extern "C" obme_status obme_pisces_add_tendencies(obme_pisces* h, const void* const* tracers, void* const* G,
                                                  double t, void* stream)
{
    if (!h || !tracers || !G) return obme_fail(OBME_ERR_INVALID, "null argument");
    for (int n = 0; n < NTR; n++)
        if (!tracers[n]) return obme_fail(OBME_ERR_INVALID, "null tracer pointer");
    for (int n = 0; n < NW; n++)
        if (!G[n]) return obme_fail(OBME_ERR_INVALID, "null tendency pointer");
    return pisces_stage_any(h, tracers, G, t, 0.0, 0, 1, 0, (hipStream_t)stream);
}

// nsteps of the fused FP32-mixed step (k_pisces_fused), one launch per step,
// in place on the caller's tracers
static obme_status pisces_fused_steps(obme_pisces* h, void* const* tracers, int nsteps, double dt, double t0,
                                      hipStream_t s)
{
    PiscesArgs<float> a = pisces_args<float>(h, t0);
    for (int n = 0; n < NTR; n++) a.in[n] = (const float*)tracers[n];
    for (int n = 0; n < NW; n++) a.out[n] = (float*)tracers[n];
    PiscesFused f;
    std::memset(&f, 0, sizeof(f));
    if (h->scaling) f.ns = pisces_scaling<float>(h->cfg.scaling);
    const double gam[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0}, zet[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
    const double cst[3] = {0.0, 8.0 / 15.0, 2.0 / 3.0};
    for (int q = 0; q < 3; q++) {
        f.A[q] = (float)(dt * gam[q]);
        f.B[q] = (float)(dt * zet[q]);
        f.div[q] = (float)h->cfg.band_division[q];
    }
    const double lat = h->cfg.params[OBME_PP_SCALARS + OBME_PS_LATITUDE];
    if (h->cfg.co2.enabled || h->cfg.o2.enabled) {
        if (!h->d_surfc && hipMalloc(&h->d_surfc, (size_t)SURF_PACK * h->ncol * sizeof(double)) != hipSuccess)
            return obme_fail(OBME_ERR_NOMEM, "PISCES surface constants");
        hipLaunchKernelGGL(k_pisces_surface_prep<float>, dim3((unsigned)((h->ncol + 255) / 256)), dim3(256), 0, s, a,
                           (double*)h->d_surfc);
        OBME_HIP_TRY(hipGetLastError());
        f.surf = (const double*)h->d_surfc;
    }
    const unsigned nbx = (unsigned)((h->cfg.grid.Nx + PF_WAVES - 1) / PF_WAVES);
    const unsigned nb = nbx * (unsigned)h->cfg.grid.Ny;
    for (int s0 = 0; s0 < nsteps; s0 += h->fused_steps) {
        f.nsteps = std::min(h->fused_steps, nsteps - s0);
        for (int sn = 0; sn < f.nsteps; sn++) {
            const double t = t0 + (s0 + sn) * dt;
            for (int q = 0; q < 3; q++) {
                const double tq = t + cst[q] * dt;
                f.sk[sn][q].dl_growth = (float)host_day_length(lat, tq);  // day_length(φ, t), growth_rate.jl:30
                f.sk[sn][q].dl_chl = (float)host_day_length(tq, lat);     // day_length(t, φ), growth_rate.jl:128
                f.PAR0[sn][q] = (float)pisces_surface_value(h->cfg, tq);
            }
        }
        f.have_H = s0 > 0 ? 1 : 0;
        if (h->scaling)
            hipLaunchKernelGGL(k_pisces_fused<true>, dim3(nb), dim3(PF_WAVES * 64), 0, s, a, f);
        else
            hipLaunchKernelGGL(k_pisces_fused<false>, dim3(nb), dim3(PF_WAVES * 64), 0, s, a, f);
        OBME_HIP_TRY(hipGetLastError());
    }
    return OBME_OK;
}

static obme_status pisces_step(obme_pisces* h, void* const* tracers, double* storage, int32_t nsteps, double dt,
                               double t0, void* stream);

extern "C" obme_status obme_pisces_step(obme_pisces* h, void* const* tracers, int32_t nsteps, double dt, double t0,
                                        void* stream)
{
    return pisces_step(h, tracers, nullptr, nsteps, dt, t0, stream);
}

extern "C" obme_status obme_pisces_step_with_sediment(obme_pisces* h, void* const* tracers, double* storage,
                                                      int32_t nsteps, double dt, double t0, void* stream)
{
    if (!h) return obme_fail(OBME_ERR_INVALID, "null handle");
    if (!h->sediment) return obme_fail(OBME_ERR_INVALID, "the PISCES handle has no sediment");
    if (!storage) return obme_fail(OBME_ERR_INVALID, "null sediment storage");
    return pisces_step(h, tracers, storage, nsteps, dt, t0, stream);
}

extern "C" obme_status obme_pisces_sediment_update_state(obme_pisces* h, const void* const* tracers, double* storage,
                                                         double dt, void* stream)
{
    if (!h || !tracers) return obme_fail(OBME_ERR_INVALID, "null argument");
    if (!h->sediment) return obme_fail(OBME_ERR_INVALID, "the PISCES handle has no sediment");
    const obme_sediment& sd = h->cfg.sediment;
    for (int q = 0; q < sd.n_sinking; q++)
        if (!tracers[sd.sinking[q]]) return obme_fail(OBME_ERR_INVALID, "null tracer pointer");
    // the tracked fields only: no step before the model's first stage
    if (!std::isfinite(dt)) return OBME_OK;
    const bool euler = !(dt == h->sed_last_dt);
    const int init = h->sed_iter == 0 ? 1 : 0;
    const unsigned nb = (unsigned)((h->ncol + 255) / 256);
    if (h->prec == OBME_FP64) {
        PiscesArgs<double> a = pisces_args<double>(h, 0.0);
        for (int n = 0; n < NTR; n++) a.in[n] = (const double*)tracers[n];
        hipLaunchKernelGGL(k_pisces_sediment_state<double>, dim3(nb), dim3(256), 0, (hipStream_t)stream, a, storage,
                           dt, euler ? 1 : 0, init);
    } else {
        PiscesArgs<float> a = pisces_args<float>(h, 0.0);
        for (int n = 0; n < NTR; n++) a.in[n] = (const float*)tracers[n];
        hipLaunchKernelGGL(k_pisces_sediment_state<float>, dim3(nb), dim3(256), 0, (hipStream_t)stream, a, storage,
                           dt, euler ? 1 : 0, init);
    }
    OBME_HIP_TRY(hipGetLastError());
    h->sed_iter += 1;
    h->sed_last_dt = dt;
    return OBME_OK;
}

// nsteps steps; with the sediment's storage, its steps in the same passes
// (the handle's clock and G⁻ advanced), without, its coupling alone
static obme_status pisces_step(obme_pisces* h, void* const* tracers, double* storage, int32_t nsteps, double dt,
                               double t0, void* stream)
{
    if (!h || !tracers) return obme_fail(OBME_ERR_INVALID, "null argument");
    if (nsteps < 0) return obme_fail(OBME_ERR_INVALID, "negative nsteps");
    for (int n = 0; n < NTR; n++)
        if (!tracers[n]) return obme_fail(OBME_ERR_INVALID, "null tracer pointer");
    if (nsteps == 0) return OBME_OK;
    hipStream_t s = (hipStream_t)stream;
    if (!h->d_oc && hipMalloc(&h->d_oc, (size_t)OMEGA_CONST_FIELDS * h->ncell * h->es) != hipSuccess)
        return obme_fail(OBME_ERR_NOMEM, "PISCES carbonate constants");
    if (!h->d_H && hipMalloc(&h->d_H, (size_t)h->ncell * sizeof(float)) != hipSuccess)
        return obme_fail(OBME_ERR_NOMEM, "PISCES [H+] store");
    if (h->prec == OBME_FP64 && !h->d_tfac &&
        hipMalloc(&h->d_tfac, (size_t)TFAC_FIELDS * h->ncell * sizeof(double)) != hipSuccess)
        return obme_fail(OBME_ERR_NOMEM, "PISCES temperature factors");
    {
        obme_status r = (h->prec == OBME_FP64) ? pisces_prep<double>(h, tracers, s) : pisces_prep<float>(h, tracers, s);
        if (r != OBME_OK) return r;
    }
    if (h->fused) return pisces_fused_steps(h, tracers, nsteps, dt, t0, s);
    for (int n = 0; n < NW; n++)
        if (!h->d_copy[n] && hipMalloc(&h->d_copy[n], h->field_bytes) != hipSuccess)
            return obme_fail(OBME_ERR_NOMEM, "PISCES stage copy");
    // Stage inputs and outputs alternate between the caller's arrays (U) and
    // library copies (C), which start as full copies of U so that their halos
    // match; the kernels write interiors only.  A step ends on the other side
    // of where it started; an odd number of steps ends with a copy back.
    for (int n = 0; n < NW; n++)
        OBME_HIP_TRY(hipMemcpyAsync(h->d_copy[n], tracers[n], h->field_bytes, hipMemcpyDeviceToDevice, s));
    const void* inU[NTR];
    const void* inC[NTR];
    void* outU[NW];
    void* outC[NW];
    for (int n = 0; n < NTR; n++) {
        inU[n] = tracers[n];
        inC[n] = (n < NW) ? h->d_copy[n] : tracers[n];
    }
    for (int n = 0; n < NW; n++) {
        outU[n] = tracers[n];
        outC[n] = h->d_copy[n];
    }
    const double cst[3] = {0.0, 8.0 / 15.0, 2.0 / 3.0};
    // the sediment of this call: its first step ever (the lazy first
    // tendency) and whether the first sediment step repeats the previous Δt
    const bool sed_init = h->sed_iter == 0;
    const bool sed_ab2 = !sed_init && (8.0 / 15.0) * dt == h->sed_last_dt;
    struct CallSediment {
        obme_pisces* h;
        ~CallSediment() { h->call_storage = nullptr; h->call_init = h->call_ab2 = 0; }
    } keep{h};
    h->call_storage = h->sediment ? storage : nullptr;
    bool inCopy = false;  // where the current state lives
    for (int st = 0; st < nsteps; st++) {
        const double t = t0 + st * dt;
        h->call_init = (sed_init && st == 0) ? 1 : 0;
        h->call_ab2 = (sed_ab2 && st == 0) ? 1 : 0;
        for (int q = 0; q < 3; q++) {
            const void* const* in = inCopy ? inC : inU;
            void* const* out = inCopy ? outU : outC;
            obme_status r = pisces_stage_any(h, in, out, t + cst[q] * dt, dt, q, 0, (st > 0 || q > 0) ? 1 : 0, s);
            if (r != OBME_OK) return r;
            inCopy = !inCopy;
        }
    }
    if (inCopy)
        for (int n = 0; n < NW; n++)
            OBME_HIP_TRY(hipMemcpyAsync(tracers[n], h->d_copy[n], h->field_bytes, hipMemcpyDeviceToDevice, s));
    if (h->sediment && storage) {
        // three sediment steps per step, the last of Δt (γ3 + ζ3)
        h->sed_iter += 3 * (long long)nsteps;
        h->sed_last_dt = (3.0 / 4.0 + -5.0 / 12.0) * dt;
    }
    return OBME_OK;
}
