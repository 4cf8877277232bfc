// obme_npd_kernels.h — the fused column-tile kernels of the NPZD / LOBSTER
// biogeochemistry step on gfx950.
//
// Work decomposition.  A workgroup owns TX consecutive i-columns of one j
// row and the whole depth: thread (tx, kg) holds levels
// k = kg*LPT .. kg*LPT+LPT-1 (0-based, bottom to top) of column i, so every
// global load/store of one level is TX consecutive elements of the
// Oceananigans layout (coalesced along i).  All tracer state of the tile
// stays in registers through the three RK3 stages of a step; HBM sees one
// read and one write of each written tracer per step.
//
// Column coupling goes through LDS:
//  - the TwoBand chlorophyll integral (Light/2band.jl:18-32) is a suffix sum
//    over levels: each thread sums its own increments, exchanges the thread
//    total, and continues the reference recurrence ∫(k) = ∫(k+1) + inc(k)
//    from the sum of the totals above it;
//  - upwind sinking needs c(k+1) for a thread's top level: the bottom value
//    of the thread above is exchanged.
// Two workgroup barriers per stage.
#pragma once

#include "obme_device.h"

namespace obme {

constexpr int MAX_SINK = 2;

template <typename FT>
struct KahanSum {
    FT s, c;
    __device__ __forceinline__ void init(FT v) { s = v; c = FT(0); }
    __device__ __forceinline__ void add(FT v)
    {
        if constexpr (sizeof(FT) == 4) {
            FT y = v - c;
            FT t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            s = s + v;
        }
    }
};

// Per-model tracer layout (tracer order of required_biogeochemical_tracers).
template <int MODEL, bool CARB>
struct ModelTraits;

template <>
struct ModelTraits<OBME_NPZD, false> {
    static constexpr int NT = 5;         // N, P, Z, T, D
    static constexpr int P_IDX = 1;
    static constexpr int RO_IDX = 3;     // T: read-only tracer
    static constexpr int NSINK = 2;
    static constexpr int SINK[2] = {1, 4};  // P, D
};
template <>
struct ModelTraits<OBME_LOBSTER, false> {
    static constexpr int NT = 7;
    static constexpr int P_IDX = 2;
    static constexpr int RO_IDX = -1;
    static constexpr int NSINK = 2;
    static constexpr int SINK[2] = {4, 5};  // sPOM, bPOM
};
template <>
struct ModelTraits<OBME_LOBSTER, true> {
    static constexpr int NT = 9;
    static constexpr int P_IDX = 2;
    static constexpr int RO_IDX = -1;
    static constexpr int NSINK = 2;
    static constexpr int SINK[2] = {4, 5};
};

// Kernel arguments (passed by value).
template <typename FT>
struct StepArgs {
    FT* tr[OBME_MAX_TRACERS];      // tracer parent arrays
    FT* G[OBME_MAX_TRACERS];       // tendency parent arrays (hook kernel)
    FT* PAR;                       // PAR parent array (may be null for the stepper)
    const FT* lev;                 // per-level geometry, see LEV_* below
    const FT* wsink;               // [MAX_SINK][Nz+1] face velocities (<= 0)
    const FT* surf;                // surface PAR array (Nx*Ny) or null
    const FT* T3;                  // 3-D temperature (LOBSTER gas exchange) or null
    const FT* S3;                  // 3-D salinity or null
    const double* Tprof;           // temperature profile [Nz] or null
    const double* Sprof;
    double Tconst, Sconst;
    int Nx, Ny, Nz, Hx, Hy, Hz;
    long sx, sxy, base0;           // strides and interior origin offset
    NPDParams<FT> q;
    FT kr, kb, xr, xb, er, eb, Rr;  // Rr = R_chl / r
    int surface_mode;
    FT surface_value;
    double dt, t0;
    int co2_on, co2_mode;
    obme_co2_exchange co2;
    obme_carbon_solver solver;
    double* flux_out;              // optional diagnostics: last-stage CO2 flux (Nx*Ny)
    double* pH_out;
};

// Per-level geometry rows in `lev` (each Nz entries).
enum { LEV_ZC = 0, LEV_UP, LEV_DN, LEV_INVDZ, LEV_COUNT };
// LEV_UP[k] = zc[k+1] - zf[k+1], LEV_DN[k] = zf[k+1] - zc[k] for k < Nz-1;
// for k = Nz-1: LEV_UP = 0 and LEV_DN = zf[Nz] - zc[Nz-1] (the top half cell).

__device__ __forceinline__ double default_surface_par(double t)
{   // Light/Light.jl:219
    double c = cos(t * 3.141592653589793 / (12 * 3600.0));
    return 100 * jmax(0.0, c);
}

template <typename FT>
__device__ __forceinline__ FT surface_par(const StepArgs<FT>& a, int i, int j, double t)
{
    if (a.surface_mode == OBME_SURFACE_ARRAY) return a.surf[i + (long)a.Nx * j];
    if (a.surface_mode == OBME_SURFACE_DIURNAL) return (FT)default_surface_par(t);
    return a.surface_value;
}

// Modes of the column-tile routine.
enum { TILE_STEP = 0, TILE_TENDENCY = 1, TILE_PAR = 2 };

template <int MODEL, bool CARB, typename FT, int LPT, int MODE>
__device__ __forceinline__ void column_tile(const StepArgs<FT>& a, FT* lds)
{
    using TR = ModelTraits<MODEL, CARB>;
    constexpr int NT = TR::NT;
    const int TX = blockDim.x, KG = blockDim.y;
    const int tx = threadIdx.x, kg = threadIdx.y;
    const int i = blockIdx.x * TX + tx;
    const int j = blockIdx.y;
    const bool col_ok = i < a.Nx;
    const int Nz = a.Nz;
    const int k0 = kg * LPT;
    const int ktop = Nz - 1;

    FT* s_pr = lds;                  // [KG][TX]
    FT* s_pb = s_pr + KG * TX;
    FT* s_c = s_pb + KG * TX;        // [MAX_SINK][KG][TX]
    FT* s_tr = s_c + MAX_SINK * KG * TX;
    FT* s_tb = s_tr + KG * TX;
    const int me = kg * TX + tx;
    const int above = (kg + 1) * TX + tx;

    const long colbase = a.base0 + (long)(col_ok ? i : 0) + a.sx * j;

    FT U[NT][LPT];
    FT Gm[NT][LPT];
    bool lv[LPT];
#pragma unroll
    for (int l = 0; l < LPT; l++) lv[l] = col_ok && (k0 + l) < Nz;
#pragma unroll
    for (int n = 0; n < NT; n++)
#pragma unroll
        for (int l = 0; l < LPT; l++) {
            U[n][l] = lv[l] ? a.tr[n][colbase + a.sxy * (k0 + l)] : FT(0);
            Gm[n][l] = FT(0);
        }

    // T and S are fixed through a step, so everything of the CO2 flux that
    // depends on them only is computed once per step into LDS (the same
    // values the per-stage evaluation would produce).
    const bool owns_top = col_ok && (k0 <= ktop) && (ktop < k0 + LPT);
    SurfaceCarbon* s_cc = reinterpret_cast<SurfaceCarbon*>(s_tb + KG * TX);
    if constexpr (CARB) {
        if (MODE != TILE_PAR && owns_top && a.co2_on) {
            long ts = colbase + a.sxy * ktop;
            double Ts = a.T3 ? (double)a.T3[ts] : (a.Tprof ? a.Tprof[ktop] : a.Tconst);
            double Ss = a.S3 ? (double)a.S3[ts] : (a.Sprof ? a.Sprof[ktop] : a.Sconst);
            SurfaceCarbon sc;
            surface_carbon_setup(a.co2, Ts, Ss, a.co2.wind_speed, sc);
            s_cc[tx] = sc;
        }
    }

    const FT* zc = a.lev + LEV_ZC * Nz;
    const FT* up = a.lev + LEV_UP * Nz;
    const FT* dn = a.lev + LEV_DN * Nz;
    const FT* invdz = a.lev + LEV_INVDZ * Nz;

    const int nstage = (MODE == TILE_STEP) ? 3 : 1;
    const double g1 = 8.0 / 15.0, g2 = 5.0 / 12.0, g3 = 3.0 / 4.0, z2 = -17.0 / 60.0, z3 = -5.0 / 12.0;

    for (int st = 0; st < nstage; st++) {
        const double t = a.t0 + (st == 0 ? 0.0 : (st == 1 ? g1 : g1 + g2 + z2)) * a.dt;

        // ---- phase A: chlorophyll terms (P R/r)^e of the own levels
        FT pr[LPT], pb[LPT];
#pragma unroll
        for (int l = 0; l < LPT; l++) {
            FT x = U[TR::P_IDX][l] * a.Rr;
            if constexpr (sizeof(FT) == 4) {
                FT lg = m_log2(x);
                pr[l] = m_exp2(a.er * lg);
                pb[l] = m_exp2(a.eb * lg);
            } else {
                pr[l] = m_pow(x, a.er);
                pb[l] = m_pow(x, a.eb);
            }
        }
        s_pr[me] = pr[0];
        s_pb[me] = pb[0];
#pragma unroll
        for (int s = 0; s < TR::NSINK; s++) s_c[s * KG * TX + me] = U[TR::SINK[s]][0];
        __syncthreads();

        // ---- phase B: increments, thread totals, CO2 flux at the top
        const bool has_above = (kg + 1) < KG;
        FT pr_up = has_above ? s_pr[above] : FT(0);
        FT pb_up = has_above ? s_pb[above] : FT(0);
        FT c_up[MAX_SINK];
#pragma unroll
        for (int s = 0; s < TR::NSINK; s++) c_up[s] = has_above ? s_c[s * KG * TX + above] : FT(0);

        FT incr[LPT], incb[LPT];
#pragma unroll
        for (int l = 0; l < LPT; l++) {
            const int k = k0 + l;
            if (k < ktop) {
                FT prn = (l + 1 < LPT) ? pr[(l + 1) < LPT ? l + 1 : l] : pr_up;
                FT pbn = (l + 1 < LPT) ? pb[(l + 1) < LPT ? l + 1 : l] : pb_up;
                incr[l] = up[k] * prn + dn[k] * pr[l];
                incb[l] = up[k] * pbn + dn[k] * pb[l];
            } else if (k == ktop) {
                incr[l] = dn[k] * pr[l];
                incb[l] = dn[k] * pb[l];
            } else {
                incr[l] = FT(0);
                incb[l] = FT(0);
            }
        }
        KahanSum<FT> totr, totb;
        totr.init(incr[LPT - 1]);
        totb.init(incb[LPT - 1]);
#pragma unroll
        for (int l = LPT - 2; l >= 0; l--) {
            totr.add(incr[l]);
            totb.add(incb[l]);
        }
        s_tr[me] = totr.s;
        s_tb[me] = totb.s;

        double F = 0.0;
        if constexpr (CARB) {
            if (MODE != TILE_PAR && owns_top && a.co2_on) {
                const int lt = ktop - k0;
                FT dic = FT(0), alk = FT(0);
#pragma unroll
                for (int l = 0; l < LPT; l++)
                    if (l == lt) {
                        dic = U[7][l];
                        alk = U[8][l];
                    }
                double pH;
                int ev = 0;
                if constexpr (sizeof(FT) == 4)
                    F = surface_flux<false>(s_cc[tx], a.solver, (double)dic, (double)alk, a.co2.air_concentration,
                                            1, pH, ev);
                else
                    F = surface_flux<true>(s_cc[tx], a.solver, (double)dic, (double)alk, a.co2.air_concentration,
                                           0, pH, ev);
                if (a.flux_out && (MODE == TILE_TENDENCY || st == 2)) {
                    a.flux_out[i + (long)a.Nx * j] = F;
                    if (a.pH_out) a.pH_out[i + (long)a.Nx * j] = pH;
                }
            }
        }
        __syncthreads();

        // ---- phase C: integral from the top, PAR, transitions, update
        KahanSum<FT> runr, runb;
        runr.init(FT(0));
        runb.init(FT(0));
        for (int m = KG - 1; m > kg; m--) {
            runr.add(s_tr[m * TX + tx]);
            runb.add(s_tb[m * TX + tx]);
        }
        const FT PAR0 = col_ok ? surface_par(a, i, j, t) : FT(0);
        FT PARv[LPT];
#pragma unroll
        for (int l = LPT - 1; l >= 0; l--) {
            runr.add(incr[l]);
            runb.add(incb[l]);
            const int k = k0 + l;
            FT z = lv[l] ? zc[k] : FT(0);
            PARv[l] = PAR0 * (m_exp(a.kr * z - a.xr * runr.s) + m_exp(a.kb * z - a.xb * runb.s)) / FT(2);
        }

        if constexpr (MODE == TILE_PAR) {
#pragma unroll
            for (int l = 0; l < LPT; l++)
                if (lv[l]) a.PAR[colbase + a.sxy * (k0 + l)] = PARv[l];
            return;
        }

        FT Lt_arr[LPT];
#pragma unroll
        for (int l = 0; l < LPT; l++) {
            Lt_arr[l] = FT(1);
            if constexpr (MODEL == OBME_NPZD) {
                if (a.q.use_Q10) Lt_arr[l] = m_pow(a.q.Q10, U[3][l] / FT(10));
            }
        }

#pragma unroll
        for (int l = 0; l < LPT; l++) {
            const int k = k0 + l;
            FT G[NT];
            if constexpr (MODEL == OBME_NPZD) {
                npzd_tendencies(a.q, U[0][l], U[1][l], U[2][l], U[3][l], U[4][l], PARv[l], Lt_arr[l], G[0], G[1],
                                G[2], G[4]);
                G[3] = FT(0);
            } else {
                FT v[NT];
#pragma unroll
                for (int n = 0; n < NT; n++) v[n] = U[n][l];
                lobster_tendencies<FT, CARB>(a.q, v, PARv[l], G);
            }
            // minus the upwind drift flux divergence: w(face) c(upwind cell)
#pragma unroll
            for (int s = 0; s < TR::NSINK; s++) {
                const int n = TR::SINK[s];
                const FT* w = a.wsink + s * (Nz + 1);
                FT cnext = (l + 1 < LPT) ? U[n][(l + 1) < LPT ? l + 1 : l] : c_up[s];
                FT Fhi = (k < ktop) ? w[k + 1] * cnext : FT(0);
                FT Flo = w[k] * U[n][l];
                FT div = lv[l] ? (Fhi - Flo) * invdz[k] : FT(0);
                G[n] = G[n] - div;
            }
            if constexpr (CARB) {
                if (k == ktop && a.co2_on) G[7] = G[7] - (FT)F * invdz[k];
            }

            if constexpr (MODE == TILE_TENDENCY) {
                if (lv[l]) {
                    long idx = colbase + a.sxy * k;
#pragma unroll
                    for (int n = 0; n < NT; n++)
                        if (n != TR::RO_IDX) a.G[n][idx] += G[n];
                    if (a.PAR) a.PAR[idx] = PARv[l];
                }
            } else {
                // RK3 substep (BoxModel/timesteppers.jl:79-92 form)
                const FT dtf = (FT)a.dt;
#pragma unroll
                for (int n = 0; n < NT; n++) {
                    if (n == TR::RO_IDX) continue;
                    if (st == 0) U[n][l] += dtf * (FT)g1 * G[n];
                    else if (st == 1) U[n][l] += dtf * ((FT)g2 * G[n] + (FT)z2 * Gm[n][l]);
                    else U[n][l] += dtf * ((FT)g3 * G[n] + (FT)z3 * Gm[n][l]);
                    Gm[n][l] = G[n];
                }
            }
        }
    }

    if constexpr (MODE == TILE_STEP) {
#pragma unroll
        for (int n = 0; n < NT; n++) {
            if (n == TR::RO_IDX) continue;
#pragma unroll
            for (int l = 0; l < LPT; l++)
                if (lv[l]) a.tr[n][colbase + a.sxy * (k0 + l)] = U[n][l];
        }
    }
}

}  // namespace obme
