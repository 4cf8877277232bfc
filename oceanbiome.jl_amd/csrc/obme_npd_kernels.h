// obme_npd_kernels.h — the column-sweep and surface kernels of the NPZD /
// LOBSTER biogeochemistry step on gfx950.
//
// Column sweep (column_lane): lane per water column.  A workgroup of NW
// waves owns 64 NW consecutive i-columns of one j row; each lane walks its
// column from the surface down, one level at a time, and runs the three RK3
// stages at each level before it moves on.  That order is legal because
// stage s at level k reads only stage-s values of the levels above it:
//  - the TwoBand chlorophyll integral (Light/2band.jl:18-32) is the running
//    sum ∫(k) = ∫(k+1) + inc(k), carried down per stage (in the reference's
//    own top-down order; Kahan-compensated in FP32);
//  - upwind sinking needs c(k+1) of the same stage, one level of history.
// Every load and store of a level is one contiguous run of 64 values per
// wave, straight from and to HBM (no LDS staging, no cross-lane traffic);
// the workgroup's waves pass a barrier (no fence) before each level, so
// their runs reach the memory together as one run of 64 NW values.  The next
// level's loads are issued before the current level's arithmetic.  Each
// tracer is read and written once per step.
//  - FP32-mixed: the step's increment D is accumulated apart from the state
//    at the start of the step U0 (stage inputs U0 + D, one rounding at the
//    store), so no stage rounds an increment against the magnitude of the
//    state; the tendencies use host-folded constant products.  FP64 keeps the
//    reference's expressions and update order.
//
// Surface kernel (k_surface).  DIC and Alk feed no other tendency
// (NutrientsPlanktonDetritus/carbonate_system.jl: one-way coupled), so the
// only place the CO2 flux enters is the DIC update of the top cell.  The
// sweep leaves that one value to a per-column kernel: it records, per
// column and stage, the surface Alk and the DIC transition without the
// flux; the surface kernel (one thread per column, no tracer state, high
// occupancy) solves the carbonate system at each stage and integrates the
// top-cell DIC with the identical update expressions.  The latency-bound
// FP64 root solve thereby leaves the bandwidth-bound sweep.
#pragma once

#include <type_traits>

#include "obme_device.h"

namespace obme {

constexpr int MAX_SINK = 2;

// Per-model tracer layout (tracer order of required_biogeochemical_tracers).
// CARB is a bit set of the optional components: 1 CarbonateSystem (LOBSTER:
// DIC, Alk at 7, 8), 2 Oxygen (O₂, last).
template <int MODEL, int CARB>
struct ModelTraits;

template <>
struct ModelTraits<OBME_NPZD, false> {
    static constexpr int NT = 5;            // N, P, Z, T, D
    static constexpr int P_IDX = 1;
    static constexpr int RO_IDX = 3;        // T: read-only tracer
    static constexpr int SINK0 = 1, SINK1 = 4;  // P, D
};
template <>
struct ModelTraits<OBME_NPZD, 2> {
    static constexpr int NT = 6;            // N, P, Z, T, D, O₂
    static constexpr int P_IDX = 1;
    static constexpr int RO_IDX = 3;
    static constexpr int SINK0 = 1, SINK1 = 4;
};
template <>
struct ModelTraits<OBME_LOBSTER, false> {
    static constexpr int NT = 7;
    static constexpr int P_IDX = 2;
    static constexpr int RO_IDX = -1;
    static constexpr int SINK0 = 4, SINK1 = 5;  // sPOM, bPOM
};
template <>
struct ModelTraits<OBME_LOBSTER, 1> {
    static constexpr int NT = 9;
    static constexpr int P_IDX = 2;
    static constexpr int RO_IDX = -1;
    static constexpr int SINK0 = 4, SINK1 = 5;
};
template <>
struct ModelTraits<OBME_LOBSTER, 2> {
    static constexpr int NT = 8;            // ..., DOM, O₂
    static constexpr int P_IDX = 2;
    static constexpr int RO_IDX = -1;
    static constexpr int SINK0 = 4, SINK1 = 5;
};
template <>
struct ModelTraits<OBME_LOBSTER, 3> {
    static constexpr int NT = 10;           // ..., DOM, DIC, Alk, O₂
    static constexpr int P_IDX = 2;
    static constexpr int RO_IDX = -1;
    static constexpr int SINK0 = 4, SINK1 = 5;
};

// Surface record written by the sweep for the surface kernel, per column:
// [s] DIC transition without the flux at stage s, [3+s] Alk at the start of
// stage s (s = 0, 1, 2).
constexpr int SURF_REC = 6;

// SimpleMultiG's nitrogen form (Models/Sediments/simple_multi_G.jl), FP64:
// decay rates λs, λf, fractions fs, ff, fr, the sinking C:N, the anoxia
// half saturation, the solid deposition fraction 0.223 r^0.336 of the
// sedimentation rate r (:427), the three fractions' A..F, and the tracer
// indices of NO3, NH4, O2.
struct SmgParams {
    double ls, lf, fs, ff, fr, R, kO2, ps;
    double qn[6], qd[6], qa[6];
    int no3, nh4, o2;
};

// The tendencies of Ns, Nf, Nr from the fields u and the sinking nitrogen
// flux f (simple_multi_G.jl:165-193).
__device__ __forceinline__ void smg_tendencies(const SmgParams& p, const double* u, double f, double* G)
{
    G[0] = (1 - p.fr) * p.fs * f - p.ls * u[0];
    G[1] = (1 - p.fr) * p.ff * f - p.lf * u[1];
    G[2] = p.fr * f;
}

__device__ __forceinline__ double smg_finite(double x) { return isfinite(x) ? x : 0.0; }

// The coupled tendencies of NH4, NO3 and O2 per unit area from the fields
// (Ns, Nf) and the bottom cell's NO3, NH4, O2 (simple_multi_G.jl:195-247;
// reactivity and the fractions :372-427; day = 86400 s; a non-finite
// fraction is zero).
__device__ __forceinline__ void smg_coupled(const SmgParams& p, const double* u, double NO3, double NH4, double O2,
                                            double& dNH4, double& dNO3, double& dO2)
{
    constexpr double DAY = 86400.0;
    const double Nr = p.ls * u[0] + p.lf * u[1];
    const double Cr = Nr * p.R;
    const double Cs = u[0] * p.R, Cf = u[1] * p.R;
    const double k = (p.ls * Cs + p.lf * Cf) / (Cs + Cf + 4.9406564584124654e-324);
    const double lC = log(Cr * DAY), lk = log(k), lkd = log(k * DAY), lO = log(O2), lN = log(NO3), lH = log(NH4);
    const double ln_n = (p.qn[0] + p.qn[1] * lC * lO + p.qn[2] * (lC * lC) + p.qn[3] * lkd * lH + p.qn[4] * lC
                         + p.qn[5] * lC * lH);
    const double pn = smg_finite(exp(ln_n) / (Nr * DAY) * O2 / (p.kO2 + O2));
    const double ln_d = (p.qd[0] + p.qd[1] * lC + p.qd[2] * (lN * lN) + p.qd[3] * (lC * lC) + p.qd[4] * (lkd * lkd)
                         + p.qd[5] * lO * lk);
    const double pd = smg_finite(exp(ln_d) / (Cr * DAY) * O2 / (p.kO2 + O2));
    const double ln_a = (p.qa[0] + p.qa[1] * lC + p.qa[2] * (lC * lC) + p.qa[3] * lkd + p.qa[4] * lO * lk
                         + p.qa[5] * (lN * lN));
    const double pa = smg_finite(exp(ln_a) / (Cr * DAY));
    dNH4 = (1 - pn) * Nr + 0.8 * pd * Cr;
    dNO3 = pn * Nr - 0.8 * pd * Cr;
    dO2 = -(1 - pa * p.ps - pd) * O2 / (p.kO2 + O2) * Cr - 2 * pn * Nr;
}

// Kernel arguments (passed by value).
template <typename FT>
struct StepArgs {
    FT* tr[OBME_MAX_TRACERS];      // tracer parent arrays
    FT* G[OBME_MAX_TRACERS];       // tendency parent arrays (hook kernel)
    FT* PAR;                       // PAR parent array (may be null)
    const FT* lev;                 // per-level geometry rows, see LEV_*
    const FT* wsink;               // [MAX_SINK][Nz+1] face velocities (<= 0)
    const FT* surf;                // surface PAR array (Nx*Ny) or null
    FT* srec;                      // surface records [SURF_REC][Nx*Ny] (CO2 on)
    // FP32 fused step with the CO2 exchange: the TwoSum residuals of the
    // top-cell DIC [0, Nx*Ny) (surface kernel) and Alk [Nx*Ny, 2 Nx*Ny)
    // (sweep), carried from step to step within a call, or null
    float* topres;
    int Nx, Ny, Nz;
    unsigned sx, sxy, base0;       // strides and interior origin offset
    unsigned field_bytes;          // byte size of one parent array
    NPDParams<FT> q;
    FT kr, kb, xr, xb, er, eb, Rr;  // Rr = R_chl / r
    int surface_mode;
    FT surface_value;
    FT PAR0_stage[3];              // diurnal surface PAR at the three stage clocks
    FT dtf;                        // Δt in the arithmetic type
    FT dz_u, zc0_u, w0_u, w1_u;    // uniform grid: Δz, zc[0], sinking w
    FT invdz_u;                    // uniform grid: 1/Δz
    FastParams f;                  // FP32 path: derived constants
    float rk_a[3], rk_b[3];        // FP32 path: Δt γ and Δt ζ per stage
    float kr2, kb2, xr2, xb2;      // FP32 path: attenuations × log2(e)
    float cr, cb;                  // FP32 path: e log2(R_chl / r) per band
    float xr2h, xb2h;              // FP32 uniform grid: xr2 dz/2, xb2 dz/2
    float wi0, wi1;                // FP32 uniform grid: w / dz of the two sinking tracers
    int co2_on;                    // CO2 flux by the surface kernel (surface records written)
    NegScale<FT, 16> ns;           // negative scaling at the start of every stage (EXTRA & 1)
    // CO2 flux computed in the sweep by the top lane (EXTRA & 2: with the
    // scaling, the top cell's DIC feeds back into the other tracers)
    const FT* T3;
    const FT* S3;
    const double* Tprof;
    const double* Sprof;
    double Tconst, Sconst;
    double* Hprev;                 // FP32-mixed warm start of [H+] per column
    FT invdz_top;
    obme_co2_exchange co2;
    obme_carbon_solver solver;
    // LOBSTER's O2 gas exchange at the top cell (CARB & 2): per column
    // k / Δz_top and the air-side concentration C_air α, formed from the
    // surface T and S once per call (k_o2_prep); applied by the top lane
    int o2_on;
    const double* o2c;             // [2 * (i + Nx j)]
    // InstantRemineralisation sediment (EXTRA & 4), worked by the lane that
    // holds a column's bottom cell
    int sed_recv;                  // receiver tracer
    FT sed_w0, sed_w1;             // -w at the bottom face of SINK0 / SINK1 when listed, else 0
    FT sed_a, sed_b, sed_k;        // burial efficiency a + b (f / (k + f))^2
    FT sed_dz0;                    // Δz at the bottom face
    double* storage;               // the storage field (Nx*Ny) or null
    double* sedG;                  // the sediment's tendencies [2][Nx*Ny]: Gⁿ, G⁻
    double sed_dt[3];              // Δt of the sediment step after each stage
    double sed_cn, sed_cm;         // AB2 weights 1.5 + χ, 0.5 + χ of the first stage's step
    int sed_ab2;                   // the first stage's step is an AB2 (not an Euler) step
    int sed_init;                  // the sediment's first step ever is in this launch
    // the sediment model (OBME_SEDIMENT_*), its fields (storage, sedG and
    // sed_fields hold nf × Nx*Ny: InstantRemineralisation 1, SimpleMultiG 3)
    int sed_model, sed_nf;
    SmgParams smg;
    const double* sed_fields;      // the hook's SimpleMultiG fields, or null (an empty sediment)
    // one RK3 stage per launch (TILE_STAGE): the stage, the CO2 flux computed
    // by the top lane, and the per-cell scratch of the step (compact, tracer
    // n at n * ncell, cell col + Nx*Ny*k): FP32 the step's increment so far,
    // both the previous stage's tendencies
    int stage;
    int co2_stage;
    const double* sflux;           // the stage's top-cell CO2 flux per column
    unsigned ncell;
    FT* Dsc;
    FT* Gsc;
};

// Per-level geometry rows in `lev` (each Nz entries).
enum { LEV_ZC = 0, LEV_UP, LEV_DN, LEV_INVDZ, LEV_COUNT };
// LEV_UP[k] = zc[k+1] - zf[k+1], LEV_DN[k] = zf[k+1] - zc[k] for k < Nz-1;
// for k = Nz-1: LEV_UP = 0 and LEV_DN = zf[Nz] - zc[Nz-1] (the top half cell).

// Modes of the column sweep: the fused RK3 step, the tendencies (hook), the
// PAR (hook), one RK3 stage with its substep (the step of the sinking schemes
// whose face values need the levels below).
enum { TILE_STEP = 0, TILE_TENDENCY = 1, TILE_PAR = 2, TILE_STAGE = 3 };

// The value at face k of a sinking tracer reconstructed from above (w <= 0:
// the upwind cell is k): fifth-order WENO from the cells k-2 .. k+2, third
// order from k-1 .. k+1, with WENO-Z weights (Borges et al. 2008) on the
// Jiang-Shu smoothness indicators, eps = 1e-8: the reconstructions of
// Oceananigans' WENO(order = 5), whose order drops near a Bounded
// boundary (WENO5 on faces 3 .. Nz-3, WENO3 on 2 and Nz-2, upwind on 1 and
// Nz-1).
template <typename FT>
__device__ __forceinline__ FT weno5_from_above(FT cm2, FT cm1, FT c0, FT cp1, FT cp2)
{
    const FT e = FT(1e-8);
    const FT q0 = (FT(2) * cp2 - FT(7) * cp1 + FT(11) * c0) / FT(6);
    const FT q1 = (-cp1 + FT(5) * c0 + FT(2) * cm1) / FT(6);
    const FT q2 = (FT(2) * c0 + FT(5) * cm1 - cm2) / FT(6);
    const FT s0 = cp2 - FT(2) * cp1 + c0, t0 = cp2 - FT(4) * cp1 + FT(3) * c0;
    const FT s1 = cp1 - FT(2) * c0 + cm1, t1 = cp1 - cm1;
    const FT s2 = c0 - FT(2) * cm1 + cm2, t2 = FT(3) * c0 - FT(4) * cm1 + cm2;
    const FT b0 = FT(13.0 / 12) * (s0 * s0) + FT(0.25) * (t0 * t0);
    const FT b1 = FT(13.0 / 12) * (s1 * s1) + FT(0.25) * (t1 * t1);
    const FT b2 = FT(13.0 / 12) * (s2 * s2) + FT(0.25) * (t2 * t2);
    const FT tau = fabs(b0 - b2);
    const FT r0 = tau / (b0 + e), r1 = tau / (b1 + e), r2 = tau / (b2 + e);
    const FT a0 = FT(0.1) * (FT(1) + r0 * r0);
    const FT a1 = FT(0.6) * (FT(1) + r1 * r1);
    const FT a2 = FT(0.3) * (FT(1) + r2 * r2);
    return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}

template <typename FT>
__device__ __forceinline__ FT weno3_from_above(FT cm1, FT c0, FT cp1)
{
    const FT e = FT(1e-8);
    const FT q0 = (-cp1 + FT(3) * c0) / FT(2);
    const FT q1 = (c0 + cm1) / FT(2);
    const FT d0 = c0 - cp1, d1 = cm1 - c0;
    const FT b0 = d0 * d0, b1 = d1 * d1;
    const FT tau = fabs(b0 - b1);
    const FT r0 = tau / (b0 + e), r1 = tau / (b1 + e);
    const FT a0 = FT(1.0 / 3) * (FT(1) + r0 * r0);
    const FT a1 = FT(2.0 / 3) * (FT(1) + r1 * r1);
    return (a0 * q0 + a1 * q1) / (a0 + a1);
}

// The face value at the bottom face k of cell k from the column's values
// cm2 = c(k-2) .. cp2 = c(k+2) (entries past the column unused).
template <int SCHEME, typename FT>
__device__ __forceinline__ FT sink_face(int k, int Nz, FT cm2, FT cm1, FT c0, FT cp1, FT cp2)
{
    if (k <= 0) return c0;
    if constexpr (SCHEME == OBME_CENTERED2) {
        return (cm1 + c0) / FT(2);
    } else {
        if (k == 1 || k >= Nz - 1) return c0;
        if (k == 2 || k == Nz - 2) return weno3_from_above(cm1, c0, cp1);
        return weno5_from_above(cm2, cm1, c0, cp1, cp2);
    }
}

// The column sweep.  Lane = water column: a workgroup of NW waves owns
// NW*64 consecutive i-columns of one j row, so every load and store of a
// level is one contiguous run of 64 values per wave (coalesced straight
// from HBM, no staging).  Each lane walks its column from the surface down,
// one level at a time, and at each level runs the RK3 stages 1, 2, 3 in
// turn: stage s at level k needs only stage-s values of the levels above
// it (the chlorophyll integral, the upwind sinking inflow), which the lane
// carries down as running sums and one level of history.  The loads of the
// next level are issued before the current level's arithmetic.
//
// EXTRA (fused step only), a bit set: 1 negative scaling at the start of
// every stage; 2 the CO2 flux computed here at the top level (with 1);
// 4 the InstantRemineralisation sediment at the bottom level.  MASKX: some
// lanes lie past the last column (Nx is not a multiple of the workgroup's
// width); false compiles the stores without a mask.  PARK: the FP32
// chlorophyll integrals are Kahan-compensated running sums.  SCHEME: the
// sinking face value, upwind (the fused step), or centred / WENO5, whose
// face values need the levels below: those run one stage per launch
// (TILE_STAGE) or as the tendency hook, reading the levels below a level
// ahead (the inputs of the stage, not yet overwritten) and keeping the two
// above in registers.
template <int MODEL, int CARB, typename FT, int MODE, bool UNIFORM, int FM = -1, int EXTRA = 0, bool MASKX = true,
          bool PARK = true, int SCHEME = OBME_UPWIND1>
__device__ __forceinline__ void column_lane(const StepArgs<FT>& a, const int cw)
{
    using TR = ModelTraits<MODEL, CARB>;
    constexpr int NT = TR::NT;
    constexpr int RO = TR::RO_IDX;
    constexpr int S0 = TR::SINK0, S1 = TR::SINK1, PI = TR::P_IDX;
    constexpr bool F32 = sizeof(FT) == 4;
    constexpr int NSTAGE = (MODE == TILE_STEP) ? 3 : 1;
    constexpr bool TS = MODE == TILE_STAGE;
    constexpr bool WIN = SCHEME != OBME_UPWIND1;
    // the CO2 flux computed here at the top level (with the scaling)
    constexpr bool INLINE_CO2 = (EXTRA & 2) != 0;
    static_assert(EXTRA == 0 || MODE == TILE_STEP, "scaling, inline CO2, sediment: fused step");
    static_assert(!(EXTRA & 2) || (CARB & 1), "inline CO2 needs the carbonate system");
    static_assert(!WIN || MODE == TILE_STAGE || MODE == TILE_TENDENCY, "centred and WENO sinking: stage or hook mode");
    static_assert(!TS || WIN, "the stage mode serves the centred and WENO sinking");

    const unsigned nbx = ((unsigned)a.Nx + (unsigned)cw - 1) / (unsigned)cw;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int j = (int)(tile / nbx);
    const int i_raw = (int)(tile % nbx) * cw + (int)threadIdx.x;
    // (MASKX false: the launch guarantees every lane holds a column)
    const bool col_ok = !MASKX || i_raw < a.Nx;
    // a lane past the last column computes on the last column and stores nothing
    const int i = col_ok ? i_raw : a.Nx - 1;
    const int Nz = a.Nz;
    const int ktop = Nz - 1;
    constexpr unsigned ES = sizeof(FT);
    const unsigned dk = a.sxy * ES;
    const unsigned off_top = (a.base0 + (unsigned)i + a.sx * (unsigned)j + a.sxy * (unsigned)ktop) * ES;
    const unsigned col2d = (unsigned)i + (unsigned)a.Nx * (unsigned)j;
    const unsigned ncol2d = (unsigned)a.Nx * (unsigned)a.Ny;
    // byte offset of level k of this column in every parent array (the
    // same for all of them): a 32-bit per-lane offset from each array's
    // base address, which stays in two scalar registers
    auto lev_off = [&](int k) -> unsigned { return off_top - (unsigned)(ktop - k) * dk; };

    // loads of one level: the PAR mode reads P alone
    auto load_level = [&](FT (&dst)[NT], int k) {
        const unsigned off = lev_off(k);
#pragma unroll
        for (int n = 0; n < NT; n++) {
            if (MODE == TILE_PAR && n != PI) dst[n] = FT(0);
            else dst[n] = gload(a.tr[n], off);
        }
    };

    FT PAR0_const = a.surface_value;
    if (a.surface_mode == OBME_SURFACE_ARRAY) PAR0_const = a.surf[col2d];

    // carried down the column, per stage: the stage values of the sinking
    // tracers and the chlorophyll terms (P R/r)^e of the level above, and
    // the running chlorophyll integrals (with their compensation)
    FT Xup[NSTAGE][2], pup[NSTAGE][2], Ir[NSTAGE], Ib[NSTAGE], Cr[NSTAGE], Cb[NSTAGE];
#pragma unroll
    for (int s = 0; s < NSTAGE; s++) {
        Xup[s][0] = Xup[s][1] = FT(0);
        pup[s][0] = pup[s][1] = FT(0);
        Ir[s] = Ib[s] = Cr[s] = Cb[s] = FT(0);
    }

    // inline CO2: [H+] of the previous stage (the Newton's warm start)
    double Hc = 0.0;
    if constexpr (INLINE_CO2) {
        if (col_ok && a.Hprev) Hc = a.Hprev[col2d];
    }
    // the stage mode: this launch's stage, and the scratch of level k at
    // byte offset coff(k) from each tracer's scratch row
    const int stg = TS ? a.stage : 0;
    auto coff = [&](int k) -> unsigned { return (col2d + ncol2d * (unsigned)k) * ES; };
    auto load_scratch = [&](FT (&dd)[NT], FT (&gg)[NT], int k) {
        const unsigned off = coff(k);
#pragma unroll
        for (int n = 0; n < NT; n++) {
            dd[n] = (F32 && stg > 0 && n != RO) ? gload(a.Dsc + (size_t)n * a.ncell, off) : FT(0);
            gg[n] = (stg > 0 && n != RO) ? gload(a.Gsc + (size_t)n * a.ncell, off) : FT(0);
        }
    };
    // centred / WENO sinking: the sinking tracers' stage inputs of the two
    // levels above (hs1: k+1, hs2: k+2), the face value of the face above,
    // and the loads of the level two below
    FT hs1[2] = {FT(0), FT(0)}, hs2[2] = {FT(0), FT(0)}, fup[2] = {FT(0), FT(0)};
    auto load_sink_inputs = [&](FT (&dst)[2], int k) {
        const unsigned off = lev_off(k);
        dst[0] = gload(a.tr[S0], off);
        dst[1] = gload(a.tr[S1], off);
        if constexpr (TS && F32) {
            if (stg > 0) {
                dst[0] += gload(a.Dsc + (size_t)S0 * a.ncell, coff(k));
                dst[1] += gload(a.Dsc + (size_t)S1 * a.ncell, coff(k));
            }
        }
    };
    // sediment (EXTRA & 4): the storage value, the tendency e f of the last
    // stage start and the G⁻ of the sediment's previous step
    // (fields: InstantRemineralisation's storage, SimpleMultiG's Ns, Nf, Nr),
    // loaded just before the bottom level, the only one that uses them
    double sed_u[3] = {0.0, 0.0, 0.0}, sed_ge[3] = {0.0, 0.0, 0.0}, sed_gm[3] = {0.0, 0.0, 0.0};
    auto load_sediment = [&] {
        if (col_ok) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
                if (q < a.sed_nf && a.storage) sed_u[q] = a.storage[q * ncol2d + col2d];
                if (q < a.sed_nf && a.sed_ab2) sed_gm[q] = a.sedG[(a.sed_nf + q) * ncol2d + col2d];
            }
        }
    };

    const double g1 = 8.0 / 15.0, g2 = 5.0 / 12.0, g3 = 3.0 / 4.0, z2 = -17.0 / 60.0, z3 = -5.0 / 12.0;

    FT nxt[NT];
    FT nxtD[TS ? NT : 1], nxtG[TS ? NT : 1];
    load_level(nxt, ktop);
    if constexpr (TS) load_scratch(nxtD, nxtG, ktop);
    // one level, all stages; top: the surface level, bot: the bottom level.
    // The plain fused step passes them as compile-time constants and
    // compiles the surface and the bottom level apart, so the levels between
    // carry none of their special cases and prefetch the next level
    // unconditionally; the other variants pass run-time flags.
    auto level = [&](const int k, auto top_c, auto bot_c) {
        const bool top = top_c;
        const bool bot = bot_c;
        // FP64: U is the state, updated in place (the reference's
        // expressions).  FP32: U is the state at the start of the step and D
        // the step's accumulated increment; stage inputs are U + D and the
        // sum is rounded once at the store (the increments are never rounded
        // against the magnitude of U).
        FT U[NT], D[F32 ? NT : 1], Gm[NT];
#pragma unroll
        for (int n = 0; n < NT; n++) {
            U[n] = nxt[n];
            Gm[n] = FT(0);
            if constexpr (F32) D[n] = FT(0);
            if constexpr (TS) {
                // the step's increment so far and the previous stage's tendencies
                if constexpr (F32) D[n] = nxtD[n];
                Gm[n] = nxtG[n];
            }
        }
        // centred / WENO: the sinking tracers' stage inputs of the level
        // two below (the level below comes with the prefetch)
        FT ah2[2] = {FT(0), FT(0)};
        if constexpr (WIN) {
            if (k >= 2) load_sink_inputs(ah2, k - 2);
        }
        if (!bot) {
            load_level(nxt, k - 1);
            if constexpr (TS) load_scratch(nxtD, nxtG, k - 1);
        }

        // geometry of level k (uniform across the wave)
        FT zc, up, dn, invdz, wlo0, whi0, wlo1, whi1;
        if constexpr (UNIFORM) {
            zc = a.zc0_u + FT(k) * a.dz_u;
            up = dn = a.dz_u * FT(0.5);
            invdz = a.invdz_u;
            wlo0 = whi0 = a.w0_u;
            wlo1 = whi1 = a.w1_u;
        } else {
            zc = a.lev[LEV_ZC * Nz + k];
            up = a.lev[LEV_UP * Nz + k];
            dn = a.lev[LEV_DN * Nz + k];
            invdz = a.lev[LEV_INVDZ * Nz + k];
            wlo0 = a.wsink[k];
            whi0 = a.wsink[k + 1];
            wlo1 = a.wsink[(Nz + 1) + k];
            whi1 = a.wsink[(Nz + 1) + k + 1];
        }
        float ezr = 0.0f, ezb = 0.0f;   // FP32: kr zc, kb zc (log2 scaled)
        if constexpr (F32) {
            ezr = a.kr2 * (float)zc;
            ezb = a.kb2 * (float)zc;
        }
        if constexpr ((CARB & 1) && MODE == TILE_STEP) {
            // the top-cell Alk carries its residual from the previous step
            // (the top-cell DIC is the surface kernel's: it reads the step's
            // start value from the tracer, which the sweep does not write)
            if constexpr (F32) {
                if (top && a.topres) D[8] = a.topres[ncol2d + col2d];
            }
        }
        if constexpr ((CARB & 1) && TS && F32) {
            // the stage mode's top-cell DIC and Alk: the first stage starts
            // from the residuals the previous step left
            if (top && a.topres && stg == 0) {
                D[7] = a.topres[col2d];
                D[8] = a.topres[ncol2d + col2d];
            }
        }

        FT PARv = FT(0);
        FT Gout[NT];   // TILE_TENDENCY: the transitions
#pragma unroll
        for (int si = 0; si < NSTAGE; si++) {
            // si: the carried per-stage sums; st: the RK3 stage (one per launch in the stage mode)
            const int st = TS ? stg : si;
            FT PAR0 = PAR0_const;
            if (a.surface_mode == OBME_SURFACE_DIURNAL) PAR0 = a.PAR0_stage[st];

            if constexpr ((EXTRA & 1) != 0) {
                // ScaleNegativeTracers first (update_biogeochemical_state!
                // order, src/OceanBioME.jl:161-167).  FP64 applies it to every
                // cell as the reference does; FP32-mixed where a member is
                // negative or not finite (elsewhere it is the identity up to
                // rounding), and the scaled value becomes the new U with no
                // pending increment.
                unsigned members = 0;
#pragma unroll 1
                for (int g = 0; g < a.ns.n_groups; g++) members |= a.ns.mask[g];
                FT x[NT];
#pragma unroll
                for (int n = 0; n < NT; n++) x[n] = (F32 && st > 0) ? U[n] + D[F32 ? n : 0] : U[n];
                if constexpr (F32) {
                    bool need = false;
#pragma unroll
                    for (int n = 0; n < NT; n++)
                        if (members & (1u << n)) need = need || !(x[n] >= 0.0f) || isinf(x[n]);
                    if (need) {
                        scale_negatives_cell<NT>(a.ns, x);
#pragma unroll
                        for (int n = 0; n < NT; n++)
                            if (members & (1u << n)) {
                                U[n] = x[n];
                                D[F32 ? n : 0] = FT(0);
                            }
                    }
                } else {
                    scale_negatives_cell<NT>(a.ns, x);
#pragma unroll
                    for (int n = 0; n < NT; n++)
                        if (members & (1u << n)) U[n] = x[n];
                }
            }

            // stage values of this level
            FT X[NT];
#pragma unroll
            for (int n = 0; n < NT; n++) {
                // (the top-cell Alk includes its residual from the first stage on)
                const bool inc = F32 && (st > 0 || ((CARB & 1) && MODE == TILE_STEP && top && n == 8) ||
                                         ((CARB & 1) && TS && top && (n == 7 || n == 8)));
                X[n] = inc ? U[n] + D[F32 ? n : 0] : U[n];
            }

            // ---- TwoBand PAR (Light/2band.jl:18-32): the chlorophyll
            // integral from the surface, ∫(k) = ∫(k+1) + up(k) p(k+1) + dn(k) p(k)
            // (the top level: dn p alone), PAR = PAR0 (e^(kr zc - χr ∫r) + e^(kb zc - χb ∫b)) / 2.
            FT pr, pb;
            if constexpr (F32) {
                const float lg = m_log2(X[PI]);
                pr = m_exp2(fmaf(a.er, lg, a.cr));
                pb = m_exp2(fmaf(a.eb, lg, a.cb));
            } else {
                // both bands' powers from one logarithm
                pow_pair(X[PI] * a.Rr, a.er, a.eb, pr, pb);
            }
            FT ir, ib;
            if constexpr (F32 && UNIFORM) {
                // dz/2 is folded into the attenuations (xr2h, xb2h)
                ir = top ? pr : pup[si][0] + pr;
                ib = top ? pb : pup[si][1] + pb;
            } else {
                ir = top ? dn * pr : up * pup[si][0] + dn * pr;
                ib = top ? dn * pb : up * pup[si][1] + dn * pb;
            }
            if constexpr (F32 && PARK) {
                // Kahan summation: every increment is >= 0, so the running
                // sum dominates and its lost low part is recovered exactly
                const FT yr = ir - Cr[si], yb = ib - Cb[si];
                const FT tr = Ir[si] + yr, tb = Ib[si] + yb;
                Cr[si] = (tr - Ir[si]) - yr;
                Cb[si] = (tb - Ib[si]) - yb;
                Ir[si] = tr;
                Ib[si] = tb;
            } else {
                Ir[si] = Ir[si] + ir;
                Ib[si] = Ib[si] + ib;
            }
            pup[si][0] = pr;
            pup[si][1] = pb;
            if constexpr (F32 && UNIFORM) {
                PARv = (PAR0 * 0.5f) * (m_exp2(fmaf(-a.xr2h, Ir[si], ezr)) + m_exp2(fmaf(-a.xb2h, Ib[si], ezb)));
            } else if constexpr (F32) {
                PARv = (PAR0 * 0.5f) * (m_exp2(fmaf(-a.xr2, Ir[si], ezr)) + m_exp2(fmaf(-a.xb2, Ib[si], ezb)));
            } else {
                PARv = PAR0 * (m_exp(a.kr * zc - a.xr * Ir[si]) + m_exp(a.kb * zc - a.xb * Ib[si])) / FT(2);
            }
            if constexpr (MODE == TILE_PAR) continue;

            // ---- the transitions
            FT Gv[NT];
            {
                FT v[NT];
#pragma unroll
                for (int n = 0; n < NT; n++) v[n] = (CARB != 0 && n >= 7) ? FT(0) : X[n];
                if constexpr (MODEL == OBME_NPZD) {
                    if constexpr (F32) {
                        const float Lt = fm_q10<FM>(a.f) ? m_exp2(v[3] * a.f.log2Q10_10) : 1.0f;
                        npzd_fast<FM>(a.f, v[0], v[1], v[2], v[4], PARv, Lt, Gv[0], Gv[1], Gv[2], Gv[4],
                                      (CARB & 2) ? &Gv[NT - 1] : nullptr);
                    } else {
                        FT Lt = FT(1);
                        if (a.q.use_Q10) Lt = m_pow(a.q.Q10, v[3] / FT(10));
                        npzd_tendencies(a.q, v[0], v[1], v[2], v[3], v[4], PARv, Lt, Gv[0], Gv[1], Gv[2], Gv[4],
                                        (CARB & 2) ? &Gv[NT - 1] : nullptr);
                    }
                    Gv[3] = FT(0);
                } else {
                    if constexpr (F32) lobster_fast<CARB, FM>(a.f, v, PARv, Gv);
                    else lobster_tendencies<FT, CARB>(a.q, v, PARv, Gv);
                }
            }
            // minus the upwind drift flux divergence, w(face) c(upwind cell):
            // the inflow through the top face is w c of the level above
            // (no flux through the surface: w = 0 there, and Xup = 0)
            if constexpr (WIN) {
                // centred / WENO: the face value of the bottom face from the
                // stage inputs of the levels k-2 .. k+2; the top face's is
                // the level above's bottom face
                FT cm1[2] = {nxt[S0], nxt[S1]};
                if constexpr (TS && F32) {
                    if (stg > 0) {
                        cm1[0] += nxtD[S0];
                        cm1[1] += nxtD[S1];
                    }
                }
                const FT c0[2] = {X[S0], X[S1]};
                FT f[2];
#pragma unroll
                for (int q = 0; q < 2; q++) f[q] = sink_face<SCHEME>(k, Nz, ah2[q], cm1[q], c0[q], hs1[q], hs2[q]);
                const FT Fhi0 = top ? FT(0) : whi0 * fup[0];
                const FT Fhi1 = top ? FT(0) : whi1 * fup[1];
                Gv[S0] = Gv[S0] - (Fhi0 - wlo0 * f[0]) * invdz;
                Gv[S1] = Gv[S1] - (Fhi1 - wlo1 * f[1]) * invdz;
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    fup[q] = f[q];
                    hs2[q] = hs1[q];
                    hs1[q] = c0[q];
                }
            } else if constexpr (F32 && UNIFORM) {
                Gv[S0] = fmaf(-a.wi0, (top ? 0.0f : Xup[si][0]) - X[S0], Gv[S0]);
                Gv[S1] = fmaf(-a.wi1, (top ? 0.0f : Xup[si][1]) - X[S1], Gv[S1]);
            } else {
                const FT Fhi0 = top ? FT(0) : whi0 * Xup[si][0];
                const FT Fhi1 = top ? FT(0) : whi1 * Xup[si][1];
                Gv[S0] = Gv[S0] - (Fhi0 - wlo0 * X[S0]) * invdz;
                Gv[S1] = Gv[S1] - (Fhi1 - wlo1 * X[S1]) * invdz;
            }
            Xup[si][0] = X[S0];
            Xup[si][1] = X[S1];

            if constexpr (INLINE_CO2) {
                // the top-cell CO2 flux (GasExchange, gas_exchange.jl:26-38) of
                // this stage's DIC and Alk, the surface kernel's expressions
                if (top) {
                    const unsigned tc = a.base0 + (unsigned)i + a.sx * (unsigned)j + a.sxy * (unsigned)ktop;
                    const double T = a.T3 ? (double)a.T3[tc] : (a.Tprof ? a.Tprof[ktop] : a.Tconst);
                    const double S = a.S3 ? (double)a.S3[tc] : (a.Sprof ? a.Sprof[ktop] : a.Sconst);
                    SurfaceCarbon sc;
                    surface_carbon_setup(a.co2, T, S, a.co2.wind_speed, sc);
                    double pH = 0.0;
                    int ev = 0;
                    const double F = surface_flux_warm(sc, a.solver, (double)X[7], (double)X[8],
                                                       a.co2.air_concentration, 1, Hc, pH, ev,
                                                       F32 ? FP32_NEWTON_TOL : 1e-14);
                    Gv[7] = Gv[7] - (FT)F * a.invdz_top;
                }
            }
            if constexpr (TS && (CARB & 1)) {
                // the stage mode: the top-cell CO2 flux of this stage's DIC
                // and Alk, solved per column before the launch (k_surface)
                if (top && a.co2_stage) Gv[7] = Gv[7] - (FT)a.sflux[col2d] * a.invdz_top;
            }
            if constexpr ((CARB & 2) != 0) {
                // the top-cell O2 flux (GasExchange for O2, gas_exchange.jl:26-38,
                // GasExchange.jl:127-146) of this stage's O2: G -= F / Δz
                if (a.o2_on && top) {
                    const FT kdz = (FT)a.o2c[2 * col2d], air = (FT)a.o2c[2 * col2d + 1];
                    Gv[NT - 1] = Gv[NT - 1] - kdz * (X[NT - 1] - air);
                }
            }
            if constexpr ((EXTRA & 4) != 0) {
                // the sediment at the bottom cell: the sinking flux through
                // the bottom face, f = Σ -w c (Sediments/tracked_fields.jl:53-70),
                // burial e = a + b (f / (k + f))^2 and the receiver's coupled
                // tendency (1 - e) f / Δz (instant_remineralisation.jl:101-123,
                // Sediments/tracer_coupling.jl:30-38)
                if (bot && a.sed_model == OBME_SEDIMENT_SIMPLE_MULTI_G) {
                    // SimpleMultiG: the sediment step after this stage's
                    // substep takes the tendencies evaluated after the
                    // previous step: from the fields now and this stage's
                    // flux.  The first step ever (after the first substep)
                    // evaluates them from the second stage's flux, and the
                    // coupled tendencies of the second stage see its result.
                    const double fd = (double)(a.sed_w0 * X[S0] + a.sed_w1 * X[S1]);
                    double G[3];
                    if (st == 1 && a.sed_init) {
                        smg_tendencies(a.smg, sed_u, fd, G);
#pragma unroll
                        for (int q = 0; q < 3; q++) sed_u[q] += a.sed_dt[0] * G[q];
                    }
                    double vno3 = 0.0, vnh4 = 0.0, vo2 = 0.0;
#pragma unroll
                    for (int n = 0; n < NT; n++) {
                        if (n == a.smg.no3) vno3 = (double)X[n];
                        if (n == a.smg.nh4) vnh4 = (double)X[n];
                        if (n == a.smg.o2) vo2 = (double)X[n];
                    }
                    double cnh4, cno3, co2;
                    smg_coupled(a.smg, sed_u, vno3, vnh4, vo2, cnh4, cno3, co2);
                    const double dz0 = (double)a.sed_dz0;
#pragma unroll
                    for (int n = 0; n < NT; n++) {
                        if (n == a.smg.nh4) Gv[n] = Gv[n] + (FT)(cnh4 / dz0);
                        if (n == a.smg.no3) Gv[n] = Gv[n] + (FT)(cno3 / dz0);
                        if (n == a.smg.o2) Gv[n] = Gv[n] + (FT)(co2 / dz0);
                    }
                    smg_tendencies(a.smg, sed_u, fd, G);
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        if (st == 0) {
                            if (!a.sed_init)
                                sed_u[q] += a.sed_dt[0] * (a.sed_ab2 ? a.sed_cn * G[q] - a.sed_cm * sed_gm[q] : G[q]);
                        } else {
                            sed_u[q] += a.sed_dt[st] * G[q];
                        }
                        sed_ge[q] = G[q];
                    }
                } else if (bot) {
                    const FT f = a.sed_w0 * X[S0] + a.sed_w1 * X[S1];
                    const FT qf = f / (a.sed_k + f);
                    const FT e = a.sed_a + a.sed_b * (qf * qf);
                    const FT cpl = ((FT(1) - e) * f) / a.sed_dz0;
#pragma unroll
                    for (int n = 0; n < NT; n++)
                        if (n == a.sed_recv) Gv[n] = Gv[n] + cpl;
                    // the sediment step after this stage's substep
                    // (update_state.jl:6-17) takes the tendency evaluated
                    // after the previous step, from this stage-start state:
                    // e f here.  Its first step ever evaluates the tendency
                    // lazily, after the first substep: the tendency of the
                    // second stage start serves both.
                    const double ge = (double)(e * f);
                    if (st == 0) {
                        if (!a.sed_init)
                            sed_u[0] += a.sed_dt[0] * (a.sed_ab2 ? a.sed_cn * ge - a.sed_cm * sed_gm[0] : ge);
                    } else {
                        if (st == 1 && a.sed_init) sed_u[0] += a.sed_dt[0] * ge;
                        sed_u[0] += a.sed_dt[st] * ge;
                    }
                    sed_ge[0] = ge;
                }
            }
            if constexpr (MODE == TILE_TENDENCY) {
#pragma unroll
                for (int n = 0; n < NT; n++) Gout[n] = Gv[n];
            } else {
                if constexpr (CARB & 1) {
                    // surface record: DIC transition without the flux, Alk at the stage start
                    if (top && col_ok && a.co2_on) {
                        a.srec[st * ncol2d + col2d] = Gv[7];
                        a.srec[(3 + st) * ncol2d + col2d] = X[8];
                    }
                }
                // RK3 substep (BoxModel/timesteppers.jl:79-92 form):
                // stage 1 U += (Δt γ1) G; stages 2, 3 U += Δt (γ G + ζ G⁻)
#pragma unroll
                for (int n = 0; n < NT; n++) {
                    if (n == RO) continue;
                    if constexpr (F32) {
                        if (st == 0) D[n] = fmaf(a.rk_a[0], Gv[n], D[n]);
                        else D[n] = fmaf(a.rk_a[st], Gv[n], fmaf(a.rk_b[st], Gm[n], D[n]));
                    } else {
                        const double gs = (st == 1) ? g2 : g3;
                        const double zs = (st == 1) ? z2 : z3;
                        const double inc = (st == 0) ? (a.dtf * g1) * Gv[n] : a.dtf * (gs * Gv[n] + zs * Gm[n]);
                        U[n] += inc;
                    }
                    Gm[n] = Gv[n];
                }
            }
        }

        // ---- the level's outputs.  A lane past the last column stores
        // nothing: with MASKX its stores go through a buffer resource at an
        // offset past the array's end, which drops them.  No store sits behind
        // a branch, so the next level's wait for its prefetched loads never
        // has to wait for these stores as well.
        const unsigned off = lev_off(k);
        auto put = [&](FT* p, FT v) {
            if constexpr (MASKX) buf_store(make_rsrc(p, a.field_bytes), col_ok ? off : a.field_bytes, v);
            else gstore(p, off, v);
        };
        if constexpr (MODE == TILE_PAR) {
            put(a.PAR, PARv);
        } else if constexpr (TS) {
            // the stage's substep: FP64 the state in place; FP32 the step's
            // increment to the scratch, the state U + D after the last stage;
            // this stage's tendencies for the next
            const unsigned co = col_ok ? coff(k) : a.ncell * ES;
            auto putc = [&](FT* p, FT v) { buf_store(make_rsrc(p, a.ncell * ES), co, v); };
#pragma unroll
            for (int n = 0; n < NT; n++) {
                if (n == RO) continue;
                if (stg < 2) putc(a.Gsc + (size_t)n * a.ncell, Gm[n]);
                if constexpr (F32) {
                    if (stg < 2) {
                        putc(a.Dsc + (size_t)n * a.ncell, D[n]);
                    } else if ((CARB & 1) && top && a.topres && (n == 7 || n == 8)) {
                        // U + D = s + e exactly (TwoSum): s stored, e to the next step
                        const float u = U[n], d = D[n], sum = u + d, dv = sum - u;
                        if (col_ok) a.topres[(n == 8 ? ncol2d : 0u) + col2d] = (u - (sum - dv)) + (d - dv);
                        put(a.tr[n], sum);
                    } else {
                        put(a.tr[n], U[n] + D[n]);
                    }
                } else {
                    put(a.tr[n], U[n]);
                }
            }
        } else if constexpr (MODE == TILE_TENDENCY) {
            FT g0[NT];
#pragma unroll
            for (int n = 0; n < NT; n++)
                if (n != RO) g0[n] = gload(a.G[n], off);
#pragma unroll
            for (int n = 0; n < NT; n++)
                if (n != RO) put(a.G[n], g0[n] + Gout[n]);
            if (a.PAR) put(a.PAR, PARv);
        } else {
#pragma unroll
            for (int n = 0; n < NT; n++) {
                if (n == RO) continue;
                // the top-cell DIC is written by the surface kernel
                if constexpr ((CARB & 1) != 0) {
                    if (top && n == 7 && a.co2_on) continue;
                    // the top-cell Alk: U + D = s + e exactly (TwoSum), e to the next step
                    if constexpr (F32) {
                        if (top && n == 8 && a.topres) {
                            const float u = U[8], d = D[8], sum = u + d, dv = sum - u;
                            if (col_ok) a.topres[ncol2d + col2d] = (u - (sum - dv)) + (d - dv);
                            put(a.tr[8], sum);
                            continue;
                        }
                    }
                }
                put(a.tr[n], F32 ? U[n] + D[F32 ? n : 0] : U[n]);
            }
            if constexpr ((EXTRA & 4) != 0) {
                // the sediment after the last stage: storage, G⁻ = the tendency
                // of its last step, Gⁿ = e f of the final state
                if (bot && col_ok) {
                    const FT x0 = F32 ? U[S0] + D[F32 ? S0 : 0] : U[S0];
                    const FT x1 = F32 ? U[S1] + D[F32 ? S1 : 0] : U[S1];
                    const FT f = a.sed_w0 * x0 + a.sed_w1 * x1;
                    if (a.sed_model == OBME_SEDIMENT_SIMPLE_MULTI_G) {
                        double G[3];
                        smg_tendencies(a.smg, sed_u, (double)f, G);
#pragma unroll
                        for (int q = 0; q < 3; q++) {
                            if (a.storage) a.storage[q * ncol2d + col2d] = sed_u[q];
                            a.sedG[q * ncol2d + col2d] = G[q];
                            a.sedG[(3 + q) * ncol2d + col2d] = sed_ge[q];
                        }
                    } else {
                        const FT qf = f / (a.sed_k + f);
                        const FT e = a.sed_a + a.sed_b * (qf * qf);
                        if (a.storage) a.storage[col2d] = sed_u[0];
                        a.sedG[col2d] = (double)(e * f);
                        a.sedG[ncol2d + col2d] = sed_ge[0];
                    }
                }
            }
        }
    };
    using T_ = std::integral_constant<bool, true>;
    using F_ = std::integral_constant<bool, false>;
    // The workgroup's waves walk their columns level by level in step: at
    // each level their loads of a tracer (256 B per wave) then reach the
    // memory together as one run of 256 NW contiguous bytes.  (A barrier
    // only: no memory fence, so the prefetched loads stay in flight.  Every
    // wave of the workgroup runs every level: no wave leaves early.)
    auto level_barrier = [] { __builtin_amdgcn_s_barrier(); };
    if constexpr (MODE == TILE_STEP && EXTRA == 0) {
        if (Nz == 1) {
            level(0, T_{}, T_{});
        } else {
            level(ktop, T_{}, F_{});
            // (unrolled by two: the prefetched and the current level's
            // registers alternate instead of being copied)
#pragma unroll 2
            for (int k = ktop - 1; k >= 1; k--) {
                level_barrier();
                level(k, F_{}, F_{});
            }
            level_barrier();
            level(0, F_{}, T_{});
        }
    } else if constexpr ((EXTRA & 4) != 0 && !((EXTRA & 1) != 0 && (CARB & 1) != 0)) {
        // the levels above the bottom compiled apart from it: the sediment's
        // state and code are not carried through the column walk (with the
        // scaling and the carbonate system together this form spills more
        // than the single loop below, which those variants keep)
#pragma unroll 1
        for (int k = ktop; k >= 1; k--) {
            level_barrier();
            level(k, k == ktop, F_{});
        }
        load_sediment();
        level_barrier();
        level(0, ktop == 0, T_{});
    } else {
        if constexpr ((EXTRA & 4) != 0) load_sediment();
#pragma unroll 1
        for (int k = ktop; k >= 0; k--) {
            level_barrier();
            level(k, k == ktop, k == 0);
        }
    }
    if constexpr (INLINE_CO2) {
        if (col_ok && a.Hprev) a.Hprev[col2d] = Hc;
    }
}

// ------------------------------------------------------------------ surface

template <typename FT>
struct SurfaceArgs {
    FT* DIC;                       // DIC parent array (top cell written, step mode)
    FT* GDIC;                      // DIC tendency parent array (tendency mode)
    const FT* Alk;                 // Alk parent array (tendency mode)
    const FT* srec;                // surface records (step mode)
    float* topres;                 // FP32 step mode: the top-cell DIC residuals [Nx*Ny], or null
    double* Hprev;                 // [Nx*Ny] warm start of [H+] or null
    const double* surfc;           // [SURF_PACK][surfc_cols] surface constants of the call (step mode) or null
    int surfc_cols;                // Nx*Ny, or 1 where T and S are the same in every column (profile, constant)
    const FT* T3;                  // 3-D temperature / salinity or null
    const FT* S3;
    const double* Tprof;
    const double* Sprof;
    double Tconst, Sconst;
    double* flux_out;              // optional diagnostics (last stage)
    double* pH_out;
    int Nx, Ny, Nz;
    unsigned sx, sxy, base0;
    FT dtf, invdz_top;
    float rk_a[3], rk_b[3];        // FP32 path: Δt γ and Δt ζ per stage
    int mode;                      // 0: reference damped Newton; 1: warm-started Newton (fallback: 0)
    obme_co2_exchange co2;
    obme_carbon_solver solver;
    // the stage mode (TILE_STAGE): the stage, and FP32 the step's increments so far
    int stage;
    const FT* Dsc;
    unsigned ncell;
};

// The surface constants of every column (T and S of the top cell), once per
// stepper call, for k_surface: a.surfc_cols columns (one where T and S do
// not vary across the columns).
template <typename FT>
__global__ void __launch_bounds__(256) k_surface_prep(SurfaceArgs<FT> a, double* out)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = a.surfc_cols;
    if (col >= ncol) return;
    const int i = col % a.Nx, j = col / a.Nx;
    const int ktop = a.Nz - 1;
    const unsigned top = a.base0 + (unsigned)i + a.sx * (unsigned)j + a.sxy * (unsigned)ktop;
    const double T = a.T3 ? (double)a.T3[top] : (a.Tprof ? a.Tprof[ktop] : a.Tconst);
    const double S = a.S3 ? (double)a.S3[top] : (a.Sprof ? a.Sprof[ktop] : a.Sconst);
    SurfaceCarbon sc;
    surface_carbon_setup(a.co2, T, S, a.co2.wind_speed, sc);
    double o[SURF_PACK];
    surface_pack(sc, 0.0, 0.0, o);
#pragma unroll
    for (int q = 0; q < SURF_PACK; q++) out[(long)q * ncol + col] = o[q];
}

template <typename FT, int MODE>
__global__ void __launch_bounds__(256) k_surface(SurfaceArgs<FT> a)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = a.Nx * a.Ny;
    if (col >= ncol) return;
    const int i = col % a.Nx, j = col / a.Nx;
    const int ktop = a.Nz - 1;
    const unsigned top = a.base0 + (unsigned)i + a.sx * (unsigned)j + a.sxy * (unsigned)ktop;
    SurfaceCarbon sc;
    if (a.surfc) {
        // T and S are fixed within a stepper call: the constants were set up once
        double o[SURF_PACK], kO2, airO2;
        const long cidx = (a.surfc_cols == 1) ? 0 : col;
#pragma unroll
        for (int q = 0; q < SURF_PACK; q++) o[q] = a.surfc[(long)q * a.surfc_cols + cidx];
        surface_unpack(o, sc, kO2, airO2);
    } else {
        const double T = a.T3 ? (double)a.T3[top] : (a.Tprof ? a.Tprof[ktop] : a.Tconst);
        const double S = a.S3 ? (double)a.S3[top] : (a.Sprof ? a.Sprof[ktop] : a.Sconst);
        surface_carbon_setup(a.co2, T, S, a.co2.wind_speed, sc);
    }
    double H = a.Hprev ? a.Hprev[col] : 0.0;
    double pH = 0.0, F = 0.0;
    int ev = 0;
    if constexpr (MODE == TILE_TENDENCY) {
        F = surface_flux_warm(sc, a.solver, (double)a.DIC[top], (double)a.Alk[top], a.co2.air_concentration, a.mode,
                              H, pH, ev, sizeof(FT) == 4 ? FP32_NEWTON_TOL : 1e-14);
        a.GDIC[top] = a.GDIC[top] - (FT)F * a.invdz_top;
    } else if constexpr (MODE == TILE_STAGE) {
        // the flux of the stage's top-cell DIC and Alk (FP32: U + D exactly;
        // the first stage's D is the residual the previous step left), into
        // flux_out for the sweep
        double dic = (double)a.DIC[top], alk = (double)a.Alk[top];
        if constexpr (sizeof(FT) == 4) {
            const unsigned ct = (unsigned)col + (unsigned)ncol * (unsigned)ktop;
            if (a.stage > 0) {
                dic += (double)a.Dsc[7 * (size_t)a.ncell + ct];
                alk += (double)a.Dsc[8 * (size_t)a.ncell + ct];
            } else if (a.topres) {
                dic += (double)a.topres[col];
                alk += (double)a.topres[ncol + col];
            }
        }
        F = surface_flux_warm(sc, a.solver, dic, alk, a.co2.air_concentration, a.mode, H, pH, ev,
                              sizeof(FT) == 4 ? FP32_NEWTON_TOL : 1e-14);
    } else {
        // the identical update expressions of the sweep, with the flux term;
        // the step starts from the top-cell DIC the previous step's surface
        // kernel wrote (the sweep leaves it alone)
        FT U = a.DIC[top];
        FT Gm = FT(0);
        if constexpr (sizeof(FT) == 4) {
            // the increment starts from the residual the previous step left
            // (DIC = U + D carried to about twice FP32's precision in a call)
            float D = a.topres ? a.topres[col] : 0.0f;
#pragma unroll
            for (int st = 0; st < 3; st++) {
                const float Gb = a.srec[st * ncol + col];
                const float Alk = a.srec[(3 + st) * ncol + col];
                F = surface_flux_warm(sc, a.solver, (double)U + (double)D, (double)Alk, a.co2.air_concentration,
                                      a.mode, H, pH, ev, FP32_NEWTON_TOL);
                const float G = Gb - (float)F * a.invdz_top;
                D = fmaf(a.rk_a[st], G, fmaf(a.rk_b[st], Gm, D));
                Gm = G;
            }
            // U + D = s + e exactly (TwoSum): s stored, e to the next step
            const float sum = U + D, dv = sum - U;
            if (a.topres) a.topres[col] = (U - (sum - dv)) + (D - dv);
            U = sum;
        } else {
            const FT g1 = FT(8.0 / 15.0), g2 = FT(5.0 / 12.0), g3 = FT(3.0 / 4.0);
            const FT z2 = FT(-17.0 / 60.0), z3 = FT(-5.0 / 12.0);
#pragma unroll
            for (int st = 0; st < 3; st++) {
                FT Gb = a.srec[st * ncol + col];
                FT Alk = a.srec[(3 + st) * ncol + col];
                F = surface_flux_warm(sc, a.solver, (double)U, (double)Alk, a.co2.air_concentration, a.mode, H, pH,
                                      ev);
                FT G = Gb - (FT)F * a.invdz_top;
                if (st == 0) U += a.dtf * g1 * G;
                else if (st == 1) U += a.dtf * (g2 * G + z2 * Gm);
                else U += a.dtf * (g3 * G + z3 * Gm);
                Gm = G;
            }
        }
        a.DIC[top] = U;
    }
    if (a.Hprev) a.Hprev[col] = H;
    if (a.flux_out) a.flux_out[col] = F;
    if (a.pH_out) a.pH_out[col] = pH;
}

}  // namespace obme
