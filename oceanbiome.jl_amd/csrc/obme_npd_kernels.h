// obme_npd_kernels.h — the column-sweep and surface kernels of the NPZD /
// LOBSTER biogeochemistry step on gfx950.
//
// Column sweep (column_wave): wave per column.  One water column lives
// across the lanes of one wavefront segment of LPC = 64/CPW lanes; lane r
// holds L consecutive levels counted from the surface,
//   k = Nz-1 - (r*L + m),  m = 0..L-1  (0-based, k = 0 at the bottom),
// so lane 0 holds the top cell.  A workgroup of NW waves owns TXB = NW*CPW
// consecutive i-columns of one j row.  Tracers are staged through LDS once
// per step: coalesced loads along i into padded [Nz][TXB+1] tiles, then
// each lane gathers its levels (and the reverse for the stores), so HBM
// sees one coalesced read and write of each written tracer per step.  All
// state stays in registers through the three RK3 stages and the stages need
// no workgroup barrier:
//  - the TwoBand chlorophyll integral (Light/2band.jl:18-32) is a suffix sum
//    from the surface: each lane sums its own increments top-down, a
//    segmented prefix scan over the lanes (DPP row shifts and broadcasts)
//    gives the sum above it, and the reference recurrence
//    ∫(k) = ∫(k+1) + inc(k) continues from there;
//  - upwind sinking needs c(k+1): the level above a lane's top level is the
//    bottom level of the lane before it, one lane shift;
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
// [0] DIC at the start of the step, [1+s] DIC transition without the flux
// at stage s, [4+s] Alk at the start of stage s (s = 0, 1, 2).
constexpr int SURF_REC = 7;

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
};

// Per-level geometry rows in `lev` (each Nz entries).
enum { LEV_ZC = 0, LEV_UP, LEV_DN, LEV_INVDZ, LEV_COUNT };
// LEV_UP[k] = zc[k+1] - zf[k+1], LEV_DN[k] = zf[k+1] - zc[k] for k < Nz-1;
// for k = Nz-1: LEV_UP = 0 and LEV_DN = zf[Nz] - zc[Nz-1] (the top half cell).

// Modes of the column sweep.
enum { TILE_STEP = 0, TILE_TENDENCY = 1, TILE_PAR = 2 };

// s_waitcnt with only the vector-memory count constrained (gfx9 encoding:
// vmcnt bits [3:0] and [15:14], expcnt [6:4] = 7, lgkmcnt [11:8] = 15).
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}
// s_waitcnt lgkmcnt(0) (LDS and scalar memory), vector memory unconstrained.
__device__ __forceinline__ void wait_lgkm0()
{
    __builtin_amdgcn_s_waitcnt(15 | (3 << 14) | (7 << 4));
}

// Pipelined sweep (PIPE): a persistent workgroup walks a run of tiles and
// loads tile n+1 straight into LDS (buffer_load ... lds, no registers)
// while it computes tile n.  LDS layout of one buffer, per tracer: blocks of
// 64/TXB levels x TXB columns, 64 elements each (one load instruction of one
// wave: lane l holds level l / TXB, column l % TXB of the block, so the
// global reads are coalesced along i), block kb at element kb * (64 + 2): the
// skew of 2 per block puts the lanes of a column gather (levels two apart,
// two columns) in distinct banks.
constexpr int PIPE_SKEW = 2;

// One wave-wide load of 4-byte elements straight into LDS: lane l's element
// (byte offset voff in the buffer r; past its end reads 0) lands at
// dst + 4 l.  dst must be the same for every lane of the wave.
__device__ __forceinline__ void lds_load_b32(rsrc_t r, void* dst, unsigned voff)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)dst, 4, voff, 0, 0, 0);
}

// EXTRA (fused step only), a bit set: 1 negative scaling at the start of
// every stage; 2 the CO2 flux computed here by the top lane (with 1);
// 4 the InstantRemineralisation sediment at the bottom cell.
template <int MODEL, int CARB, typename FT, int L, int CPW, int NW, int MODE, bool UNIFORM, bool FULL,
          bool PIPE = false, int FM = -1, int EXTRA = 0>
__device__ __forceinline__ void column_wave(const StepArgs<FT>& a, FT* lds)
{
    using TR = ModelTraits<MODEL, CARB>;
    constexpr int NT = TR::NT;
    constexpr int RO = TR::RO_IDX;
    constexpr int LPC = 64 / CPW;          // lanes per column
    static_assert(LPC == 16 || LPC == 32 || LPC == 64, "segments of 16, 32 or 64 lanes");
    constexpr int TXB = NW * CPW;          // columns per workgroup
    constexpr int ROW = TXB + 1;           // padded LDS row: conflict-free gathers
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int r = lane % LPC;              // rank in the column, 0 = top
    const int c = w * CPW + lane / LPC;    // column in the workgroup
    const unsigned nbx = ((unsigned)a.Nx + TXB - 1) / TXB;
    const int Nz = FULL ? L * LPC : a.Nz;
    const int ktop = Nz - 1;

    // ---- PIPE geometry and the workgroup's run of tiles
    constexpr int PB = (64 % TXB == 0) ? 64 / TXB : 1;    // levels per LDS block
    constexpr int NBLK = (L * LPC) / PB;                   // blocks per tracer (FULL)
    constexpr int BLK = 64 + PIPE_SKEW;
    constexpr int TRS = NBLK * BLK;                        // elements per tracer
    constexpr int LPW = (ModelTraits<MODEL, CARB>::NT * NBLK) / NW;   // block loads per wave per tile
    static_assert(!PIPE || (FULL && MODE == TILE_STEP && sizeof(FT) == 4 && 64 % TXB == 0 && NBLK % NW == 0 &&
                            LPW * 2 < 64),
                  "pipelined sweep: FP32 fused step, FULL columns, TXB | 64, NW | blocks");
    unsigned t_first, t_step, t_end;
    if constexpr (PIPE) {
        // XCD x (workgroups b = x mod 8) takes a contiguous chunk of tiles;
        // its workgroups walk it with stride = their number, so the tiles in
        // flight on one XCD at a time are neighbours.
        const unsigned ntiles = nbx * (unsigned)a.Ny;
        const unsigned b = blockIdx.x, nb = gridDim.x, X = 8;
        const unsigned x = b % X, s = b / X;
        const unsigned nwx = nb / X + (x < nb % X ? 1u : 0u);
        const unsigned tq = ntiles / X, trem = ntiles % X;
        const unsigned start = x * tq + (x < trem ? x : trem);
        t_first = start + s;
        t_step = nwx;
        t_end = start + tq + (x < trem ? 1u : 0u);
    } else {
        t_first = xcd_tile(blockIdx.x, gridDim.x);
        t_step = 1;
        t_end = t_first + 1;
    }
    auto pipe_buf = [&](int n) -> FT* { return lds + (n & 1) * (ModelTraits<MODEL, CARB>::NT * TRS); };
    // byte offset of (ii, jj, k) in a parent array
    auto goff2 = [&](int ii, int jj, int k) -> unsigned {
        return (a.base0 + (unsigned)ii + a.sx * (unsigned)jj + a.sxy * (unsigned)k) * (unsigned)sizeof(FT);
    };
    // issue the LDS loads of tile t into buffer B (LPW instructions per wave)
    auto pipe_issue = [&](unsigned t, FT* B) {
        const int jj = (int)(t / nbx);
        const int ib = (int)(t % nbx) * TXB;
        const int kl = lane / TXB, cc = lane % TXB;
        const int ii = (ib + cc < a.Nx) ? ib + cc : a.Nx - 1;
#pragma unroll
        for (int n = 0; n < ModelTraits<MODEL, CARB>::NT; n++) {
            const rsrc_t rn = make_rsrc(a.tr[n], a.field_bytes);
#pragma unroll
            for (int q = 0; q < NBLK / NW; q++) {
                const int kb = w + NW * q;
                lds_load_b32(rn, B + n * TRS + kb * BLK, goff2(ii, jj, kb * PB + kl));
            }
        }
    };
    bool prev_tile = false;                // a tile was stored before this one
    bool next_issued = false;              // the tile after this one has been issued
    if constexpr (PIPE) {
        if (t_first < t_end) pipe_issue(t_first, pipe_buf(0));
        if (t_first + t_step < t_end) {
            pipe_issue(t_first + t_step, pipe_buf(1));
            next_issued = true;
        }
    }
    const FT surface_value0 = a.surface_value;
    int it = 0;
#pragma unroll 1
    for (unsigned wg_tile = t_first; wg_tile < t_end; wg_tile += t_step, it++) {
    const int j = (int)(wg_tile / nbx);
    const int ibase = (int)(wg_tile % nbx) * TXB;
    const int i_raw = ibase + c;
    const bool col_ok = i_raw < a.Nx;
    const int i = col_ok ? i_raw : a.Nx - 1;

    auto kof = [&](int m) -> int { return ktop - (r * L + m); };
    // FULL: Nz == L * LPC, every lane level is a cell of the column
    auto has = [&](int m) -> bool { return FULL || kof(m) >= 0; };
    auto rs = [&](const FT* p) { return make_rsrc(p, a.field_bytes); };
    // byte offset of (ii, j, k) in a parent array
    auto goff = [&](int ii, int k) -> unsigned {
        return (a.base0 + (unsigned)ii + a.sx * (unsigned)j + a.sxy * (unsigned)k) * (unsigned)sizeof(FT);
    };

    // ---- stage in: coalesced loads along i into [Nz][ROW] tiles, gather.
    // Element e of a tile is level e / TXB, column e % TXB; thread t handles
    // e = t + q * NW*64 for q < L (Nz * TXB <= L * NW * 64).  All loads of
    // all tracers are issued before the first LDS write.
    constexpr int NLOAD = (MODE == TILE_PAR) ? 1 : NT;
    constexpr int BS = NW * 64;
    auto tile = [&](int t) -> FT* { return lds + t * (Nz * ROW); };
    const int nel = Nz * TXB;
    // byte offset of staging element q (beyond the buffer when outside the
    // tile: reads 0) and its LDS index (-1 outside the tile)
    auto e_off = [&](int q) -> unsigned {
        const int e = threadIdx.x + q * BS;
        const int k = e / TXB, cc = e % TXB;
        const int ii = (ibase + cc < a.Nx) ? ibase + cc : a.Nx - 1;
        return (e < nel) ? goff(ii, k) : a.field_bytes;
    };
    auto e_idx = [&](int q) -> int {
        const int e = threadIdx.x + q * BS;
        return (e < nel) ? (e / TXB) * ROW + e % TXB : -1;
    };
    // PIPE: element (tracer n, level k, column cc) of the current buffer
    FT* const pbuf = pipe_buf(it);
    auto pidx = [&](int n, int k, int cc) -> int { return n * TRS + (k / PB) * BLK + (k % PB) * TXB + cc; };
    if constexpr (PIPE) {
        // this tile's loads are complete once at most the later ones are
        // outstanding: the previous tile's LPW stores and, if issued, the
        // next tile's LPW loads (each wave issues exactly these counts)
        const int later = (prev_tile ? 1 : 0) + (next_issued ? 1 : 0);
        if (later == 2) wait_vmcnt<2 * LPW>();
        else if (later == 1) wait_vmcnt<LPW>();
        else wait_vmcnt<0>();
        __builtin_amdgcn_s_barrier();
    } else {
        FT stg[NLOAD][L];
#pragma unroll
        for (int t = 0; t < NLOAD; t++) {
            const int n = (MODE == TILE_PAR) ? TR::P_IDX : t;
            const rsrc_t rn = rs(a.tr[n]);
#pragma unroll
            for (int q = 0; q < L; q++) buf_load(rn, e_off(q), stg[t][q]);
        }
#pragma unroll
        for (int t = 0; t < NLOAD; t++)
#pragma unroll
            for (int q = 0; q < L; q++)
                if (e_idx(q) >= 0) tile(t)[e_idx(q)] = stg[t][q];
        __syncthreads();
    }

    // FP64: U is the state, updated in place (the reference's expressions).
    // FP32: U is the state at the start of the step and D the step's
    // accumulated increment; stage inputs are U + D and the sum is rounded
    // once at the store (a compensated update: the increments are never
    // rounded against the magnitude of U).
    constexpr bool F32 = sizeof(FT) == 4;
    FT U[NT][L];
    FT Gm[NT][L];
    FT D[F32 ? NT : 1][L];
#pragma unroll
    for (int n = 0; n < NT; n++)
#pragma unroll
        for (int m = 0; m < L; m++) {
            const int k = kof(m);
            const bool ok = has(m);
            int t = (MODE == TILE_PAR) ? (n == TR::P_IDX ? 0 : -1) : n;
            if constexpr (PIPE) U[n][m] = pbuf[pidx(n, k, c)];
            else U[n][m] = (ok && t >= 0) ? tile(t)[k * ROW + c] : FT(0);
            Gm[n][m] = FT(0);
            if constexpr (F32) D[n][m] = FT(0);
        }
    // stage value of tracer n at level m
    auto val = [&](int n, int m) -> FT {
        if constexpr (F32) return (MODE == TILE_STEP) ? U[n][m] + D[n < NT ? n : 0][m] : U[n][m];
        else return U[n][m];
    };

    // geometry of level k: computed (uniform grid) or read from the per-level rows
    auto g_zc = [&](int k) -> FT {
        if constexpr (UNIFORM) return a.zc0_u + FT(k) * a.dz_u;
        else return a.lev[LEV_ZC * Nz + k];
    };
    auto g_up = [&](int k) -> FT {
        if constexpr (UNIFORM) return a.dz_u * FT(0.5);
        else return a.lev[LEV_UP * Nz + k];
    };
    auto g_dn = [&](int k) -> FT {
        if constexpr (UNIFORM) return a.dz_u * FT(0.5);
        else return a.lev[LEV_DN * Nz + k];
    };
    auto g_invdz = [&](int k) -> FT {
        if constexpr (UNIFORM) return a.invdz_u;
        else return a.lev[LEV_INVDZ * Nz + k];
    };
    // face velocities below (face k) and above (face k+1) cell k, sinking tracer s
    auto g_wlo = [&](int s, int k) -> FT {
        if constexpr (UNIFORM) return s == 0 ? a.w0_u : a.w1_u;
        else return a.wsink[s * (Nz + 1) + k];
    };
    auto g_whi = [&](int s, int k) -> FT {
        if constexpr (UNIFORM) return s == 0 ? a.w0_u : a.w1_u;
        else return a.wsink[s * (Nz + 1) + k + 1];
    };

    // (the constant is read before the tile loop.  The pipelined sweep is
    // not used with a surface PAR array: the wait for a vector load here
    // would also wait for the prefetch of the next tile.)
    FT PAR0_const = surface_value0;
    if constexpr (!PIPE) {
        if (a.surface_mode == OBME_SURFACE_ARRAY) PAR0_const = a.surf[i + a.Nx * j];
    }

    const unsigned col2d = (unsigned)i + (unsigned)a.Nx * (unsigned)j;
    const unsigned ncol2d = (unsigned)a.Nx * (unsigned)a.Ny;
    const bool top_lane = col_ok && r == 0;
    if constexpr ((CARB & 1) && MODE == TILE_STEP) {
        if (top_lane && a.co2_on) a.srec[col2d] = U[7][0];
    }

    // FP32 on a uniform grid: the water terms kr zc, kb zc of the own levels (log2 scaled)
    float ezr[L], ezb[L];
    if constexpr (F32 && UNIFORM) {
#pragma unroll
        for (int m = 0; m < L; m++) {
            const float z = g_zc(has(m) ? kof(m) : 0);
            ezr[m] = a.kr2 * z;
            ezb[m] = a.kb2 * z;
        }
    }

    constexpr int NSTAGE = (MODE == TILE_STEP) ? 3 : 1;
    const double g1 = 8.0 / 15.0, g2 = 5.0 / 12.0, g3 = 3.0 / 4.0, z2 = -17.0 / 60.0, z3 = -5.0 / 12.0;
    FT PARv[L];   // PAR of the own levels (output of the PAR and tendency modes)
    static_assert(EXTRA == 0 || (MODE == TILE_STEP && !PIPE), "scaling, inline CO2, sediment: fused step, direct sweep");
    static_assert(!(EXTRA & 2) || (CARB & 1), "inline CO2 needs the carbonate system");
    // inline CO2 (EXTRA & 2): [H+] of the previous stage (FP32-mixed warm start)
    double Hc = 0.0;
    if constexpr ((EXTRA & 2) && F32) {
        if (top_lane && a.Hprev) Hc = a.Hprev[col2d];
    }
    // sediment (EXTRA & 4): the bottom lane's storage value, the tendency e f
    // of the last stage start and the G⁻ of the sediment's previous step
    const bool sed_lane = col_ok && r * L <= ktop && ktop < r * L + L;
    double sed_st = 0.0, sed_ge = 0.0, sed_gm = 0.0;
    if constexpr ((EXTRA & 4) != 0) {
        if (sed_lane) {
            if (a.storage) sed_st = a.storage[col2d];
            if (a.sed_ab2) sed_gm = a.sedG[ncol2d + col2d];
        }
    }

    auto run_stage = [&](const int st) {
        FT PAR0 = PAR0_const;
        if (a.surface_mode == OBME_SURFACE_DIURNAL) PAR0 = a.PAR0_stage[st];

        if constexpr ((EXTRA & 1) != 0) {
            // ScaleNegativeTracers first (update_biogeochemical_state! order,
            // src/OceanBioME.jl:161-167).  FP64 applies it to every cell as
            // the reference does; FP32-mixed where a member is negative or
            // not finite (elsewhere it is the identity up to rounding), and
            // the scaled value becomes the new U with no pending increment.
            unsigned members = 0;
#pragma unroll 1
            for (int g = 0; g < a.ns.n_groups; g++) members |= a.ns.mask[g];
#pragma unroll
            for (int m = 0; m < L; m++) {
                FT x[NT];
#pragma unroll
                for (int n = 0; n < NT; n++) x[n] = val(n, m);
                if constexpr (F32) {
                    bool need = false;
#pragma unroll
                    for (int n = 0; n < NT; n++)
                        if (members & (1u << n)) need = need || !(x[n] >= 0.0f) || isinf(x[n]);
                    if (need && has(m)) {
                        scale_negatives_cell<NT>(a.ns, x);
#pragma unroll
                        for (int n = 0; n < NT; n++)
                            if (members & (1u << n)) {
                                U[n][m] = x[n];
                                D[n][m] = 0.0f;
                            }
                    }
                } else {
                    scale_negatives_cell<NT>(a.ns, x);
#pragma unroll
                    for (int n = 0; n < NT; n++)
                        if (members & (1u << n)) U[n][m] = x[n];
                }
            }
        }

        // chlorophyll terms (P R/r)^e of the own levels
        FT pr[L], pb[L];
#pragma unroll
        for (int m = 0; m < L; m++) {
            if constexpr (F32) {
                const FT lg = m_log2(val(TR::P_IDX, m));
                pr[m] = m_exp2(fmaf(a.er, lg, a.cr));
                pb[m] = m_exp2(fmaf(a.eb, lg, a.cb));
            } else {
                const FT x = val(TR::P_IDX, m) * a.Rr;
                pr[m] = m_pow(x, a.er);
                pb[m] = m_pow(x, a.eb);
            }
        }
        // the level above a lane's top level is the previous lane's bottom level
        const FT pr_up = lane_shr1(pr[L - 1]);
        const FT pb_up = lane_shr1(pb[L - 1]);
        const FT c0_up = lane_shr1(val(TR::SINK0, L - 1));
        const FT c1_up = lane_shr1(val(TR::SINK1, L - 1));

        // increments top-down: inc(k) = up(k) p(k+1) + dn(k) p(k); the top
        // level has none above: dn p alone.  FP32 on a uniform grid sums
        // p(k+1) + p(k) and folds dz/2 into the attenuation.
        FT incr[L], incb[L];
        FT totr = FT(0), totb = FT(0);
#pragma unroll
        for (int m = 0; m < L; m++) {
            const int k = kof(m);
            const FT prn = (m == 0) ? pr_up : pr[m > 0 ? m - 1 : 0];
            const FT pbn = (m == 0) ? pb_up : pb[m > 0 ? m - 1 : 0];
            const bool below_top = (m > 0) || (k < ktop);
            FT ir, ib;
            if constexpr (F32 && UNIFORM) {
                ir = below_top ? prn + pr[m] : pr[m];
                ib = below_top ? pbn + pb[m] : pb[m];
            } else {
                const int kc = has(m) ? k : 0;
                const FT upk = g_up(kc), dnk = g_dn(kc);
                ir = below_top ? upk * prn + dnk * pr[m] : dnk * pr[m];
                ib = below_top ? upk * pbn + dnk * pb[m] : dnk * pb[m];
            }
            incr[m] = has(m) ? ir : FT(0);
            incb[m] = has(m) ? ib : FT(0);
            totr = totr + incr[m];
            totb = totb + incb[m];
        }
        // sum of the increments above the lane: exclusive segmented prefix
        // (a lane tree, error growing as log2 of the depth)
        FT runr = lane_shr1(seg_prefix<LPC>(totr));
        FT runb = lane_shr1(seg_prefix<LPC>(totb));
        if (r == 0) {
            runr = FT(0);
            runb = FT(0);
        }
        // PAR of the own levels
        const FT PAR0h = PAR0 * FT(0.5);
#pragma unroll
        for (int m = 0; m < L; m++) {
            runr = runr + incr[m];
            runb = runb + incb[m];
            if constexpr (F32 && UNIFORM) {
                PARv[m] = PAR0h * (m_exp2(fmaf(-a.xr2h, runr, ezr[m])) + m_exp2(fmaf(-a.xb2h, runb, ezb[m])));
            } else {
                const int k = kof(m);
                const FT z = g_zc(has(m) ? k : 0);
                if constexpr (F32)
                    PARv[m] = PAR0h * (m_exp2(fmaf(a.kr2, z, -a.xr2 * runr)) + m_exp2(fmaf(a.kb2, z, -a.xb2 * runb)));
                else
                    PARv[m] = PAR0 * (m_exp(a.kr * z - a.xr * runr) + m_exp(a.kb * z - a.xb * runb)) / FT(2);
            }
        }

        if constexpr (MODE != TILE_PAR) {
            // bottom-up within the lane (FP64 updates in place: level m-1,
            // above m, must still hold its stage value when m reads it)
#pragma unroll
            for (int m = L - 1; m >= 0; m--) {
                const int k = kof(m);
                const int kc = has(m) ? k : 0;
                FT Gv[NT];
                FT v[NT];
#pragma unroll
                for (int n = 0; n < NT; n++) v[n] = (CARB != 0 && n >= 7) ? FT(0) : val(n, m);
                if constexpr (MODEL == OBME_NPZD) {
                    if constexpr (F32) {
                        const float Lt = fm_q10<FM>(a.f) ? m_exp2(v[3] * a.f.log2Q10_10) : 1.0f;
                        npzd_fast<FM>(a.f, v[0], v[1], v[2], v[4], PARv[m], Lt, Gv[0], Gv[1], Gv[2], Gv[4],
                                      (CARB & 2) ? &Gv[NT - 1] : nullptr);
                    } else {
                        FT Lt = FT(1);
                        if (a.q.use_Q10) Lt = m_pow(a.q.Q10, v[3] / FT(10));
                        npzd_tendencies(a.q, v[0], v[1], v[2], v[3], v[4], PARv[m], Lt, Gv[0], Gv[1], Gv[2], Gv[4],
                                        (CARB & 2) ? &Gv[NT - 1] : nullptr);
                    }
                    Gv[3] = FT(0);
                } else {
                    if constexpr (F32) lobster_fast<CARB, FM>(a.f, v, PARv[m], Gv);
                    else lobster_tendencies<FT, CARB>(a.q, v, PARv[m], Gv);
                }
                // minus the upwind drift flux divergence: w(face) c(upwind cell);
                // no flux through the top face (w = 0 there)
                {
                    const bool below_top = (m > 0) || (k < ktop);
                    const FT cn0 = (m == 0) ? c0_up : val(TR::SINK0, m > 0 ? m - 1 : 0);
                    const FT cn1 = (m == 0) ? c1_up : val(TR::SINK1, m > 0 ? m - 1 : 0);
                    if constexpr (F32 && UNIFORM) {
                        // w (c(k+1) - c(k)) / dz, c(k+1) = 0 above the top cell
                        const float d0 = (below_top ? cn0 : 0.0f) - v[TR::SINK0];
                        const float d1 = (below_top ? cn1 : 0.0f) - v[TR::SINK1];
                        Gv[TR::SINK0] = fmaf(-a.wi0, d0, Gv[TR::SINK0]);
                        Gv[TR::SINK1] = fmaf(-a.wi1, d1, Gv[TR::SINK1]);
                    } else {
                        const FT idz = g_invdz(kc);
                        const FT Fhi0 = below_top ? g_whi(0, kc) * cn0 : FT(0);
                        const FT Fhi1 = below_top ? g_whi(1, kc) * cn1 : FT(0);
                        Gv[TR::SINK0] = Gv[TR::SINK0] - (Fhi0 - g_wlo(0, kc) * v[TR::SINK0]) * idz;
                        Gv[TR::SINK1] = Gv[TR::SINK1] - (Fhi1 - g_wlo(1, kc) * v[TR::SINK1]) * idz;
                    }
                }
                if constexpr ((EXTRA & 2) != 0) {
                    // the top-cell CO2 flux (GasExchange, gas_exchange.jl:26-38) of
                    // this stage's DIC and Alk, the surface kernel's expressions
                    if (top_lane && m == 0) {
                        const unsigned top = (unsigned)a.base0 + (unsigned)i + a.sx * (unsigned)j +
                                             a.sxy * (unsigned)ktop;
                        const double T = a.T3 ? (double)a.T3[top] : (a.Tprof ? a.Tprof[ktop] : a.Tconst);
                        const double S = a.S3 ? (double)a.S3[top] : (a.Sprof ? a.Sprof[ktop] : a.Sconst);
                        SurfaceCarbon sc;
                        surface_carbon_setup(a.co2, T, S, a.co2.wind_speed, sc);
                        double pH = 0.0;
                        int ev = 0;
                        const double F = surface_flux_warm(sc, a.solver, (double)val(7, 0), (double)val(8, 0),
                                                           a.co2.air_concentration, F32 ? 1 : 0, Hc, pH, ev,
                                                           F32 ? FP32_NEWTON_TOL : 1e-14);
                        Gv[7] = Gv[7] - (FT)F * a.invdz_top;
                    }
                }
                if constexpr ((CARB & 2) != 0) {
                    // the top-cell O2 flux (GasExchange for O2, gas_exchange.jl:26-38,
                    // GasExchange.jl:127-146) of this stage's O2: G -= F / Δz
                    if (a.o2_on && top_lane && m == 0) {
                        const FT kdz = (FT)a.o2c[2 * col2d], air = (FT)a.o2c[2 * col2d + 1];
                        Gv[NT - 1] = Gv[NT - 1] - kdz * (val(NT - 1, 0) - air);
                    }
                }
                if constexpr ((EXTRA & 4) != 0) {
                    // the sediment at the bottom cell: the sinking flux through
                    // the bottom face, f = Σ -w c (Sediments/tracked_fields.jl:53-70),
                    // burial e = a + b (f / (k + f))^2 and the receiver's coupled
                    // tendency (1 - e) f / Δz (instant_remineralisation.jl:101-123,
                    // Sediments/tracer_coupling.jl:30-38)
                    if (col_ok && k == 0 && has(m)) {
                        const FT f = a.sed_w0 * v[TR::SINK0] + a.sed_w1 * v[TR::SINK1];
                        const FT qf = f / (a.sed_k + f);
                        const FT e = a.sed_a + a.sed_b * (qf * qf);
                        const FT cpl = ((FT(1) - e) * f) / a.sed_dz0;
#pragma unroll
                        for (int n = 0; n < NT; n++)
                            if (n == a.sed_recv) Gv[n] = Gv[n] + cpl;
                        // the sediment step after this stage's substep
                        // (update_state.jl:6-17) takes the tendency evaluated
                        // after the previous step, from this stage-start
                        // state: e f here.  Its first step ever evaluates
                        // the tendency lazily, after the first substep: the
                        // tendency of the second stage start serves both.
                        const double ge = (double)(e * f);
                        if (st == 0) {
                            if (!a.sed_init)
                                sed_st += a.sed_dt[0] * (a.sed_ab2 ? a.sed_cn * ge - a.sed_cm * sed_gm : ge);
                        } else {
                            if (st == 1 && a.sed_init) sed_st += a.sed_dt[0] * ge;
                            sed_st += a.sed_dt[st] * ge;
                        }
                        sed_ge = ge;
                    }
                }
                if constexpr (MODE == TILE_TENDENCY) {
#pragma unroll
                    for (int n = 0; n < NT; n++) Gm[n][m] = Gv[n];   // Gm carries G out
                } else {
                    if constexpr (CARB & 1) {
                        // surface record: DIC transition without the flux, Alk at the stage start
                        if (top_lane && a.co2_on && m == 0) {
                            a.srec[(1 + st) * ncol2d + col2d] = Gv[7];
                            a.srec[(4 + st) * ncol2d + col2d] = val(8, 0);
                        }
                    }
                    // RK3 substep (BoxModel/timesteppers.jl:79-92 form):
                    // stage 1 U += (Δt γ1) G; stages 2, 3 U += Δt (γ G + ζ G⁻)
#pragma unroll
                    for (int n = 0; n < NT; n++) {
                        if (n == RO) continue;
                        if constexpr (F32) {
                            const float A = (st == 0) ? a.rk_a[0] : (st == 1 ? a.rk_a[1] : a.rk_a[2]);
                            const float B = (st == 0) ? a.rk_b[0] : (st == 1 ? a.rk_b[1] : a.rk_b[2]);
                            D[n][m] = fmaf(A, Gv[n], fmaf(B, Gm[n][m], D[n][m]));
                        } else {
                            const double gs = (st == 1) ? g2 : g3;
                            const double zs = (st == 1) ? z2 : z3;
                            const double inc = (st == 0) ? (a.dtf * g1) * Gv[n] : a.dtf * (gs * Gv[n] + zs * Gm[n][m]);
                            U[n][m] += inc;
                        }
                        Gm[n][m] = Gv[n];
                    }
                }
            }
        }
    };
    // (the stages stay a loop: unrolled, the FP32 step measured 6 % slower)
#pragma unroll 1
    for (int st = 0; st < NSTAGE; st++) run_stage(st);
    if constexpr ((EXTRA & 2) && F32) {
        if (top_lane && a.Hprev) a.Hprev[col2d] = Hc;
    }
    if constexpr ((EXTRA & 4) != 0) {
        // the sediment after the last step: storage, G⁻ = the tendency of its
        // last step, Gⁿ = e f of the final state
#pragma unroll
        for (int m = 0; m < L; m++) {
            if (sed_lane && kof(m) == 0) {
                const FT f = a.sed_w0 * val(TR::SINK0, m) + a.sed_w1 * val(TR::SINK1, m);
                const FT qf = f / (a.sed_k + f);
                const FT e = a.sed_a + a.sed_b * (qf * qf);
                if (a.storage) a.storage[col2d] = sed_st;
                a.sedG[col2d] = (double)(e * f);
                a.sedG[ncol2d + col2d] = sed_ge;
            }
        }
    }

    if constexpr (PIPE) {
        // ---- stage out (PIPE): each lane scatters to the elements it
        // gathered, then every wave reads its blocks in load order
        // (coalesced), the buffer is refilled with the tile after next, and
        // the stores are issued; an element that must not be written (past
        // Nx, or the top-cell DIC left to the surface kernel) gets an offset
        // past the buffer, which the buffer store drops, so every wave
        // issues exactly LPW stores.
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int m = 0; m < L; m++) pbuf[pidx(n, kof(m), c)] = val(n, m);
        wait_lgkm0();
        __builtin_amdgcn_s_barrier();
        FT v[NT][NBLK / NW];
        unsigned off[NT][NBLK / NW];
        const int kl = lane / TXB, cc = lane % TXB;
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int q = 0; q < NBLK / NW; q++) {
                const int kb = w + NW * q;
                const int k = kb * PB + kl;
                v[n][q] = pbuf[n * TRS + kb * BLK + lane];
                bool ok = ibase + cc < a.Nx && n != RO;
                if constexpr (CARB & 1) {
                    if (n == 7 && a.co2_on && k == ktop) ok = false;
                }
                off[n][q] = ok ? goff2(ibase + cc, j, k) : a.field_bytes;
            }
        wait_lgkm0();
        __builtin_amdgcn_s_barrier();
        const unsigned t_next = wg_tile + 2 * t_step;
        if (t_next < t_end) pipe_issue(t_next, pbuf);
        // seen from the next tile: a tile was stored before it, and the tile
        // after it has been issued iff it exists
        prev_tile = true;
        next_issued = t_next < t_end;
#pragma unroll
        for (int n = 0; n < NT; n++) {
            const rsrc_t rd = rs(a.tr[n]);
#pragma unroll
            for (int q = 0; q < NBLK / NW; q++) buf_store(rd, off[n][q], v[n][q]);
        }
        continue;
    }

    // ---- stage out: scatter to the tiles, coalesced stores along i.
    // Outputs: STEP the written tracers; PAR the PAR field; TENDENCY the
    // written tendencies (added to G) and PAR (tile NT) when requested.
    __syncthreads();
    constexpr int NOUT = (MODE == TILE_PAR) ? 1 : (MODE == TILE_TENDENCY ? NT + 1 : NT);
#pragma unroll
    for (int n = 0; n < NOUT; n++) {
        if (MODE != TILE_PAR && n == RO) continue;
        FT* T = tile(n);
#pragma unroll
        for (int m = 0; m < L; m++) {
            const int k = kof(m);
            FT v;
            if constexpr (MODE == TILE_PAR) v = PARv[m];
            else if constexpr (MODE == TILE_TENDENCY) v = (n == NT) ? PARv[m] : Gm[n < NT ? n : 0][m];
            else v = val(n < NT ? n : 0, m);
            if (has(m)) T[k * ROW + c] = v;
        }
    }
    __syncthreads();
#pragma unroll
    for (int n = 0; n < NOUT; n++) {
        if (MODE != TILE_PAR && n == RO) continue;
        FT* T = tile(n);
        FT* dst;
        if constexpr (MODE == TILE_PAR) dst = a.PAR;
        else if constexpr (MODE == TILE_TENDENCY) dst = (n == NT) ? a.PAR : a.G[n];
        else dst = a.tr[n];
        if (dst == nullptr) continue;
        const rsrc_t rd = rs(dst);
        const bool accumulate = (MODE == TILE_TENDENCY) && n < NT;
        FT v[L];
#pragma unroll
        for (int q = 0; q < L; q++) v[q] = (e_idx(q) >= 0) ? T[e_idx(q)] : FT(0);
        if (accumulate) {
            FT g0[L];
#pragma unroll
            for (int q = 0; q < L; q++) buf_load(rd, e_off(q), g0[q]);
#pragma unroll
            for (int q = 0; q < L; q++) v[q] = g0[q] + v[q];
        }
#pragma unroll
        for (int q = 0; q < L; q++) {
            const int e = threadIdx.x + q * BS;
            const int k = e / TXB, cc = e % TXB;
            bool ok = e_idx(q) >= 0 && ibase + cc < a.Nx;
            if constexpr ((CARB & 1) && MODE == TILE_STEP) {
                // the top-cell DIC is written by the surface kernel
                if (n == 7 && a.co2_on && k == ktop) ok = false;
            }
            // an element not to be written gets an offset past the buffer,
            // which the buffer store drops (no branch around the store)
            buf_store(rd, ok ? e_off(q) : a.field_bytes, v[q]);
        }
    }
    }   // tile loop
}

// ------------------------------------------------------------------ surface

template <typename FT>
struct SurfaceArgs {
    FT* DIC;                       // DIC parent array (top cell written, step mode)
    FT* GDIC;                      // DIC tendency parent array (tendency mode)
    const FT* Alk;                 // Alk parent array (tendency mode)
    const FT* srec;                // surface records (step mode)
    double* Hprev;                 // [Nx*Ny] warm start of [H+] (mixed path) or null
    const double* surfc;           // [SURF_PACK][Nx*Ny] surface constants of the call (step mode) or null
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
    int mode;                      // 0: reference damped Newton; 1: mixed
    obme_co2_exchange co2;
    obme_carbon_solver solver;
};

// The surface constants of every column (T and S of the top cell), once per
// stepper call, for k_surface.
template <typename FT>
__global__ void __launch_bounds__(256) k_surface_prep(SurfaceArgs<FT> a, double* out)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = a.Nx * a.Ny;
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
#pragma unroll
        for (int q = 0; q < SURF_PACK; q++) o[q] = a.surfc[(long)q * ncol + col];
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
    } else {
        // the identical update expressions of the sweep, with the flux term
        FT U = a.srec[col];
        FT Gm = FT(0);
        if constexpr (sizeof(FT) == 4) {
            float D = 0.0f;
#pragma unroll
            for (int st = 0; st < 3; st++) {
                const float Gb = a.srec[(1 + st) * ncol + col];
                const float Alk = a.srec[(4 + st) * ncol + col];
                F = surface_flux_warm(sc, a.solver, (double)(U + D), (double)Alk, a.co2.air_concentration, a.mode,
                                      H, pH, ev, FP32_NEWTON_TOL);
                const float G = Gb - (float)F * a.invdz_top;
                D = fmaf(a.rk_a[st], G, fmaf(a.rk_b[st], Gm, D));
                Gm = G;
            }
            U = U + D;
        } else {
            const FT g1 = FT(8.0 / 15.0), g2 = FT(5.0 / 12.0), g3 = FT(3.0 / 4.0);
            const FT z2 = FT(-17.0 / 60.0), z3 = FT(-5.0 / 12.0);
#pragma unroll
            for (int st = 0; st < 3; st++) {
                FT Gb = a.srec[(1 + st) * ncol + col];
                FT Alk = a.srec[(4 + st) * ncol + col];
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
