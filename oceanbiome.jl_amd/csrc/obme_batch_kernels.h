// obme_batch_kernels.h — the batched ensemble of independent NPZD water
// columns: ragged depths, per-column parameters and surface PAR, every
// column advanced many RK3 steps with its state held in registers.
//
// Reference semantics: one NPZD column per entry, the column stepper of the
// fused sweep (TwoBand PAR, Light/2band.jl:1-33; upwind sinking with an open
// bottom, Utils/sinking_velocity_fields.jl:10-35; NPZD transitions,
// NutrientsPlanktonDetritus/*.jl; the RK3 substep, BoxModel/timesteppers.jl:79-92),
// with the column's parameters in place of the model defaults.
//
// FP32-mixed: the state is U plus an increment D.  Within a step D gathers
// the stage increments (stage inputs U + D); at the end of the step U + D is
// split exactly (TwoSum) into its rounding s and the residual, U = s and D =
// the residual, so no step's increment is lost to the rounding of U: the
// state is carried to about twice FP32's precision for as long as it stays
// in registers.  A call split over several launches hands the residual to
// the next launch through a residual array; the stored state is rounded
// once, at the end of the call.
//
// Layout: tracer n (N, P, Z, D) is one packed array; column c occupies
// [offsets[c], offsets[c+1]) with k = 0 at the bottom.  Columns are packed
// into wave segments of LPC = 64/CPW lanes, lane r of a segment holding the
// L levels counted from the top, j = r*L + m (k = Nz_c - 1 - j).  The host
// sorts columns into classes by depth (capacity L*LPC >= Nz_c) and launches
// one persistent kernel per class over that class's column list.  Cell
// centres depend only on j: zc = -(j + 1/2) dz.
#pragma once

#include "obme_npd_kernels.h"

namespace obme {

// Column constants beyond the NPZD parameters.
struct BatchColumn {
    double PAR0;         // surface PAR (constant)
    double wP, wD;       // face velocities (<= 0) of P and D
    double Lt;           // Q10^(T/10) at the ensemble's temperature
};

template <typename FT>
struct BatchArgs {
    FT* U[4];                       // N, P, Z, D packed arrays
    FT* C[4];                       // FP32: the residuals of U, packed alike (launch to launch)
    int c_in, c_out;                // FP32: start from the stored residuals; store them at the end
    const int* offsets;             // [n_columns + 1]
    const int* cols;                // this class's column list
    int ncols;                      // entries of `cols`
    const FastParams* fpar;         // FP32: per-column derived constants
    const NPDParams<double>* dpar;  // FP64: per-column parameters
    const BatchColumn* bcol;        // per-column constants
    int nsteps;
    FT dtf, dz, invdz;
    float rk_a[3], rk_b[3];         // FP32: Δt γ, Δt ζ per stage
    // TwoBand coefficients: FP64 reference form and FP32 log2-scaled form
    double kr, kb, xr, xb, er, eb, Rr;
    float kr2, kb2, er_f, eb_f, cr, cb, xr2h, xb2h;
    // InstantRemineralisation sediment under every column (sed_on; storage
    // null: the coupling alone, the storage not advanced).  The flux through
    // the bottom face f = sedP (-wP) P + sedD (-wD) D
    // of the bottom cell, burial e = a + b (f / (k + f))^2, the receiver's
    // coupled tendency (1 - e) f / dz; after each stage's substep the storage
    // steps by the stage Δt with the tendency e f of the stage's start
    // (update_state.jl:6-17; instant_remineralisation.jl:101-123).
    int sed_on;
    double* storage;                // [n_columns] or null
    double* sedGm;                  // [n_columns] G⁻: the tendency of the sediment's last step
    int sed_init;                   // the sediment's first step ever is this launch's first step
    int sed_ab2;                    // this launch's first sediment step is a QuasiAdamsBashforth2 step
    int sed_recv;                   // receiver: 0 N, 1 P, 2 Z, 3 D
    double sed_dt[3];               // Δt γ1, Δt (γ2 + ζ2), Δt (γ3 + ζ3)
    double sed_cn, sed_cm;          // 3/2 + χ, 1/2 + χ
    FT sed_a, sed_b, sed_k;         // burial efficiency constants
    FT sedP, sedD;                  // 1 where P / D is one of the sinking tracers, else 0
};

// One persistent workgroup loops over tiles of NW*CPW columns.  SED: the
// sediment's code compiled in (the ensemble without one keeps its registers).
template <typename FT, int L, int CPW, int NW, bool SED>
__device__ __forceinline__ void batch_columns(const BatchArgs<FT>& a)
{
    constexpr int LPC = 64 / CPW;
    static_assert(LPC == 16 || LPC == 32 || LPC == 64, "segments of 16, 32 or 64 lanes");
    constexpr bool F32 = sizeof(FT) == 4;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int r = lane % LPC;
    const int seg = w * CPW + lane / LPC;
    const int ntiles = (a.ncols + NW * CPW - 1) / (NW * CPW);

#pragma unroll 1
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int g = tile * (NW * CPW) + seg;
        const int col = (g < a.ncols) ? a.cols[g] : -1;
        const int off = (col >= 0) ? a.offsets[col] : 0;
        const int Nzc = (col >= 0) ? a.offsets[col + 1] - off : 0;
        // level m of this lane: j = r*L + m from the top, k = Nzc-1-j
        auto has = [&](int m) -> bool { return r * L + m < Nzc; };
        auto idx = [&](int m) -> int { return off + Nzc - 1 - (r * L + m); };

        FT U[4][L], D[F32 ? 4 : 1][L], Gm[4][L];
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int m = 0; m < L; m++) {
                U[n][m] = has(m) ? a.U[n][idx(m)] : FT(0);
                Gm[n][m] = FT(0);
                if constexpr (F32) D[n][m] = (a.c_in && has(m)) ? a.C[n][idx(m)] : FT(0);
            }
        auto val = [&](int n, int m) -> FT {
            if constexpr (F32) return U[n][m] + D[n][m];
            else return U[n][m];
        };

        // column constants (every lane of a segment reads the same entry)
        const int cc = (col >= 0) ? col : 0;
        const BatchColumn bc = a.bcol[cc];
        const FT PAR0 = (FT)bc.PAR0;
        const FT Lt = (FT)bc.Lt;
        const FT wP = (FT)bc.wP, wD = (FT)bc.wD;
        FastParams fp;
        NPDParams<double> dp;
        if constexpr (F32) fp = a.fpar[cc];
        else dp = a.dpar[cc];
        const float wiP = (float)(bc.wP / (double)a.dz), wiD = (float)(bc.wD / (double)a.dz);

        // the sediment of this column, held by the lane of its bottom cell
        // (j = Nzc - 1: lane rb, level mb)
        const int jb = Nzc - 1;
        const bool sed = SED && a.sed_on && col >= 0 && r == jb / L;
        const int mb = jb % L;
        // the storage and the last tendency in LDS, one slot per segment
        // (not held in registers through the launch)
        __shared__ double s_sed[2][NW * CPW];
        if (sed) {
            s_sed[0][seg] = a.storage ? a.storage[col] : 0.0;
            s_sed[1][seg] = (a.storage && a.sed_ab2) ? a.sedGm[col] : 0.0;
        }

        // water terms of the own levels
        float ezr[L], ezb[L];
        FT zc[L];
#pragma unroll
        for (int m = 0; m < L; m++) {
            zc[m] = -(FT(r * L + m) + FT(0.5)) * a.dz;
            if constexpr (F32) {
                ezr[m] = a.kr2 * zc[m];
                ezb[m] = a.kb2 * zc[m];
            }
        }
        const double g1 = 8.0 / 15.0, g2 = 5.0 / 12.0, g3 = 3.0 / 4.0, z2 = -17.0 / 60.0, z3 = -5.0 / 12.0;

#pragma unroll 1
        for (int s = 0; s < a.nsteps; s++) {
#pragma unroll 1
            for (int st = 0; st < 3; st++) {
                // chlorophyll terms and the light integral (as in the sweep)
                FT pr[L], pb[L];
#pragma unroll
                for (int m = 0; m < L; m++) {
                    if constexpr (F32) {
                        const float lg = m_log2(val(1, m));
                        pr[m] = m_exp2(fmaf(a.er_f, lg, a.cr));
                        pb[m] = m_exp2(fmaf(a.eb_f, lg, a.cb));
                    } else {
                        const double x = val(1, m) * a.Rr;
                        pow_pair(x, a.er, a.eb, pr[m], pb[m]);
                    }
                }
                const FT pr_up = lane_shr1(pr[L - 1]);
                const FT pb_up = lane_shr1(pb[L - 1]);
                const FT cP_up = lane_shr1(val(1, L - 1));
                const FT cD_up = lane_shr1(val(3, L - 1));
                FT incr[L], incb[L];
                FT totr = FT(0), totb = FT(0);
#pragma unroll
                for (int m = 0; m < L; m++) {
                    const bool top = (m == 0) && (r == 0);
                    const FT prn = (m == 0) ? pr_up : pr[m > 0 ? m - 1 : 0];
                    const FT pbn = (m == 0) ? pb_up : pb[m > 0 ? m - 1 : 0];
                    FT ir, ib;
                    if constexpr (F32) {
                        ir = top ? pr[m] : prn + pr[m];
                        ib = top ? pb[m] : pbn + pb[m];
                    } else {
                        const double h = a.dz / 2;
                        ir = top ? h * pr[m] : h * prn + h * pr[m];
                        ib = top ? h * pb[m] : h * pbn + h * pb[m];
                    }
                    incr[m] = has(m) ? ir : FT(0);
                    incb[m] = has(m) ? ib : FT(0);
                    totr = totr + incr[m];
                    totb = totb + incb[m];
                }
                FT runr = lane_shr1(seg_prefix<LPC>(totr));
                FT runb = lane_shr1(seg_prefix<LPC>(totb));
                if (r == 0) {
                    runr = FT(0);
                    runb = FT(0);
                }
                FT PARv[L];
#pragma unroll
                for (int m = 0; m < L; m++) {
                    runr = runr + incr[m];
                    runb = runb + incb[m];
                    if constexpr (F32)
                        PARv[m] = (PAR0 * 0.5f) * (m_exp2(fmaf(-a.xr2h, runr, ezr[m])) +
                                                   m_exp2(fmaf(-a.xb2h, runb, ezb[m])));
                    else
                        PARv[m] = PAR0 * (m_exp(a.kr * zc[m] - a.xr * runr) + m_exp(a.kb * zc[m] - a.xb * runb)) / 2;
                }
                // the sediment: the bottom cell's flux and burial, the
                // receiver's coupled tendency, the sediment step
                FT cpl = FT(0);
                if (sed) {
                    FT Pb = val(1, 0), Db = val(3, 0);
#pragma unroll
                    for (int m = 1; m < L; m++)
                        if (m == mb) {
                            Pb = val(1, m);
                            Db = val(3, m);
                        }
                    const FT f = a.sedP * (-wP) * Pb + a.sedD * (-wD) * Db;
                    FT qf, e;
                    if constexpr (F32) {
                        // FP32-mixed: the hardware reciprocal (1 ulp) and 1/Δz
                        qf = f * __builtin_amdgcn_rcpf(a.sed_k + f);
                        e = a.sed_a + a.sed_b * (qf * qf);
                        cpl = ((FT(1) - e) * f) * a.invdz;
                    } else {
                        qf = f / (a.sed_k + f);
                        e = a.sed_a + a.sed_b * (qf * qf);
                        cpl = ((FT(1) - e) * f) / a.dz;
                    }
                    const double ge = (double)(e * f);
                    const bool init = a.sed_init && s == 0;
                    if (st == 0) {
                        if (!init) {
                            const bool ab2 = a.sed_ab2 && s == 0;
                            s_sed[0][seg] += a.sed_dt[0] * (ab2 ? a.sed_cn * ge - a.sed_cm * s_sed[1][seg] : ge);
                        }
                    } else {
                        if (st == 1 && init) s_sed[0][seg] += a.sed_dt[0] * ge;
                        s_sed[0][seg] += a.sed_dt[st] * ge;
                    }
                    s_sed[1][seg] = ge;
                }
                // tendencies bottom-up within the lane, sinking, RK3 substep
#pragma unroll
                for (int m = L - 1; m >= 0; m--) {
                    const bool top = (m == 0) && (r == 0);
                    const FT N = val(0, m), P = val(1, m), Z = val(2, m), Dd = val(3, m);
                    FT G[4];
                    if constexpr (F32) npzd_fast(fp, N, P, Z, Dd, PARv[m], Lt, G[0], G[1], G[2], G[3]);
                    else npzd_tendencies(dp, N, P, Z, FT(0), Dd, PARv[m], Lt, G[0], G[1], G[2], G[3]);
                    const FT cPn = (m == 0) ? cP_up : val(1, m > 0 ? m - 1 : 0);
                    const FT cDn = (m == 0) ? cD_up : val(3, m > 0 ? m - 1 : 0);
                    if constexpr (F32) {
                        G[1] = fmaf(-wiP, (top ? 0.0f : cPn) - P, G[1]);
                        G[3] = fmaf(-wiD, (top ? 0.0f : cDn) - Dd, G[3]);
                    } else {
                        const double FhiP = top ? 0.0 : wP * cPn, FhiD = top ? 0.0 : wD * cDn;
                        G[1] = G[1] - (FhiP - wP * P) * a.invdz;
                        G[3] = G[3] - (FhiD - wD * Dd) * a.invdz;
                    }
                    if (sed && m == mb) {
#pragma unroll
                        for (int n = 0; n < 4; n++)
                            if (n == a.sed_recv) G[n] = G[n] + cpl;
                    }
#pragma unroll
                    for (int n = 0; n < 4; n++) {
                        if constexpr (F32) {
                            const float A = (st == 0) ? a.rk_a[0] : (st == 1 ? a.rk_a[1] : a.rk_a[2]);
                            const float B = (st == 0) ? a.rk_b[0] : (st == 1 ? a.rk_b[1] : a.rk_b[2]);
                            D[n][m] = fmaf(A, G[n], fmaf(B, Gm[n][m], D[n][m]));
                        } else {
                            const double gs = (st == 1) ? g2 : g3;
                            const double zs = (st == 1) ? z2 : z3;
                            U[n][m] += (st == 0) ? (a.dtf * g1) * G[n] : a.dtf * (gs * G[n] + zs * Gm[n][m]);
                        }
                        Gm[n][m] = G[n];
                    }
                }
            }
            if constexpr (F32) {
                // U + D = s + e exactly (TwoSum): U = s, and e carries on
#pragma unroll
                for (int n = 0; n < 4; n++)
#pragma unroll
                    for (int m = 0; m < L; m++) {
                        const float u = U[n][m], d = D[n][m];
                        const float sum = u + d;
                        const float dv = sum - u;
                        D[n][m] = (u - (sum - dv)) + (d - dv);
                        U[n][m] = sum;
                    }
            }
        }
        if (sed && a.storage) {
            a.storage[col] = s_sed[0][seg];
            a.sedGm[col] = s_sed[1][seg];
        }
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int m = 0; m < L; m++)
                if (has(m)) {
                    a.U[n][idx(m)] = U[n][m];
                    if constexpr (F32) {
                        if (a.c_out) a.C[n][idx(m)] = D[n][m];
                    }
                }
    }
}

}  // namespace obme
