// obme_api.hip — kernels instantiation and the C ABI of libobme_hip.so.
//
// Entry points (declared in include/obme.h) and the reference interface
// each one stands in for are listed there.  All device work is queued on
// the caller's stream.  The batched ensemble records multi-launch calls into
// a hipGraph and replays it.  PISCES lives in obme_pisces.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "obme_batch_kernels.h"
#include "obme_internal.h"
#include "obme_npd_kernels.h"

using namespace obme;

// ------------------------------------------------------------------ errors

static thread_local std::string g_last_error;

obme_status obme_fail(obme_status code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

static obme_status fail(obme_status code, const std::string& msg) { return obme_fail(code, msg); }

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(OBME_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));   \
    } while (0)

extern "C" const char* obme_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* obme_version(void) { return "obme 0.1.0 gfx950"; }

// ------------------------------------------------------------------ kernels

// The sweep: lane per water column, workgroups of blockDim.x = 64 NW
// consecutive columns of one row (obme_npd_kernels.h).  FP32 is held to 128
// VGPRs (4 waves per SIMD, the occupancy at which the 512x512 grid's 4096
// waves fill the chip once).
constexpr int SWEEP_MAX_THREADS = 512;
template <int MODEL, int CARB, typename FT, int MODE, bool UNIFORM, int FM = -1, int EXTRA = 0, bool MASKX = true,
          int SCHEME = OBME_UPWIND1>
__global__ void __launch_bounds__(SWEEP_MAX_THREADS, (sizeof(FT) == 4 && (EXTRA & 2) == 0) ? 4 : 1)
k_column_lane(StepArgs<FT> a)
{
    column_lane<MODEL, CARB, FT, MODE, UNIFORM, FM, EXTRA, MASKX, true, SCHEME>(a, (int)blockDim.x);
}

// The O2 gas exchange's per-column coefficients (GasExchange.jl:127-146):
// k / Δz_top and the air-side concentration C_air α at the column's surface
// T and S (T and S are fixed during a call), one thread per column.
template <typename FT>
__global__ void __launch_bounds__(256) k_o2_prep(StepArgs<FT> a, obme_o2_exchange g, double invdz_top, double* out)
{
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (long)a.Nx * a.Ny) return;
    const int i = (int)(col % a.Nx), j = (int)(col / a.Nx), ktop = a.Nz - 1;
    const unsigned top = a.base0 + (unsigned)i + a.sx * (unsigned)j + a.sxy * (unsigned)ktop;
    const double T = a.T3 ? (double)a.T3[top] : (a.Tprof ? a.Tprof[ktop] : a.Tconst);
    const double S = a.S3 ? (double)a.S3[top] : (a.Sprof ? a.Sprof[ktop] : a.Sconst);
    double k, air;
    o2_exchange_coefficients(g, T, S, k, air);
    out[2 * col] = k * invdz_top;
    out[2 * col + 1] = air;
}

// ScaleNegativeTracers on every interior cell, in place (the modifier that
// runs first in update_biogeochemical_state!, src/OceanBioME.jl:161-167;
// src/Utils/negative_tracers.jl:198-251): one thread per cell.
template <typename FT, int NT>
__global__ void __launch_bounds__(256) k_scale_grid(StepArgs<FT> a)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ncol = (long)a.Nx * a.Ny;
    if (c >= ncol * a.Nz) return;
    const int k = (int)(c / ncol);
    const int col = (int)(c % ncol);
    const unsigned p = a.base0 + (unsigned)(col % a.Nx) + a.sx * (unsigned)(col / a.Nx) + a.sxy * (unsigned)k;
    unsigned members = 0;
    for (int g = 0; g < a.ns.n_groups; g++) members |= a.ns.mask[g];
    FT x[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) x[n] = (members & (1u << n)) ? a.tr[n][p] : FT(0);
    scale_negatives_cell<NT>(a.ns, x);
#pragma unroll
    for (int n = 0; n < NT; n++)
        if (members & (1u << n)) a.tr[n][p] = x[n];
}

// The InstantRemineralisation sediment at the bottom cell of column col:
// the sinking flux f through the bottom face and the tendency e f
// (instant_remineralisation.jl:101-123; tracked_fields.jl:53-70).
template <typename FT>
__device__ __forceinline__ void sediment_flux(const StepArgs<FT>& a, int s0, int s1, unsigned col, FT& f, FT& e)
{
    const unsigned p = a.base0 + col % (unsigned)a.Nx + a.sx * (col / (unsigned)a.Nx);
    f = a.sed_w0 * a.tr[s0][p] + a.sed_w1 * a.tr[s1][p];
    const FT qf = f / (a.sed_k + f);
    e = a.sed_a + a.sed_b * (qf * qf);
}

// update_tendencies!(bgc, sediment, model) (Sediments/tracer_coupling.jl:3-38):
// G[receiver] of the bottom cell += (1 - e) f / Δz, one thread per column.
template <typename FT>
__global__ void __launch_bounds__(256) k_sediment_couple(StepArgs<FT> a, int s0, int s1)
{
    const unsigned col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (unsigned)a.Nx * (unsigned)a.Ny) return;
    const unsigned p = a.base0 + col % (unsigned)a.Nx + a.sx * (col / (unsigned)a.Nx);
    if (a.sed_model == OBME_SEDIMENT_SIMPLE_MULTI_G) {
        // NH4, NO3 and O2 from the sediment's fields (none given: an empty
        // sediment) and the bottom cell (simple_multi_G.jl:195-247)
        const unsigned ncol = (unsigned)a.Nx * (unsigned)a.Ny;
        double u[3] = {0.0, 0.0, 0.0};
        if (a.sed_fields)
            for (int q = 0; q < 3; q++) u[q] = a.sed_fields[q * ncol + col];
        double cnh4, cno3, co2;
        smg_coupled(a.smg, u, (double)a.tr[a.smg.no3][p], (double)a.tr[a.smg.nh4][p], (double)a.tr[a.smg.o2][p],
                    cnh4, cno3, co2);
        const double dz0 = (double)a.sed_dz0;
        a.G[a.smg.nh4][p] += (FT)(cnh4 / dz0);
        a.G[a.smg.no3][p] += (FT)(cno3 / dz0);
        a.G[a.smg.o2][p] += (FT)(co2 / dz0);
        return;
    }
    FT f, e;
    sediment_flux(a, s0, s1, col, f, e);
    a.G[a.sed_recv][p] += ((FT(1) - e) * f) / a.sed_dz0;
}

// One time step of the sediment's storage (Sediments/update_state.jl:6-17;
// QuasiAdamsBashforth2, timesteppers.jl:16-43): the tendency Gⁿ (at the
// first step, e f of the tracked flux), a forward Euler step or the
// χ-weighted step with G⁻, then G⁻ = Gⁿ and Gⁿ = e f.  One thread per column.
template <typename FT>
__global__ void __launch_bounds__(256) k_sediment_state(StepArgs<FT> a, int s0, int s1, double dt, int euler)
{
    const unsigned col = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned ncol = (unsigned)a.Nx * (unsigned)a.Ny;
    if (col >= ncol) return;
    FT f, e;
    sediment_flux(a, s0, s1, col, f, e);
    if (a.sed_model == OBME_SEDIMENT_SIMPLE_MULTI_G) {
        // the three fields Ns, Nf, Nr, their Gⁿ from the fields and the flux
        double u[3], Gn[3], Gnew[3];
        for (int q = 0; q < 3; q++) u[q] = a.storage ? a.storage[q * ncol + col] : 0.0;
        if (a.sed_init) smg_tendencies(a.smg, u, (double)f, Gn);
        else
            for (int q = 0; q < 3; q++) Gn[q] = a.sedG[q * ncol + col];
        for (int q = 0; q < 3; q++) {
            const double Gu = euler ? Gn[q] : a.sed_cn * Gn[q] - a.sed_cm * a.sedG[(3 + q) * ncol + col];
            u[q] = u[q] + dt * Gu;
            a.sedG[(3 + q) * ncol + col] = Gn[q];
        }
        smg_tendencies(a.smg, u, (double)f, Gnew);
        for (int q = 0; q < 3; q++) {
            if (a.storage) a.storage[q * ncol + col] = u[q];
            a.sedG[q * ncol + col] = Gnew[q];
        }
        return;
    }
    const double ge = (double)(e * f);
    const double Gn = a.sed_init ? ge : a.sedG[col];
    // Euler: (1.5 + χ) Gⁿ with χ = -1/2, and G⁻ multiplied by a strong zero
    const double Gu = euler ? Gn : a.sed_cn * Gn - a.sed_cm * a.sedG[ncol + col];
    if (a.storage) a.storage[col] = a.storage[col] + dt * Gu;
    a.sedG[ncol + col] = Gn;
    a.sedG[col] = ge;
}

// The model's default formulations (fm_code): LOBSTER Mondo light,
// quadratic mortality and grazing; NPZD analytical light, linear
// mortality, quadratic grazing, Q10 on.  The FP32 fused step on full
// columns is compiled for these with the choices folded in.
template <int MODEL>
constexpr int default_fm()
{
    return MODEL == OBME_LOBSTER ? (0 | (1 << 1) | (1 << 2)) : (1 | (0 << 1) | (1 << 2) | (1 << 3));
}

// CO2 flux and pH on n points (SoA, FP64).
__global__ void __launch_bounds__(256) k_co2_flux(const double* __restrict__ DIC, const double* __restrict__ Alk,
                                                  const double* __restrict__ T, const double* __restrict__ S,
                                                  const double* __restrict__ u10, double pCO2_air,
                                                  double* __restrict__ out_pH, double* __restrict__ out_flux,
                                                  long n, int mode, obme_co2_exchange g, obme_carbon_solver s,
                                                  unsigned long long* __restrict__ evals)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int ev = 0;
    if (idx < n) {
        double u = u10 ? u10[idx] : g.wind_speed;
        double pH;
        double F = co2_flux(g, s, DIC[idx], Alk[idx], T[idx], S[idx], u, pCO2_air, mode, pH, ev);
        if (out_pH) out_pH[idx] = pH;
        if (out_flux) out_flux[idx] = F;
    }
    if (evals) {
        // wave-level reduction of the evaluation count, one atomic per wave
        unsigned long long v = (unsigned long long)ev;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(evals, v);
    }
}

// ------------------------------------------------------------------ handle

struct obme_handle {
    obme_config cfg;
    int device = 0;            // the device current at obme_create (its buffers live there)
    int ntracers;
    int prec;
    std::vector<double> zc, zf, dz, Tprof, Sprof;
    void* d_lev = nullptr;     // LEV_COUNT * Nz of FT
    void* d_wsink = nullptr;   // MAX_SINK * (Nz+1) of FT
    double* d_Tprof = nullptr;
    double* d_Sprof = nullptr;
    void* d_srec = nullptr;    // SURF_REC * Nx*Ny of FT (CO2 exchange on)
    double* d_o2c = nullptr;   // 2*Nx*Ny: k / Δz_top and C_air α of the O2 exchange
    bool o2c_ready = false;    // d_o2c holds the current obme_step call's coefficients
    double* d_Hprev = nullptr; // Nx*Ny warm starts of [H+]
    float* d_topres = nullptr; // 2*Nx*Ny residuals of the top-cell DIC and Alk (mixed path, CO2 exchange)
    double* d_surfc = nullptr; // [SURF_PACK][surfc_cols] surface constants of the current stepper call
    int surfc_cols = 0;        // Nx*Ny, or 1 when the call's T and S are the same in every column
    bool surfc_ready = false;  // d_surfc holds the constants of this call's T and S
    int sink_idx[MAX_SINK];
    int sweep_waves;           // waves per workgroup of the sweep (1, 2, 4 or 8)
    bool uniform;
    bool scaling;              // ScaleNegativeTracers groups configured
    bool co2_inline;           // scaling with the CO2 flux: the flux is computed in the sweep
    bool sediment;             // a sediment configured (InstantRemineralisation or SimpleMultiG)
    int sed_nf = 0;            // its fields: 1 (storage) or 3 (Ns, Nf, Nr)
    const double* hook_sed_fields = nullptr;   // the SimpleMultiG fields of the current hook call
    int scheme;                // the sinking face value (OBME_UPWIND1, OBME_CENTERED2, OBME_WENO5)
    bool staged;               // centred / WENO sinking: steps run one stage per launch
    int cur_stage = 0;         // the stage of the next TILE_STAGE launch
    void* d_stageD = nullptr;  // staged FP32: NT * Nx*Ny*Nz, the step's increment so far
    void* d_stageG = nullptr;  // staged: NT * Nx*Ny*Nz, the previous stage's tendencies
    double* d_sflux = nullptr; // staged with the CO2 exchange: Nx*Ny top-cell fluxes of the stage
    double* d_sedG = nullptr;  // the sediment time stepper's Gⁿ, G⁻ (2 * Nx*Ny)
    long long sed_iter = 0;    // sediment steps taken (its clock's iteration)
    double sed_last_dt = HUGE_VAL;   // Δt of its last step (Inf before the first)
    double dz_u, zc0_u, w0_u, w1_u;
    // kernel timing: event pairs per kernel kind, and a pool of free events
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timed[OBME_KERNEL_COUNT];
    std::vector<hipEvent_t> event_pool;
};

// Brackets one launch with events on stream s when timing is enabled.
struct LaunchTimer {
    obme_handle* h;
    int which;
    hipStream_t s;
    hipEvent_t e0 = nullptr;
    LaunchTimer(obme_handle* h_, int which_, hipStream_t s_) : h(h_), which(which_), s(s_)
    {
        if (h->timing && take(e0)) (void)hipEventRecord(e0, s);
    }
    bool take(hipEvent_t& e)
    {
        if (!h->event_pool.empty()) {
            e = h->event_pool.back();
            h->event_pool.pop_back();
            return true;
        }
        return hipEventCreate(&e) == hipSuccess;
    }
    void done()
    {
        if (!e0) return;
        hipEvent_t e1;
        if (take(e1) && hipEventRecord(e1, s) == hipSuccess) h->timed[which].push_back({e0, e1});
        else h->event_pool.push_back(e0);
        e0 = nullptr;
    }
};

// InstantRemineralisation: receiver a written tracer, sinking tracers valid
// and distinct, the default time stepper.
static obme_status validate_sediment(const obme_sediment& s, int ntracers, int readonly)
{
    if (s.model == OBME_SEDIMENT_NONE) return OBME_OK;
    if (s.model != OBME_SEDIMENT_INSTANT_REMINERALISATION && s.model != OBME_SEDIMENT_SIMPLE_MULTI_G)
        return fail(OBME_ERR_INVALID, "unknown sediment model");
    auto written = [&](int t) { return t >= 0 && t < ntracers && t != readonly; };
    if (s.model == OBME_SEDIMENT_INSTANT_REMINERALISATION && !written(s.receiver))
        return fail(OBME_ERR_INVALID, "sediment receiver is not a written tracer");
    if (s.model == OBME_SEDIMENT_SIMPLE_MULTI_G) {
        if (!written(s.ammonia) || !written(s.nitrate) || !written(s.oxygen) || s.ammonia == s.nitrate ||
            s.ammonia == s.oxygen || s.nitrate == s.oxygen)
            return fail(OBME_ERR_INVALID, "SimpleMultiG needs three distinct written tracers NH4, NO3 and O2");
        if (!(s.sedimentation_rate > 0.0))
            return fail(OBME_ERR_INVALID, "SimpleMultiG sedimentation rate must be positive");
    }
    if (s.n_sinking < 1 || s.n_sinking > OBME_MAX_SEDIMENT_TRACERS)
        return fail(OBME_ERR_INVALID, "bad number of sediment sinking tracers");
    unsigned seen = 0;
    for (int q = 0; q < s.n_sinking; q++) {
        const int t = s.sinking[q];
        if (t < 0 || t >= ntracers) return fail(OBME_ERR_INVALID, "sediment sinking tracer is not a tracer");
        if (seen & (1u << t)) return fail(OBME_ERR_INVALID, "sediment sinking tracer listed twice");
        seen |= 1u << t;
    }
    if (s.timestepper != OBME_SEDIMENT_AB2)
        return fail(OBME_ERR_UNSUPPORTED, "only the QuasiAdamsBashforth2 sediment time stepper is supported");
    return OBME_OK;
}

// ScaleNegativeTracers groups: counts within the limits, members valid
// tracer indices, no member twice in a group.
static obme_status validate_scaling(const obme_negative_scaling& s, int ntracers)
{
    if (s.n_groups < 0 || s.n_groups > OBME_MAX_SCALE_GROUPS) return fail(OBME_ERR_INVALID, "bad number of scaling groups");
    for (int g = 0; g < s.n_groups; g++) {
        if (s.n_members[g] < 1 || s.n_members[g] > OBME_MAX_SCALE_MEMBERS)
            return fail(OBME_ERR_INVALID, "bad number of scaling group members");
        unsigned long long seen = 0;
        for (int m = 0; m < s.n_members[g]; m++) {
            const int t = s.member[g][m];
            if (t < 0 || t >= ntracers || t >= 64) return fail(OBME_ERR_INVALID, "scaling member is not a tracer");
            if (seen & (1ull << t)) return fail(OBME_ERR_INVALID, "scaling member listed twice in a group");
            seen |= 1ull << t;
        }
    }
    return OBME_OK;
}

template <typename FT, int NTMAX>
static NegScale<FT, NTMAX> convert_scaling(const obme_negative_scaling& s)
{
    NegScale<FT, NTMAX> d;
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

static int ntracers_of(const obme_config& c)
{
    if (c.model == OBME_NPZD) return c.oxygen.enabled ? 6 : 5;
    if (c.model == OBME_LOBSTER) return (c.carbonate_system ? 9 : 7) + (c.oxygen.enabled ? 1 : 0);
    return 0;
}

template <typename FT>
static obme_status upload_geometry(obme_handle* h)
{
    const int Nz = h->cfg.grid.Nz;
    std::vector<FT> lev((size_t)LEV_COUNT * Nz);
    for (int k = 0; k < Nz; k++) {
        lev[LEV_ZC * Nz + k] = (FT)h->zc[k];
        if (k < Nz - 1) {
            lev[LEV_UP * Nz + k] = (FT)(h->zc[k + 1] - h->zf[k + 1]);
            lev[LEV_DN * Nz + k] = (FT)(h->zf[k + 1] - h->zc[k]);
        } else {
            lev[LEV_UP * Nz + k] = (FT)0;
            lev[LEV_DN * Nz + k] = (FT)(h->zf[Nz] - h->zc[Nz - 1]);
        }
        lev[LEV_INVDZ * Nz + k] = (FT)(1.0 / h->dz[k]);
    }
    HIP_TRY(hipMalloc(&h->d_lev, lev.size() * sizeof(FT)));
    HIP_TRY(hipMemcpy(h->d_lev, lev.data(), lev.size() * sizeof(FT), hipMemcpyHostToDevice));
    // drift velocities (Utils/sinking_velocity_fields.jl:10-35), faces 1..Nz+1
    std::vector<FT> w((size_t)MAX_SINK * (Nz + 1), (FT)0);
    const obme_sinking& sk = h->cfg.sinking;
    double d = sk.smoothing_distance > 0 ? sk.smoothing_distance : 2.0;
    for (int s = 0; s < MAX_SINK; s++) {
        int n = h->sink_idx[s];
        double speed = (n >= 0) ? sk.speed[n] : 0.0;
        for (int k = 1; k <= Nz; k++)
            w[(size_t)s * (Nz + 1) + (k - 1)] = (FT)(-speed * (sk.open_bottom ? 1.0 : (1 - std::exp((1 - k) / d))));
        w[(size_t)s * (Nz + 1) + Nz] = (FT)0;
    }
    HIP_TRY(hipMalloc(&h->d_wsink, w.size() * sizeof(FT)));
    HIP_TRY(hipMemcpy(h->d_wsink, w.data(), w.size() * sizeof(FT), hipMemcpyHostToDevice));

    // Uniform-grid fast path: the kernel forms zc = zc[0] + k dz, the half
    // cells dz/2, 1/dz and constant face velocities itself.  Taken only when
    // those reproduce every per-level value exactly in the arithmetic type
    // (with or without contraction into an FMA).
    const FT dzf = (FT)h->dz[0];
    const FT zc0 = lev[LEV_ZC * Nz + 0];
    bool uni = sk.open_bottom != 0;
    for (int k = 0; k < Nz && uni; k++) {
        FT zk = zc0 + (FT)k * dzf;
        FT zk_fma = (FT)std::fma((FT)k, dzf, zc0);
        uni = uni && (zk == lev[LEV_ZC * Nz + k]) && (zk_fma == zk);
        uni = uni && (lev[LEV_DN * Nz + k] == dzf / (FT)2);
        uni = uni && (lev[LEV_UP * Nz + k] == ((k < Nz - 1) ? dzf / (FT)2 : (FT)0));
        uni = uni && (lev[LEV_INVDZ * Nz + k] == (FT)1 / dzf);
    }
    h->uniform = uni;
    h->dz_u = (double)dzf;
    h->zc0_u = (double)zc0;
    h->w0_u = (double)w[0];
    h->w1_u = (double)w[(size_t)(Nz + 1)];
    return OBME_OK;
}

// Waves per workgroup of the sweep: a workgroup covers 64 NW consecutive
// columns of a row.  4 measured fastest on the 512x512x64 LOBSTER step;
// obme_autotune times every choice.
static const int SWEEP_WAVES[] = {1, 2, 4, 8};
static const int N_SWEEP_WAVES = sizeof(SWEEP_WAVES) / sizeof(SWEEP_WAVES[0]);
static bool valid_sweep_waves(int nw)
{
    for (int q = 0; q < N_SWEEP_WAVES; q++)
        if (SWEEP_WAVES[q] == nw) return true;
    return false;
}

extern "C" obme_status obme_create(const obme_config* cfg, obme_handle** out)
{
    if (!cfg || !out) return fail(OBME_ERR_INVALID, "null argument");
    *out = nullptr;
    const obme_grid& g = cfg->grid;
    if (cfg->model != OBME_NPZD && cfg->model != OBME_LOBSTER) return fail(OBME_ERR_INVALID, "unknown model");
    if (cfg->precision != OBME_FP64 && cfg->precision != OBME_FP32_MIXED)
        return fail(OBME_ERR_INVALID, "unknown precision");
    if (cfg->model == OBME_NPZD && cfg->carbonate_system)
        return fail(OBME_ERR_UNSUPPORTED, "the carbonate system is supported with LOBSTER");
    if (cfg->o2.enabled && !(cfg->model == OBME_LOBSTER && cfg->oxygen.enabled))
        return fail(OBME_ERR_INVALID, "the O2 gas exchange needs LOBSTER with the oxygen component");
    if (g.Nx < 1 || g.Ny < 1 || g.Nz < 1 || g.Hx < 0 || g.Hy < 0 || g.Hz < 0)
        return fail(OBME_ERR_INVALID, "invalid grid size or halo");
    {
        size_t elems = (size_t)(g.Nx + 2 * g.Hx) * (g.Ny + 2 * g.Hy) * (g.Nz + 2 * g.Hz);
        size_t es = (cfg->precision == OBME_FP64) ? 8 : 4;
        if (elems * es >= ((size_t)1 << 31))
            return fail(OBME_ERR_UNSUPPORTED, "a field of 2 GiB or more exceeds the 32-bit buffer offsets");
    }
    if (!g.zc || !g.zf || !g.dz) return fail(OBME_ERR_INVALID, "grid coordinates missing");
    {
        obme_status sv = validate_scaling(cfg->scaling, ntracers_of(*cfg));
        if (sv != OBME_OK) return sv;
        sv = validate_sediment(cfg->sediment, ntracers_of(*cfg), cfg->model == OBME_NPZD ? 3 : -1);
        if (sv != OBME_OK) return sv;
    }
    if (cfg->sinking.scheme != OBME_UPWIND1 && cfg->sinking.scheme != OBME_CENTERED2 &&
        cfg->sinking.scheme != OBME_WENO5)
        return fail(OBME_ERR_INVALID, "unknown sinking scheme");
    if (cfg->sinking.scheme != OBME_UPWIND1 && (cfg->scaling.n_groups > 0 || cfg->sediment.model != OBME_SEDIMENT_NONE))
        return fail(OBME_ERR_UNSUPPORTED,
                    "the centred and WENO sinking schemes run without negative scaling and without a sediment");
    if (cfg->par.surface_mode == OBME_SURFACE_ARRAY && !cfg->par.surface_array)
        return fail(OBME_ERR_INVALID, "surface PAR array missing");

    obme_handle* h = new obme_handle();
    h->cfg = *cfg;
    h->ntracers = ntracers_of(*cfg);
    h->prec = cfg->precision;
    h->zc.assign(g.zc, g.zc + g.Nz);
    h->zf.assign(g.zf, g.zf + g.Nz + 1);
    h->dz.assign(g.dz, g.dz + g.Nz);
    h->cfg.grid.zc = h->zc.data();
    h->cfg.grid.zf = h->zf.data();
    h->cfg.grid.dz = h->dz.data();
    // sinking tracers must match the kernels' compile-time sets
    int s0, s1;
    if (cfg->model == OBME_NPZD) { s0 = 1; s1 = 4; }
    else { s0 = 4; s1 = 5; }
    for (int n = 0; n < h->ntracers; n++) {
        if (cfg->sinking.speed[n] < 0) { delete h; return fail(OBME_ERR_INVALID, "negative sinking speed"); }
        if (cfg->sinking.speed[n] != 0 && n != s0 && n != s1) {
            delete h;
            return fail(OBME_ERR_UNSUPPORTED, "drift velocity on a tracer the kernels do not sink");
        }
    }
    h->sink_idx[0] = s0;
    h->sink_idx[1] = s1;
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return fail(OBME_ERR_HIP, "hipGetDevice failed");
    }
    // field sources
    auto copy_prof = [&](const obme_field_source& f, std::vector<double>& v, double** d) -> obme_status {
        if (f.mode == OBME_FIELD_PROFILE) {
            if (!f.profile) return fail(OBME_ERR_INVALID, "profile missing");
            v.assign(f.profile, f.profile + g.Nz);
            HIP_TRY(hipMalloc((void**)d, g.Nz * sizeof(double)));
            HIP_TRY(hipMemcpy(*d, v.data(), g.Nz * sizeof(double), hipMemcpyHostToDevice));
        }
        return OBME_OK;
    };
    obme_status st = copy_prof(cfg->temperature, h->Tprof, &h->d_Tprof);
    if (st == OBME_OK) st = copy_prof(cfg->salinity, h->Sprof, &h->d_Sprof);
    if (st == OBME_OK) st = (h->prec == OBME_FP64) ? upload_geometry<double>(h) : upload_geometry<float>(h);
    if (st != OBME_OK) {
        obme_destroy(h);
        return st;
    }
    // a workgroup's waves walk in step, so wider workgroups make longer
    // contiguous runs per level: the widest that a row fills
    h->sweep_waves = g.Nx >= 512 ? 8 : g.Nx >= 256 ? 4 : g.Nx >= 128 ? 2 : 1;
    if (const char* e = std::getenv("OBME_SWEEP_WAVES")) {
        const int nw = std::atoi(e);
        if (valid_sweep_waves(nw)) h->sweep_waves = nw;
    }
    h->scaling = cfg->scaling.n_groups > 0;
    h->scheme = cfg->sinking.scheme;
    h->staged = h->scheme != OBME_UPWIND1;
    h->co2_inline = h->scaling && cfg->model == OBME_LOBSTER && cfg->carbonate_system && cfg->co2.enabled;
    if (cfg->model == OBME_LOBSTER && cfg->oxygen.enabled && cfg->o2.enabled) {
        if (hipMalloc((void**)&h->d_o2c, 2 * (size_t)g.Nx * g.Ny * sizeof(double)) != hipSuccess) {
            obme_destroy(h);
            return fail(OBME_ERR_HIP, "hipMalloc of the O2 exchange coefficients failed");
        }
    }
    if (cfg->model == OBME_LOBSTER && cfg->carbonate_system && cfg->co2.enabled) {
        size_t ncol = (size_t)g.Nx * g.Ny;
        size_t es = (h->prec == OBME_FP64) ? 8 : 4;
        hipError_t e1 = hipMalloc(&h->d_srec, SURF_REC * ncol * es);
        hipError_t e2 = hipMalloc((void**)&h->d_Hprev, ncol * sizeof(double));
        hipError_t e3 = (h->prec == OBME_FP64) ? hipSuccess : hipMalloc((void**)&h->d_topres, 2 * ncol * sizeof(float));
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
            obme_destroy(h);
            return fail(OBME_ERR_NOMEM, "surface buffers");
        }
        if (hipMemset(h->d_Hprev, 0, ncol * sizeof(double)) != hipSuccess) {
            obme_destroy(h);
            return fail(OBME_ERR_HIP, "hipMemset");
        }
    }
    h->sediment = cfg->sediment.model != OBME_SEDIMENT_NONE;
    h->sed_nf = cfg->sediment.model == OBME_SEDIMENT_SIMPLE_MULTI_G ? 3 : 1;
    if (h->sediment) {
        size_t bytes = 2 * (size_t)h->sed_nf * g.Nx * g.Ny * sizeof(double);
        if (hipMalloc((void**)&h->d_sedG, bytes) != hipSuccess) {
            obme_destroy(h);
            return fail(OBME_ERR_NOMEM, "sediment buffers");
        }
        if (hipMemset(h->d_sedG, 0, bytes) != hipSuccess) {
            obme_destroy(h);
            return fail(OBME_ERR_HIP, "hipMemset");
        }
    }
    *out = h;
    return OBME_OK;
}

extern "C" obme_status obme_destroy(obme_handle* h)
{
    if (!h) return OBME_OK;
    if (h->d_lev) (void)hipFree(h->d_lev);
    if (h->d_wsink) (void)hipFree(h->d_wsink);
    if (h->d_Tprof) (void)hipFree(h->d_Tprof);
    if (h->d_Sprof) (void)hipFree(h->d_Sprof);
    if (h->d_srec) (void)hipFree(h->d_srec);
    if (h->d_Hprev) (void)hipFree(h->d_Hprev);
    if (h->d_topres) (void)hipFree(h->d_topres);
    if (h->d_o2c) (void)hipFree(h->d_o2c);
    if (h->d_sedG) (void)hipFree(h->d_sedG);
    if (h->d_surfc) (void)hipFree(h->d_surfc);
    if (h->d_stageD) (void)hipFree(h->d_stageD);
    if (h->d_stageG) (void)hipFree(h->d_stageG);
    if (h->d_sflux) (void)hipFree(h->d_sflux);
    for (auto& v : h->timed)
        for (auto& p : v) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
    for (hipEvent_t e : h->event_pool) (void)hipEventDestroy(e);
    delete h;
    return OBME_OK;
}

extern "C" int32_t obme_required_tracers(const obme_handle* h, char* buf, int32_t cap)
{
    if (!h) return -1;
    const char* names;
    if (h->cfg.model == OBME_NPZD) names = h->cfg.oxygen.enabled ? "N;P;Z;T;D;O₂" : "N;P;Z;T;D";
    else if (h->cfg.carbonate_system)
        names = h->cfg.oxygen.enabled ? "NO₃;NH₄;P;Z;sPOM;bPOM;DOM;DIC;Alk;O₂" : "NO₃;NH₄;P;Z;sPOM;bPOM;DOM;DIC;Alk";
    else names = h->cfg.oxygen.enabled ? "NO₃;NH₄;P;Z;sPOM;bPOM;DOM;O₂" : "NO₃;NH₄;P;Z;sPOM;bPOM;DOM";
    if (buf && cap > 0) {
        std::strncpy(buf, names, cap - 1);
        buf[cap - 1] = 0;
    }
    return h->ntracers;
}

// the sweep takes its arguments by value: within the 4 KiB kernel-argument limit
static_assert(sizeof(StepArgs<double>) <= 4096 && sizeof(StepArgs<float>) <= 4096, "StepArgs exceeds 4 KiB");

template <typename FT>
static StepArgs<FT> make_args(const obme_handle* h, void* const* tracers)
{
    const obme_config& c = h->cfg;
    const obme_grid& g = c.grid;
    StepArgs<FT> a;
    std::memset(&a, 0, sizeof(a));
    for (int n = 0; n < h->ntracers; n++) a.tr[n] = tracers ? (FT*)tracers[n] : nullptr;
    a.lev = (const FT*)h->d_lev;
    a.wsink = (const FT*)h->d_wsink;
    a.surf = (const FT*)c.par.surface_array;
    a.srec = (FT*)h->d_srec;
    a.Nx = g.Nx; a.Ny = g.Ny; a.Nz = g.Nz;
    a.sx = (unsigned)(g.Nx + 2 * g.Hx);
    a.sxy = a.sx * (unsigned)(g.Ny + 2 * g.Hy);
    a.base0 = (unsigned)g.Hx + a.sx * (unsigned)g.Hy + a.sxy * (unsigned)g.Hz;
    a.field_bytes = (unsigned)((size_t)a.sxy * (size_t)(g.Nz + 2 * g.Hz) * sizeof(FT));
    a.q = convert_params<FT>(c.npd);
    if (c.oxygen.enabled) {
        a.q.Rp = (FT)c.oxygen.respiration_oxygen_nitrogen_ratio;
        a.q.Rn = (FT)c.oxygen.nitrification_oxygen_nitrogen_ratio;
    }
    a.kr = (FT)c.par.water_red_attenuation;
    a.kb = (FT)c.par.water_blue_attenuation;
    a.xr = (FT)c.par.chlorophyll_red_attenuation;
    a.xb = (FT)c.par.chlorophyll_blue_attenuation;
    a.er = (FT)c.par.chlorophyll_red_exponent;
    a.eb = (FT)c.par.chlorophyll_blue_exponent;
    // (P R_chl / r) is formed as P * (R_chl / r): a reassociation of the
    // reference's (P * R_chl) / r (Light/2band.jl:20).
    a.Rr = (FT)(c.par.phytoplankton_chlorophyll_ratio / c.par.pigment_ratio);
    a.surface_mode = c.par.surface_mode;
    a.surface_value = (FT)c.par.surface_value;
    a.dz_u = (FT)h->dz_u;
    a.zc0_u = (FT)h->zc0_u;
    a.w0_u = (FT)h->w0_u;
    a.w1_u = (FT)h->w1_u;
    a.invdz_u = (FT)1 / (FT)h->dz_u;
    a.co2_on = (c.model == OBME_LOBSTER && c.carbonate_system && c.co2.enabled && !h->co2_inline) ? 1 : 0;
    a.ns = convert_scaling<FT, 16>(c.scaling);
    a.o2_on = (c.model == OBME_LOBSTER && c.oxygen.enabled && c.o2.enabled) ? 1 : 0;
    a.o2c = h->d_o2c;
    if (h->co2_inline || a.o2_on || h->staged) {
        a.Tprof = (c.temperature.mode == OBME_FIELD_PROFILE) ? h->d_Tprof : nullptr;
        a.Sprof = (c.salinity.mode == OBME_FIELD_PROFILE) ? h->d_Sprof : nullptr;
        a.Tconst = c.temperature.value;
        a.Sconst = c.salinity.value;
        a.Hprev = h->d_Hprev;
        a.invdz_top = (FT)(1.0 / h->dz[g.Nz - 1]);
        a.co2 = c.co2;
        a.solver = c.solver;
    }
    if (h->sediment) {
        // -w at the bottom face (Utils/sinking_velocity_fields.jl:16, k = 1):
        // the speed with an open bottom, 0 with a closed one
        const obme_sediment& sd = c.sediment;
        auto listed = [&](int t) {
            for (int q = 0; q < sd.n_sinking; q++)
                if (sd.sinking[q] == t) return true;
            return false;
        };
        const int s0 = h->sink_idx[0], s1 = h->sink_idx[1];
        a.sed_w0 = (FT)((listed(s0) && c.sinking.open_bottom) ? c.sinking.speed[s0] : 0.0);
        a.sed_w1 = (FT)((listed(s1) && c.sinking.open_bottom) ? c.sinking.speed[s1] : 0.0);
        a.sed_recv = sd.receiver;
        a.sed_a = (FT)sd.burial_efficiency_constant1;
        a.sed_b = (FT)sd.burial_efficiency_constant2;
        a.sed_k = (FT)sd.burial_efficiency_half_saturation;
        a.sed_dz0 = (FT)h->dz[0];
        a.sedG = h->d_sedG;
        a.sed_cn = 1.5 + sd.ab2_chi;
        a.sed_cm = 0.5 + sd.ab2_chi;
        a.sed_model = sd.model;
        a.sed_nf = h->sed_nf;
        a.sed_fields = h->hook_sed_fields;
        if (sd.model == OBME_SEDIMENT_SIMPLE_MULTI_G) {
            SmgParams& q = a.smg;
            q.ls = sd.slow_decay_rate;
            q.lf = sd.fast_decay_rate;
            q.fs = sd.slow_fraction;
            q.ff = sd.fast_fraction;
            q.fr = sd.refactory_fraction;
            q.R = sd.sinking_redfield;
            q.kO2 = sd.anoxia_half_saturation;
            q.ps = 0.223 * std::pow(sd.sedimentation_rate, 0.336);   // solid_deposition_fraction
            for (int i = 0; i < 6; i++) {
                q.qn[i] = sd.nitrate_oxidation_params[i];
                q.qd[i] = sd.denitrification_params[i];
                q.qa[i] = sd.anoxic_params[i];
            }
            q.no3 = sd.nitrate;
            q.nh4 = sd.ammonia;
            q.o2 = sd.oxygen;
        }
    }
    if (sizeof(FT) == 4) {
        a.f = make_fast_params(c.npd);
        if (c.oxygen.enabled) {
            a.f.oRp = (float)c.oxygen.respiration_oxygen_nitrogen_ratio;
            a.f.oRpn = (float)(c.oxygen.respiration_oxygen_nitrogen_ratio - c.oxygen.nitrification_oxygen_nitrogen_ratio);
        }
        const double log2e = 1.4426950408889634;
        a.kr2 = (float)(c.par.water_red_attenuation * log2e);
        a.kb2 = (float)(c.par.water_blue_attenuation * log2e);
        a.xr2 = (float)(c.par.chlorophyll_red_attenuation * log2e);
        a.xb2 = (float)(c.par.chlorophyll_blue_attenuation * log2e);
        const double lRr = std::log2(c.par.phytoplankton_chlorophyll_ratio / c.par.pigment_ratio);
        a.cr = (float)(c.par.chlorophyll_red_exponent * lRr);
        a.cb = (float)(c.par.chlorophyll_blue_exponent * lRr);
        a.xr2h = (float)(c.par.chlorophyll_red_attenuation * log2e * h->dz_u / 2);
        a.xb2h = (float)(c.par.chlorophyll_blue_attenuation * log2e * h->dz_u / 2);
        a.wi0 = (float)(h->w0_u / h->dz_u);
        a.wi1 = (float)(h->w1_u / h->dz_u);
    }
    return a;
}

// RK3 coefficients Δt γ_s and Δt ζ_s (BoxModel/timesteppers.jl:79-92) for the FP32 update
static void rk3_products(double dt, float* A, float* B)
{
    A[0] = (float)(dt * (8.0 / 15.0));
    A[1] = (float)(dt * (5.0 / 12.0));
    A[2] = (float)(dt * (3.0 / 4.0));
    B[0] = 0.0f;
    B[1] = (float)(dt * (-17.0 / 60.0));
    B[2] = (float)(dt * (-5.0 / 12.0));
}

template <typename FT>
static SurfaceArgs<FT> make_surface_args(const obme_handle* h, void* const* tracers, const void* T, const void* S)
{
    const obme_config& c = h->cfg;
    const obme_grid& g = c.grid;
    SurfaceArgs<FT> s;
    std::memset(&s, 0, sizeof(s));
    s.DIC = (FT*)tracers[7];
    s.Alk = (const FT*)tracers[8];
    s.srec = (const FT*)h->d_srec;
    s.Hprev = h->d_Hprev;
    s.T3 = (const FT*)T;
    s.S3 = (const FT*)S;
    s.Tprof = (c.temperature.mode == OBME_FIELD_PROFILE) ? h->d_Tprof : nullptr;
    s.Sprof = (c.salinity.mode == OBME_FIELD_PROFILE) ? h->d_Sprof : nullptr;
    s.Tconst = c.temperature.value;
    s.Sconst = c.salinity.value;
    s.Nx = g.Nx; s.Ny = g.Ny; s.Nz = g.Nz;
    s.sx = (unsigned)(g.Nx + 2 * g.Hx);
    s.sxy = s.sx * (unsigned)(g.Ny + 2 * g.Hy);
    s.base0 = (unsigned)g.Hx + s.sx * (unsigned)g.Hy + s.sxy * (unsigned)g.Hz;
    s.invdz_top = (FT)(1.0 / h->dz[g.Nz - 1]);
    // both precisions: Newton warm-started from the column's previous [H+]
    // (FP32 to FP32_NEWTON_TOL, FP64 to a 1e-14 relative step), the
    // reference's damped Newton as the fallback
    s.mode = 1;
    s.co2 = c.co2;
    s.solver = c.solver;
    return s;
}

template <int MODEL, int CARB, typename FT, int MODE>
static obme_status launch_lane(obme_handle* h, const StepArgs<FT>& a, hipStream_t s)
{
    const int cw = 64 * h->sweep_waves;
    // one-dimensional: tile = xcd_tile(blockIdx.x), row-major over (i tile, j)
    const unsigned n_tiles = (unsigned)(((h->cfg.grid.Nx + cw - 1) / cw) * h->cfg.grid.Ny);
    using Kernel = void (*)(StepArgs<FT>);
    Kernel k = nullptr;
    if constexpr (MODE == TILE_STAGE) {
        // centred / WENO sinking, one stage per launch (the general-grid form)
        k = (h->scheme == OBME_WENO5) ? k_column_lane<MODEL, CARB, FT, MODE, false, -1, 0, true, OBME_WENO5>
                                      : k_column_lane<MODEL, CARB, FT, MODE, false, -1, 0, true, OBME_CENTERED2>;
    } else {
        k = h->uniform ? k_column_lane<MODEL, CARB, FT, MODE, true> : k_column_lane<MODEL, CARB, FT, MODE, false>;
    }
    // rows of whole workgroups (Nx a multiple of 64 NW): stores without a
    // mask; the FP32 step and tendencies compiled with the model's default
    // formulations
    const bool whole = h->cfg.grid.Nx % cw == 0;
    if constexpr (MODE == TILE_TENDENCY) {
        if (h->scheme == OBME_WENO5) {
            k = k_column_lane<MODEL, CARB, FT, MODE, false, -1, 0, true, OBME_WENO5>;
        } else if (h->scheme == OBME_CENTERED2) {
            k = k_column_lane<MODEL, CARB, FT, MODE, false, -1, 0, true, OBME_CENTERED2>;
        } else if (whole && h->uniform) {
            if constexpr (sizeof(FT) == 4) {
                if (fm_code(a.f) == default_fm<MODEL>())
                    k = k_column_lane<MODEL, CARB, FT, MODE, true, default_fm<MODEL>(), 0, false>;
            } else {
                k = k_column_lane<MODEL, CARB, FT, MODE, true, -1, 0, false>;
            }
        }
    }
    if constexpr (MODE == TILE_PAR) {
        if (whole && h->uniform) k = k_column_lane<MODEL, CARB, FT, MODE, true, -1, 0, false>;
    }
    if constexpr (MODE == TILE_STEP) {
        if constexpr (sizeof(FT) == 4) {
            if (fm_code(a.f) == default_fm<MODEL>()) {
                constexpr int FM = default_fm<MODEL>();
                k = whole ? (h->uniform ? k_column_lane<MODEL, CARB, FT, MODE, true, FM, 0, false>
                                        : k_column_lane<MODEL, CARB, FT, MODE, false, FM, 0, false>)
                          : (h->uniform ? k_column_lane<MODEL, CARB, FT, MODE, true, FM>
                                        : k_column_lane<MODEL, CARB, FT, MODE, false, FM>);
            }
        } else if (whole) {
            k = h->uniform ? k_column_lane<MODEL, CARB, FT, MODE, true, -1, 0, false>
                           : k_column_lane<MODEL, CARB, FT, MODE, false, -1, 0, false>;
        }
    }
    if constexpr (MODE == TILE_STEP) {
        // negative scaling (1), the inline CO2 flux (2, with 1) and the
        // sediment (4) in the fused step, with the extra work compiled in
        const int ex = (h->scaling ? 1 : 0) | (h->co2_inline ? 2 : 0) | (h->sediment ? 4 : 0);
#define OBME_EXTRA_KERNEL(E)                                                               \
    (h->uniform ? k_column_lane<MODEL, CARB, FT, MODE, true, -1, E>                       \
                : k_column_lane<MODEL, CARB, FT, MODE, false, -1, E>)
        if (ex == 1) k = OBME_EXTRA_KERNEL(1);
        if (ex == 4) k = OBME_EXTRA_KERNEL(4);
        if (ex == 5) k = OBME_EXTRA_KERNEL(5);
        if constexpr ((CARB & 1) != 0) {
            if (ex == 3) k = OBME_EXTRA_KERNEL(3);
            if (ex == 7) k = OBME_EXTRA_KERNEL(7);
        }
#undef OBME_EXTRA_KERNEL
    }
    LaunchTimer tm(h, OBME_KERNEL_SWEEP, s);
    hipLaunchKernelGGL(k, dim3(n_tiles), dim3(cw), 0, s, a);
    tm.done();
    HIP_TRY(hipGetLastError());
    return OBME_OK;
}

template <typename FT, int MODE>
static obme_status dispatch_model(obme_handle* h, const StepArgs<FT>& a, hipStream_t s)
{
    if (h->cfg.oxygen.enabled) {
        if (h->cfg.model == OBME_NPZD) return launch_lane<OBME_NPZD, 2, FT, MODE>(h, a, s);
        if (h->cfg.carbonate_system) return launch_lane<OBME_LOBSTER, 3, FT, MODE>(h, a, s);
        return launch_lane<OBME_LOBSTER, 2, FT, MODE>(h, a, s);
    }
    if (h->cfg.model == OBME_NPZD) return launch_lane<OBME_NPZD, 0, FT, MODE>(h, a, s);
    // LOBSTER's components: 1 CarbonateSystem, 2 Oxygen (above)
    if (h->cfg.carbonate_system) return launch_lane<OBME_LOBSTER, 1, FT, MODE>(h, a, s);
    return launch_lane<OBME_LOBSTER, 0, FT, MODE>(h, a, s);
}

template <typename FT, int MODE>
static obme_status launch_surface(obme_handle* h, const SurfaceArgs<FT>& sa, hipStream_t s)
{
    const int bs = 256;
    const long ncol = (long)sa.Nx * sa.Ny;
    LaunchTimer tm(h, OBME_KERNEL_SURFACE, s);
    hipLaunchKernelGGL((k_surface<FT, MODE>), dim3((ncol + bs - 1) / bs), dim3(bs), 0, s, sa);
    tm.done();
    HIP_TRY(hipGetLastError());
    return OBME_OK;
}

template <typename FT, int MODE>
static obme_status run_typed(obme_handle* h, void* const* tracers, void* const* G, void* PAR, const void* T,
                             const void* S, double t, double dt, hipStream_t s, double* storage = nullptr)
{
    StepArgs<FT> a = make_args<FT>(h, tracers);
    if (G) for (int n = 0; n < h->ntracers; n++) a.G[n] = (FT*)G[n];
    a.PAR = (FT*)PAR;
    a.dtf = (FT)dt;
    rk3_products(dt, a.rk_a, a.rk_b);
    if (h->cfg.par.surface_mode == OBME_SURFACE_DIURNAL) {
        // default_surface_PAR (Light/Light.jl:219) at the stage clocks
        const double cst[3] = {0.0, 8.0 / 15.0, 8.0 / 15.0 + 5.0 / 12.0 - 17.0 / 60.0};
        for (int q = 0; q < 3; q++) {
            double tq = t + cst[q] * dt;
            double cv = std::cos(tq * 3.141592653589793 / (12 * 3600.0));
            double d = 0.0 - cv;  // Julia max(0, cos): NaN propagates, ±0 ordered
            double v = std::isnan(cv) ? d : (std::signbit(d) ? cv : 0.0);
            a.PAR0_stage[q] = (FT)(100 * v);
        }
    }
    if (MODE != TILE_STEP) {
        // the hooks: the flux by the surface kernel in tendency mode
        const obme_config& c = h->cfg;
        a.co2_on = (c.model == OBME_LOBSTER && c.carbonate_system && c.co2.enabled) ? 1 : 0;
    }
    a.T3 = (const FT*)T;
    a.S3 = (const FT*)S;
    if constexpr (MODE == TILE_STAGE) {
        // one stage: the CO2 flux by the top lane, the step's scratch
        const obme_config& c = h->cfg;
        a.stage = h->cur_stage;
        a.co2_stage = (c.model == OBME_LOBSTER && c.carbonate_system && c.co2.enabled) ? 1 : 0;
        a.co2_on = 0;
        a.ncell = (unsigned)((size_t)c.grid.Nx * c.grid.Ny * c.grid.Nz);
        a.Dsc = (FT*)h->d_stageD;
        a.Gsc = (FT*)h->d_stageG;
        a.sflux = h->d_sflux;
    }
    // the top-cell DIC and Alk residuals (FP32 fused step, CO2 by the surface kernel)
    float* topres = ((MODE == TILE_STEP && a.co2_on) || (MODE == TILE_STAGE && a.co2_stage)) ? h->d_topres : nullptr;
    a.topres = topres;
    if (MODE != TILE_PAR && a.o2_on && !h->o2c_ready) {
        // the O2 exchange coefficients of this call's surface T and S (once
        // per obme_step call, which sets o2c_ready around its steps)
        const long ncol = (long)h->cfg.grid.Nx * h->cfg.grid.Ny;
        hipLaunchKernelGGL(k_o2_prep<FT>, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, s, a, h->cfg.o2,
                           1.0 / h->dz[h->cfg.grid.Nz - 1], h->d_o2c);
        HIP_TRY(hipGetLastError());
        if (MODE == TILE_STEP) h->o2c_ready = true;
    }
    if (MODE == TILE_PAR && h->scaling) {
        // update_biogeochemical_state!: the scaling modifier first, in place
        const int bs = 256;
        const long ncell = (long)h->cfg.grid.Nx * h->cfg.grid.Ny * h->cfg.grid.Nz;
        auto ks = k_scale_grid<FT, 5>;
        if (h->ntracers == 6) ks = k_scale_grid<FT, 6>;
        if (h->ntracers == 7) ks = k_scale_grid<FT, 7>;
        if (h->ntracers == 8) ks = k_scale_grid<FT, 8>;
        if (h->ntracers == 9) ks = k_scale_grid<FT, 9>;
        if (h->ntracers == 10) ks = k_scale_grid<FT, 10>;
        hipLaunchKernelGGL(ks, dim3((unsigned)((ncell + bs - 1) / bs)), dim3(bs), 0, s, a);
        HIP_TRY(hipGetLastError());
    }
    double sed_last = 0.0;
    if (MODE == TILE_STEP && h->sediment) {
        // the sediment steps after the three stages: Δt γ1, Δt (γ2 + ζ2),
        // Δt (γ3 + ζ3) (the stage Δt of Oceananigans' RK3, its clock's
        // last_stage_Δt); an AB2 step only where a step's Δt repeats the
        // previous one, which within a step never happens
        a.storage = storage;
        a.sed_dt[0] = (8.0 / 15.0) * dt;
        a.sed_dt[1] = (5.0 / 12.0 + -17.0 / 60.0) * dt;
        a.sed_dt[2] = (3.0 / 4.0 + -5.0 / 12.0) * dt;
        a.sed_init = h->sed_iter == 0 ? 1 : 0;
        a.sed_ab2 = (!a.sed_init && a.sed_dt[0] == h->sed_last_dt) ? 1 : 0;
        sed_last = a.sed_dt[2];
    }
    if constexpr (MODE == TILE_STAGE) {
        if (a.co2_stage) {
            // the stage's top-cell CO2 flux per column, from the stage's inputs
            SurfaceArgs<FT> sa = make_surface_args<FT>(h, tracers, T, S);
            if (h->surfc_ready) {
                sa.surfc = h->d_surfc;
                sa.surfc_cols = h->surfc_cols;
            }
            sa.topres = a.topres;
            sa.stage = a.stage;
            sa.Dsc = a.Dsc;
            sa.ncell = a.ncell;
            sa.flux_out = h->d_sflux;
            obme_status rs = launch_surface<FT, MODE>(h, sa, s);
            if (rs != OBME_OK) return rs;
        }
    }
    obme_status r = dispatch_model<FT, MODE>(h, a, s);
    if (r == OBME_OK && MODE == TILE_STEP && h->sediment) {
        h->sed_iter += 3;
        h->sed_last_dt = sed_last;
    }
    if (r == OBME_OK && MODE == TILE_TENDENCY && h->sediment) {
        const unsigned ncol = (unsigned)(h->cfg.grid.Nx * h->cfg.grid.Ny);
        hipLaunchKernelGGL(k_sediment_couple<FT>, dim3((ncol + 255) / 256), dim3(256), 0, s, a, h->sink_idx[0],
                           h->sink_idx[1]);
        HIP_TRY(hipGetLastError());
    }
    if (r != OBME_OK || MODE == TILE_PAR || !a.co2_on) return r;
    SurfaceArgs<FT> sa = make_surface_args<FT>(h, tracers, T, S);
    if (MODE == TILE_STEP && h->surfc_ready) {
        sa.surfc = h->d_surfc;
        sa.surfc_cols = h->surfc_cols;
    }
    sa.topres = topres;
    sa.dtf = (FT)dt;
    rk3_products(dt, sa.rk_a, sa.rk_b);
    if (MODE == TILE_TENDENCY) sa.GDIC = (FT*)G[7];
    return launch_surface<FT, MODE>(h, sa, s);
}

template <int MODE>
static obme_status run_tile(obme_handle* h, void* const* tracers, void* const* G, void* PAR, const void* T,
                            const void* S, double t, double dt, hipStream_t s, double* storage = nullptr)
{
    if (h->prec == OBME_FP64) return run_typed<double, MODE>(h, tracers, G, PAR, T, S, t, dt, s, storage);
    return run_typed<float, MODE>(h, tracers, G, PAR, T, S, t, dt, s, storage);
}

extern "C" obme_status obme_update_state(obme_handle* h, void* const* tracers, void* PAR, double t, void* stream)
{
    if (!h || !tracers || !PAR) return fail(OBME_ERR_INVALID, "null argument");
    if (h->scaling)
        for (int n = 0; n < h->ntracers; n++)
            if (!tracers[n]) return fail(OBME_ERR_INVALID, "null tracer pointer");
    return run_tile<TILE_PAR>(h, (void* const*)tracers, nullptr, PAR, nullptr, nullptr, t, 0.0,
                              (hipStream_t)stream);
}

extern "C" obme_status obme_add_tendencies(obme_handle* h, const void* const* tracers, void* const* G,
                                           const void* PAR, const void* T, const void* S, double t, void* stream)
{
    if (!h || !tracers || !G) return fail(OBME_ERR_INVALID, "null argument");
    for (int n = 0; n < h->ntracers; n++)
        if (!tracers[n] || !G[n]) return fail(OBME_ERR_INVALID, "null tracer or tendency pointer");
    return run_tile<TILE_TENDENCY>(h, (void* const*)tracers, G, (void*)PAR, T, S, t, 0.0, (hipStream_t)stream);
}

extern "C" obme_status obme_add_tendencies_with_sediment(obme_handle* h, const void* const* tracers, void* const* G,
                                                         const void* PAR, const void* T, const void* S,
                                                         const double* storage, double t, void* stream)
{
    if (!h) return fail(OBME_ERR_INVALID, "null argument");
    // the fields reach the coupling kernel through the launch's arguments
    struct Restore {
        obme_handle* h;
        ~Restore() { h->hook_sed_fields = nullptr; }
    } restore{h};
    h->hook_sed_fields = storage;
    return obme_add_tendencies(h, tracers, G, PAR, T, S, t, stream);
}

// p[0 .. n) = 0 (a kernel rather than hipMemsetAsync, which HIP graph
// capture does not take reliably)
__global__ void __launch_bounds__(256) k_zero_f64(double* p, long n)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = 0.0;
}

extern "C" obme_status obme_step_with_sediment(obme_handle* h, void* const* tracers, const void* T, const void* S,
                                               double* storage, int32_t nsteps, double dt, double t0, void* stream)
{
    if (!h || !tracers) return fail(OBME_ERR_INVALID, "null argument");
    if (nsteps < 0) return fail(OBME_ERR_INVALID, "negative nsteps");
    for (int n = 0; n < h->ntracers; n++)
        if (!tracers[n]) return fail(OBME_ERR_INVALID, "null tracer pointer");
    if (storage && !h->sediment) return fail(OBME_ERR_INVALID, "no sediment configured");
    if (h->sediment && !std::isfinite(dt)) return fail(OBME_ERR_INVALID, "non-finite time step with a sediment");
    hipStream_t s = (hipStream_t)stream;
    // the CO2 surface kernel's constants depend on the top cell's T and S
    // alone, which hold for the whole call: set them up once
    const obme_config& c = h->cfg;
    h->surfc_ready = false;
    if ((nsteps > 1 || (h->staged && nsteps > 0)) && c.model == OBME_LOBSTER && c.carbonate_system &&
        c.co2.enabled && !h->co2_inline) {
        const size_t ncol = (size_t)c.grid.Nx * c.grid.Ny;
        if (!h->d_surfc && hipMalloc((void**)&h->d_surfc, SURF_PACK * ncol * sizeof(double)) != hipSuccess)
            return fail(OBME_ERR_NOMEM, "surface constants");
        // T and S from the configuration (a profile or a constant) are the
        // same in every column: one column of constants serves them all
        h->surfc_cols = (!T && !S) ? 1 : (int)ncol;
        const unsigned nb = (unsigned)((h->surfc_cols + 255) / 256);
        if (h->prec == OBME_FP64) {
            SurfaceArgs<double> sa = make_surface_args<double>(h, tracers, T, S);
            sa.surfc_cols = h->surfc_cols;
            hipLaunchKernelGGL(k_surface_prep<double>, dim3(nb), dim3(256), 0, s, sa, h->d_surfc);
        } else {
            SurfaceArgs<float> sa = make_surface_args<float>(h, tracers, T, S);
            sa.surfc_cols = h->surfc_cols;
            hipLaunchKernelGGL(k_surface_prep<float>, dim3(nb), dim3(256), 0, s, sa, h->d_surfc);
        }
        HIP_TRY(hipGetLastError());
        h->surfc_ready = true;
    }
    if (nsteps > 0 && h->d_topres && c.model == OBME_LOBSTER && c.carbonate_system && c.co2.enabled &&
        !h->co2_inline) {
        // the top-cell DIC and Alk residuals start at zero in every call
        const size_t ncol = (size_t)c.grid.Nx * c.grid.Ny;
        hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, s, (double*)h->d_topres,
                           (long)ncol);
        HIP_TRY(hipGetLastError());
    }
    obme_status r = OBME_OK;
    h->o2c_ready = false;
    if (h->staged && nsteps > 0) {
        // centred / WENO sinking: each stage's face values need the stage
        // inputs of the levels below, so a step is three launches of one stage
        const size_t ncell = (size_t)c.grid.Nx * c.grid.Ny * c.grid.Nz;
        const size_t es = (h->prec == OBME_FP64) ? 8 : 4;
        if (!h->d_stageG && hipMalloc(&h->d_stageG, (size_t)h->ntracers * ncell * es) != hipSuccess)
            return fail(OBME_ERR_NOMEM, "stage tendencies");
        if (h->prec != OBME_FP64 && !h->d_stageD &&
            hipMalloc(&h->d_stageD, (size_t)h->ntracers * ncell * es) != hipSuccess)
            return fail(OBME_ERR_NOMEM, "stage increments");
        if (c.model == OBME_LOBSTER && c.carbonate_system && c.co2.enabled && !h->d_sflux &&
            hipMalloc((void**)&h->d_sflux, (size_t)c.grid.Nx * c.grid.Ny * sizeof(double)) != hipSuccess)
            return fail(OBME_ERR_NOMEM, "stage fluxes");
        for (int st = 0; st < nsteps && r == OBME_OK; st++)
            for (int q = 0; q < 3 && r == OBME_OK; q++) {
                h->cur_stage = q;
                r = run_tile<TILE_STAGE>(h, tracers, nullptr, nullptr, T, S, t0 + st * dt, dt, s);
            }
        h->o2c_ready = false;
        return r;
    }
    for (int st = 0; st < nsteps && r == OBME_OK; st++)
        r = run_tile<TILE_STEP>(h, tracers, nullptr, nullptr, T, S, t0 + st * dt, dt, s, storage);
    h->surfc_ready = false;
    h->o2c_ready = false;
    return r;
}

extern "C" obme_status obme_step(obme_handle* h, void* const* tracers, const void* T, const void* S, int32_t nsteps,
                                 double dt, double t0, void* stream)
{
    return obme_step_with_sediment(h, tracers, T, S, nullptr, nsteps, dt, t0, stream);
}

// ------------------------------------------------------------------ allocation-free stepper

// obme_step_noalloc keeps obme_step's per-call scratch (the surface records,
// the [H+] warm starts, the surface constants, the O2 exchange coefficients
// and the top-cell residuals) in this static device array instead of buffers the library
// allocates.  Each device the library runs on holds its own copy until the
// process ends.  In host code `scratch` names a host-side stand-in: its
// device address comes from hipGetSymbolAddress, once per device, and only
// that address is ever handed to a kernel.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_addr[NOALLOC_MAX_DEVICES] = {};

static size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }

// The device address of this device's scratch array (set up on the first
// call on the device, with the device current).
static obme_status scratch_address(int dev, unsigned char** out)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(OBME_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(OBME_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_addr[dev] = p;
    }
    *out = (unsigned char*)g_scratch_addr[dev];
    return OBME_OK;
}

extern "C" obme_status obme_step_noalloc(obme_handle* h, void* const* tracers, const void* T, const void* S,
                                         int32_t nsteps, double dt, double t0, void* stream)
{
    if (!h || !tracers) return fail(OBME_ERR_INVALID, "null argument");
    if (nsteps < 0) return fail(OBME_ERR_INVALID, "negative nsteps");
    for (int n = 0; n < h->ntracers; n++)
        if (!tracers[n]) return fail(OBME_ERR_INVALID, "null tracer pointer");
    if (h->staged)
        return fail(OBME_ERR_UNSUPPORTED, "obme_step_noalloc: the upwind sinking scheme only (the centred and WENO "
                                          "steps keep per-cell scratch)");
    const obme_config& c = h->cfg;
    const size_t ncol = (size_t)c.grid.Nx * c.grid.Ny;
    const size_t es = (h->prec == OBME_FP64) ? 8 : 4;
    const bool co2 = c.model == OBME_LOBSTER && c.carbonate_system && c.co2.enabled;
    const bool o2 = c.model == OBME_LOBSTER && c.oxygen.enabled && c.o2.enabled;
    // the regions of this call: surface records, warm starts, surface
    // constants (one column when T and S come from the configuration), O2
    const size_t b_srec = co2 ? round256(SURF_REC * ncol * es) : 0;
    const size_t b_hprev = co2 ? round256(ncol * sizeof(double)) : 0;
    const size_t b_surfc = co2 ? round256(SURF_PACK * ((!T && !S) ? 1 : ncol) * sizeof(double)) : 0;
    const size_t b_o2c = o2 ? round256(2 * ncol * sizeof(double)) : 0;
    const size_t b_topres = (co2 && h->prec != OBME_FP64) ? round256(2 * ncol * sizeof(float)) : 0;
    const size_t need = b_srec + b_hprev + b_surfc + b_o2c + b_topres;
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(OBME_ERR_INVALID, "obme_step_noalloc: this grid needs " + std::to_string(need) +
                                          " bytes of scratch; the static scratch holds " +
                                          std::to_string(NOALLOC_SCRATCH_BYTES) + " (64 MiB)");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != h->device) return fail(OBME_ERR_INVALID, "obme_step_noalloc: not the device the handle was created on");
    unsigned char* base = nullptr;
    obme_status st = scratch_address(dev, &base);
    if (st != OBME_OK) return st;
    hipStream_t s = (hipStream_t)stream;

    // run obme_step with the handle's scratch pointers on the static array
    // (restored on every return), and without kernel timing, whose events
    // the library would have to create
    struct Swap {
        obme_handle* h;
        void* srec;
        double *hprev, *surfc, *o2c;
        float* topres;
        bool timing;
        ~Swap()
        {
            h->d_srec = srec;
            h->d_Hprev = hprev;
            h->d_surfc = surfc;
            h->d_o2c = o2c;
            h->d_topres = topres;
            h->timing = timing;
        }
    } keep{h, h->d_srec, h->d_Hprev, h->d_surfc, h->d_o2c, h->d_topres, h->timing};
    h->timing = false;
    size_t off = 0;
    if (co2) {
        h->d_srec = base + off;
        off += b_srec;
        h->d_Hprev = (double*)(base + off);
        off += b_hprev;
        h->d_surfc = (double*)(base + off);
        off += b_surfc;
        // the array keeps the last call's contents: the warm starts begin at
        // zero (the FP32-mixed solve starts from its FP32 guess)
        hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, s, h->d_Hprev, (long)ncol);
        HIP_TRY(hipGetLastError());
    }
    if (o2) {
        h->d_o2c = (double*)(base + off);
        off += b_o2c;
    }
    // (obme_step_with_sediment clears the residuals)
    h->d_topres = b_topres ? (float*)(base + off) : nullptr;
    return obme_step_with_sediment(h, tracers, T, S, nullptr, nsteps, dt, t0, s);
}

// ------------------------------------------------------------------ sweep launch shape

extern "C" obme_status obme_sweep_configure(obme_handle* h, int32_t waves)
{
    if (!h) return fail(OBME_ERR_INVALID, "null handle");
    if (!valid_sweep_waves(waves)) return fail(OBME_ERR_INVALID, "waves per workgroup must be 1, 2, 4 or 8");
    h->sweep_waves = waves;
    return OBME_OK;
}

extern "C" int32_t obme_sweep_waves(const obme_handle* h) { return h ? h->sweep_waves : -1; }

// Times obme_step with every launch shape on a scratch copy of the state and
// keeps the fastest.  The buffers a step carries forward in the handle (the
// sediment stepper's G, the [H+] warm starts) are swapped for scratch ones
// during the trials, so the next real step sees them as they were.
extern "C" obme_status obme_autotune(obme_handle* h, void* const* tracers, const void* T, const void* S,
                                     int32_t trial_steps, double dt, void* stream)
{
    if (!h || !tracers) return fail(OBME_ERR_INVALID, "null argument");
    if (trial_steps < 1) return fail(OBME_ERR_INVALID, "trial_steps must be positive");
    if (!std::isfinite(dt)) return fail(OBME_ERR_INVALID, "non-finite time step");
    for (int n = 0; n < h->ntracers; n++)
        if (!tracers[n]) return fail(OBME_ERR_INVALID, "null tracer pointer");
    hipStream_t s = (hipStream_t)stream;
    const obme_grid& g = h->cfg.grid;
    const size_t es = (h->prec == OBME_FP64) ? 8 : 4;
    const size_t bytes = (size_t)(g.Nx + 2 * g.Hx) * (g.Ny + 2 * g.Hy) * (g.Nz + 2 * g.Hz) * es;
    const size_t ncol = (size_t)g.Nx * g.Ny;

    struct Keep {
        obme_handle* h;
        double *hprev, *sedG;
        long long sed_iter;
        double sed_last_dt;
        bool timing;
        int waves;
        std::vector<void*> owned;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Keep()
        {
            h->d_Hprev = hprev;
            h->d_sedG = sedG;
            h->sed_iter = sed_iter;
            h->sed_last_dt = sed_last_dt;
            h->timing = timing;
            h->sweep_waves = waves;
            for (void* p : owned) (void)hipFree(p);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } keep{h, h->d_Hprev, h->d_sedG, h->sed_iter, h->sed_last_dt, h->timing, h->sweep_waves, {}};
    h->timing = false;
    auto alloc = [&](size_t b, void** p) -> bool {
        if (hipMalloc(p, b) != hipSuccess) return false;
        keep.owned.push_back(*p);
        return true;
    };
    std::vector<void*> scratch(h->ntracers, nullptr);
    for (int n = 0; n < h->ntracers; n++)
        if (!alloc(bytes, &scratch[n])) return fail(OBME_ERR_NOMEM, "autotune scratch state");
    if (h->d_Hprev) {
        void* p = nullptr;
        if (!alloc(ncol * sizeof(double), &p)) return fail(OBME_ERR_NOMEM, "autotune scratch warm starts");
        HIP_TRY(hipMemcpyAsync(p, keep.hprev, ncol * sizeof(double), hipMemcpyDeviceToDevice, s));
        h->d_Hprev = (double*)p;
    }
    if (h->d_sedG) {
        void* p = nullptr;
        const size_t sb = 2 * (size_t)h->sed_nf * ncol * sizeof(double);
        if (!alloc(sb, &p)) return fail(OBME_ERR_NOMEM, "autotune scratch sediment");
        HIP_TRY(hipMemcpyAsync(p, keep.sedG, sb, hipMemcpyDeviceToDevice, s));
        h->d_sedG = (double*)p;
    }
    if (hipEventCreate(&keep.e0) != hipSuccess || hipEventCreate(&keep.e1) != hipSuccess)
        return fail(OBME_ERR_HIP, "autotune events");

    float best_ms[N_SWEEP_WAVES];
    for (int q = 0; q < N_SWEEP_WAVES; q++) best_ms[q] = 1e30f;
    for (int round = 0; round < 2; round++)
        for (int q = 0; q < N_SWEEP_WAVES; q++) {
            h->sweep_waves = SWEEP_WAVES[q];
            for (int n = 0; n < h->ntracers; n++)
                HIP_TRY(hipMemcpyAsync(scratch[n], tracers[n], bytes, hipMemcpyDeviceToDevice, s));
            obme_status st = obme_step_with_sediment(h, scratch.data(), T, S, nullptr, 1, dt, 0.0, s);
            if (st != OBME_OK) return st;
            HIP_TRY(hipEventRecord(keep.e0, s));
            st = obme_step_with_sediment(h, scratch.data(), T, S, nullptr, trial_steps, dt, dt, s);
            if (st != OBME_OK) return st;
            HIP_TRY(hipEventRecord(keep.e1, s));
            HIP_TRY(hipEventSynchronize(keep.e1));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, keep.e0, keep.e1));
            best_ms[q] = std::min(best_ms[q], ms);
        }
    int best = 0;
    for (int q = 1; q < N_SWEEP_WAVES; q++)
        if (best_ms[q] < best_ms[best]) best = q;
    keep.waves = SWEEP_WAVES[best];
    // the scratch buffers are freed once the trials have finished
    HIP_TRY(hipStreamSynchronize(s));
    return OBME_OK;
}

extern "C" obme_status obme_sediment_update_state(obme_handle* h, const void* const* tracers, double* storage,
                                                  double dt, void* stream)
{
    if (!h || !tracers) return fail(OBME_ERR_INVALID, "null argument");
    if (!h->sediment) return fail(OBME_ERR_INVALID, "no sediment configured");
    const int s0 = h->sink_idx[0], s1 = h->sink_idx[1];
    if (!tracers[s0] || !tracers[s1]) return fail(OBME_ERR_INVALID, "null tracer pointer");
    const obme_sediment& sd = h->cfg.sediment;
    if (sd.model == OBME_SEDIMENT_SIMPLE_MULTI_G && (!tracers[sd.ammonia] || !tracers[sd.nitrate] || !tracers[sd.oxygen]))
        return fail(OBME_ERR_INVALID, "null tracer pointer");
    // the tracked fields only: no step before the model's first stage
    if (!std::isfinite(dt)) return OBME_OK;
    const unsigned ncol = (unsigned)(h->cfg.grid.Nx * h->cfg.grid.Ny);
    const bool euler = !(dt == h->sed_last_dt);
    if (h->prec == OBME_FP64) {
        StepArgs<double> a = make_args<double>(h, (void* const*)tracers);
        a.storage = storage;
        a.sed_init = h->sed_iter == 0 ? 1 : 0;
        hipLaunchKernelGGL(k_sediment_state<double>, dim3((ncol + 255) / 256), dim3(256), 0, (hipStream_t)stream, a,
                           s0, s1, dt, euler ? 1 : 0);
    } else {
        StepArgs<float> a = make_args<float>(h, (void* const*)tracers);
        a.storage = storage;
        a.sed_init = h->sed_iter == 0 ? 1 : 0;
        hipLaunchKernelGGL(k_sediment_state<float>, dim3((ncol + 255) / 256), dim3(256), 0, (hipStream_t)stream, a,
                           s0, s1, dt, euler ? 1 : 0);
    }
    HIP_TRY(hipGetLastError());
    h->sed_iter += 1;
    h->sed_last_dt = dt;
    return OBME_OK;
}

extern "C" obme_status obme_co2_flux(obme_handle* h, const double* DIC, const double* Alk, const double* T,
                                     const double* S, const double* u10, double pCO2_air, double* out_pH,
                                     double* out_flux, int64_t n, int32_t mode, void* stream)
{
    if (!h || !DIC || !Alk || !T || !S) return fail(OBME_ERR_INVALID, "null argument");
    if (n < 0) return fail(OBME_ERR_INVALID, "negative n");
    if (mode != 0 && mode != 1) return fail(OBME_ERR_INVALID, "mode must be 0 or 1");
    if (n == 0) return OBME_OK;
    const int bs = 256;
    long nb = (long)((n + bs - 1) / bs);
    hipLaunchKernelGGL(k_co2_flux, dim3(nb), dim3(bs), 0, (hipStream_t)stream, DIC, Alk, T, S, u10, pCO2_air,
                       out_pH, out_flux, (long)n, mode, h->cfg.co2, h->cfg.solver, nullptr);
    HIP_TRY(hipGetLastError());
    return OBME_OK;
}

extern "C" obme_status obme_set_kernel_timing(obme_handle* h, int32_t enable)
{
    if (!h) return fail(OBME_ERR_INVALID, "null handle");
    h->timing = enable != 0;
    return OBME_OK;
}

extern "C" obme_status obme_kernel_timing(obme_handle* h, int32_t which, double* total_ms, int64_t* launches)
{
    if (!h || !total_ms || !launches) return fail(OBME_ERR_INVALID, "null argument");
    if (which < 0 || which >= OBME_KERNEL_COUNT) return fail(OBME_ERR_INVALID, "unknown kernel");
    double sum = 0.0;
    int64_t n = 0;
    obme_status st = OBME_OK;
    for (auto& p : h->timed[which]) {
        float ms = 0.0f;
        if (st == OBME_OK) {
            hipError_t e = hipEventSynchronize(p.second);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.first, p.second);
            if (e != hipSuccess) st = fail(OBME_ERR_HIP, hipGetErrorString(e));
        }
        sum += ms;
        n++;
        h->event_pool.push_back(p.first);
        h->event_pool.push_back(p.second);
    }
    h->timed[which].clear();
    *total_ms = sum;
    *launches = n;
    return st;
}

// ------------------------------------------------------------------ batched ensemble

template <typename FT, int L, int CPW, int NW, bool SED = false>
__global__ void __launch_bounds__(NW * 64, (SED && sizeof(FT) == 4 && L == 2) ? 5 : 1) k_batch(BatchArgs<FT> a)
{
    batch_columns<FT, L, CPW, NW, SED>(a);
}

// Tile shapes of the ensemble kernel: capacity L*64/CPW levels per column.
struct BatchTile {
    int L, CPW, NW;
};
static const BatchTile BATCH_TILES[] = {
    {1, 4, 4}, {2, 4, 4}, {1, 2, 4}, {2, 2, 4}, {1, 1, 4}, {2, 1, 4}, {4, 2, 4}, {4, 4, 4},
};
static int batch_capacity(int t) { return BATCH_TILES[t].L * 64 / BATCH_TILES[t].CPW; }
// Packing plans: the tile shapes a plan sorts columns into, by capacity.
static const std::vector<std::vector<int>> BATCH_PLANS = {
    {5}, {3, 5}, {1, 3, 5}, {0, 1, 3, 5}, {2, 4, 5}, {7, 6}, {0, 2, 4, 5},
};

struct obme_batch {
    obme_config cfg;
    int prec;
    int n_columns;
    std::vector<int> offsets;
    double dz, T;
    int* d_offsets = nullptr;
    FastParams* d_fpar = nullptr;
    NPDParams<double>* d_dpar = nullptr;
    BatchColumn* d_bcol = nullptr;
    float* d_comp = nullptr;              // FP32: the residual arrays, 4 x offsets[n_columns]
    int plan = -1;
    int steps_per_launch = 0;             // 0: every step of a call in one launch per class
    std::vector<int> class_tile, class_n; // per class of the plan
    std::vector<int*> class_cols;
    // captured multi-launch step
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int g_nsteps = -1, g_spl = -1, g_plan = -1;
    double g_dt = 0.0;
    void* g_ptrs[4] = {nullptr, nullptr, nullptr, nullptr};
    double* g_storage = nullptr;
    int g_sed_init = -1, g_sed_ab2 = -1;
    // an InstantRemineralisation sediment under every column (cfg.sediment)
    bool sediment = false;
    int sed_recv = 0;                     // receiver in the ensemble's order N, P, Z, D
    bool sed_P = false, sed_D = false;    // P / D among the sinking tracers
    double* d_sedGm = nullptr;            // [n_columns] G⁻ of the sediment's last step
    long long sed_iter = 0;               // sediment steps taken (its clock's iteration)
    double sed_last_dt = HUGE_VAL;        // Δt of its last step (Inf before the first)
};

static void batch_free_classes(obme_batch* b)
{
    for (int* p : b->class_cols)
        if (p) (void)hipFree(p);
    b->class_cols.clear();
    b->class_tile.clear();
    b->class_n.clear();
}

static void batch_drop_graph(obme_batch* b)
{
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->g_nsteps = -1;
}

// Sorts the columns into the classes of plan p (smallest capacity that holds
// the column), deepest columns first within a class.
static obme_status batch_apply_plan(obme_batch* b, int p)
{
    if (p < 0 || p >= (int)BATCH_PLANS.size()) return fail(OBME_ERR_INVALID, "unknown packing plan");
    std::vector<int> tiles = BATCH_PLANS[p];
    std::sort(tiles.begin(), tiles.end(), [](int x, int y) { return batch_capacity(x) < batch_capacity(y); });
    std::vector<std::vector<int>> lists(tiles.size());
    for (int c = 0; c < b->n_columns; c++) {
        const int nz = b->offsets[c + 1] - b->offsets[c];
        int q = -1;
        for (size_t t = 0; t < tiles.size() && q < 0; t++)
            if (nz <= batch_capacity(tiles[t])) q = (int)t;
        if (q < 0) return fail(OBME_ERR_UNSUPPORTED, "a column is deeper than the plan's largest tile");
        lists[q].push_back(c);
    }
    batch_free_classes(b);
    batch_drop_graph(b);
    for (size_t t = 0; t < tiles.size(); t++) {
        std::vector<int>& l = lists[t];
        if (l.empty()) continue;
        std::stable_sort(l.begin(), l.end(), [&](int x, int y) {
            return b->offsets[x + 1] - b->offsets[x] > b->offsets[y + 1] - b->offsets[y];
        });
        int* d = nullptr;
        HIP_TRY(hipMalloc((void**)&d, l.size() * sizeof(int)));
        HIP_TRY(hipMemcpy(d, l.data(), l.size() * sizeof(int), hipMemcpyHostToDevice));
        b->class_cols.push_back(d);
        b->class_tile.push_back(tiles[t]);
        b->class_n.push_back((int)l.size());
    }
    b->plan = p;
    return OBME_OK;
}

template <typename FT>
static BatchArgs<FT> batch_args(const obme_batch* b, void* const* tracers, double dt)
{
    BatchArgs<FT> a;
    std::memset(&a, 0, sizeof(a));
    for (int n = 0; n < 4; n++) a.U[n] = (FT*)tracers[n];
    a.offsets = b->d_offsets;
    a.fpar = b->d_fpar;
    a.dpar = b->d_dpar;
    a.bcol = b->d_bcol;
    a.dtf = (FT)dt;
    a.dz = (FT)b->dz;
    a.invdz = (FT)(1.0 / b->dz);
    rk3_products(dt, a.rk_a, a.rk_b);
    const obme_two_band_par& par = b->cfg.par;
    a.kr = par.water_red_attenuation;
    a.kb = par.water_blue_attenuation;
    a.xr = par.chlorophyll_red_attenuation;
    a.xb = par.chlorophyll_blue_attenuation;
    a.er = par.chlorophyll_red_exponent;
    a.eb = par.chlorophyll_blue_exponent;
    a.Rr = par.phytoplankton_chlorophyll_ratio / par.pigment_ratio;
    const double log2e = 1.4426950408889634;
    const double lRr = std::log2(a.Rr);
    a.kr2 = (float)(a.kr * log2e);
    a.kb2 = (float)(a.kb * log2e);
    a.er_f = (float)a.er;
    a.eb_f = (float)a.eb;
    a.cr = (float)(a.er * lRr);
    a.cb = (float)(a.eb * lRr);
    a.xr2h = (float)(a.xr * log2e * b->dz / 2);
    a.xb2h = (float)(a.xb * log2e * b->dz / 2);
    return a;
}

template <typename FT>
static obme_status batch_launch_class(const obme_batch* b, int q, BatchArgs<FT> a, int nsteps, hipStream_t s)
{
    a.cols = b->class_cols[q];
    a.ncols = b->class_n[q];
    a.nsteps = nsteps;
    const BatchTile& t = BATCH_TILES[b->class_tile[q]];
    const int per_wg = t.NW * t.CPW;
    dim3 grid((a.ncols + per_wg - 1) / per_wg), block(t.NW * 64);
    using Kernel = void (*)(BatchArgs<FT>);
    Kernel k = nullptr;
#define OBME_BATCH_KERNEL(L_, CPW_) (b->sediment ? k_batch<FT, L_, CPW_, 4, true> : k_batch<FT, L_, CPW_, 4, false>)
    switch (b->class_tile[q]) {
    case 0: k = OBME_BATCH_KERNEL(1, 4); break;
    case 1: k = OBME_BATCH_KERNEL(2, 4); break;
    case 2: k = OBME_BATCH_KERNEL(1, 2); break;
    case 3: k = OBME_BATCH_KERNEL(2, 2); break;
    case 4: k = OBME_BATCH_KERNEL(1, 1); break;
    case 5: k = OBME_BATCH_KERNEL(2, 1); break;
    case 6: k = OBME_BATCH_KERNEL(4, 2); break;
    case 7: k = OBME_BATCH_KERNEL(4, 4); break;
    default: return fail(OBME_ERR_INVALID, "bad batch tile");
    }
#undef OBME_BATCH_KERNEL
    hipLaunchKernelGGL(k, grid, block, 0, s, a);
    HIP_TRY(hipGetLastError());
    return OBME_OK;
}

// The sediment's stepping of a call: its first step ever (the lazy first
// tendency) and whether the call's first sediment step is an AB2 step (its
// Δt repeats the previous step's; between RK3 stages it never does).
struct BatchSedimentCall {
    double* storage = nullptr;
    double* Gm = nullptr;
    int init = 0, ab2 = 0;
};

static BatchSedimentCall batch_sediment_call(const obme_batch* b, double* storage, double* Gm, double dt)
{
    BatchSedimentCall c;
    c.storage = storage;
    c.Gm = Gm;
    c.init = b->sed_iter == 0 ? 1 : 0;
    c.ab2 = (!c.init && (8.0 / 15.0) * dt == b->sed_last_dt) ? 1 : 0;
    return c;
}

// All launches of nsteps steps: per chunk of steps_per_launch steps, one
// launch per class.
template <typename FT>
static obme_status batch_enqueue(const obme_batch* b, void* const* tracers, int nsteps, double dt, hipStream_t s,
                                 const BatchSedimentCall& sc)
{
    BatchArgs<FT> a = batch_args<FT>(b, tracers, dt);
    const int chunk = (b->steps_per_launch > 0) ? b->steps_per_launch : nsteps;
    const size_t ncell = (size_t)b->offsets[b->n_columns];
    if (b->d_comp)
        for (int n = 0; n < 4; n++) a.C[n] = (FT*)(b->d_comp + n * ncell);
    if (b->sediment) {
        const obme_sediment& sd = b->cfg.sediment;
        a.sed_on = 1;
        a.storage = sc.storage;
        a.sedGm = sc.Gm;
        a.sed_recv = b->sed_recv;
        a.sed_dt[0] = (8.0 / 15.0) * dt;
        a.sed_dt[1] = (5.0 / 12.0 + -17.0 / 60.0) * dt;
        a.sed_dt[2] = (3.0 / 4.0 + -5.0 / 12.0) * dt;
        a.sed_cn = 1.5 + sd.ab2_chi;
        a.sed_cm = 0.5 + sd.ab2_chi;
        a.sed_a = (FT)sd.burial_efficiency_constant1;
        a.sed_b = (FT)sd.burial_efficiency_constant2;
        a.sed_k = (FT)sd.burial_efficiency_half_saturation;
        a.sedP = b->sed_P ? FT(1) : FT(0);
        a.sedD = b->sed_D ? FT(1) : FT(0);
    }
    for (int done = 0; done < nsteps; done += chunk) {
        const int n = std::min(chunk, nsteps - done);
        // FP32: the residuals start at zero with each call and pass from
        // launch to launch within it
        a.c_in = (done > 0 && b->d_comp) ? 1 : 0;
        a.c_out = (done + n < nsteps && b->d_comp) ? 1 : 0;
        a.sed_init = (done == 0) ? sc.init : 0;
        a.sed_ab2 = (done == 0) ? sc.ab2 : 0;
        for (size_t q = 0; q < b->class_tile.size(); q++) {
            obme_status r = batch_launch_class<FT>(b, (int)q, a, n, s);
            if (r != OBME_OK) return r;
        }
    }
    return OBME_OK;
}

static obme_status batch_enqueue_any(const obme_batch* b, void* const* tracers, int nsteps, double dt, hipStream_t s,
                                     const BatchSedimentCall& sc)
{
    return (b->prec == OBME_FP64) ? batch_enqueue<double>(b, tracers, nsteps, dt, s, sc)
                                  : batch_enqueue<float>(b, tracers, nsteps, dt, s, sc);
}

extern "C" obme_status obme_batch_destroy(obme_batch* b)
{
    if (!b) return OBME_OK;
    batch_drop_graph(b);
    batch_free_classes(b);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    if (b->d_offsets) (void)hipFree(b->d_offsets);
    if (b->d_fpar) (void)hipFree(b->d_fpar);
    if (b->d_dpar) (void)hipFree(b->d_dpar);
    if (b->d_bcol) (void)hipFree(b->d_bcol);
    if (b->d_comp) (void)hipFree(b->d_comp);
    if (b->d_sedGm) (void)hipFree(b->d_sedGm);
    delete b;
    return OBME_OK;
}

extern "C" obme_status obme_batch_create(const obme_config* cfg, int32_t n_columns, const int32_t* offsets, double dz,
                                         const obme_batch_params* params, double temperature, obme_batch** out)
{
    if (!cfg || !offsets || !params || !out) return fail(OBME_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->model != OBME_NPZD) return fail(OBME_ERR_UNSUPPORTED, "the batched ensemble runs NPZD columns");
    if (cfg->scaling.n_groups != 0)
        return fail(OBME_ERR_UNSUPPORTED, "negative scaling is not built for the batched ensemble");
    if (cfg->sediment.model != OBME_SEDIMENT_NONE && cfg->sediment.model != OBME_SEDIMENT_INSTANT_REMINERALISATION)
        return fail(OBME_ERR_UNSUPPORTED, "the batched ensemble's sediment is InstantRemineralisation (SimpleMultiG "
                                          "needs NH4, NO3 and O2, which NPZD does not have)");
    if (cfg->sediment.model == OBME_SEDIMENT_INSTANT_REMINERALISATION) {
        // NPZD's tracers N, P, Z, T, D: T is read-only
        const obme_sediment& sd = cfg->sediment;
        if (sd.receiver < 0 || sd.receiver > 4 || sd.receiver == 3)
            return fail(OBME_ERR_INVALID, "the sediment's receiver must be N, P, Z or D");
        if (sd.n_sinking < 1 || sd.n_sinking > OBME_MAX_SEDIMENT_TRACERS)
            return fail(OBME_ERR_INVALID, "the sediment needs sinking tracers");
        for (int q = 0; q < sd.n_sinking; q++) {
            if (sd.sinking[q] < 0 || sd.sinking[q] > 4)
                return fail(OBME_ERR_INVALID, "a sinking tracer of the sediment is not an NPZD tracer");
            for (int q2 = 0; q2 < q; q2++)
                if (sd.sinking[q2] == sd.sinking[q])
                    return fail(OBME_ERR_INVALID, "a sinking tracer of the sediment is listed twice");
        }
        if (sd.timestepper != OBME_SEDIMENT_AB2)
            return fail(OBME_ERR_UNSUPPORTED, "only the QuasiAdamsBashforth2 sediment time stepper is built");
    }
    if (cfg->oxygen.enabled || cfg->o2.enabled)
        return fail(OBME_ERR_UNSUPPORTED, "the oxygen component is not built for the batched ensemble");
    if (cfg->precision != OBME_FP64 && cfg->precision != OBME_FP32_MIXED)
        return fail(OBME_ERR_INVALID, "unknown precision");
    if (n_columns < 1) return fail(OBME_ERR_INVALID, "no columns");
    if (!(dz > 0)) return fail(OBME_ERR_INVALID, "dz must be positive");
    if (!cfg->sinking.open_bottom) return fail(OBME_ERR_UNSUPPORTED, "the batched ensemble has open-bottom sinking");
    if (cfg->sinking.scheme != OBME_UPWIND1) return fail(OBME_ERR_UNSUPPORTED, "only upwind sinking on the GPU");
    if (offsets[0] != 0) return fail(OBME_ERR_INVALID, "offsets[0] must be 0");
    for (int c = 0; c < n_columns; c++) {
        const int nz = offsets[c + 1] - offsets[c];
        if (nz < 1) return fail(OBME_ERR_INVALID, "a column has no cells");
        if (nz > 128) return fail(OBME_ERR_UNSUPPORTED, "columns deeper than 128 cells");
    }
    if ((long long)offsets[n_columns] * 8 >= (1ll << 31)) return fail(OBME_ERR_UNSUPPORTED, "ensemble too large");

    obme_batch* b = new obme_batch();
    b->cfg = *cfg;
    b->cfg.grid.zc = b->cfg.grid.zf = b->cfg.grid.dz = nullptr;
    b->prec = cfg->precision;
    b->n_columns = n_columns;
    b->offsets.assign(offsets, offsets + n_columns + 1);
    b->dz = dz;
    b->T = temperature;
    // per-column parameters: device SoA (double) -> host -> derived constants
    std::vector<std::vector<double>> P(OBME_BP_COUNT);
    for (int f = 0; f < OBME_BP_COUNT; f++) {
        if (!params->fields[f]) continue;
        P[f].resize(n_columns);
        if (hipMemcpy(P[f].data(), params->fields[f], n_columns * sizeof(double), hipMemcpyDeviceToHost) !=
            hipSuccess) {
            obme_batch_destroy(b);
            return fail(OBME_ERR_HIP, "reading the per-column parameters");
        }
    }
    std::vector<FastParams> fpar(n_columns);
    std::vector<NPDParams<double>> dpar(n_columns);
    std::vector<BatchColumn> bcol(n_columns);
    for (int c = 0; c < n_columns; c++) {
        obme_npd_params q = cfg->npd;
        auto get = [&](int f, double dflt) { return P[f].empty() ? dflt : P[f][c]; };
        q.nitrate_half_saturation = get(OBME_BP_NITRATE_HALF_SATURATION, q.nitrate_half_saturation);
        q.light_half_saturation = get(OBME_BP_LIGHT_HALF_SATURATION, q.light_half_saturation);
        q.phytoplankton_maximum_growth_rate = get(OBME_BP_MAX_GROWTH_RATE, q.phytoplankton_maximum_growth_rate);
        q.phytoplankton_exudation_fraction = get(OBME_BP_EXUDATION_FRACTION, q.phytoplankton_exudation_fraction);
        q.ammonia_fraction_of_exudate = get(OBME_BP_AMMONIA_FRACTION_OF_EXUDATE, q.ammonia_fraction_of_exudate);
        q.temperature_coefficient = get(OBME_BP_TEMPERATURE_COEFFICIENT, q.temperature_coefficient);
        q.phytoplankton_mortality_rate = get(OBME_BP_PHYTO_MORTALITY_RATE, q.phytoplankton_mortality_rate);
        q.zooplankton_mortality_rate = get(OBME_BP_ZOO_MORTALITY_RATE, q.zooplankton_mortality_rate);
        q.zooplankton_excretion_rate = get(OBME_BP_ZOO_EXCRETION_RATE, q.zooplankton_excretion_rate);
        q.phytoplankton_solid_waste_fraction =
            get(OBME_BP_PHYTO_SOLID_WASTE_FRACTION, q.phytoplankton_solid_waste_fraction);
        q.excretion_inorganic_fraction = get(OBME_BP_EXCRETION_INORGANIC_FRACTION, q.excretion_inorganic_fraction);
        q.preference_for_phytoplankton = get(OBME_BP_PREFERENCE_FOR_PHYTO, q.preference_for_phytoplankton);
        q.maximum_grazing_rate = get(OBME_BP_MAX_GRAZING_RATE, q.maximum_grazing_rate);
        q.grazing_half_saturation = get(OBME_BP_GRAZING_HALF_SATURATION, q.grazing_half_saturation);
        q.zooplankton_assimilation_fraction = get(OBME_BP_ZOO_ASSIMILATION_FRACTION, q.zooplankton_assimilation_fraction);
        q.remineralisation_rate = get(OBME_BP_REMINERALISATION_RATE, q.remineralisation_rate);
        q.small_particle_fraction = get(OBME_BP_SMALL_PARTICLE_FRACTION, q.small_particle_fraction);
        fpar[c] = make_fast_params(q);
        dpar[c] = convert_params<double>(q);
        BatchColumn& bc = bcol[c];
        bc.PAR0 = get(OBME_BP_SURFACE_PAR, cfg->par.surface_value);
        bc.wP = -get(OBME_BP_P_SINKING_SPEED, cfg->sinking.speed[1]);
        bc.wD = -get(OBME_BP_D_SINKING_SPEED, cfg->sinking.speed[4]);
        // Q10^(T/10) (plankton.jl:210-214) at the ensemble's uniform temperature
        bc.Lt = q.use_temperature_coefficient ? std::pow(q.temperature_coefficient, temperature / 10) : 1.0;
        if (bc.wP > 0 || bc.wD > 0) {
            obme_batch_destroy(b);
            return fail(OBME_ERR_INVALID, "negative sinking speed");
        }
    }
    auto up = [&](void** d, const void* h, size_t bytes) -> bool {
        return hipMalloc(d, bytes) == hipSuccess && hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    bool ok = up((void**)&b->d_offsets, b->offsets.data(), b->offsets.size() * sizeof(int)) &&
              up((void**)&b->d_bcol, bcol.data(), bcol.size() * sizeof(BatchColumn));
    if (ok && b->prec == OBME_FP64) ok = up((void**)&b->d_dpar, dpar.data(), dpar.size() * sizeof(NPDParams<double>));
    if (ok && b->prec != OBME_FP64) ok = up((void**)&b->d_fpar, fpar.data(), fpar.size() * sizeof(FastParams));
    if (ok && b->prec != OBME_FP64)
        ok = hipMalloc((void**)&b->d_comp, 4 * (size_t)b->offsets[n_columns] * sizeof(float)) == hipSuccess;
    if (ok && cfg->sediment.model == OBME_SEDIMENT_INSTANT_REMINERALISATION) {
        const obme_sediment& sd = cfg->sediment;
        b->sediment = true;
        b->sed_recv = sd.receiver == 4 ? 3 : sd.receiver;
        for (int q = 0; q < sd.n_sinking; q++) {
            // only P and D sink in the ensemble: the other tracers' flux is 0
            if (sd.sinking[q] == 1) b->sed_P = true;
            if (sd.sinking[q] == 4) b->sed_D = true;
        }
        ok = hipMalloc((void**)&b->d_sedGm, (size_t)n_columns * sizeof(double)) == hipSuccess &&
             hipMemset(b->d_sedGm, 0, (size_t)n_columns * sizeof(double)) == hipSuccess;
    }
    if (!ok) {
        obme_batch_destroy(b);
        return fail(OBME_ERR_NOMEM, "batch buffers");
    }
    obme_status st = batch_apply_plan(b, 1);
    if (st != OBME_OK) st = batch_apply_plan(b, 0);
    if (st != OBME_OK) {
        obme_batch_destroy(b);
        return st;
    }
    *out = b;
    return OBME_OK;
}

extern "C" obme_status obme_batch_configure(obme_batch* b, int32_t plan, int32_t steps_per_launch)
{
    if (!b) return fail(OBME_ERR_INVALID, "null handle");
    if (steps_per_launch < 0) return fail(OBME_ERR_INVALID, "negative steps_per_launch");
    if (plan != b->plan) {
        obme_status st = batch_apply_plan(b, plan);
        if (st != OBME_OK) return st;
    }
    if (steps_per_launch != b->steps_per_launch) batch_drop_graph(b);
    b->steps_per_launch = steps_per_launch;
    return OBME_OK;
}

// nsteps steps (with the sediment when the ensemble has one: storage null
// applies its coupling alone)
static obme_status batch_step(obme_batch* b, void* const* tracers, double* storage, int32_t nsteps, double dt,
                              void* stream)
{
    if (!b || !tracers) return fail(OBME_ERR_INVALID, "null argument");
    for (int n = 0; n < 4; n++)
        if (!tracers[n]) return fail(OBME_ERR_INVALID, "null tracer pointer");
    if (nsteps < 0) return fail(OBME_ERR_INVALID, "negative nsteps");
    if (nsteps == 0) return OBME_OK;
    hipStream_t s = (hipStream_t)stream;
    const BatchSedimentCall sc = batch_sediment_call(b, storage, b->d_sedGm, dt);
    const bool multi = b->steps_per_launch > 0 && b->steps_per_launch < nsteps;
    obme_status r = OBME_OK;
    if (!multi) {
        r = batch_enqueue_any(b, tracers, nsteps, dt, s, sc);
    } else {
        // several launches per call: recorded once into a hipGraph per
        // (nsteps, steps_per_launch, dt, tracers, plan, sediment call) and replayed
        const bool same = b->graph && b->g_nsteps == nsteps && b->g_spl == b->steps_per_launch && b->g_dt == dt &&
                          b->g_plan == b->plan && std::memcmp(b->g_ptrs, tracers, sizeof(b->g_ptrs)) == 0 &&
                          b->g_storage == storage && b->g_sed_init == sc.init && b->g_sed_ab2 == sc.ab2;
        if (!same) {
            batch_drop_graph(b);
            if (!b->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
            HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
            obme_status rc = batch_enqueue_any(b, tracers, nsteps, dt, b->cap_stream, sc);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
            if (rc != OBME_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess)
                return fail(OBME_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                b->graph = nullptr;
                return fail(OBME_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            }
            b->g_nsteps = nsteps;
            b->g_spl = b->steps_per_launch;
            b->g_dt = dt;
            b->g_plan = b->plan;
            std::memcpy(b->g_ptrs, tracers, sizeof(b->g_ptrs));
            b->g_storage = storage;
            b->g_sed_init = sc.init;
            b->g_sed_ab2 = sc.ab2;
        }
        if (hipGraphLaunch(b->graph, s) != hipSuccess) r = fail(OBME_ERR_HIP, "hipGraphLaunch");
    }
    if (r == OBME_OK && b->sediment && storage) {
        // three sediment steps per step, the last of Δt (γ3 + ζ3)
        b->sed_iter += 3 * (long long)nsteps;
        b->sed_last_dt = (3.0 / 4.0 + -5.0 / 12.0) * dt;
    }
    return r;
}

extern "C" obme_status obme_batch_step(obme_batch* b, void* const* tracers, int32_t nsteps, double dt, double t0,
                                       void* stream)
{
    (void)t0;  // surface PAR is constant per column
    return batch_step(b, tracers, nullptr, nsteps, dt, stream);
}

extern "C" obme_status obme_batch_step_with_sediment(obme_batch* b, void* const* tracers, double* storage,
                                                     int32_t nsteps, double dt, double t0, void* stream)
{
    (void)t0;
    if (!b) return fail(OBME_ERR_INVALID, "null handle");
    if (!b->sediment) return fail(OBME_ERR_INVALID, "the ensemble has no sediment");
    if (!storage) return fail(OBME_ERR_INVALID, "null sediment storage");
    return batch_step(b, tracers, storage, nsteps, dt, stream);
}

// Times each packing plan and steps-per-launch choice on a scratch copy of
// the state (the caller's tracers are only read) and keeps the fastest.
extern "C" obme_status obme_batch_autotune(obme_batch* b, void* const* tracers, int32_t trial_steps, double dt,
                                           void* stream)
{
    if (!b || !tracers) return fail(OBME_ERR_INVALID, "null argument");
    if (trial_steps < 1) return fail(OBME_ERR_INVALID, "trial_steps must be positive");
    hipStream_t s = (hipStream_t)stream;
    const size_t es = (b->prec == OBME_FP64) ? 8 : 4;
    const size_t bytes = (size_t)b->offsets[b->n_columns] * es;
    void* scratch[4] = {nullptr, nullptr, nullptr, nullptr};
    double* sed_scratch = nullptr;   // the sediment's storage and G⁻ (the caller's are not touched)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    obme_status st = OBME_OK;
    auto cleanup = [&]() {
        for (void* p : scratch)
            if (p) (void)hipFree(p);
        if (sed_scratch) (void)hipFree(sed_scratch);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    for (int n = 0; n < 4 && st == OBME_OK; n++)
        if (hipMalloc(&scratch[n], bytes) != hipSuccess) st = fail(OBME_ERR_NOMEM, "autotune scratch");
    BatchSedimentCall sc;
    if (st == OBME_OK && b->sediment) {
        const size_t sb = 2 * (size_t)b->n_columns * sizeof(double);
        if (hipMalloc((void**)&sed_scratch, sb) != hipSuccess || hipMemsetAsync(sed_scratch, 0, sb, s) != hipSuccess)
            st = fail(OBME_ERR_NOMEM, "autotune scratch sediment");
        sc.storage = sed_scratch;
        sc.Gm = sed_scratch + b->n_columns;
        sc.init = 1;
    }
    if (st == OBME_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        st = fail(OBME_ERR_HIP, "autotune events");
    if (st != OBME_OK) {
        cleanup();
        return st;
    }
    const int keep_plan = b->plan, keep_spl = b->steps_per_launch;
    int best_plan = keep_plan, best_spl = keep_spl;
    float best_ms = 1e30f;
    std::vector<int> spls = {0, 1, 10};
    for (int p = 0; p < (int)BATCH_PLANS.size() && st == OBME_OK; p++) {
        if (batch_apply_plan(b, p) != OBME_OK) continue;  // a plan that cannot hold the deepest column
        for (int spl : spls) {
            b->steps_per_launch = spl;
            for (int n = 0; n < 4; n++)
                if (hipMemcpyAsync(scratch[n], tracers[n], bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
                    st = fail(OBME_ERR_HIP, "autotune copy");
            if (st != OBME_OK) break;
            st = batch_enqueue_any(b, scratch, 1, dt, s, sc);  // warm-up
            if (st == OBME_OK && hipEventRecord(e0, s) != hipSuccess) st = fail(OBME_ERR_HIP, "event");
            if (st == OBME_OK) st = batch_enqueue_any(b, scratch, trial_steps, dt, s, sc);
            if (st == OBME_OK && hipEventRecord(e1, s) != hipSuccess) st = fail(OBME_ERR_HIP, "event");
            float ms = 0.0f;
            if (st == OBME_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                st = fail(OBME_ERR_HIP, "autotune timing");
            if (st != OBME_OK) break;
            if (ms < best_ms) {
                best_ms = ms;
                best_plan = p;
                best_spl = spl;
            }
        }
    }
    cleanup();
    b->steps_per_launch = keep_spl;
    obme_status r = batch_apply_plan(b, best_plan);
    if (r == OBME_OK) b->steps_per_launch = best_spl;
    (void)keep_plan;
    return st != OBME_OK ? st : r;
}

// "plan P steps_per_launch K classes: L,CPW,NW:count ..." (returns length or < 0)
extern "C" int32_t obme_batch_describe(const obme_batch* b, char* buf, int32_t cap)
{
    if (!b || !buf || cap < 1) return -1;
    std::string s = "plan " + std::to_string(b->plan) + " steps_per_launch " + std::to_string(b->steps_per_launch) +
                    " classes:";
    for (size_t q = 0; q < b->class_tile.size(); q++) {
        const BatchTile& t = BATCH_TILES[b->class_tile[q]];
        s += " " + std::to_string(t.L) + "," + std::to_string(t.CPW) + "," + std::to_string(t.NW) + ":" +
             std::to_string(b->class_n[q]);
    }
    std::snprintf(buf, cap, "%s", s.c_str());
    return (int32_t)s.size();
}

// Sizes of the ABI structs, so bindings can check their layouts.
extern "C" int64_t obme_abi_sizeof(int32_t which)
{
    switch (which) {
    case 0: return sizeof(obme_config);
    case 1: return sizeof(obme_grid);
    case 2: return sizeof(obme_npd_params);
    case 3: return sizeof(obme_two_band_par);
    case 4: return sizeof(obme_sinking);
    case 5: return sizeof(obme_carbon_solver);
    case 6: return sizeof(obme_co2_exchange);
    case 7: return sizeof(obme_field_source);
    case 8: return sizeof(obme_batch_params);
    case 9: return sizeof(obme_pisces_config);
    case 10: return sizeof(obme_negative_scaling);
    case 11: return sizeof(obme_o2_exchange);
    case 12: return sizeof(obme_sediment);
    case 13: return sizeof(obme_oxygen);
    default: return -1;
    }
}
