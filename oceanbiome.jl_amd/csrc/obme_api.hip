// obme_api.hip — kernels instantiation and the C ABI of libobme_hip.so.
//
// Entry points (declared in include/obme.h) and the reference interface
// each one stands in for are listed there.  All device work is queued on
// the caller's stream; the multi-step stepper records its launches into a
// hipGraph on first use for a given (nsteps, dt, t0) and replays it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "obme_npd_kernels.h"

using namespace obme;

// ------------------------------------------------------------------ errors

static thread_local std::string g_last_error;

static obme_status fail(obme_status code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(OBME_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));   \
    } while (0)

extern "C" const char* obme_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* obme_version(void) { return "obme 0.1.0 gfx950"; }

// ------------------------------------------------------------------ kernels

// The sweep: one workgroup of NW waves, wave-per-column segments
// (obme_npd_kernels.h).  __launch_bounds__(NW*64) lets the compiler use
// 512/(NW/4) VGPRs per lane at most (128 for 16 waves, 256 for 8).
template <int MODEL, bool CARB, typename FT, int L, int CPW, int NW, int MODE, bool UNIFORM>
__global__ void __launch_bounds__(NW * 64) k_column_wave(StepArgs<FT> a)
{
    extern __shared__ unsigned char smem_raw[];
    column_wave<MODEL, CARB, FT, L, CPW, NW, MODE, UNIFORM>(a, reinterpret_cast<FT*>(smem_raw));
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
    int ntracers;
    int prec;
    std::vector<double> zc, zf, dz, Tprof, Sprof;
    void* d_lev = nullptr;     // LEV_COUNT * Nz of FT
    void* d_wsink = nullptr;   // MAX_SINK * (Nz+1) of FT
    double* d_Tprof = nullptr;
    double* d_Sprof = nullptr;
    void* d_srec = nullptr;    // SURF_REC * Nx*Ny of FT (CO2 exchange on)
    double* d_Hprev = nullptr; // Nx*Ny warm starts of [H+] (mixed path)
    int sink_idx[MAX_SINK];
    int cfg_id;                // sweep tile configuration (TileCfg table)
    bool uniform;
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

static int ntracers_of(const obme_config& c)
{
    if (c.model == OBME_NPZD) return 5;
    if (c.model == OBME_LOBSTER) return c.carbonate_system ? 9 : 7;
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

// Sweep tile configurations: L levels per lane, CPW columns per wave, NW
// waves per workgroup (TXB = NW * CPW columns, Nz <= L * 64 / CPW).
struct TileCfg {
    int L, CPW, NW;
};
static const TileCfg TILE_CFGS[] = {
    {2, 2, 16}, {4, 4, 8}, {2, 2, 8}, {1, 1, 16}, {2, 1, 16}, {4, 2, 8}, {1, 1, 8}, {2, 1, 8},
};
static const int N_TILE_CFG = sizeof(TILE_CFGS) / sizeof(TILE_CFGS[0]);

static int choose_tile(int prec, int Nz)
{
    // preference lists by precision (indices into TILE_CFGS)
    static const int pref32[] = {0, 4, 5};
    static const int pref64[] = {3, 4, 7};
    const int* p = (prec == OBME_FP64) ? pref64 : pref32;
    for (int q = 0; q < 3; q++) {
        const TileCfg& t = TILE_CFGS[p[q]];
        if (Nz <= t.L * 64 / t.CPW) return p[q];
    }
    return -1;
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
    if (g.Nx < 1 || g.Ny < 1 || g.Nz < 1 || g.Hx < 0 || g.Hy < 0 || g.Hz < 0)
        return fail(OBME_ERR_INVALID, "invalid grid size or halo");
    if (g.Nz > 128) return fail(OBME_ERR_UNSUPPORTED, "Nz > 128 is not supported by the column-tile kernels");
    {
        size_t elems = (size_t)(g.Nx + 2 * g.Hx) * (g.Ny + 2 * g.Hy) * (g.Nz + 2 * g.Hz);
        size_t es = (cfg->precision == OBME_FP64) ? 8 : 4;
        if (elems * es >= ((size_t)1 << 31))
            return fail(OBME_ERR_UNSUPPORTED, "a field of 2 GiB or more exceeds the 32-bit buffer offsets");
    }
    if (!g.zc || !g.zf || !g.dz) return fail(OBME_ERR_INVALID, "grid coordinates missing");
    if (cfg->sinking.scheme != OBME_UPWIND1)
        return fail(OBME_ERR_UNSUPPORTED, "only the first-order upwind sinking scheme runs on the GPU");
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
    // sweep tile configuration: the first of the precision's preference list
    // whose column fits (Nz <= L * 64 / CPW); OBME_TILE="L,CPW,NW" overrides.
    h->cfg_id = choose_tile(h->prec, g.Nz);
    if (const char* e = std::getenv("OBME_TILE")) {
        int l = 0, c = 0, w = 0;
        if (std::sscanf(e, "%d,%d,%d", &l, &c, &w) == 3) {
            for (int t = 0; t < N_TILE_CFG; t++)
                if (TILE_CFGS[t].L == l && TILE_CFGS[t].CPW == c && TILE_CFGS[t].NW == w &&
                    g.Nz <= l * 64 / c)
                    h->cfg_id = t;
        }
    }
    if (h->cfg_id < 0) {
        obme_destroy(h);
        return fail(OBME_ERR_UNSUPPORTED, "no tile configuration for this Nz");
    }
    if (cfg->model == OBME_LOBSTER && cfg->carbonate_system && cfg->co2.enabled) {
        size_t ncol = (size_t)g.Nx * g.Ny;
        size_t es = (h->prec == OBME_FP64) ? 8 : 4;
        hipError_t e1 = hipMalloc(&h->d_srec, SURF_REC * ncol * es);
        hipError_t e2 = hipMalloc((void**)&h->d_Hprev, ncol * sizeof(double));
        if (e1 != hipSuccess || e2 != hipSuccess) {
            obme_destroy(h);
            return fail(OBME_ERR_NOMEM, "surface buffers");
        }
        if (hipMemset(h->d_Hprev, 0, ncol * sizeof(double)) != hipSuccess) {
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
    if (h->cfg.model == OBME_NPZD) names = "N;P;Z;T;D";
    else if (h->cfg.carbonate_system) names = "NO₃;NH₄;P;Z;sPOM;bPOM;DOM;DIC;Alk";
    else names = "NO₃;NH₄;P;Z;sPOM;bPOM;DOM";
    if (buf && cap > 0) {
        std::strncpy(buf, names, cap - 1);
        buf[cap - 1] = 0;
    }
    return h->ntracers;
}

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
    a.co2_on = (c.model == OBME_LOBSTER && c.carbonate_system && c.co2.enabled) ? 1 : 0;
    if (sizeof(FT) == 4) {
        a.f = make_fast_params(c.npd);
        const double log2e = 1.4426950408889634;
        a.kr2 = (float)(c.par.water_red_attenuation * log2e);
        a.kb2 = (float)(c.par.water_blue_attenuation * log2e);
        a.xr2 = (float)(c.par.chlorophyll_red_attenuation * log2e);
        a.xb2 = (float)(c.par.chlorophyll_blue_attenuation * log2e);
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
    s.Hprev = (h->prec == OBME_FP64) ? nullptr : h->d_Hprev;
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
    s.mode = (h->prec == OBME_FP64) ? 0 : 1;
    s.co2 = c.co2;
    s.solver = c.solver;
    return s;
}

template <int MODEL, bool CARB, typename FT, int L, int CPW, int NW, int MODE>
static obme_status launch_wave(obme_handle* h, const StepArgs<FT>& a, hipStream_t s)
{
    constexpr int NT = ModelTraits<MODEL, CARB>::NT;
    constexpr int TXB = NW * CPW;
    const int ntiles = (MODE == TILE_PAR) ? 1 : (MODE == TILE_TENDENCY ? NT + 1 : NT);
    const size_t lds = (size_t)ntiles * h->cfg.grid.Nz * (TXB + 1) * sizeof(FT);
    if (lds > 160 * 1024) return fail(OBME_ERR_UNSUPPORTED, "sweep tile exceeds the 160 KiB LDS");
    dim3 block(NW * 64);
    dim3 grid((h->cfg.grid.Nx + TXB - 1) / TXB, h->cfg.grid.Ny);
    auto k = h->uniform ? k_column_wave<MODEL, CARB, FT, L, CPW, NW, MODE, true>
                        : k_column_wave<MODEL, CARB, FT, L, CPW, NW, MODE, false>;
    HIP_TRY(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    LaunchTimer tm(h, OBME_KERNEL_SWEEP, s);
    hipLaunchKernelGGL(k, grid, block, lds, s, a);
    tm.done();
    HIP_TRY(hipGetLastError());
    return OBME_OK;
}

template <typename FT, int L, int CPW, int NW, int MODE>
static obme_status dispatch_model(obme_handle* h, const StepArgs<FT>& a, hipStream_t s)
{
    if (h->cfg.model == OBME_NPZD) return launch_wave<OBME_NPZD, false, FT, L, CPW, NW, MODE>(h, a, s);
    if (h->cfg.carbonate_system) return launch_wave<OBME_LOBSTER, true, FT, L, CPW, NW, MODE>(h, a, s);
    return launch_wave<OBME_LOBSTER, false, FT, L, CPW, NW, MODE>(h, a, s);
}

template <typename FT, int MODE>
static obme_status dispatch_tile_cfg(obme_handle* h, const StepArgs<FT>& a, hipStream_t s)
{
    switch (h->cfg_id) {
    case 0: return dispatch_model<FT, 2, 2, 16, MODE>(h, a, s);
    case 1: return dispatch_model<FT, 4, 4, 8, MODE>(h, a, s);
    case 2: return dispatch_model<FT, 2, 2, 8, MODE>(h, a, s);
    case 3: return dispatch_model<FT, 1, 1, 16, MODE>(h, a, s);
    case 4: return dispatch_model<FT, 2, 1, 16, MODE>(h, a, s);
    case 5: return dispatch_model<FT, 4, 2, 8, MODE>(h, a, s);
    case 6: return dispatch_model<FT, 1, 1, 8, MODE>(h, a, s);
    case 7: return dispatch_model<FT, 2, 1, 8, MODE>(h, a, s);
    default: return fail(OBME_ERR_INVALID, "bad tile configuration");
    }
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
                             const void* S, double t, double dt, hipStream_t s)
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
    obme_status r = dispatch_tile_cfg<FT, MODE>(h, a, s);
    if (r != OBME_OK || MODE == TILE_PAR || !a.co2_on) return r;
    SurfaceArgs<FT> sa = make_surface_args<FT>(h, tracers, T, S);
    sa.dtf = (FT)dt;
    rk3_products(dt, sa.rk_a, sa.rk_b);
    if (MODE == TILE_TENDENCY) sa.GDIC = (FT*)G[7];
    return launch_surface<FT, MODE>(h, sa, s);
}

template <int MODE>
static obme_status run_tile(obme_handle* h, void* const* tracers, void* const* G, void* PAR, const void* T,
                            const void* S, double t, double dt, hipStream_t s)
{
    if (h->prec == OBME_FP64) return run_typed<double, MODE>(h, tracers, G, PAR, T, S, t, dt, s);
    return run_typed<float, MODE>(h, tracers, G, PAR, T, S, t, dt, s);
}

extern "C" obme_status obme_update_state(obme_handle* h, const void* const* tracers, void* PAR, double t,
                                         void* stream)
{
    if (!h || !tracers || !PAR) return fail(OBME_ERR_INVALID, "null argument");
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

extern "C" obme_status obme_step(obme_handle* h, void* const* tracers, const void* T, const void* S, int32_t nsteps,
                                 double dt, double t0, void* stream)
{
    if (!h || !tracers) return fail(OBME_ERR_INVALID, "null argument");
    if (nsteps < 0) return fail(OBME_ERR_INVALID, "negative nsteps");
    for (int n = 0; n < h->ntracers; n++)
        if (!tracers[n]) return fail(OBME_ERR_INVALID, "null tracer pointer");
    hipStream_t s = (hipStream_t)stream;
    for (int st = 0; st < nsteps; st++) {
        obme_status r = run_tile<TILE_STEP>(h, tracers, nullptr, nullptr, T, S, t0 + st * dt, dt, s);
        if (r != OBME_OK) return r;
    }
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
    default: return -1;
    }
}
