// obme_api.hip — kernels instantiation and the C ABI of libobme_hip.so.
//
// Entry points (declared in include/obme.h) and the reference interface
// each one stands in for are listed there.  All device work is queued on
// the caller's stream; the multi-step stepper records its launches into a
// hipGraph on first use for a given (nsteps, dt, t0) and replays it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
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

template <int MODEL, bool CARB, typename FT, int LPT, int MODE, int MAXT>
__global__ void __launch_bounds__(MAXT) k_column_tile(StepArgs<FT> a)
{
    extern __shared__ unsigned char smem_raw[];
    column_tile<MODEL, CARB, FT, LPT, MODE>(a, reinterpret_cast<FT*>(smem_raw));
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
    int sink_idx[MAX_SINK];
    int tx, kg, lpt;
    // recorded multi-step graph
    hipGraphExec_t graph = nullptr;
    int g_nsteps = -1;
    double g_dt = 0, g_t0 = 0;
    void* g_stream = nullptr;
    void* const* g_tr = nullptr;
    std::vector<void*> g_trv;
    const void* g_T = nullptr;
    const void* g_S = nullptr;
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
    return OBME_OK;
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
    // launch shape: LPT levels per thread, KG = ceil(Nz / LPT) level groups,
    // TX columns per workgroup with TX * KG <= the kernel's bound
    h->lpt = 2;
    h->kg = (g.Nz + h->lpt - 1) / h->lpt;
    int maxt = (h->prec == OBME_FP64) ? 512 : 1024;
    if (h->kg > maxt) { obme_destroy(h); return fail(OBME_ERR_UNSUPPORTED, "too many levels"); }
    int tx = 64;
    while (tx * h->kg > maxt) tx >>= 1;
    h->tx = tx;
    *out = h;
    return OBME_OK;
}

extern "C" obme_status obme_destroy(obme_handle* h)
{
    if (!h) return OBME_OK;
    if (h->graph) hipGraphExecDestroy(h->graph);
    if (h->d_lev) (void)hipFree(h->d_lev);
    if (h->d_wsink) (void)hipFree(h->d_wsink);
    if (h->d_Tprof) (void)hipFree(h->d_Tprof);
    if (h->d_Sprof) (void)hipFree(h->d_Sprof);
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
static StepArgs<FT> make_args(const obme_handle* h, void* const* tracers, const void* T, const void* S)
{
    const obme_config& c = h->cfg;
    const obme_grid& g = c.grid;
    StepArgs<FT> a;
    std::memset(&a, 0, sizeof(a));
    for (int n = 0; n < h->ntracers; n++) a.tr[n] = tracers ? (FT*)tracers[n] : nullptr;
    a.lev = (const FT*)h->d_lev;
    a.wsink = (const FT*)h->d_wsink;
    a.surf = (const FT*)c.par.surface_array;
    a.T3 = (const FT*)T;
    a.S3 = (const FT*)S;
    a.Tprof = (c.temperature.mode == OBME_FIELD_PROFILE) ? h->d_Tprof : nullptr;
    a.Sprof = (c.salinity.mode == OBME_FIELD_PROFILE) ? h->d_Sprof : nullptr;
    a.Tconst = c.temperature.value;
    a.Sconst = c.salinity.value;
    a.Nx = g.Nx; a.Ny = g.Ny; a.Nz = g.Nz; a.Hx = g.Hx; a.Hy = g.Hy; a.Hz = g.Hz;
    a.sx = g.Nx + 2 * g.Hx;
    a.sxy = a.sx * (g.Ny + 2 * g.Hy);
    a.base0 = (long)g.Hx + a.sx * g.Hy + a.sxy * g.Hz;
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
    a.co2_on = (c.model == OBME_LOBSTER && c.carbonate_system && c.co2.enabled) ? 1 : 0;
    a.co2_mode = (h->prec == OBME_FP64) ? 0 : 1;
    a.co2 = c.co2;
    a.solver = c.solver;
    return a;
}

template <int MODEL, bool CARB, typename FT, int LPT, int MODE, int MAXT>
static obme_status launch_tile(const obme_handle* h, const StepArgs<FT>& a, hipStream_t s)
{
    dim3 block(h->tx, h->kg);
    dim3 grid((h->cfg.grid.Nx + h->tx - 1) / h->tx, h->cfg.grid.Ny);
    size_t lds = (size_t)(4 + MAX_SINK) * h->tx * h->kg * sizeof(FT) + (size_t)h->tx * sizeof(SurfaceCarbon);
    hipLaunchKernelGGL((k_column_tile<MODEL, CARB, FT, LPT, MODE, MAXT>), grid, block, lds, s, a);
    HIP_TRY(hipGetLastError());
    return OBME_OK;
}

template <typename FT, int LPT, int MODE, int MAXT>
static obme_status dispatch_tile(const obme_handle* h, const StepArgs<FT>& a, hipStream_t s)
{
    if (h->cfg.model == OBME_NPZD) return launch_tile<OBME_NPZD, false, FT, LPT, MODE, MAXT>(h, a, s);
    if (h->cfg.carbonate_system) return launch_tile<OBME_LOBSTER, true, FT, LPT, MODE, MAXT>(h, a, s);
    return launch_tile<OBME_LOBSTER, false, FT, LPT, MODE, MAXT>(h, a, s);
}

template <int MODE>
static obme_status run_tile(obme_handle* h, void* const* tracers, void* const* G, void* PAR, const void* T,
                            const void* S, double t, double dt, hipStream_t s)
{
    if (h->prec == OBME_FP64) {
        StepArgs<double> a = make_args<double>(h, tracers, T, S);
        if (G) for (int n = 0; n < h->ntracers; n++) a.G[n] = (double*)G[n];
        a.PAR = (double*)PAR;
        a.t0 = t;
        a.dt = dt;
        return dispatch_tile<double, 2, MODE, 512>(h, a, s);
    }
    StepArgs<float> a = make_args<float>(h, tracers, T, S);
    if (G) for (int n = 0; n < h->ntracers; n++) a.G[n] = (float*)G[n];
    a.PAR = (float*)PAR;
    a.t0 = t;
    a.dt = dt;
    return dispatch_tile<float, 2, MODE, 1024>(h, a, s);
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
