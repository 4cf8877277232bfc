/*
 * oracle.h — FP64 CPU restatement of OceanBioME.jl v0.17.6's biogeochemistry
 * step.  TEST INFRASTRUCTURE ONLY: this code is the oracle the HIP library
 * is tested against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; nothing in the product links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the OceanBioME.jl source tree).  Evaluation order follows the Julia
 * expressions left to right; the build uses -ffp-contract=off so no FMA is
 * formed where Julia would not form one.
 *
 * Pins: the reference's own known answers (tests/golden/reference_values.json),
 * the budgets of its tests, and closed forms at special states of every
 * NPZD / LOBSTER transition term and of the sinking reconstructions, each
 * shown to catch a planted mistake (tests/test_oracle_pins_npd.py,
 * tests/test_oracle_sinking_pins.py); SimpleMultiG's nitrogen balances, its
 * fractions where their logarithms vanish and the reference's budget
 * (tests/test_oracle_sediment_smg.py); the ensemble with its sediment
 * against the 3-D oracle's single column and the reference's budget per
 * column (tests/test_oracle_batch_sediment.py).
 *
 * Parity unpinned (no reference-held value or closed form fixes the bits):
 * orc_teos10_density away from the reference's doctest point; the RK3
 * coefficients of orc_column_run beyond the amplification polynomial and
 * the step sequence of orc_sediment_time_step (Oceananigans' RK3 and
 * QuasiAdamsBashforth2, outside the reference tree) beyond the hand-worked
 * sequence and the nitrogen budget; the WENO reconstruction against
 * Oceananigans' own (the textbook scheme is pinned).
 */
#ifndef OBME_ORACLE_H
#define OBME_ORACLE_H

#include "../include/obme.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Julia scalar semantics ---- */
double orc_julia_min(double x, double y);
double orc_julia_max(double x, double y);
double orc_julia_eps(double x);
double orc_pow_int(double x, int n);  /* Julia Float64^Integer */
double orc_pow_int4(double x);

/* ---- seawater density and carbonate chemistry ---- */
double orc_teos10_density(double T, double Sp, double Pbar);
typedef struct orc_constants {
    double K0, K1, K2, KB, KW, KS, KF, KP1, KP2, KP3, KSi, Is;
} orc_constants;
void orc_equilibrium_constants(double Tk, double S, int has_P, double P, orc_constants* out);
double orc_pressure_correction(int which, double Tk, double P);
double orc_ksp_calcite(double Tk, double S, int has_P, double P);
double orc_ksp_aragonite(double Tk, double S, int has_P, double P);
double orc_virial_B(double Tk);
double orc_virial_delta(double Tk);

typedef struct orc_carbon_params {
    double DIC, Alk, boron, sulfate, fluoride, silicate, phosphate;
    double K1, K2, KB, KW, KS, KF, KP1, KP2, KP3, KSi;
} orc_carbon_params;
double orc_alkalinity_residual(double H, const orc_carbon_params* p);
double orc_dalkalinity_residual(double H, const orc_carbon_params* p);

/* Damped Newton on the alkalinity residual; *evals counts residual calls.
 * fast_exit=1 stops at an exact fixed point (bit-identical result). */
double orc_solve_for_H(const orc_carbon_params* p, const obme_carbon_solver* s,
                       double x0, int fast_exit, int* iters, int* evals);

/* Damped Newton on one of the analytic test functions of test/test_solvers.jl
 * (which = 1..5); lower bound used iff has_lower. */
double orc_damped_newton_test(int which, double x0, double damping, int has_lower,
                              double lower, int fast_exit, int* iters);

/* CarbonChemistry functor outputs (carbon_chemistry.jl:111-211).
 * output: 0 fCO2, 1 pCO2, 2 pH free, 3 pH total, 4 pH seawater.
 * pH < -100 means "no pH given, solve for it". */
double orc_carbon_chemistry(double DIC, double Alk, double T, double S, double pH,
                            int output, const obme_carbon_solver* s, int fast_exit,
                            double silicate, double phosphate, int* evals);
double orc_carbonate_concentration(double DIC, double Alk, double T, double S,
                                   int has_P, double P, const obme_carbon_solver* s);
double orc_calcite_saturation(double DIC, double Alk, double T, double S,
                              int has_P, double P, const obme_carbon_solver* s);

double orc_calcite_saturation_si(double DIC, double Alk, double T, double S, int has_P, double P,
                                 double silicate, const obme_carbon_solver* s);
void orc_calcite_saturation_batch(const double* DIC, const double* Alk, const double* T, const double* S,
                                  const double* P, const double* Si, double* out, long n,
                                  const obme_carbon_solver* s, int nthreads);

double orc_schmidt_co2(const obme_co2_exchange* g, double T);
double orc_o2_flux(const obme_o2_exchange* g, double O2, double T, double S);
double orc_transfer_velocity(const obme_co2_exchange* g, double u10, double T, double S);
/* F = k (pCO2_w - pCO2_air); also returns the free-scale pH of the solve. */
double orc_co2_flux(const obme_co2_exchange* g, const obme_carbon_solver* s,
                    double DIC, double Alk, double T, double S, double u10,
                    double pCO2_air, int fast_exit, double* pH_out, int* evals);
void orc_co2_flux_batch(const obme_co2_exchange* g, const obme_carbon_solver* s,
                        const double* DIC, const double* Alk, const double* T,
                        const double* S, const double* u10, double pCO2_air,
                        double* out_pH, double* out_flux, long n, int nthreads);

/* ---- NutrientsPlanktonDetritus tendencies ---- */
/* vals: tracer values in tracer order (NPZD: N,P,Z,T,D; LOBSTER: NO3,NH4,P,Z,
 * sPOM,bPOM,DOM[,DIC,Alk]); out: the BGC transition of every tracer. */
void orc_npd_tendencies(int model, int carbonate, const obme_npd_params* p,
                        const double* vals, double PAR, double* out);
int orc_npd_ntracers(int model, int carbonate);
/* NPZD or LOBSTER with the Oxygen component: tracer order as above with O2 last. */
void orc_npd_tendencies_oxygen(int model, int carbonate, const obme_npd_params* p, const obme_oxygen* ox,
                               const double* vals, double PAR, double* out);

/* ---- light ---- */
/* TwoBand PAR for one column (Light/2band.jl:1-33); P and PAR of length Nz,
 * bottom to top. */
void orc_two_band_column(const obme_two_band_par* par, int Nz, const double* zc,
                         const double* zf, const double* P, double PAR0, double* PAR);
double orc_default_surface_PAR(double t);

/* ---- sinking ---- */
/* w on the Nz+1 faces of a column (bottom face first), Utils/sinking_velocity_fields.jl:10-35 */
void orc_sinking_profile(const obme_sinking* s, double speed, int Nz, double* w);
/* (F(k+1) - F(k)) / dz(k) with F = w c(face) and the face value of the
 * scheme (OBME_UPWIND1, OBME_CENTERED2, OBME_WENO5); the tendency is minus this */
void orc_sinking_divergence(int scheme, int Nz, const double* w, const double* c, const double* dz,
                            double* div);
/* The face value at face k (0 .. Nz; k = 0 the bottom) of the scheme, from above. */
double orc_sinking_face_value(int scheme, int Nz, const double* c, int k);

/* ---- one column, nsteps of RK3 ---- */
typedef struct orc_column_env {
    const obme_config* cfg;
    int Nz;
    const double* zc;   /* [Nz]   */
    const double* zf;   /* [Nz+1] */
    const double* dz;   /* [Nz]   */
    double PAR0_const;  /* used when surface_mode != DIURNAL */
    double T_surface;   /* gas exchange inputs (LOBSTER) */
    double S_surface;
    const double* Tcol; /* NPZD: T tracer column (read-only) or NULL */
    int fast_exit;
    /* the sediment of this column (cfg->sediment): its storage, Gⁿ and G⁻,
     * and the sediment clock (last Δt, iteration), or NULL */
    struct orc_sediment_column* sed;
} orc_column_env;

/* One column's share of a BiogeochemicalSediment's state: its fields
 * (InstantRemineralisation: storage; SimpleMultiG: Ns, Nf, Nr) and their
 * Gⁿ, G⁻, and the sediment clock. */
#define ORC_SEDIMENT_MAX_FIELDS 3
typedef struct orc_sediment_column {
    double u[ORC_SEDIMENT_MAX_FIELDS], Gn[ORC_SEDIMENT_MAX_FIELDS], Gm[ORC_SEDIMENT_MAX_FIELDS];
    double last_dt;
    long iteration;
} orc_sediment_column;

/* The number of sediment fields of the configured model (0 without one). */
int orc_sediment_fields(const obme_sediment* s);
/* The sinking flux through the bottom face of a column (U[n*Nz + k]
 * layout), InstantRemineralisation's burial efficiency, the tendencies of
 * the sediment's fields from their values u and the tracked flux f, and one
 * sediment time step (QuasiAdamsBashforth2) from the tracked flux f. */
double orc_sediment_flux(const obme_config* cfg, const double* U, int Nz);
double orc_burial_efficiency(const obme_sediment* s, double f);
void orc_sediment_tendencies(const obme_sediment* s, const double* u, double f, double* G);
void orc_sediment_time_step(const obme_sediment* s, orc_sediment_column* c, double f, double dt);
/* SimpleMultiG's coupled tendencies (per unit area, before the division by
 * Δz) of NH4, NO3 and O2 from the fields u = (Ns, Nf, Nr) and the bottom
 * cell's NO3, NH4, O2 (simple_multi_G.jl:193-247, 384-426). */
void orc_smg_coupled(const obme_sediment* s, const double* u, double NO3, double NH4, double O2,
                     double* dNH4, double* dNO3, double* dO2);

/* U[n*Nz + k]: tracer n at level k (0-based, bottom to top).  Advances nsteps
 * in place; if PAR_out != NULL writes the PAR of the last stage evaluated. */
void orc_column_run(const orc_column_env* env, double* U, int nsteps, double dt,
                    double t0, double* PAR_out);
/* ScaleNegativeTracers on one cell's tracer values (indexed by tracer). */
void orc_scale_negatives(const obme_negative_scaling* s, double* vals);
/* ... in place on every interior cell of a grid. */
int orc_grid_scale_negatives(const obme_config* cfg, double* const* tracers);
/* Tendencies of one column at time t (PAR scan + BGC + sinking + CO2 flux). */
void orc_column_tendencies(const orc_column_env* env, const double* U, double t,
                           double* G, double* PAR);

/* ---- 3-D driver on Oceananigans-layout arrays ---- */
/* tracers[n]: parent arrays (Nx+2Hx)(Ny+2Hy)(Nz+2Hz).  T3/S3 may be NULL.
 * Runs nthreads OpenMP threads over columns. */
/* PAR0: host Nx*Ny surface PAR (OBME_SURFACE_ARRAY) or NULL. */
int orc_grid_run(const obme_config* cfg, double* const* tracers, const double* T3,
                 const double* S3, const double* PAR0, int nsteps, double dt, double t0,
                 int nthreads, int fast_exit);
/* orc_grid_run with the sediment: storage, Gn, Gm host fields x Nx*Ny (i
 * fastest, field after field), the sediment clock in *last_dt, *iteration
 * (Inf and 0 before its first step). */
int orc_grid_run_sediment(const obme_config* cfg, double* const* tracers, const double* T3,
                          const double* S3, const double* PAR0, int nsteps, double dt, double t0,
                          int nthreads, int fast_exit, double* storage, double* Gn, double* Gm,
                          double* last_dt, long* iteration);
/* update_biogeochemical_state!(model, sediment): one sediment step of dt
 * (none for a non-finite dt) from the flux of the given tracers. */
int orc_grid_sediment_update_state(const obme_config* cfg, const double* const* tracers,
                                   double* storage, double* Gn, double* Gm, double* last_dt,
                                   long* iteration, double dt);
int orc_grid_run_columns(const obme_config* cfg, double* const* tracers, const double* T3,
                         const double* S3, const double* PAR0, int nsteps, double dt,
                         double t0, const int* col_i, const int* col_j, int ncols,
                         int nthreads, int fast_exit);
int orc_grid_update_state(const obme_config* cfg, const double* const* tracers,
                          const double* PAR0, double* PAR, double t);
int orc_grid_tendencies(const obme_config* cfg, const double* const* tracers,
                        double* const* G, const double* T3, const double* S3,
                        const double* PAR0, double t, double* PAR);
/* orc_grid_tendencies with the sediment's fields (host, fields x Nx*Ny, i
 * fastest; NULL: an empty sediment) for the coupled tendencies. */
int orc_grid_tendencies_sediment(const obme_config* cfg, const double* const* tracers,
                                 double* const* G, const double* T3, const double* S3,
                                 const double* PAR0, double t, double* PAR, const double* fields);

/* ---- batched ensemble of independent NPZD columns ---- */
/* offsets[c]..offsets[c+1]: levels of column c, k fastest, bottom to top.
 * params: OBME_BP_COUNT arrays of n_columns doubles. U: 4 tracers (N,P,Z,D)
 * each of total length offsets[n]; T is uniform T_const. */
int orc_batch_run(const obme_config* cfg, int n_columns, const int* offsets,
                  const double* const* params, double dz, double T_const,
                  double* const* U, int nsteps, double dt, double t0, int nthreads);
/* orc_batch_run with an InstantRemineralisation sediment under every column
 * (cfg->sediment; the tracer indices in NPZD's order N, P, Z, T, D):
 * storage, Gn, Gm of n_columns each and the shared sediment clock. */
int orc_batch_run_sediment(const obme_config* cfg, int n_columns, const int* offsets,
                           const double* const* params, double dz, double T_const,
                           double* const* U, int nsteps, double dt, double t0, int nthreads,
                           double* storage, double* Gn, double* Gm, double* last_dt, long* iteration);

#ifdef __cplusplus
}
#endif
#endif
