/*
 * obme.h — C ABI of libobme_hip.so, the MI355X (gfx950) biogeochemistry step.
 *
 * The library is a drop-in for the biogeochemistry step that OceanBioME.jl
 * (v0.17.6) plugs into Oceananigans: per-cell NPZD / LOBSTER(+CarbonateSystem)
 * tendencies, the two-band PAR column scan, the sinking (drift-velocity) flux
 * divergence, the CarbonChemistry pH solve and the GasExchange CO2 flux, and
 * the RK3 stage sequence that strings them together.
 *
 * Conventions
 *  - Every entry point returns an obme_status (0 = OK). obme_last_error()
 *    returns a thread-local message for the last failure on this thread.
 *  - Array arguments are DEVICE pointers unless stated otherwise.  Host
 *    arrays inside configuration structs are copied at obme_create().
 *  - 3-D fields use the Oceananigans Field layout: parent(field) data of
 *    size (Nx+2Hx, Ny+2Hy, Nz+2Hz), i fastest, k = 1 at the bottom; interior
 *    cell (i,j,k) (1-based) lives at
 *      (i-1+Hx) + (Nx+2Hx)*((j-1+Hy) + (Ny+2Hy)*(k-1+Hz)).
 *  - Element type of tracer / tendency / PAR arrays: double for OBME_FP64,
 *    float for OBME_FP32_MIXED.  Carbonate-chemistry arrays are always double.
 *  - All calls are asynchronous on the given stream (NULL = default stream),
 *    one handle per host thread; a handle is not re-entrant.  One stream per
 *    handle: the calls on a handle share its scratch buffers (surface
 *    records, warm starts, exchange coefficients, residuals), so calls on a
 *    handle made on different streams must be ordered by the caller (an
 *    event, or a stream synchronisation) or they race.
 *  - Numerical failures are not errors: NaN propagates (min/max propagate NaN
 *    as Julia's do; a non-integer power of a negative value gives NaN).
 */
#ifndef OBME_H
#define OBME_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t obme_status;
enum {
    OBME_OK = 0,
    OBME_ERR_INVALID = 1,     /* bad argument or configuration            */
    OBME_ERR_HIP = 2,         /* HIP runtime error                        */
    OBME_ERR_UNSUPPORTED = 3, /* valid request the library does not handle */
    OBME_ERR_NOMEM = 4
};

/* Model compositions (NutrientsPlanktonDetritus/constructors.jl). */
enum {
    OBME_NPZD = 1,    /* Nutrient + PhytoZoo + Detritus; tracers (N, P, Z, T, D)      */
    OBME_LOBSTER = 2  /* NitrateAmmonia + PhytoZoo + TwoParticleAndDissolved;
                         tracers (NO3, NH4, P, Z, sPOM, bPOM, DOM [, DIC, Alk])    */
};

enum { OBME_FP64 = 0, OBME_FP32_MIXED = 1 };

enum { OBME_LIGHT_MONDO = 0, OBME_LIGHT_ANALYTICAL = 1 };   /* plankton.jl:216-220 */
enum { OBME_LINEAR = 0, OBME_QUADRATIC = 1 };               /* plankton.jl:83-90   */
/* sinking face value: first-order upwind, second-order centred, fifth-order
 * WENO (the reconstructions of Oceananigans' UpwindBiased(order = 1),
 * Centered(order = 2) and WENO(order = 5) applied to the drift velocity) */
enum { OBME_UPWIND1 = 0, OBME_CENTERED2 = 1, OBME_WENO5 = 2 };

enum {
    OBME_SURFACE_CONSTANT = 0,  /* PAR0 = surface_value                          */
    OBME_SURFACE_DIURNAL = 1,   /* PAR0 = 100 max(0, cos(t pi / 12 h)), Light.jl:219 */
    OBME_SURFACE_ARRAY = 2      /* PAR0[i,j] from a device array (Nx*Ny, i fastest) */
};

enum { OBME_FIELD_CONSTANT = 0, OBME_FIELD_PROFILE = 1, OBME_FIELD_3D = 2 };

#define OBME_MAX_TRACERS 32

/* NutrientsPlanktonDetritus parameters, one flat POD struct.
 * PhytoZoo:                plankton.jl:19-58
 * NitrateAmmonia:          nutrients.jl:17-19
 * TwoParticleAndDissolved: detritus.jl:116-128
 * Detritus:                detritus.jl:355-361
 * Defaults for NPZD / LOBSTER are set by the host package (constructors.jl). */
typedef struct obme_npd_params {
    double nitrate_half_saturation;
    double ammonia_half_saturation;
    double nitrate_ammonia_inhibition;
    double light_half_saturation;
    double phytoplankton_maximum_growth_rate;
    double phytoplankton_exudation_fraction;
    double ammonia_fraction_of_exudate;
    double temperature_coefficient;          /* used when use_temperature_coefficient */
    double phytoplankton_mortality_rate;
    double zooplankton_mortality_rate;
    double zooplankton_excretion_rate;
    double phytoplankton_solid_waste_fraction;
    double excretion_inorganic_fraction;
    double preference_for_phytoplankton;
    double maximum_grazing_rate;
    double grazing_half_saturation;
    double zooplankton_assimilation_fraction;
    double redfield_ratio;                   /* PhytoZoo redfield_ratio      */
    double carbon_calcite_ratio;
    double zooplankton_gut_calcite_dissolution;
    double phytoplankton_chlorophyll_ratio;
    double nitrification_rate;               /* NitrateAmmonia               */
    double remineralisation_inorganic_fraction; /* TwoParticleAndDissolved   */
    double small_remineralisation_rate;
    double large_remineralisation_rate;
    double dissolved_remineralisation_rate;
    double small_solid_waste_fraction;
    double detritus_redfield_ratio;          /* detritus redfield_ratio      */
    double remineralisation_rate;            /* Detritus (single class)      */
    double small_particle_fraction;          /* Detritus (single class)      */
    int32_t light_limitation;                /* OBME_LIGHT_*                 */
    int32_t use_temperature_coefficient;     /* 0 = `nothing` (no Q10)       */
    int32_t phytoplankton_mortality_formulation; /* OBME_LINEAR/QUADRATIC   */
    int32_t grazing_concentration_formulation;   /* OBME_LINEAR/QUADRATIC   */
} obme_npd_params;

/* TwoBandPhotosyntheticallyActiveRadiation (Light/2band.jl:35-70, defaults :106-117). */
typedef struct obme_two_band_par {
    double water_red_attenuation;
    double water_blue_attenuation;
    double chlorophyll_red_attenuation;
    double chlorophyll_blue_attenuation;
    double chlorophyll_red_exponent;
    double chlorophyll_blue_exponent;
    double pigment_ratio;
    double phytoplankton_chlorophyll_ratio;
    int32_t surface_mode;          /* OBME_SURFACE_*                        */
    int32_t pad0;
    double surface_value;          /* PAR0 for OBME_SURFACE_CONSTANT (W/m2) */
    const void* surface_array;     /* device, Nx*Ny of the tracer element type */
} obme_two_band_par;

/* Drift velocities (Utils/sinking_velocity_fields.jl:10-35): per tracer, in
 * tracer order, a sinking speed >= 0 in m/s; 0 means no drift velocity. */
typedef struct obme_sinking {
    double speed[OBME_MAX_TRACERS];
    int32_t open_bottom;           /* 1: w = -speed on faces 1..Nz; 0: damped to 0 at the bottom */
    int32_t scheme;                /* OBME_UPWIND1, OBME_CENTERED2 or OBME_WENO5 */
    double smoothing_distance;     /* default 2 */
} obme_sinking;

/* DampedNewtonRaphsonSolver (Utils/solvers.jl:81-88) with the CarbonChemistry
 * lower bound 0 (CarbonChemistry/carbon_chemistry.jl:81). */
typedef struct obme_carbon_solver {
    int32_t max_iters;             /* 100   */
    int32_t pad0;
    double atol;                   /* 1e-20 */
    double damping;                /* 0.5   */
    double armijo_constant;        /* 0.5   */
    double min_damping;            /* 0.5^10 */
    double initial_pH_guess;       /* 8     */
} obme_carbon_solver;

/* CarbonDioxideGasExchangeBoundaryCondition (GasExchange/GasExchange.jl:104-124). */
typedef struct obme_co2_exchange {
    int32_t enabled;               /* adds the top DIC flux when the carbonate system is on */
    int32_t pad0;
    double air_concentration;      /* ppmv, 413                                      */
    double wind_speed;             /* m/s, 2                                         */
    double air_pressure;           /* atm, 1 (carbon_dioxide_concentration.jl:222)   */
    double schmidt[5];             /* Sc(T) = a + bT + cT^2 + dT^3 + eT^4 (gas_solubility.jl:211) */
    double transfer_velocity[3];   /* k660(u) = c0 + c1 u + c2 u^2, Ho06 (gas_transfer_velocity.jl:61) */
} obme_co2_exchange;

/* Grid: interior sizes, halos and vertical coordinates (host arrays, copied). */
typedef struct obme_grid {
    int32_t Nx, Ny, Nz;
    int32_t Hx, Hy, Hz;
    const double* zc;              /* [Nz]   cell centres, bottom to top      */
    const double* zf;              /* [Nz+1] faces, bottom to top             */
    const double* dz;              /* [Nz]   cell thickness (Delta z^ccc)     */
} obme_grid;

/* Read-only physical field (temperature or salinity). */
typedef struct obme_field_source {
    int32_t mode;                  /* OBME_FIELD_*                            */
    int32_t pad0;
    double value;                  /* OBME_FIELD_CONSTANT                     */
    const double* profile;         /* host [Nz], OBME_FIELD_PROFILE (copied)  */
} obme_field_source;

/* ScaleNegativeTracers (src/Utils/negative_tracers.jl:36-59, :137-276): per
 * cell and conserved group, t = Σ s·c over the members, p = Σ s·c over the
 * positive members, t := invalid_fill_value where t < 0, then every member
 * c := c·t/p where c > 0 or c is not finite, else 0.  Groups apply in order
 * (a tuple of modifiers); members are tracer indices in tracer order, in
 * the order of the reference's conserved_tracers (the order of the sums).
 * It runs first in update_biogeochemical_state! (src/OceanBioME.jl:161-167),
 * so the fused steppers apply it at the start of every RK3 stage, before the
 * light and the tendencies.  n_groups = 0: off. */
#define OBME_MAX_SCALE_GROUPS 5
#define OBME_MAX_SCALE_MEMBERS 12
typedef struct obme_negative_scaling {
    int32_t n_groups;
    int32_t n_members[OBME_MAX_SCALE_GROUPS];
    int32_t member[OBME_MAX_SCALE_GROUPS][OBME_MAX_SCALE_MEMBERS];
    int32_t pad0;
    double scalefactor[OBME_MAX_SCALE_GROUPS][OBME_MAX_SCALE_MEMBERS];
    double invalid_fill_value[OBME_MAX_SCALE_GROUPS];
} obme_negative_scaling;

/* Sediment models (src/Models/Sediments). */
enum { OBME_SEDIMENT_NONE = 0, OBME_SEDIMENT_INSTANT_REMINERALISATION = 1, OBME_SEDIMENT_SIMPLE_MULTI_G = 2 };
enum { OBME_SEDIMENT_AB2 = 0 };   /* QuasiAdamsBashforth2, the sediment's default time stepper */

/* InstantRemineralisationSediment (Models/Sediments/instant_remineralisation.jl:60-123)
 * coupled through BiogeochemicalSediment (src/Sediments): at the bottom face
 * the sinking flux of the listed tracers, f = Σ -w c (tracked_fields.jl:53-70,
 * first-order upwind: c of the bottom cell), burial efficiency
 * e = a + b (f / (k + f))^2; the receiver's bottom cell gains (1 - e) f / Δz
 * (tracer_coupling.jl:3-38, Δz at the bottom face = the bottom cell's
 * thickness) and the 2-D storage field gains e f through the sediment's own
 * time stepper (update_state.jl:6-17, timesteppers.jl:16-43). */
#define OBME_MAX_SEDIMENT_TRACERS 6
typedef struct obme_sediment {
    int32_t model;                 /* OBME_SEDIMENT_*                                   */
    int32_t receiver;              /* tracer index of remineralisation_reciever (N, NH4) */
    int32_t n_sinking;
    int32_t sinking[OBME_MAX_SEDIMENT_TRACERS];   /* tracer indices of sinking_tracers  */
    int32_t timestepper;           /* OBME_SEDIMENT_AB2                                 */
    double burial_efficiency_constant1;           /* a, 0.013                           */
    double burial_efficiency_constant2;           /* b, 0.53                            */
    double burial_efficiency_half_saturation;     /* k, 7 / 6.56                        */
    double ab2_chi;                /* χ of QuasiAdamsBashforth2, 0.1                     */
    /* SimpleMultiG, nitrogen form (Models/Sediments/simple_multi_G.jl:37-45,
     * 158-247, 338-426): fields Ns, Nf, Nr; the sinking nitrogen flux of the
     * `sinking` tracers feeds them, and the bottom cells of NH4, NO3 and O2
     * receive the coupled remineralisation.  Unused by InstantRemineralisation. */
    int32_t ammonia, nitrate, oxygen;             /* tracer indices of NH4, NO3, O2      */
    double fast_decay_rate, slow_decay_rate;      /* λf 2/day, λs 0.2/day (1/s)          */
    double fast_fraction, slow_fraction, refactory_fraction;   /* 0.74, 0.26, 0.1       */
    double sedimentation_rate;     /* cm/year: 982 |z of the bottom cell centre|^-1.548  */
    double anoxia_half_saturation; /* mmol/m³, 1.0                                      */
    double sinking_redfield;       /* C:N of the sinking nitrogen, 6.56                 */
    double nitrate_oxidation_params[6];           /* A .. F of the ammonia oxidation     */
    double denitrification_params[6];
    double anoxic_params[6];
} obme_sediment;

/* OxygenGasExchangeBoundaryCondition (GasExchange/GasExchange.jl:127-146):
 * F = k (O2 - C_air α) with k = k660(u10) / sqrt(Sc(T)/660)
 * (gas_transfer_velocity.jl:24-34; Ho06 :61), Sc the Wanninkhof (2014)
 * quartic (schmidt_number.jl:15-16), α = β/Tk the Wanninkhof (1992)
 * solubility (gas_solubility.jl:28-54, with its S (B1 + B2 T/100 +
 * B2 (T/100)^2) term as written there). */
typedef struct obme_o2_exchange {
    int32_t enabled;
    int32_t pad0;
    double air_concentration;      /* mmol O2 / m3 (9352.7) */
    double wind_speed;             /* m/s */
    double schmidt[5];             /* a, b, c, d, e */
    double transfer_velocity[3];   /* k660 = c0 + c1 u + c2 u^2 */
    double solubility[6];          /* A1, A2, A3, B1, B2, B3 */
} obme_o2_exchange;

/* Oxygen() (NutrientsPlanktonDetritus/oxygen.jl:14-46): the O₂ tracer of
 * LOBSTER or NPZD, last in the tracer order (NutrientsPlanktonDetritus.jl:69-74).
 * One-way coupled: its tendency (LOBSTER: Rp μP - (Rp - Rn) G_NH₄ - Rp
 * nitrification; NPZD: Rp μP - (Rp - Rn) × the plankton's inorganic nitrogen
 * waste) reads the other tracers and changes none of them. */
typedef struct obme_oxygen {
    int32_t enabled;               /* 1 adds O₂                                          */
    int32_t pad0;
    double respiration_oxygen_nitrogen_ratio;     /* Rp, 10.75 mol O / mol N            */
    double nitrification_oxygen_nitrogen_ratio;   /* Rn, 2.0 mol O / mol N              */
} obme_oxygen;

typedef struct obme_config {
    int32_t model;                 /* OBME_NPZD | OBME_LOBSTER                */
    int32_t precision;             /* OBME_FP64 | OBME_FP32_MIXED             */
    int32_t carbonate_system;      /* LOBSTER: 1 adds DIC and Alk             */
    int32_t pad0;
    obme_grid grid;
    obme_npd_params npd;
    obme_two_band_par par;
    obme_sinking sinking;
    obme_carbon_solver solver;
    obme_co2_exchange co2;
    obme_field_source temperature; /* LOBSTER gas exchange; NPZD reads T as a tracer */
    obme_field_source salinity;
    obme_negative_scaling scaling; /* NPZD(grid; scale_negatives = true) etc. */
    obme_sediment sediment;        /* NPZD(; grid, sediment_model) etc.          */
    obme_oxygen oxygen;            /* NPZD/LOBSTER(grid; oxygen = Oxygen())     */
    obme_o2_exchange o2;           /* LOBSTER with O₂: the O₂ gas exchange at the top */
} obme_config;

typedef struct obme_handle obme_handle;

/* ---- lifecycle ---------------------------------------------------------- */

/* Replaces the construction of NPZD(grid) / LOBSTER(grid; carbonate_system)
 * (NutrientsPlanktonDetritus/constructors.jl:65-96, :177-227) for the device. */
obme_status obme_create(const obme_config* cfg, obme_handle** out);
obme_status obme_destroy(obme_handle* h);
const char* obme_last_error(void);
/* Library version string, e.g. "obme 0.1.0 gfx950". */
const char* obme_version(void);

/* required_biogeochemical_tracers(bgc) (src/OceanBioME.jl:122;
 * NutrientsPlanktonDetritus.jl:69-74): writes up to `cap` tracer names,
 * separated by ';', into `buf`; returns the tracer count or <0 on error. */
int32_t obme_required_tracers(const obme_handle* h, char* buf, int32_t cap);

/* ---- the Oceananigans.Biogeochemistry hooks ----------------------------- */

/* update_biogeochemical_state!(bgc, model) (src/OceanBioME.jl:161-167):
 * the modifiers first — with negative scaling configured, the tracers are
 * rescaled in place (src/Utils/negative_tracers.jl:137-251) — then the
 * light step, i.e. the TwoBand PAR column scan (Light/2band.jl:1-33,
 * :148-155). Reads tracers[P], writes the interior of PAR at time t. */
obme_status obme_update_state(obme_handle* h, void* const* tracers,
                              void* PAR, double t, void* stream);

/* update_tendencies!(bgc, model) (src/OceanBioME.jl:148-155), in the form
 * of the sediment coupling hook (src/Sediments/tracer_coupling.jl:3-38):
 * ADDS into G[n] the biogeochemical transition of every tracer
 * ((bgc)(i,j,k,grid,Val(name),clock,fields), src/OceanBioME.jl:157-159),
 * minus the drift-velocity flux divergence (biogeochemical_drift_velocity,
 * src/OceanBioME.jl:126) and, with the carbonate system and gas exchange
 * enabled, minus F/dz of the top-cell CO2 flux (GasExchange/gas_exchange.jl:26-38);
 * with LOBSTER's O2 and its gas exchange, minus F/dz of the top-cell O2 flux
 * (GasExchange.jl:127-146).  T and S: optional 3-D fields (NULL = the
 * configured source). */
obme_status obme_add_tendencies(obme_handle* h, const void* const* tracers,
                                void* const* G, const void* PAR,
                                const void* T, const void* S,
                                double t, void* stream);

/* Standalone fused stepper with frozen physics: nsteps RK3 steps
 * (Oceananigans RungeKutta3: gamma 8/15, 5/12, 3/4; zeta -17/60, -5/12; the
 * substep form of src/BoxModel/timesteppers.jl:79-92) of update-state,
 * tendencies and substep, in place on `tracers`. */
obme_status obme_step(obme_handle* h, void* const* tracers,
                      const void* T, const void* S,
                      int32_t nsteps, double dt, double t0, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  obme_step with its per-call scratch (surface records,
 * [H+] warm starts, surface constants, O2 exchange coefficients, top-cell
 * DIC and Alk residuals) in a static
 * device array of 64 MiB per device instead of buffers the library
 * allocates; it calls no hipMalloc or hipFree (the handle itself comes from
 * obme_create).  A grid whose scratch exceeds 64 MiB is rejected with
 * OBME_ERR_INVALID (FP32-mixed LOBSTER + DIC/Alk with CO2 exchange takes
 * 40 B per column, about 1.6 million columns).  The warm starts begin at
 * zero on every call: where obme_step's handle carries warm starts from
 * earlier calls, results differ from it within the root solve's tolerance
 * (a relative Newton step of 1e-4 FP32-mixed, 1e-14 FP64); on a fresh
 * handle FP64 results are the same.  Every call on a
 * device shares that device's array: calls that overlap on one device race.
 * Must be called on the device the handle was created on. */
obme_status obme_step_noalloc(obme_handle* h, void* const* tracers, const void* T, const void* S,
                              int32_t nsteps, double dt, double t0, void* stream);

/* No counterpart in the reference: the launch shape of the column sweep.
 * waves = waves per workgroup (1, 2, 4 or 8; a workgroup covers 64*waves
 * consecutive columns of one row).  Results do not depend on it, bit for
 * bit.  The default is the widest of 8, 4, 2, 1 whose 64*waves columns a
 * row of Nx fills (8 from Nx = 512), or the environment's OBME_SWEEP_WAVES. */
obme_status obme_sweep_configure(obme_handle* h, int32_t waves);
/* The current waves per workgroup, or < 0 for a null handle. */
int32_t obme_sweep_waves(const obme_handle* h);
/* No counterpart in the reference: the sweep's autotuner.  Times obme_step
 * (one warm-up step, then trial_steps steps, best of two rounds) with every
 * launch shape on a scratch copy of `tracers`, which are only read (T and S
 * as in obme_step), and keeps the fastest.  The handle's carried state (the
 * sediment stepper's history, the [H+] warm starts, the kernel-timing
 * records) is left as it was.  Allocates the scratch copy; synchronises the
 * stream. */
obme_status obme_autotune(obme_handle* h, void* const* tracers, const void* T, const void* S,
                          int32_t trial_steps, double dt, void* stream);

/* ---- sediment (BiogeochemicalSediment, src/Sediments) -------------------
 * The sediment's state is its fields (device doubles, the caller's; `storage`
 * below): InstantRemineralisation's storage (Nx*Ny, i fastest), or
 * SimpleMultiG's Ns, Nf, Nr one after the other (3 Nx*Ny); and its time
 * stepper's tendencies, which the handle keeps.  With an
 * InstantRemineralisation sediment, obme_add_tendencies also adds the
 * coupled tendency of the receiver (update_tendencies!(bgc, sediment, model),
 * tracer_coupling.jl:3-38) from the sinking flux of the tracers it is given,
 * and obme_step applies it at every stage (storage not advanced).
 * SimpleMultiG's coupled tendencies depend on its fields: they are added by
 * obme_add_tendencies_with_sediment and by the steppers, from the fields
 * given (obme_step, without fields: an empty sediment, which couples
 * nothing). */

/* update_biogeochemical_state!(model, sediment) (Sediments/update_state.jl:6-17):
 * the tracked sinking flux f of the given tracers (tracked_fields.jl:8-70)
 * and, for a finite dt (the model's last stage Δt), one sediment time step:
 * QuasiAdamsBashforth2 (timesteppers.jl:16-43 in Oceananigans' AB2
 * time_step!): a forward Euler step when dt differs from the previous step's
 * (always, between RK3 stages), else χ-weighted with the previous tendency;
 * the tendency e f is evaluated after each step from the tracked flux, and
 * before the first step (iteration 0) from the flux of that call.
 * SimpleMultiG: the tendencies of Ns, Nf, Nr from the tracked nitrogen flux
 * and the fields (simple_multi_G.jl:158-189). */
obme_status obme_sediment_update_state(obme_handle* h, const void* const* tracers, double* storage,
                                       double dt, void* stream);

/* obme_add_tendencies, with the coupled tendencies of a SimpleMultiG
 * sediment from its fields `storage` (3 Nx*Ny doubles: Ns, Nf, Nr) and the
 * bottom cells' NO3, NH4, O2 (update_coupled_tendency!, tracer_coupling.jl:
 * 30-38: G[bottom] += tendency / Δz; simple_multi_G.jl:193-247).  Any other
 * configuration: the same as obme_add_tendencies. */
obme_status obme_add_tendencies_with_sediment(obme_handle* h, const void* const* tracers, void* const* G,
                                              const void* PAR, const void* T, const void* S,
                                              const double* storage, double t, void* stream);

/* obme_step with the sediment's storage advanced in the same pass: at every
 * RK3 stage the coupled tendency of the stage's flux, and after every stage
 * the sediment time step of obme_sediment_update_state with the stage Δt
 * (Δt γ1, Δt (γ2 + ζ2), Δt (γ3 + ζ3)), as the reference's
 * update_biogeochemical_state! does after each substep.  storage may be
 * NULL (the coupling alone). */
obme_status obme_step_with_sediment(obme_handle* h, void* const* tracers, const void* T, const void* S,
                                    double* storage, int32_t nsteps, double dt, double t0, void* stream);

/* ---- carbonate chemistry / gas exchange on point sets (SoA, n points) --- */

/* CarbonChemistry()(; DIC, Alk, T, S, output = Val(:pHᶠ)) and the CO2
 * GasExchange flux F = k(u10,T,S) (pCO2_w - pCO2_air)
 * (CarbonChemistry/carbon_chemistry.jl:111-165; GasExchange/gas_exchange.jl:26-38).
 * mode 0: the reference damped Newton from pH 8 in FP64;
 * mode 1: an FP32 initial guess refined by FP64 Newton.
 * All arrays are double; u10 may be NULL (use cfg wind_speed). out_pH and
 * out_flux may be NULL. */
obme_status obme_co2_flux(obme_handle* h, const double* DIC, const double* Alk,
                          const double* T, const double* S, const double* u10,
                          double pCO2_air, double* out_pH, double* out_flux,
                          int64_t n, int32_t mode, void* stream);

/* ---- batched ensemble of independent NPZD water columns ----------------- */

/* Per-column NPZD parameter set, SoA of n_columns entries each (device). */
typedef struct obme_batch_params {
    const double* fields[40];      /* see OBME_BP_* indices */
} obme_batch_params;

enum {
    OBME_BP_NITRATE_HALF_SATURATION = 0,
    OBME_BP_LIGHT_HALF_SATURATION,
    OBME_BP_MAX_GROWTH_RATE,
    OBME_BP_EXUDATION_FRACTION,
    OBME_BP_AMMONIA_FRACTION_OF_EXUDATE,
    OBME_BP_TEMPERATURE_COEFFICIENT,
    OBME_BP_PHYTO_MORTALITY_RATE,
    OBME_BP_ZOO_MORTALITY_RATE,
    OBME_BP_ZOO_EXCRETION_RATE,
    OBME_BP_PHYTO_SOLID_WASTE_FRACTION,
    OBME_BP_EXCRETION_INORGANIC_FRACTION,
    OBME_BP_PREFERENCE_FOR_PHYTO,
    OBME_BP_MAX_GRAZING_RATE,
    OBME_BP_GRAZING_HALF_SATURATION,
    OBME_BP_ZOO_ASSIMILATION_FRACTION,
    OBME_BP_REMINERALISATION_RATE,
    OBME_BP_SMALL_PARTICLE_FRACTION,
    OBME_BP_P_SINKING_SPEED,
    OBME_BP_D_SINKING_SPEED,
    OBME_BP_SURFACE_PAR,
    OBME_BP_COUNT
};

typedef struct obme_batch obme_batch;

/* An ensemble of independent NPZD water columns (the BoxModel setting of
 * src/BoxModel/boxmodel.jl and timesteppers.jl:30-92, one column per member):
 * column c holds cells [offsets[c], offsets[c+1]) of every tracer array,
 * k = 0 at the bottom, uniform thickness dz, surface at z = 0.  `cfg`
 * supplies the model (OBME_NPZD), precision, TwoBand coefficients, sinking
 * (open bottom, upwind), an optional InstantRemineralisation sediment, and
 * every parameter that `params` leaves NULL;
 * params->fields[OBME_BP_*] are device arrays of n_columns doubles (read once
 * here).  Temperature is uniform.  `offsets` is a host array of n_columns+1. */
obme_status obme_batch_create(const obme_config* cfg, int32_t n_columns, const int32_t* offsets, double dz,
                              const obme_batch_params* params, double temperature, obme_batch** out);
obme_status obme_batch_destroy(obme_batch* b);
/* nsteps RK3 steps of every column, in place on tracers[0..3] = N, P, Z, D
 * (device arrays of offsets[n_columns] elements, double for OBME_FP64, float
 * for OBME_FP32_MIXED).  Each launch advances steps_per_launch steps (0: all
 * of them) with the column state in registers; several launches per call are
 * recorded once into a hipGraph and replayed. */
obme_status obme_batch_step(obme_batch* b, void* const* tracers, int32_t nsteps, double dt, double t0,
                            void* stream);
/* obme_batch_step with the ensemble's InstantRemineralisation sediment
 * (cfg->sediment at create, tracer indices in NPZD's order N, P, Z, T, D;
 * BiogeochemicalSediment, src/Sediments, with instant_remineralisation.jl):
 * storage is a device array of n_columns doubles, one sediment per column.
 * At every stage the bottom cell's sinking flux f (P and D leave through the
 * open bottom face) returns (1 - e) f / dz to the receiver's bottom cell,
 * e = a + b (f / (k + f))^2, and after each stage's substep the storage
 * takes a sediment step of the stage Δt with the tendency e f, as
 * obme_step_with_sediment does for a grid (update_state.jl:6-17); the handle
 * keeps the sediment clock and each column's previous tendency between
 * calls.  obme_batch_step on such an ensemble applies the coupling and
 * leaves the storage alone, as obme_step does. */
obme_status obme_batch_step_with_sediment(obme_batch* b, void* const* tracers, double* storage, int32_t nsteps,
                                          double dt, double t0, void* stream);
/* Column packing: plan = index of a set of tile shapes (columns go to the
 * smallest tile that holds them), steps_per_launch as above. */
obme_status obme_batch_configure(obme_batch* b, int32_t plan, int32_t steps_per_launch);
/* Times every packing plan and steps-per-launch choice for trial_steps steps
 * on a scratch copy of `tracers` (left unchanged) and keeps the fastest. */
obme_status obme_batch_autotune(obme_batch* b, void* const* tracers, int32_t trial_steps, double dt,
                                void* stream);
/* Text description of the current packing; returns its length or < 0. */
int32_t obme_batch_describe(const obme_batch* b, char* buf, int32_t cap);


/* ---- PISCES (src/Models/AdvectedPopulations/PISCES) -------------------- */

/* Tracer order of required_biogeochemical_tracers(PISCES) (PISCES.jl:91-103):
 * the first 24 are written, T and S are read. */
enum {
    OBME_PI_P = 0, OBME_PI_PCHL, OBME_PI_PFE, OBME_PI_D, OBME_PI_DCHL, OBME_PI_DFE, OBME_PI_DSI,
    OBME_PI_Z, OBME_PI_M, OBME_PI_DOC, OBME_PI_POC, OBME_PI_GOC, OBME_PI_SFE, OBME_PI_BFE, OBME_PI_PSI,
    OBME_PI_CACO3, OBME_PI_NO3, OBME_PI_NH4, OBME_PI_PO4, OBME_PI_FE, OBME_PI_SI, OBME_PI_DIC,
    OBME_PI_ALK, OBME_PI_O2, OBME_PI_T, OBME_PI_S, OBME_PI_COUNT
};
#define OBME_PISCES_WRITTEN 24

/* Parameters: obme_pisces_config.params holds two phytoplankton blocks
 * (nano at OBME_PP_NANO, diatoms at OBME_PP_DIATOMS; MixedMondo with
 * GrowthRespirationLimitedProduction and NitrogenIronPhosphateSilicateLimitation,
 * the phytoplankton/ sources), two zooplankton blocks (QualityDependantZooplankton,
 * micro at OBME_PP_MICRO, meso at OBME_PP_MESO) and the scalars at
 * OBME_PP_SCALARS; within a block the OBME_PH_* / OBME_PZ_* / OBME_PS_*
 * offsets apply.  Rates in 1/s, concentrations in mmol/m3 (Fe in umol/m3). */
enum {
    OBME_PH_MU0 = 0,
    OBME_PH_BT,
    OBME_PH_DARK_TOLERANCE,
    OBME_PH_ALPHA0,
    OBME_PH_BETA,
    OBME_PH_BASAL_RESPIRATION,
    OBME_PH_REFERENCE_GROWTH,
    OBME_PH_KNH4,
    OBME_PH_KNO3,
    OBME_PH_KPO4,
    OBME_PH_THETA_FE_OPT,
    OBME_PH_SI_LIMITED,
    OBME_PH_KSI_MIN,
    OBME_PH_PK,
    OBME_PH_DELTA,
    OBME_PH_BETA1,
    OBME_PH_BETA2,
    OBME_PH_BETA3,
    OBME_PH_MORT_K,
    OBME_PH_MORT_LIN,
    OBME_PH_MORT_W0,
    OBME_PH_MORT_W1,
    OBME_PH_CHL_MIN,
    OBME_PH_CHL_MAX,
    OBME_PH_FE_MAX,
    OBME_PH_SI_K1,
    OBME_PH_SI_K2,
    OBME_PH_SI_THETA0,
    OBME_PH_KFE_UPTAKE,
    OBME_PH_SIZE_THRESHOLD,
    OBME_PH_SIZE_RATIO,
    OBME_PH_COUNT
};
enum {
    OBME_PZ_B = 0,
    OBME_PZ_G0,
    OBME_PZ_PREF_P,
    OBME_PZ_PREF_D,
    OBME_PZ_PREF_POC,
    OBME_PZ_PREF_Z,
    OBME_PZ_FOOD_THRESHOLD,
    OBME_PZ_J,
    OBME_PZ_K,
    OBME_PZ_FF,
    OBME_PZ_IRON_RATIO,
    OBME_PZ_E0,
    OBME_PZ_SIGMA,
    OBME_PZ_KM,
    OBME_PZ_M0,
    OBME_PZ_R,
    OBME_PZ_GAMMA,
    OBME_PZ_ETA,
    OBME_PZ_COUNT
};
enum {
    OBME_PS_RAIN_RATIO = 0,
    OBME_PS_BZ,
    OBME_PS_BM,
    OBME_PS_BACT_EXP,
    OBME_PS_BACT_MAX,
    OBME_PS_K_DOC,
    OBME_PS_K_NO3_B,
    OBME_PS_K_NH4_B,
    OBME_PS_K_PO4_B,
    OBME_PS_K_FE_B,
    OBME_PS_DOC_LAMBDA,
    OBME_PS_DOC_BACT_REF,
    OBME_PS_DOC_B,
    OBME_PS_DOC_A1,
    OBME_PS_DOC_A2,
    OBME_PS_DOC_A3,
    OBME_PS_DOC_A4,
    OBME_PS_DOC_A5,
    OBME_PS_POM_B,
    OBME_PS_POM_LAMBDA,
    OBME_PS_POM_A1,
    OBME_PS_POM_A2,
    OBME_PS_POM_A3,
    OBME_PS_POM_A4,
    OBME_PS_FE_SCAV_MIN,
    OBME_PS_FE_SCAV_LOAD,
    OBME_PS_BACT_FE_EFF,
    OBME_PS_SMALL_BACT_FE,
    OBME_PS_LARGE_BACT_FE,
    OBME_PS_CHI0,
    OBME_PS_SI_FAST,
    OBME_PS_SI_SLOW,
    OBME_PS_CACO3_LAMBDA,
    OBME_PS_CACO3_EXP,
    OBME_PS_BACT_FE_MAX,
    OBME_PS_BACT_FE_K,
    OBME_PS_BACT_MU0,
    OBME_PS_NITRIF,
    OBME_PS_FIX_MAX,
    OBME_PS_FIX_KFE,
    OBME_PS_FIX_KPO4,
    OBME_PS_FIX_E,
    OBME_PS_FE_EXCESS,
    OBME_PS_FE_LIGAND_MAX,
    OBME_PS_FE_LIGAND_RATIO,
    OBME_PS_O2_RESP,
    OBME_PS_O2_NITRIF,
    OBME_PS_ANOXIA1,
    OBME_PS_ANOXIA2,
    OBME_PS_THETAN,
    OBME_PS_THETAP,
    OBME_PS_MIXED_LAYER_SHEAR,
    OBME_PS_BACKGROUND_SHEAR,
    OBME_PS_LATITUDE,
    OBME_PS_SILICATE_CLIMATOLOGY,
    OBME_PS_MIXED_LAYER_DEPTH,
    OBME_PS_KAPPA,
    OBME_PS_POC_SPEED,
    OBME_PS_GOC_MIN_SPEED,
    OBME_PS_GOC_MAX_SPEED,
    OBME_PS_GRAVITY,
    OBME_PS_COUNT
};
enum {
    OBME_PP_NANO = 0,
    OBME_PP_DIATOMS = OBME_PH_COUNT,
    OBME_PP_MICRO = 2 * OBME_PH_COUNT,
    OBME_PP_MESO = 2 * OBME_PH_COUNT + OBME_PZ_COUNT,
    OBME_PP_SCALARS = 2 * OBME_PH_COUNT + 2 * OBME_PZ_COUNT,
    OBME_PP_COUNT = 2 * OBME_PH_COUNT + 2 * OBME_PZ_COUNT + OBME_PS_COUNT
};

/* PISCES configuration: grid, parameters (OBME_PP_*), MultiBand light
 * (MultiBandPhotosyntheticallyActiveRadiation, Light/multi_band.jl: per band
 * kʷ, e, χ and the share of the surface PAR) and the carbonate solver of the
 * calcite saturation.  The mixed-layer depth and the mean mixed-layer
 * diffusivity are prescribed constants (params).  The surface PAR is a
 * constant, default_surface_PAR (Light.jl:47, the PISCES default,
 * PISCES.jl:325-331) at the clock of each stage, or a 2-D array. */
typedef struct obme_pisces_config {
    int32_t precision;             /* OBME_FP64 | OBME_FP32_MIXED (FP64 calcite solve either way) */
    int32_t open_bottom;           /* POC sinking: 1 = -speed on every face, 0 = damped at the bottom */
    obme_grid grid;
    double params[OBME_PP_COUNT];
    double band_kw[3], band_e[3], band_chi[3], band_division[3];
    double surface_PAR;            /* W/m2 (OBME_SURFACE_CONSTANT) */
    int32_t surface_mode;          /* OBME_SURFACE_CONSTANT | _DIURNAL | _ARRAY */
    const void* surface_array;     /* OBME_SURFACE_ARRAY: Nx*Ny device values of the element
                                      type, i fastest */
    obme_carbon_solver solver;
    obme_negative_scaling scaling; /* PISCES(; scale_negatives = true): the five groups of
                                      PISCES/coupling_utils.jl:11-35 */
    obme_co2_exchange co2;         /* top DIC flux (CarbonDioxideGasExchangeBoundaryCondition, no Si/PO4) */
    obme_o2_exchange o2;           /* top O2 flux */
    obme_sediment sediment;        /* PISCES(; sediment): InstantRemineralisation or none; tracer
                                      indices in the OBME_PI_* order, sinking ones among POC, SFe,
                                      GOC, BFe, PSi, CaCO3 */
} obme_pisces_config;

typedef struct obme_pisces obme_pisces;

/* PISCES(; grid) (PISCES.jl:268-330) on the device. */
obme_status obme_pisces_create(const obme_pisces_config* cfg, obme_pisces** out);
obme_status obme_pisces_destroy(obme_pisces* h);

/* update_biogeochemical_state!(model, ::PISCES) with the MultiBand light
 * (update_state.jl:1-17; multi_band.jl:147-185; compute_euphotic_depth.jl:3-41;
 * mean_mixed_layer_properties.jl:25-47; compute_calcite_saturation.jl:9-37):
 * from tracers[OBME_PI_COUNT] at time t writes, when non-NULL, the 3-D PAR
 * (total) and PAR1..3 fields, the calcite saturation Omega (3-D), and the
 * 2-D euphotic depth and mixed-layer mean PAR (Nx*Ny, i fastest). */
obme_status obme_pisces_update_state(obme_pisces* h, const void* const* tracers, void* PAR, void* const* PARn,
                                     void* Omega, void* euphotic_depth, void* mixed_layer_PAR, double t,
                                     void* stream);
/* ADDS the 24 transitions minus the sinking flux divergence into G[0..23]
 * (the per-tracer `(bgc::PISCES)(i, j, k, grid, Val(name), ...)` methods plus
 * biogeochemical_drift_velocity of POC/SFe and GOC/BFe/PSi/CaCO3), after the
 * update-state step at time t. */
obme_status obme_pisces_add_tendencies(obme_pisces* h, const void* const* tracers, void* const* G, double t,
                                       void* stream);
/* nsteps RK3 steps (update-state, tendencies, substep per stage) in place
 * on tracers[0..23]; T and S (tracers[24], [25]) are read. */
obme_status obme_pisces_step(obme_pisces* h, void* const* tracers, int32_t nsteps, double dt, double t0,
                             void* stream);
/* obme_pisces_step with PISCES's InstantRemineralisation sediment
 * (cfg->sediment; PISCES.jl:233,344-353 takes a `sediment`,
 * instant_remineralisation.jl:60-123): storage is a device array of Nx*Ny
 * doubles (i fastest).  At every stage the bottom cell's sinking flux f of
 * the listed tracers (their bottom-face velocities; open_bottom for POC)
 * returns (1 - e) f / dz to the receiver's bottom cell, and after each
 * stage's substep the storage takes the sediment's step of the stage Δt
 * with the tendency e f (update_state.jl:6-17); the handle keeps the
 * sediment clock and each column's previous tendency.  The per-stage
 * kernels step a handle with a sediment.  obme_pisces_step (and
 * obme_pisces_add_tendencies) on such a handle apply the coupling and leave
 * the storage alone. */
obme_status obme_pisces_step_with_sediment(obme_pisces* h, void* const* tracers, double* storage, int32_t nsteps,
                                           double dt, double t0, void* stream);
/* update_biogeochemical_state!(model, sediment) for PISCES's sediment
 * (Sediments/update_state.jl:6-17), as obme_sediment_update_state: the
 * bottom-face flux of the listed tracers and, for a finite dt (the model's
 * last stage Δt), one QuasiAdamsBashforth2 step of the storage. */
obme_status obme_pisces_sediment_update_state(obme_pisces* h, const void* const* tracers, double* storage,
                                              double dt, void* stream);

/* ---- kernel timing ------------------------------------------------------ */

/* No counterpart in the reference: measurement instrumentation.  While
 * enabled, every launch of kernel `which` made by this handle is bracketed
 * by HIP events recorded on the caller's stream. */
enum { OBME_KERNEL_SWEEP = 0, OBME_KERNEL_SURFACE = 1, OBME_KERNEL_COUNT };
obme_status obme_set_kernel_timing(obme_handle* h, int32_t enable);
/* Waits for the recorded events of kernel `which`, returns their summed
 * duration in ms and the launch count since the previous call, and
 * releases them. */
obme_status obme_kernel_timing(obme_handle* h, int32_t which, double* total_ms, int64_t* launches);

/* Size in bytes of ABI struct `which` as compiled into the library
 * (0 obme_config, 1 obme_grid, 2 obme_npd_params, 3 obme_two_band_par,
 * 4 obme_sinking, 5 obme_carbon_solver, 6 obme_co2_exchange,
 * 7 obme_field_source, 8 obme_batch_params, 9 obme_pisces_config,
 * 10 obme_negative_scaling, 11 obme_o2_exchange, 12 obme_sediment,
 * 13 obme_oxygen); -1 for an
 * unknown index.  No reference
 * counterpart: it lets a binding check its struct layouts. */
int64_t obme_abi_sizeof(int32_t which);

#ifdef __cplusplus
}
#endif
#endif /* OBME_H */
