/*
 * fsi_api.h -- C ABI of libfsi_hip.so, the MI355X (gfx950) backend of the
 * EulerianFSI.jl four-stage IMEX fractional step.
 *
 * The entry points mirror the reference's Julia step API one to one:
 *
 *   build_imex_context(metrics, bcs, models; numerics)  src/TimeIntegrationIMEX.jl:468-493
 *   allocate_imex_workspace(metrics)                     src/TimeIntegrationIMEX.jl:389-441
 *   make_velocity_compatible!(state, ctx, work)          src/TimeIntegrationIMEX.jl:1088-1109
 *   integrate_step!(state, t, dt, ctx, work)            src/TimeIntegrationIMEX.jl:1152-1256
 *   project!(U, V, target, system, metrics; ...)        src/Projection.jl:170-307
 *
 * Array convention (identical to the reference's Julia matrices, so a Julia
 * caller passes pointer(A) with no transposition):
 *   column-major, x index fastest, 0-based here;
 *   U   : (nx+1) x ny   x-face velocity,  element (i,j) at U[i + j*(nx+1)]
 *   V   : nx x (ny+1)   y-face velocity,  element (i,j) at V[i + j*nx]
 *   cell: nx x ny       phi, resource, sigma_xx, sigma_yy, sigma_xy
 * All values are IEEE double.  The state is resident in device memory
 * between calls; host copies happen only in fsi_state_upload/_download.
 *
 * Error model: every call returns an fsi_status.  The codes map to the
 * exceptions the reference throws (ArgumentError, DimensionMismatch,
 * ErrorException, DomainError); fsi_last_error() returns the message of the
 * last failing call on a context.  A context, its states and workspaces use
 * one HIP stream and are not safe for concurrent use.
 */
#ifndef FSI_API_H
#define FSI_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
typedef enum {
    FSI_OK = 0,
    FSI_E_ARGUMENT = 1,    /* ArgumentError    (invalid value, failed stage-1/compatibility check) */
    FSI_E_DIMENSION = 2,   /* DimensionMismatch                                           */
    FSI_E_SOLVER = 3,      /* ErrorException   (projection residual / constraint check)   */
    FSI_E_DOMAIN = 4,      /* DomainError      (non-SPD sigma under geometric extension)  */
    FSI_E_UNSUPPORTED = 5, /* configuration outside the device path; caller uses the CPU path */
    FSI_E_HIP = 6          /* device runtime error                                        */
} fsi_status;

/* ---- boundary conditions (src/BoundaryConditions.jl:26-72) -------------- */
typedef enum {
    FSI_BC_PERIODIC = 0,
    FSI_BC_DIRICHLET = 1,
    FSI_BC_NEUMANN = 2
} fsi_bc_kind;

typedef struct {
    int32_t kind;   /* fsi_bc_kind                                   */
    double value;   /* scalar datum (Dirichlet value / outward normal derivative) */
} fsi_bc_side;

typedef struct {
    fsi_bc_side left, right, bottom, top;
} fsi_bc_setup;

/* FSIBoundaryConditions (src/TimeIntegrationIMEX.jl:127-147) */
typedef struct {
    fsi_bc_setup U, V, correction, phi, resource, sigma_xx, sigma_yy, sigma_xy;
} fsi_bcs;

/* ---- uniform MAC grid (src/MeshGeneration.jl:179-233, :uniform stretching) */
typedef struct {
    int32_t nx, ny;
    double lx, ly;
    int32_t periodic_x, periodic_y;
} fsi_grid;

/* ---- closures (FSIModels, src/TimeIntegrationIMEX.jl:156-208) ----------- */
typedef enum {
    FSI_PHASE_UPWIND = 0, /* LegacySplitPhaseFieldMethod(NoInterfaceFlux())  PhaseFieldMethods.jl:408-435 */
    FSI_PHASE_CDI = 1,    /* CurrentCDIPhaseField                            PhaseFieldMethods.jl:457-475 */
    FSI_PHASE_ACDI = 2,   /* JainACDIPhaseField                              PhaseFieldMethods.jl:652-671 */
    FSI_PHASE_UWCDI = 3   /* PirozzoliUWCDIPhaseField(:standard)             PhaseFieldMethods.jl:852-867 */
} fsi_phase_method;

typedef enum { FSI_VISC_NONE = 0, FSI_VISC_CONSTANT = 1 } fsi_viscosity_kind;
typedef enum { FSI_ELASTIC_NONE = 0, FSI_ELASTIC_LINEAR = 1 } fsi_elastic_kind;
typedef enum { FSI_RELAX_NONE = 0, FSI_RELAX_MAXWELL = 1 } fsi_relaxation_kind;
typedef enum {
    FSI_FORCE_NONE = 0,
    FSI_FORCE_CONSTANT = 1,  /* ConstantBodyForce(fx, fy)                        */
    FSI_FORCE_KOLMOGOROV = 2 /* KolmogorovBodyForce(amplitude, wavenumber; phase) */
} fsi_body_force_kind;

typedef enum {
    FSI_UWCDI_NORMAL_STANDARD = 0,  /* :standard  centred gradient of phi                 */
    FSI_UWCDI_NORMAL_YOUNGS = 1,    /* :youngs    Youngs' 3x3 stencil off the walls       */
    FSI_UWCDI_NORMAL_SMOOTHED9 = 2  /* :smoothed9 centred gradient of the 9-point-filtered phi */
} fsi_uwcdi_normal;
typedef enum { FSI_CAPILLARITY_NONE = 0, FSI_CAPILLARITY_DIFFUSE_AREA = 1 } fsi_capillarity_kind;

/* growth and resource closures (src/GrowthModels.jl, src/TransportConstitutive.jl:181-335) */
typedef enum {
    FSI_GROWTH_NONE = 0,        /* NoGrowth: G = 0                                        GrowthModels.jl:214-223 */
    FSI_GROWTH_CONSTANT = 1,    /* ConstantIntrinsicGrowth(rate): G = rate phi            :225-235 */
    FSI_GROWTH_SATURATING = 2   /* SaturatingResourceGrowth(rate, K): G = rate phi c/(K+c) :237-252 */
} fsi_growth_source;
typedef enum {
    FSI_GROWTH_TENSOR_ISOTROPIC = 0,   /* IsotropicGrowthTensor: K_g = G/2 I                  :268-289 */
    FSI_GROWTH_TENSOR_DIRECTIONAL = 1  /* DirectionalGrowthTensor(tx, ty): K_g = G t (x) t     :291-318 */
} fsi_growth_tensor;
typedef enum {
    FSI_DIFFUSION_NONE = 0,      /* NoDiffusion: the resource is not diffused (no implicit solve) */
    FSI_DIFFUSION_CONSTANT = 1,  /* ConstantDiffusivity(value): diffusivity_fluid          */
    FSI_DIFFUSION_HARMONIC = 2   /* HarmonicPhaseDiffusivity(D_f, D_s): 1/((1-phi)/D_f + phi/D_s) */
} fsi_diffusivity_kind;
typedef enum {
    FSI_CONSUMPTION_NONE = 0,       /* NoResourceConsumption                                  */
    FSI_CONSUMPTION_SATURATING = 1  /* SaturatingResourceConsumption(rate, K): phi rate c/(K+c) */
} fsi_consumption_kind;

typedef struct {
    int32_t phase_method;      /* fsi_phase_method */
    /* CDI: ConservativeDiffuseInterfaceFlux(inverse_peclet, epsilon; delta_normal) */
    double cdi_inverse_peclet, cdi_epsilon, cdi_delta_normal;
    /* ACDI: JainACDIPhaseField(gamma, epsilon; psi_floor, delta_normal, dynamic_gamma) */
    double acdi_gamma, acdi_epsilon, acdi_psi_floor, acdi_delta_normal;
    int32_t acdi_dynamic_gamma;
    /* UWCDI: PirozzoliUWCDIPhaseField(alpha; delta_normal) */
    double uwcdi_alpha, uwcdi_delta_normal;

    int32_t viscosity_kind;    double viscosity;       /* ConstantViscosity(mu)          */
    int32_t elastic_kind;      double shear_modulus;   /* LinearInternalTensorStress(G)  */
    int32_t relaxation_kind;   double relaxation_rate; /* MaxwellRelaxation(rate)        */
    int32_t body_force_kind;
    double force_fx, force_fy;                         /* ConstantBodyForce              */
    double force_amplitude, force_wavenumber, force_phase; /* KolmogorovBodyForce        */
    /* UWCDI normal_reconstruction (PhaseFieldMethods.jl:96-120, :246-404): fsi_uwcdi_normal */
    int32_t uwcdi_normal_method;
    /* capillarity (Capillarity.jl:21-41): fsi_capillarity_kind;
     * DiffuseAreaCapillarity(coefficient; delta_normal, smoothing_passes) */
    int32_t capillarity_kind;
    double capillary_coefficient, capillary_delta_normal;
    int32_t capillary_smoothing_passes;
    /* growth and resource (FSIModels growth_source, growth_tensor, diffusivity,
     * resource_consumption, TimeIntegrationIMEX.jl:156-208).  Any of them
     * active selects the growth step: G_s and K_g,s from phi_s and c_s, the
     * projection onto div v = G_s, the S v convection term, phi S, the K_g terms
     * of the structure tensor, the growth viscous stress -2 mu K_g, and the
     * resource stage (upwind advection, sink, implicit diffusion solve). */
    int32_t growth_source_kind;     /* fsi_growth_source */
    double growth_rate, growth_half_saturation;
    int32_t growth_tensor_kind;     /* fsi_growth_tensor */
    double growth_direction_x, growth_direction_y;  /* normalised by the library, as the constructor does */
    int32_t diffusivity_kind;       /* fsi_diffusivity_kind */
    double diffusivity_fluid, diffusivity_structure;  /* ConstantDiffusivity: diffusivity_fluid */
    int32_t consumption_kind;       /* fsi_consumption_kind */
    double consumption_rate, consumption_half_saturation;
} fsi_models;

/* ---- numerics (IMEXNumerics, src/TimeIntegrationIMEX.jl:262-301) -------- */
typedef enum {
    FSI_EXT_NONE = 0,      /* NoStructureExtension                        */
    FSI_EXT_IDENTITY = 1,  /* IdentityOutsideStructure(cutoff)            */
    FSI_EXT_GEOMETRIC = 2  /* GeometricStructureExtension(rate, exponent; safety_cutoff) */
} fsi_extension_kind;

/* FSI_PRECISION_MIXED: the velocity predictor by FP64 iterative refinement
 * around an FP32 CG (FP32 r, p, q; FP64 solution, residual checks and dot
 * products; the stopping rules unchanged), and the multigrid V-cycle of the
 * projection PCG in FP32.  State, tendencies, stage sums, the direct
 * projection solve and every reduction stay FP64. */
typedef enum { FSI_PRECISION_FP64 = 0, FSI_PRECISION_MIXED = 1 } fsi_precision;

/* LinearSolverOptions.velocity_method (src/TimeIntegrationIMEX.jl:211-252) */
typedef enum {
    FSI_SOLVER_DIRECT = 0, /* :direct (the reference default): the exact solve A \ b; on the device
                              CG to a relative residual of 1e-14 (the system is SPD with a
                              condition number 1 + O(dt mu / h^2))                          */
    FSI_SOLVER_GMRES = 1   /* :gmres: stop at ||b - A x|| <= reltol ||b - A uv_n|| (x0 = uv_n, :807),
                              at most maxiter iterations; then accepted when the relative
                              residual is <= fallback_reltol, else the direct solve
                              (direct_fallback) or ErrorException (:602-641)               */
} fsi_solver_method;

typedef struct {
    int32_t extension_kind;
    double ext_max_rate, ext_exponent, ext_cutoff;
    /* velocity predictor solve (LinearSolverOptions.reltol / maxiter) */
    double velocity_rtol;
    int32_t velocity_maxiter;
    /* projection (projection_reltol / projection_maxiter) */
    double projection_rtol;
    int32_t projection_maxiter;
    /* stage-1 divergence check (stage1_constraint_atol/rtol, check_stage1_constraint) */
    double stage1_atol, stage1_rtol;
    int32_t check_stage1;
    int32_t precision;         /* fsi_precision */
    /* LinearSolverOptions.velocity_method / direct_fallback / fallback_reltol */
    int32_t velocity_method;   /* fsi_solver_method */
    int32_t direct_fallback;
    double fallback_rtol;
    /* LinearSolverOptions.resource_method (the resource stage solve, :742-774):
     * fsi_solver_method with the same reltol, maxiter and fallback rules */
    int32_t resource_method;
} fsi_numerics;

/* ---- results ------------------------------------------------------------ */
/* StepReport, field order of src/TimeIntegrationIMEX.jl:318-327 */
typedef struct {
    double maximum_projection_constraint_linf;
    double maximum_growth_trace_error;
    double minimum_phi;
    double maximum_phi;
    double minimum_resource;
    double minimum_sigma_eigenvalue;
    double maximum_abs_sigma_determinant_error;
    int64_t non_spd_cells;
} fsi_step_report;

/* ProjectionResult scalars, src/Projection.jl:31-37 (the correction field
 * itself stays in the workspace; fsi_workspace_download_correction reads it) */
typedef struct {
    double linear_relative_residual;
    double constraint_linf;
    double constraint_l2;
    double compatibility_defect;
    int32_t iterations;
} fsi_projection_result;

/* Per-step solver counters, for measurement (no reference counterpart). */
typedef struct {
    int32_t velocity_iterations[3];   /* stages 2..4 */
    int32_t projection_iterations[3]; /* stages 2..4 */
    int32_t planned;                  /* 1: the step ran with planned iteration counts and one read-back */
    /* max |U| over the x-faces and max |V| over the y-faces of the new
     * state (NaN if any face is NaN), as the driver's timestep bound reads
     * them (scripts/benchmarks/run_sugiyama_lid_driven_cavity.jl:862-866);
     * formed by the step's last projection check, so they cost no launch */
    double max_abs_u;
    double max_abs_v;
    /* per predictor solve (stages 2..4): the final residual |b - A x| / max(|b|, eps)
     * and the initial residual from uv_n, |b - A uv_n| / max(|b|, eps), which
     * the :gmres rule scales reltol by (0 for :direct); 0 without viscosity */
    double velocity_residual[3];
    double velocity_reference_residual[3];
    /* FSI_PRECISION_MIXED: FP64 refinement cycles of each predictor solve
     * (velocity_iterations then counts its FP32 inner CG iterations); 0 in FP64 */
    int32_t velocity_cycles[3];
} fsi_step_stats;

typedef struct fsi_context fsi_context;
typedef struct fsi_state fsi_state;
typedef struct fsi_workspace fsi_workspace;

/* ---- entry points --------------------------------------------------------- */

/* build_imex_context, src/TimeIntegrationIMEX.jl:468-493 (periodicity
 * validation :447-466, projection system + multigrid hierarchy
 * src/Projection.jl:149-166).  Returns FSI_E_UNSUPPORTED for closures the
 * device path does not cover. */
fsi_status fsi_context_create(const fsi_grid* grid, const fsi_bcs* bcs,
                              const fsi_models* models, const fsi_numerics* numerics,
                              fsi_context** out);
void fsi_context_destroy(fsi_context* ctx);
const char* fsi_last_error(const fsi_context* ctx);

/* FSIState, src/TimeIntegrationIMEX.jl:103-118: device-resident 7 fields. */
fsi_status fsi_state_create(fsi_context* ctx, fsi_state** out);
void fsi_state_destroy(fsi_state* st);
/* Copy host arrays (column-major, shapes above) to/from the device.  Any
 * pointer may be NULL on download to skip that field. */
fsi_status fsi_state_upload(fsi_context* ctx, fsi_state* st,
                            const double* U, const double* V, const double* phi,
                            const double* resource, const double* sigma_xx,
                            const double* sigma_yy, const double* sigma_xy);
fsi_status fsi_state_download(fsi_context* ctx, const fsi_state* st,
                              double* U, double* V, double* phi, double* resource,
                              double* sigma_xx, double* sigma_yy, double* sigma_xy);

/* IMEXWorkspace, src/TimeIntegrationIMEX.jl:333-441: stage buffers,
 * histories, solver scratch and the projection warm start `correction`. */
fsi_status fsi_workspace_create(fsi_context* ctx, fsi_workspace** out);
void fsi_workspace_destroy(fsi_workspace* ws);
fsi_status fsi_workspace_download_correction(fsi_context* ctx, const fsi_workspace* ws,
                                             double* correction);
fsi_status fsi_workspace_upload_correction(fsi_context* ctx, fsi_workspace* ws,
                                           const double* correction);

/* make_velocity_compatible!, src/TimeIntegrationIMEX.jl:1088-1109 */
fsi_status fsi_make_velocity_compatible(fsi_context* ctx, fsi_state* st, fsi_workspace* ws,
                                        fsi_projection_result* result);

/* integrate_step!, src/TimeIntegrationIMEX.jl:1152-1256.  On any error the
 * state is left unchanged. `stats` may be NULL. */
fsi_status fsi_integrate_step(fsi_context* ctx, fsi_state* st, fsi_workspace* ws,
                              double t, double dt, fsi_step_report* report,
                              fsi_step_stats* stats);

/* Batched ensemble: integrate_step! (src/TimeIntegrationIMEX.jl:1152-1256)
 * applied to n independent instances, instance k being (ctx[k], st[k],
 * ws[k], t[k], dt[k]) with pairwise distinct contexts.  The reference steps
 * instances one after another from its drivers; this entry point has no
 * single counterpart there: it enqueues every instance's planned step on its
 * context's stream before waiting for any, so small instances run
 * concurrently on the device, the launches shared among up to
 * FSI_ENSEMBLE_THREADS (default 8) host threads.  Each instance's result equals that of
 * fsi_integrate_step on it; status[k] receives its status (the message via
 * fsi_last_error(ctx[k])), and the first non-OK status is returned.
 * `reports` and `stats` (n entries each) may be NULL. */
fsi_status fsi_ensemble_step(int32_t n, fsi_context* const* ctx, fsi_state* const* st,
                             fsi_workspace* const* ws, const double* t, const double* dt,
                             fsi_step_report* reports, fsi_step_stats* stats, int32_t* status);

/* project! on the state's velocity with target divergence G = 0 (NoGrowth),
 * src/Projection.jl:170-307, warm-started from the workspace correction. */
fsi_status fsi_project(fsi_context* ctx, fsi_state* st, fsi_workspace* ws,
                       fsi_projection_result* result);

/* Stage tendencies of the state, as _evaluate_stage_tendencies! computes
 * them for one stage (src/TimeIntegrationIMEX.jl:944-1071), for kernel
 * parity: E_u (nx+1)x ny and E_v nx x (ny+1) are the explicit momentum
 * terms (-material convection + body force + elastic force,
 * :962-988), I_u / I_v the viscous operator applied to the state velocity
 * (:1057-1061) on the compact faces (other faces are written as 0),
 * E_phi the phase_field_rhs (:991-999), E_sxx/E_syy/E_sxy the
 * structure_tensor_rhs (:1032-1052).  Output pointers may be NULL. */
fsi_status fsi_eval_tendencies(fsi_context* ctx, const fsi_state* st, fsi_workspace* ws,
                               double t,
                               double* E_u, double* E_v, double* I_u, double* I_v,
                               double* E_phi, double* E_sxx, double* E_syy, double* E_sxy);

/* uwcdi_strang_step (src/PhaseFieldMethods.jl:793-850): the Strang-split
 * directional UWCDI update of the state's phi by the state's velocity,
 * Euler sweeps x (dt/2), y (dt), x (dt/2) of uwcdi_directional_rhs
 * (:768-791); first_direction 0 = :x, 1 = :y (y, x, y).  phi is replaced by
 * the result (the reference returns it as a new array).  FSI_E_ARGUMENT when
 * the context's phase field is not UWCDI, dt is negative or not finite, or
 * first_direction is neither.  Uses the workspace's stage E_phi buffers. */
fsi_status fsi_uwcdi_strang_step(fsi_context* ctx, fsi_state* st, fsi_workspace* ws, double dt,
                                 int32_t first_direction);

/* ---- driver diagnostics ---------------------------------------------------
 * The per-step and per-interval reductions of the reference's cavity driver
 * (scripts/benchmarks/run_sugiyama_lid_driven_cavity.jl, "DIAGNOSTICS" and
 * "TIMESTEP SELECTION", :483-934), computed on the device-resident state so
 * that a driver loop never copies fields to the host.  Uniform grids with the
 * origin at (0, 0): x_p[i] = lx (i + 1/2) / nx (MeshGeneration.jl:215).
 * Sums are reduced in a fixed order (deterministic); a NaN seen by an
 * extremum makes that extremum NaN, as Julia's minimum/maximum do. */

/* max |U| over every x-face and max |V| over every y-face of the state: the
 * data comparison_timestep_bounds reads before every step (:862-866). */
fsi_status fsi_velocity_extrema(fsi_context* ctx, const fsi_state* st, double* max_abs_u,
                                double* max_abs_v);

/* phase_centroid (:496-509).  FSI_E_SOLVER (the reference's error()) when
 * the phase integral is not above eps(Float64). */
fsi_status fsi_phase_centroid(fsi_context* ctx, const fsi_state* st, double* x_centroid,
                              double* y_centroid, double* phase_integral);

/* The phase centroid of fsi_phase_centroid computed without allocating
 * anything: no counterpart in the reference.  For callers that cannot
 * allocate device memory or synchronise (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The reductions use the
 * library's static device scratch instead of the context's buffers, and the
 * call enqueues its two kernels on `stream` (a hipStream_t; NULL: the
 * context's stream) and returns: `result` (device memory, 3 doubles) receives
 * x_centroid, y_centroid, phase_integral once the stream reaches it, the
 * centroid NaN where fsi_phase_centroid reports FSI_E_SOLVER.  Grids of at
 * most FSI_NOALLOC_MAX_TILES tiles of 64 x 16 cells (4096 x 4096, or any
 * ceil(nx/64) * ceil(ny/16) <= 16384); larger ones give FSI_E_ARGUMENT.  The
 * scratch (about 2.2 MB per GPU, held until the process ends) is shared by
 * every call on a GPU: calls that overlap on one GPU race. */
#define FSI_NOALLOC_MAX_TILES 16384
fsi_status fsi_phase_centroid_noalloc(fsi_context* ctx, const fsi_state* st, double* result, void* stream);

#define FSI_DIAG_MAX_MODES 32

/* BenchmarkConfig fields the interval diagnostics read (:40-128). */
typedef struct {
    double density;                 /* flow_diagnostics kinetic energy */
    double elastic_modulus;         /* tensor_cell_diagnostics_2d stress */
    double sigma_core_cutoff;       /* tensor_diagnostics_detailed regions :824-849 */
    double sigma_interface_cutoff;
    double sigma_safety_cutoff;
    double h_interface;             /* effective_interface_cells = width / h_interface */
    int32_t deformation_mode_max;   /* nmax of the Fourier modes, 0..FSI_DIAG_MAX_MODES */
} fsi_diag_params;

/* tensor_region_diagnostics (:753-822), fields in TENSOR_REGION_FIELDS
 * order (:1069-1081); an empty region has NaN extrema and means. */
typedef struct {
    int64_t cells;
    double area;
    double determinant_min;
    double determinant_max;
    double max_abs_determinant_error;
    double mean_abs_determinant_error;
    double lambda_min;
    double lambda_max;
    int64_t non_spd_cells;
    double max_effective_stress;
    double mean_effective_stress;
} fsi_tensor_region;

typedef struct {
    /* phase_geometry_diagnostics (:632-707) */
    double x_centroid;
    double y_centroid;
    double phase_integral;
    double perimeter_proxy;
    double mixed_layer_measure;
    double effective_interface_width;
    double effective_interface_cells;
    double covariance_major;
    double covariance_minor;
    double aspect_ratio;
    double orientation_radians;
    double R0;
    int32_t nmax;
    double Rc[FSI_DIAG_MAX_MODES];
    double Rs[FSI_DIAG_MAX_MODES];
    double amplitudes[FSI_DIAG_MAX_MODES];
    /* phase_bound_diagnostics (:709-722) */
    double phi_min;
    double phi_max;
    double lower_violation;
    double upper_violation;
    int64_t out_of_bounds_cells;
    /* flow_diagnostics (:724-751) */
    double maximum_speed;
    double kinetic_energy;
    double divergence_linf;
    double divergence_l2;
    /* the velocity extrema of fsi_velocity_extrema */
    double max_abs_u;
    double max_abs_v;
    /* tensor_diagnostics_detailed (:824-849): every cell; phi >= core;
     * interface < phi < core; safety < phi <= interface */
    fsi_tensor_region global_region;
    fsi_tensor_region core;
    fsi_tensor_region interface_region;
    fsi_tensor_region tail;
} fsi_diagnostics;

/* Every interval diagnostic of the driver in one call (two passes over the
 * state, the second about the centroid).  FSI_E_ARGUMENT for parameters the
 * driver's validation rejects (:205-236) or nmax outside 0..32; FSI_E_SOLVER
 * when the phase integral cannot define a centroid. */
fsi_status fsi_diagnostics_compute(fsi_context* ctx, const fsi_state* st, const fsi_diag_params* params,
                                   fsi_diagnostics* out);

/* Measurement hook; no counterpart in the reference (whose code is not
 * instrumented).  Launches the named kernel `reps` times on the context's
 * stream with the buffers of the workspace (after at least one step), timed
 * with HIP events on that stream; returns the mean duration of one launch in
 * milliseconds and the kernel's algorithmic HBM bytes per launch.  Names:
 * "copy" (calibration), "vel_iter", "vel_iter_f32" (the FP32 inner CG
 * iteration of a mixed-precision context), "spec_solve" (the projection's direct
 * solve, three kernels), "pcg_dir", "pcg_upd", "mg_down0", "mg_up0",
 * "mg_coarse", "pcg_iter", "mg_jacobi", "momentum", "cells". */
fsi_status fsi_benchmark_kernel(fsi_context* ctx, fsi_state* st, fsi_workspace* ws, const char* name,
                                int32_t reps, double* avg_ms, double* bytes_per_launch);

/* Library identification: device name, architecture and build. */
const char* fsi_version(void);

/* ---- discrete adjoint ----------------------------------------------------
 * The reverse-mode derivative of integrate_step! (src/TimeIntegrationIMEX.jl
 * :1152-1256) and of a run of steps, for the reference's differentiation
 * use (gradients of an objective of the final state with respect to the
 * initial state and the shear modulus, LinearInternalTensorStress(G),
 * StructuralConstitutive.jl:100-113).  The reference has no hand-written
 * adjoint: a Julia caller would differentiate integrate_step! with an AD
 * package through ChainRulesCore; these entry points are what its rrule
 * binds.  Covered: walled domains (no periodic side),
 * ACDI with a fixed gamma, FP64, the closures of the step otherwise;
 * FSI_E_UNSUPPORTED elsewhere.  The step is differentiated along the
 * branches it took (upwind sides, the psi clamp, the safety reset: frozen
 * switches) with its linear solves taken as exact, so a run meant for
 * derivatives uses tight solver settings (velocity_method :direct).
 * Cotangents are fsi_state objects of the same context (their resource
 * field is ignored). */
typedef struct fsi_adjoint fsi_adjoint;
typedef struct {
    int64_t steps;          /* steps of the last run                               */
    int64_t advances;       /* steps recomputed from checkpoints                   */
    int64_t recordings;     /* steps run with their stage fields taped             */
    int64_t adjoint_steps;  /* transposed steps                                    */
    int32_t snapshots_used; /* most checkpoints held at once                       */
    int32_t repetitions;    /* the schedule's repetition number                    */
    double device_bytes;    /* device memory the adjoint object holds              */
    int64_t replanned;      /* transposed steps whose planned solves fell short and ran again polled */
    int32_t last_shortfall; /* which: bit k velocity solve of stage k + 2, bit 3 + k its projection */
    int32_t solve_iterations[3]; /* iterations of the last transposed predictor solves */
} fsi_adjoint_stats;

/* The adjoint of runs on (ctx, ws), holding up to `snapshots` checkpoints of
 * the state (binomial checkpointing, Griewank & Walther's revolve; the
 * recomputed steps run without the velocity warm-start history, so they
 * reproduce the recorded ones exactly).  snapshots = 0 keeps instead every
 * step's stage fields (19 fields per step, allocated by the forward run):
 * nothing is recomputed. */
fsi_status fsi_adjoint_create(fsi_context* ctx, fsi_workspace* ws, int32_t snapshots, fsi_adjoint** out);
void fsi_adjoint_destroy(fsi_adjoint* adj);
/* The forward half of the rrule of n_steps integrate_step! calls from t0 by
 * dt: advances st to the final state, keeping checkpoints. */
fsi_status fsi_adjoint_forward(fsi_adjoint* adj, fsi_state* st, double t0, double dt, int32_t n_steps);
/* The pullback: `bar` holds the cotangent of the final state on entry and
 * that of the initial state on return; *d_shear_modulus receives dJ/dG.
 * Once per fsi_adjoint_forward. */
fsi_status fsi_adjoint_backward(fsi_adjoint* adj, fsi_state* bar, double* d_shear_modulus);
fsi_status fsi_adjoint_get_stats(const fsi_adjoint* adj, fsi_adjoint_stats* out);
/* One step: fsi_adjoint_record_step runs integrate_step! on st (advancing
 * it) and tapes it; fsi_adjoint_step maps the cotangent of its output
 * (bar, in place) to that of its input. */
fsi_status fsi_adjoint_record_step(fsi_adjoint* adj, fsi_state* st, double t, double dt);
fsi_status fsi_adjoint_step(fsi_adjoint* adj, fsi_state* bar, double* d_shear_modulus);
/* J = |x_c - (tx, ty)|^2 of the phase centroid (phase_centroid,
 * scripts/benchmarks/run_sugiyama_lid_driven_cavity.jl:496-509) and its
 * cotangent with respect to the state, written to bar. */
fsi_status fsi_adjoint_centroid_objective(fsi_context* ctx, const fsi_state* st, double tx, double ty,
                                          fsi_state* bar, double* J);
/* Kernel parity.  fsi_adjoint_tendencies: the transposed linearisation of
 * the stage tendencies of fsi_eval_tendencies at the state; bar_in[8] =
 * cotangents of E_u, E_v, I_u, I_v, E_phi, E_sxx, E_syy, E_sxy (host arrays
 * in their shapes, NULL = zero), grad_out[6] = cotangents of U, V, phi,
 * sigma_xx, sigma_yy, sigma_xy (NULL = skipped), *d_shear_modulus the
 * shear-modulus part.  fsi_adjoint_primal_tendencies: the same tendencies as
 * fsi_eval_tendencies from the functions the adjoint differentiates
 * (out[8] in the order of bar_in). */
fsi_status fsi_adjoint_tendencies(fsi_context* ctx, const fsi_state* st, const double* const* bar_in,
                                  double* const* grad_out, double* d_shear_modulus);
fsi_status fsi_adjoint_primal_tendencies(fsi_context* ctx, const fsi_state* st, double* const* out);
/* The checkpoint schedule of a run (host only, no device): forward steps
 * recomputed, most snapshots held, repetition number, and optionally the
 * actions (kind, step, slot, count; kinds 0 advance, 1 store, 2 restore,
 * 3 record, 4 adjoint, 5 free). */
fsi_status fsi_adjoint_plan(int32_t n_steps, int32_t snapshots, int64_t* advances, int32_t* most_held,
                            int32_t* repetitions, int32_t* actions, int64_t max_actions, int64_t* n_actions);

#ifdef __cplusplus
}
#endif

#endif /* FSI_API_H */
