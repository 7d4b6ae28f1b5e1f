/*
 * fsi_oracle.h -- CPU restatement of EulerianFSI.jl's integrate_step! path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP backend is
 * compared against and the CPU baseline that is timed beside it.  Nothing in
 * the product (eulerianfsi_amd/, libfsi_hip.so) links, loads or calls it.
 *
 * Arrays are host double arrays in the layout of include/fsi_api.h
 * (column-major, x fastest).  Compact velocity vectors follow
 * _extract_compact_velocity! (src/TimeIntegrationIMEX.jl:529-541):
 * uv[k] = U[i,j], uv[N+k] = V[i,j], k = i + j*nx, N = nx*ny.
 */
#ifndef FSI_ORACLE_H
#define FSI_ORACLE_H

#include "../include/fsi_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fsio_ctx fsio_ctx;
/* scalar arguments (times, steps, shifts): double in every build of the
 * oracle, the tangent build (fsi_tangent.cpp) included */
typedef double fsio_scalar;

int fsio_create(const fsi_grid* grid, const fsi_bcs* bcs, const fsi_models* models,
                const fsi_numerics* numerics, fsio_ctx** out);
void fsio_destroy(fsio_ctx* c);
const char* fsio_last_error(const fsio_ctx* c);

/* integrate_step!: the state arrays are updated in place; `correction` is the
 * workspace warm start (nx*ny), updated as the reference updates it. */
int fsio_integrate_step(fsio_ctx* c, double* U, double* V, double* phi, double* resource,
                        double* sxx, double* syy, double* sxy, double* correction,
                        fsio_scalar t, fsio_scalar dt, fsi_step_report* rep, fsi_step_stats* stats);
int fsio_make_velocity_compatible(fsio_ctx* c, double* U, double* V, const double* phi,
                                  const double* resource, double* correction,
                                  fsi_projection_result* res);
/* project! with target divergence G = 0. */
int fsio_project(fsio_ctx* c, double* U, double* V, double* correction,
                 fsi_projection_result* res);
int fsio_eval_tendencies(fsio_ctx* c, const double* U, const double* V, const double* phi, /* sigma after the safety reset */
                         const double* sxx, const double* syy, const double* sxy, fsio_scalar t,
                         double* E_u, double* E_v, double* I_u, double* I_v, double* E_phi,
                         double* E_sxx, double* E_syy, double* E_sxy);

/* Operator-level entry points (for closed-form and invariant checks). */
void fsio_divergence(fsio_ctx* c, const double* U, const double* V, double* out);
/* which: 0 = correction BCs, 1 = phi BCs */
void fsio_scalar_gradient(fsio_ctx* c, int which, const double* q, double* gx, double* gy);
int fsio_velocity_gradient(fsio_ctx* c, const double* U, const double* V, double* dudx,
                           double* dudy, double* dvdx, double* dvdy);
void fsio_stress_divergence(fsio_ctx* c, const double* Txx, const double* Tyy,
                            const double* Txy, double* Fx, double* Fy);
void fsio_material_convection(fsio_ctx* c, const double* U, const double* V, const double* S,
                              double* Cu, double* Cv);
int fsio_phase_rhs(fsio_ctx* c, const double* U, const double* V, const double* phi,
                   const double* S, double* out);
int fsio_structure_rhs(fsio_ctx* c, const double* sxx, const double* syy, const double* sxy,
                       const double* U, const double* V, const double* phi, double* oxx,
                       double* oyy, double* oxy);
int fsio_structure_extension_rhs(fsio_ctx* c, const double* sxx, const double* syy,
                                 const double* sxy, const double* phi, int n, double* oxx,
                                 double* oyy, double* oxy);
/* apply_operator(viscous, uv) = M uv + b (with_boundary=1) or M uv (0); compact 2N. */
void fsio_viscous_apply(fsio_ctx* c, const double* uv, double* out, int with_boundary);
/* Unit-coefficient projection Laplacian A q (+ b_bc when with_boundary). */
void fsio_laplacian_apply(fsio_ctx* c, const double* q, double* out, int with_boundary);
/* Solve (I - alpha M) uv = rhs + alpha b with Dirichlet rows, as
 * _build_velocity_stage! (src/TimeIntegrationIMEX.jl:794-812).  `rhs` is the
 * stage sum before the alpha*b term.  Returns CG iterations in *iters. */
int fsio_velocity_solve(fsio_ctx* c, fsio_scalar alpha, const double* rhs, const double* x0,
                        double* x, int* iters);
/* uwcdi_strang_step (PhaseFieldMethods.jl:793-850): phi replaced by the
 * Strang-split directional step; first_direction 0 = :x, 1 = :y. */
int fsio_uwcdi_strang_step(fsio_ctx* c, const double* U, const double* V, double* phi, fsio_scalar dt,
                           int first_direction);
/* The UWCDI centre normals of the context's normal_reconstruction, and the
 * 9-point filter they and the capillary geometry use. */
int fsio_uwcdi_center_normals(fsio_ctx* c, const double* phi, double* nxc, double* nyc);
void fsio_smooth9(fsio_ctx* c, const double* f, double* out);
/* capillary_force of the context's capillarity model (zero for NoCapillarity). */
void fsio_capillary_force(fsio_ctx* c, const double* phi, double* fx, double* fy);
/* Metric vectors: which 0..15 in MeshMetrics field order. */
const double* fsio_metric(const fsio_ctx* c, int which, int* len);
/* growth and resource closures: G of (phi, c) with the growth tensor's trace
 * error as the return value; -div(upwind c v) - sink; L q (+ b) of the
 * diffusion operator of phi; the resource stage solve (I - alpha L) x = rhs + alpha b */
double fsio_growth_source(fsio_ctx* c, const double* phi, const double* res, double* G);
int fsio_resource_explicit_rhs(fsio_ctx* c, const double* U, const double* V, const double* phi,
                               const double* res, double* out);
int fsio_eval_tendencies_growth(fsio_ctx* c, const double* U, const double* V, const double* phi,
                                const double* sxx, const double* syy, const double* sxy, const double* res,
                                const double* G, double* E_u, double* E_v);
int fsio_diffusion_apply(fsio_ctx* c, const double* phi, const double* q, double* out, int with_boundary);
int fsio_resource_solve(fsio_ctx* c, fsio_scalar alpha, const double* phi, const double* rhs, const double* x0,
                        double* x, int* iters);

#ifdef __cplusplus
}
#endif
#endif
