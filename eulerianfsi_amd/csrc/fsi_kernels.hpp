// fsi_kernels.hpp -- kernel launchers shared by the physics, solver and API
// translation units of libfsi_hip.
#pragma once

#include <vector>

#include "fsi_device.hpp"

namespace fsi {

// One level of the projection multigrid hierarchy (cell-centred, uniform).
struct MgLevel {
    int nx, ny, P;
    double hx, hy;
    double* x;  // smoothed iterate / correction (ping)
    double* t;  // pong
    double* b;  // right-hand side
    double* r;  // residual
    float *xf, *tf, *bf;  // FP32 x, t, b of the mixed-precision V-cycle (null in FP64 mode)
};

// Level geometry passed to the multigrid kernels.
struct MgGeom {
    int nx, ny, P;
    int px, py;
    int dl, dr, db, dt;  // 1 where the side is Dirichlet (else Neumann/periodic)
    double ihx2, ihy2;
    double om1, om2;     // weights of the two Jacobi-preconditioned smoothing sweeps
    double d0, id0;      // interior diagonal of -A (2/hx^2 + 2/hy^2) and 1 / d0
};

// Coarse levels solved in one workgroup (fsi_mg.hip): a top level of at most
// coarse_max_cells() cells (1024 by default) and its halvings, with a bottom
// of at most CO_DENSE_MAX cells solved by a dense (pseudo-)inverse.
constexpr int CO_DENSE_MAX = 64;
struct CoarsePlan {
    int levels;             // levels in the workgroup, top included
    int dense_n;            // cells of the bottom when solved densely, else 0
    size_t lds_cells;       // dynamic LDS of the launch, in values
    const double* dense;    // device copy of the dense inverse (dense_n^2)
    const float* dense_f;   // the same in FP32 (mixed-precision V-cycle)
};

// ---- physics (fsi_physics.hip) ----
// Running sums of the velocity right-hand sides of the later stages
// (_build_velocity_stage!, TimeIntegrationIMEX.jl:778-800): the momentum
// kernel of stage q adds dt (aE[s][q] E_q + aI[s][q] I_q) to the sum of every
// stage s > q (starting from u^n at q = 0, a term skipped when both
// coefficients are zero) and, at s = q + 1, the viscous lifting
// alpha_s b_visc; the sum of stage s is then that stage's right-hand side,
// its terms added in the same order as a separate combination would.
struct RhsSums {
    double* ru[3];  // stages s = 1..3 at index s - 1
    double* rv[3];
    double ce[3], ci[3];  // aE[s][q], aI[s][q] at index s - 1 (s > q)
    double alpha;         // dt aI[q+1][q+1]
    double dt;
    const double* bu;  // b_visc, or null without viscosity
    const double* bv;
    double* mu;  // q = 0: M u^n (the viscous operator without its lifting) stored here, or null
    double* mv;
    int q;  // the stage of the tendencies (0..2); -1: no sums
};
// E_u, E_v, I_u, I_v may be null (not stored)
void launch_momentum(hipStream_t s, const Params& p, const double* U, const double* V, const double* phi,
                     const double* sxx, const double* syy, const double* sxy, int reset, double* Eu, double* Ev,
                     double* Iu, double* Iv, const RhsSums* sums = nullptr);
// Running sums of the explicit stage values of phi and sigma
// (_build_explicit_stage!, TimeIntegrationIMEX.jl:716-740): the sigma kernel
// of stage q adds dt aE[s][q] (E_phi, E_sigma) of stage q to the sum of every
// later stage s (from phi^n and sigma^n with the safety reset at q = 0, a
// term skipped when its coefficient is zero); at s = q + 1 the sum is final:
// the safety reset where phi_s < cutoff and, for s = 3, the StepReport
// partials.  The terms are added in the order a separate combination would.
struct CellSums {
    double* phi[3];  // stages s = 1..3 at index s - 1
    double* sxx[3];
    double* syy[3];
    double* sxy[3];
    const double* Ephi;  // E_phi of stage q (written by the phase kernel)
    double cf[3];        // dt aE[s][q] at index s - 1 (s > q)
    int reset;           // the safety reset of the stage values
    int q;               // the stage of the tendencies (0..2); -1: no sums
};
// with `sums`, E_sigma is not stored (Exx, Eyy, Exy may be null)
void launch_cells(hipStream_t s, const Params& p, const BcSet& bpsi, const double* U, const double* V,
                  const double* phi, const double* sxx, const double* syy, const double* sxy, int reset,
                  double* Ephi, double* Exx, double* Eyy, double* Exy, ErrSlot* err, int stage1,
                  const CellSums* sums = nullptr);
void launch_visc_boundary(hipStream_t s, const Params& p, const double* zero, double* bu, double* bv);
// uwcdi_directional_rhs (PhaseFieldMethods.jl:768-791) of phi: dir 1 the
// x-fluxes, 2 the y-fluxes; then phi_out = phi + h rhs (the Euler sweep
// :793-810).  rhs and phi_out may not alias phi.
void launch_uwcdi_sweep(hipStream_t s, const Params& p, const double* U, const double* V, const double* phi, int dir,
                        double h, double* rhs, double* phi_out, ErrSlot* err);
// StepReport extrema and counts, and the stage-1 checks, from the per-block
// partials of the step's last combination and of its stage-1 tendencies
void launch_ext_final(hipStream_t s, const Params& p, Status* st);

// ---- driver diagnostics (fsi_physics.hip) ----
// Each launcher runs a per-block reduction into buf (rows of p.exnb partials)
// and a one-block final reduction into out: the sums, then the maxima (a
// minimum negated), in the order documented at each kernel.
struct DiagRegion {
    int all;           // every cell
    double lo, hi;     // else lo < phi < hi, each bound inclusive when *_in
    int lo_in, hi_in;
};
constexpr int DIAG_STATE_VALUES = 11 + 6, DIAG_TENSOR_VALUES = 6 + 6, DIAG_GEOMETRY_VALUES = 4 + 2 * 8 + 1;
void launch_diag_state(hipStream_t s, const Params& p, const double* U, const double* V, const double* phi,
                       double lx, double ly, double density, double* buf, double* out);
void launch_diag_tensor(hipStream_t s, const Params& p, const double* phi, const double* sxx, const double* syy,
                        const double* sxy, double G, const DiagRegion& rg, double* buf, double* out);
void launch_diag_geometry(hipStream_t s, const Params& p, const double* phi, double lx, double ly, double xc,
                          double yc, int n0, int nmax, double* buf, double* out);
int diag_modes_per_launch();
// (x, y, phase integral) of the centroid from the reduced values of
// launch_diag_state, the centroid NaN when the integral is not above eps
void launch_centroid_result(hipStream_t s, const double* reduced, double* result);
void launch_status_reset(hipStream_t s, Status* st);

// ---- velocity predictor CG (fsi_solvers.hip) ----
struct VelBufs {
    const double* Un;  // warm start: uv_n (state velocity) or the previous step's stage solution
    const double* Vn;
    const double* Pu;  // the stage solution of the step before
    const double* Pv;
    const double* Ou;  // and of the step before that
    const double* Ov;
    const double* Wu;  // and of the one before
    const double* Wv;
    int extrap;        // 0: x0 = Un; 1: x0 = 2 Un - Pu; 2: x0 = 3 (Un - Pu) + Ou;
                       // 3: x0 = 4 (Un + Ou) - 6 Pu - Wu
    int x0_in_place;   // x0 (with its Dirichlet values) already in xu, xv: the start only reads it
    const double* Su;  // the state velocity u^n and M u^n (the stage-1 momentum kernel's):
    const double* Sv;  //   the gmres rule's reference residual ||b - A u^n||, formed
    const double* Mu;  //   on the own cells (null: not needed, the :direct method)
    const double* Mv;
    const double* ru;  // right-hand side (stage sum + alpha*b)
    const double* rv;
    double* xu;  // solution, written into the stage velocity arrays
    double* xv;
    double* r_u[2];  // residual (ping-pong for the fused iteration)
    double* r_v[2];
    double* p_u[2];
    double* p_v[2];
    double* q_u[2];  // q = A p
    double* q_v[2];
    double* partials;
    // mixed precision: FP32 r, p, q of the inner CG (null in FP64 mode); x
    // stays FP64 and accumulates the inner updates
    float *fr_u[2], *fr_v[2], *fp_u[2], *fp_v[2], *fq_u[2], *fq_v[2];
    int restart;  // k_vel_start: r = b - A x of the current x, for the FP64 check of a cycle
};
// x0 of a predictor solve formed ahead of it into b.xu, b.xv (the history
// of b.Un .. b.Wv extrapolated as b.extrap says, the Dirichlet values set)
void launch_vel_predict(hipStream_t s, const Params& p, const VelBufs& b);
// the start of a predictor solve (x0, r0, q0 and the stopping thresholds of `tol`)
void launch_vel_init(hipStream_t s, const Params& p, const VelBufs& b, double alpha, Status* st, int slot,
                     const VelTol& tol);
// kernels m0 .. m0+count-1; those with m >= max(check_from, 1) test the
// previous iterate before loading any field
void launch_vel_iters(hipStream_t s, const Params& p, const VelBufs& b, double alpha, Status* st, int m0, int count,
                      int slot, const VelTol& tol, int check_from);
void launch_vel_copy(hipStream_t s, const Params& p, const VelBufs& b);
// mixed precision (FP32 inner CG inside FP64 iterative refinement):
// FP32 CG iterations m0 .. m0+count-1 of the current cycle
void launch_vel_iters_f32(hipStream_t s, const Params& p, const VelBufs& b, double alpha, Status* st, int m0,
                          int count, int slot, const VelTol& tol);
// the end of a cycle: r = b - A x in FP64 (FP32 copy for the next cycle) and
// the outer stopping rule on it; `next`: the next cycle's q0 as well
void launch_vel_refine(hipStream_t s, const Params& p, const VelBufs& b, double alpha, Status* st, int slot,
                       const VelTol& tol, int next);

// ---- projection (fsi_solvers.hip) ----
struct ProjBufs {
    double* U;  // the projected velocity (state or stage arrays)
    double* V;
    double* Uin;  // the velocity to project (== U, V when projected in place); its
    double* Vin;  // injected boundary faces are written here
    int inject;            // write the injected boundary faces first (_inject_compact_predictor!)
    const double* lap_bc;  // b_bc of the Laplacian
    double* rhs;
    double* x;      // PCG iterate (warm start in, solution out)
    double* q_out;  // mean-free correction written here
    double* r;      // PCG residual = right-hand side of the V-cycle
    double* z;      // preconditioned residual = V-cycle result
    double* p[2];
    double* w;
    double* partials;
    // With the direct solve the right-hand side is formed by its first
    // transform pass and the mean of x summed by its last: their per-workgroup
    // partials (rhs: 4 rows of rhs_np, x: one row of q_np), or null (the
    // per-tile partial sets of the kernels below)
    double* rhs_part;
    int rhs_np;
    double* q_part;
    int q_np;
    const double* G;  // target divergence (cells), or null for G = 0
};
void launch_proj_rhs(hipStream_t s, const Params& p, const ProjBufs& b);
// x = 2 cur - prev (extrap) or x = cur: the initial guess of a projection
void launch_extrapolate(hipStream_t s, const Params& p, double* x, const double* cur, const double* prev,
                        int extrap);
void launch_pcg_init(hipStream_t s, const Params& p, const MgGeom& g0, const ProjBufs& b, Status* st, int has_null,
                     double rtol, int maxiter);
void launch_pcg_iters(hipStream_t s, const Params& p, const ProjBufs& b, const MgLevel* lv, const MgGeom* geo,
                      int nlev, const CoarsePlan& cp, Status* st, int m0, int count, int slot);
void launch_proj_finish(hipStream_t s, const Params& p, const MgGeom& g0, const ProjBufs& b, Status* st,
                        int has_null, double lxly, int slot);
// The whole projection after the right-hand side when the warm start x is
// only tested (PCG iteration 0; a miss is left to the caller): one pass over
// x for the convergence test, the true residual and the mean, then the
// velocity update, its constraint check and the final reduction (which makes
// the convergence decision).
void launch_proj_tested(hipStream_t s, const Params& p, const MgGeom& g0, const ProjBufs& b, Status* st,
                        int has_null, double rtol, int maxiter, double lxly, int slot);
void launch_pcg_check(hipStream_t s, Status* st, const double* rr, int nb, int m, int slot);
// the convergence test of PCG iteration m alone (the first step of that
// iteration), on the r.r partials of the previous one
void launch_pcg_test(hipStream_t s, const Params& p, const ProjBufs& b, Status* st, int m, int slot);
void launch_rz(hipStream_t s, const Params& p, const ProjBufs& b, double* rz, int nb, const int* done);

int grid_blocks(const Params& p);

// ---- fused multigrid levels (fsi_mg.hip) ----
// The finest down-kernel (st != nullptr) first sums the r.r partials of the
// previous PCG kernel and ends the solve when converged; the finest up-kernel
// (rz != nullptr) writes the r.z block partials.
// FP64 levels, and the mixed-precision variants (FP32 levels; the finest
// level reads the FP64 residual and writes the FP64 z).
void launch_mg_down(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const double* b, double* x2, double* bc,
                    const int* done, Status* st, const double* rr, int nb, int m, int slot);
void launch_mg_down(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const double* b, float* x2, float* bc,
                    const int* done, Status* st, const double* rr, int nb, int m, int slot);
void launch_mg_down(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const float* b, float* x2, float* bc,
                    const int* done, Status* st, const double* rr, int nb, int m, int slot);
void launch_mg_up(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const double* x2, const double* b,
                  const double* ec, double* xout, const int* done, double* rz, int nb);
void launch_mg_up(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const float* x2, const double* b,
                  const float* ec, double* xout, const int* done, double* rz, int nb);
void launch_mg_up(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const float* x2, const float* b,
                  const float* ec, float* xout, const int* done, double* rz, int nb);
// Plan of the single-workgroup coarse solve for a top level (false when the
// level is too large); `dense` receives the bottom's inverse, row-major.
bool coarse_plan(const MgGeom& top, int max_cells, CoarsePlan& cp, std::vector<double>& dense);
int coarse_max_cells();
void launch_mg_coarse(hipStream_t s, const MgGeom& top, const CoarsePlan& cp, const double* b, double* t,
                      const int* done);
void launch_mg_coarse(hipStream_t s, const MgGeom& top, const CoarsePlan& cp, const float* b, float* t,
                      const int* done);
// ---- direct Poisson solve by fast real transforms (fsi_spectral.hip) ----
// Per axis the 5-point operator -A is diagonalised by
//   SPEC_COS  (Neumann on both sides)   the DCT-II basis cos(pi k (i + 1/2) / n)
//   SPEC_ALT  (Dirichlet on both sides) the same basis applied to (-1)^i q
//   SPEC_RDFT (periodic)                the real DFT in half-complex order
enum SpecKind { SPEC_NONE = -1, SPEC_COS = 0, SPEC_ALT = 1, SPEC_RDFT = 2 };
constexpr int SPEC_MIN_N = 8, SPEC_MAX_N = 2048;
struct SpecAxis {
    int n, m, kind;        // length 2^m (r3 = 0) or 3 * 2^m (r3 = 1), SpecKind
    int r3;
    const double2* wfft;   // e^{-2 pi i t / n}, t < n/2
    const double2* wpost;  // e^{-i pi k / (2n)}, k < n (cosine kinds)
    const double* mu;      // eigenvalue of the 1-D operator per coefficient index
};
struct SpecPlan {
    int enabled;
    SpecAxis x, y;
};
// The kind of one axis from its geometry (SPEC_NONE if not diagonalisable
// here: mixed Neumann/Dirichlet sides, or a length outside [SPEC_MIN_N,
// SPEC_MAX_N] that is neither 2^m nor 3 * 2^m).
int spec_kind(int n, int periodic, int dlo, int dhi);
// Host tables of one axis, packed as wfft (n/2 complex), wpost (n complex),
// mu (n); h2inv = 1 / h^2.
std::vector<double> spec_tables(int n, int kind, double h2inv);
// x = solution of (-A) x = -rhs with the (0, 0) mode set to zero when A is
// singular (the zero-mean gauge); rhs and x are cell fields of pitch P.
void launch_spectral_solve(hipStream_t s, const SpecPlan& sp, int P, const double* rhs, double* x);
// The projection's right-hand side and direct solve in the three transform
// passes: the first forms rhs = div(v*) - b_bc from b.Uin, b.Vin (injecting
// the boundary faces when b.inject) with the compatibility partials into
// b.rhs_part, the last adds the partials of the area-weighted sum of x into
// b.q_part; the solution goes to b.x.
void launch_spectral_proj(hipStream_t s, const SpecPlan& sp, const Params& p, const ProjBufs& b);
// partials (rows of (ny + 1) / 2 doubles) that launch_spectral_proj writes
constexpr int SPEC_PART_ROWS = 5;
// kernel attributes the plan's launches need (LDS above 64 KiB); false if the
// device cannot run them
bool spec_prepare(const SpecPlan& sp);

size_t partials_size(const Params& p);
double bench_launch(hipStream_t s, const char* name, const Params& p, const VelBufs& vb, const ProjBufs& pb,
                    const MgLevel* lv, const MgGeom* geo, int nlev, const CoarsePlan& cp, Status* st,
                    const SpecPlan* sp = nullptr);

// ---- growth and resource path (fsi_growth.hip) ----
// Closure parameters: G(phi, c), K_g = G (kxx, kxy; kxy, kyy), the diffusivity
// and the consumption law, the resource boundary conditions.
struct GrowthP {
    int src;           // fsi_growth_source
    double rate, K;
    double kxx, kxy, kyy;
    int diff;          // fsi_diffusivity_kind
    double Df, Ds;
    int cons;          // fsi_consumption_kind
    double crate, cK;
    int visc;          // the growth viscous stress -2 mu K_g is added
    double mu;
    BcSet bC;          // resource
};
// G of (phi, c) into G; the trace error (first: overwrite, else max with the
// earlier stages); stage1: max |div v - G|, max |G| and the non-finite count
// of the input state; the sum of G^2 dA into st->proj.tgt_l2sq
struct GrSrc {
    const double *phi, *c, *U, *V, *sxx, *syy, *sxy;
    double* G;
    double* part;
    int first, stage1;
};
void launch_gr_source(hipStream_t s, const Params& p, const GrowthP& g, const GrSrc& a, Status* st);
// stage values: phi_s, sigma_s (safety reset), the resource right-hand side and
// the velocity right-hand side (+ alpha b_visc); mu, mv: M u^n = I_0 - b_visc
struct GrStage {
    const double *phin, *sxxn, *syyn, *sxyn, *cn, *Un, *Vn;
    const double *Ep[3], *Exx[3], *Eyy[3], *Exy[3], *Ec[3], *Ic[3], *Eu[3], *Ev[3], *Iu[3], *Iv[3];
    double ce[3], ci[3];
    double dt, alpha;
    const double *bu, *bv;
    double *phi, *sxx, *syy, *sxy, *crhs, *ru, *rv, *mu, *mv;
    int ns, reset_n, reset;
};
void launch_gr_stage(hipStream_t s, const Params& p, const GrStage& a);
// face diffusivities of phi (and the diffusivity's validation)
void launch_gr_diff(hipStream_t s, const Params& p, const GrowthP& g, const double* phi, double* Dx, double* Dy,
                    Status* st);
// the resource stage solve (I - alpha L) x = b + alpha b_L by CG (x0 = c_n)
struct ResCg {
    const double *Dx, *Dy, *b, *x0;
    double *x, *r, *p, *q, *part;
    double alpha;
    double rtol, fb, maxiter_d;
    int method, fallback, maxiter;
};
void launch_rcg_start(hipStream_t s, const Params& p, const GrowthP& g, const ResCg& a, Status* st);
void launch_rcg_iters(hipStream_t s, const Params& p, const GrowthP& g, const ResCg& a, Status* st, int m0,
                      int count);
// growth terms added to the stored tendencies of a stage, and the resource
// tendencies E_c, I_c
struct GrTend {
    const double *U, *V, *phi, *sxx, *syy, *sxy, *G, *c, *Dx, *Dy;
    int reset;
    double *Eu, *Ev, *Ep, *Exx, *Eyy, *Exy, *Ec, *Ic;
};
void launch_gr_tend(hipStream_t s, const Params& p, const GrowthP& g, const GrTend& a, Status* st);
// the StepReport partials of the new state (EX_* slots)
void launch_gr_report(hipStream_t s, const Params& p, const double* phi, const double* sxx, const double* syy,
                      const double* sxy, const double* c);

// ---- discrete adjoint (fsi_adjoint.hip) ----
// The linearisation point: one stage's fields (sigma after the safety reset,
// or with `reset` the reset applied on load, as the first stage reads sigma^n)
struct AdjStage {
    const double *U, *V, *phi, *sxx, *syy, *sxy;
    int reset;
};
// adjoints of the stage's tendencies E_u, E_v, I_u, I_v, E_phi, E_sigma (null: zero)
struct AdjBar {
    const double *Eu, *Ev, *Iu, *Iv, *Ephi, *Exx, *Eyy, *Exy;
};
// adjoints of the stage's fields, accumulated (+=); k_adj_primal writes its
// tendencies here instead (E_u, E_v in U, V; I_u, I_v in aux_u, aux_v)
struct AdjGrads {
    double *U, *V, *phi, *sxx, *syy, *sxy;
    double *aux_u, *aux_v;
};
int adj_block_count(const Params& p);
void launch_adj_primal(hipStream_t s, const Params& p, const BcSet& bpsi, const AdjStage& st, const AdjGrads& out);
// out += (d tendencies / d stage fields)^T b; *dG += b . d tendencies / d G
// (partials: adj_block_count(p) values)
void launch_adj_tendencies(hipStream_t s, const Params& p, const BcSet& bpsi, const AdjStage& st, const AdjBar& b,
                           const AdjGrads& out, double* partials, double* dG);
void launch_adj_axpy(hipStream_t s, size_t n, double* y, double a, const double* x);
// zero each of up to 48 arrays of n values
struct AdjPtrs {
    double* p[48];
    int n;
};
void launch_adj_zero(hipStream_t s, size_t n, const AdjPtrs& z);
// dst[d] += coef[d] * src[from[d]] (d < ndst <= 16, nsrc <= 4); sources with
// bit q of `masked` read as zero where phi < cutoff (phi null: never)
struct AdjCombine {
    const double* src[4];
    int nsrc;
    double* dst[16];
    double coef[16];
    int from[16];
    int ndst;
    const double* phi;
    double cutoff;
    int masked;
};
void launch_adj_combine(hipStream_t s, size_t n, const AdjCombine& c);
void launch_adj_mask(hipStream_t s, const Params& p, const double* phi, double* gxx, double* gyy, double* gxy);
void launch_adj_walls(hipStream_t s, const Params& p, double* U, double* V);
void launch_adj_affine(hipStream_t s, const Params& p, double* phi, double A, double B, double C);

}  // namespace fsi
