// fsi_internal.hpp -- host-side types of libfsi_hip shared by its C-ABI
// translation units (fsi_api.hip: context, state, workspace and the step;
// fsi_adjoint_api.hip: the discrete adjoint): the objects behind the opaque
// handles of fsi_api.h and the step pieces the adjoint reuses.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "fsi_kernels.hpp"

using namespace fsi;

namespace {

// the IMEX tableau (src/TimeIntegrationIMEX.jl:72-90)
const double AE[4][4] = {{0, 0, 0, 0},
                         {8.0 / 15.0, 0, 0, 0},
                         {1.0 / 4.0, 5.0 / 12.0, 0, 0},
                         {1.0 / 4.0, 0, 3.0 / 4.0, 0}};
const double AI[4][4] = {{0, 0, 0, 0},
                         {4.0 / 15.0, 4.0 / 15.0, 0, 0},
                         {4.0 / 15.0, 1.0 / 3.0, 1.0 / 15.0, 0},
                         {4.0 / 15.0, 1.0 / 3.0, 7.0 / 30.0, 1.0 / 6.0}};

constexpr int kMaxLevels = 24;

}  // namespace

struct fsi_context {
    fsi_grid grid{};
    fsi_bcs bcs{};
    fsi_models models{};
    fsi_numerics num{};
    Params prm{};
    BcSet bpsi{};
    int P = 0;
    size_t felems = 0;
    int has_null = 0;
    int has_reset = 0;
    hipStream_t stream = nullptr;
    // The phase and sigma kernels of stage q run on `side` while `stream`
    // solves stage q + 1 (they share no buffer: the stage-2 velocity has a
    // buffer of its own); events join the two before each stage's tendencies.
    hipStream_t side = nullptr;
    hipEvent_t ev_start = nullptr, ev_proj[3] = {}, ev_cells[3] = {}, ev_x0[3] = {};
    int dual = 0;
    std::string err;
    int device = -1;  // the HIP device current at creation
    double* force_row = nullptr;
    // argument tables of the batched ensemble steps this context leads
    // (pinned host staging and its device copy)
    char* batch_host = nullptr;
    char* batch_dev = nullptr;
    size_t batch_cap = 0;
    double* uw_smooth = nullptr;  // UWCDI :smoothed9: the filtered phi of the stage being evaluated
    double* cap_buf = nullptr;    // DiffuseAreaCapillarity: force x, force y, then 5 work fields
    double* ext = nullptr;  // per-block extrema of the report and checks (EX_* slots)
    double* diag = nullptr;      // per-block partials of the driver diagnostics
    double* diag_out = nullptr;  // their reduced values
    double* lap_bc = nullptr;
    double* bvisc_u = nullptr;
    double* bvisc_v = nullptr;
    double* partials = nullptr;
    Status* st = nullptr;       // device
    Status* st_host = nullptr;  // pinned mirror
    std::vector<MgGeom> geo;
    CoarsePlan cplan{};         // single-workgroup coarse solve of geo.back()
    double* co_dense = nullptr;  // its dense bottom inverse
    float* co_dense_f = nullptr; // the same in FP32 (mixed precision)
    // Direct solve of the projection system by fast transforms (power-of-two
    // axes with Neumann/Neumann, Dirichlet/Dirichlet or periodic sides): it
    // is the PCG's initial guess, which then usually meets the tolerance at
    // iteration 0.  FSI_SPECTRAL=0 keeps the multigrid-PCG solve alone.
    int use_spectral = 1;
    SpecPlan spec{};
    double* spec_tab = nullptr;  // device tables of both axes
    double* spec_part = nullptr;  // partials of the fused transform passes (inside spec_tab)
    int last_vel_iters[3] = {8, 8, 8};
    int last_proj_iters[3] = {6, 6, 6};
    // Planned steps: each solve enqueued with a planned iteration count and no
    // read-back until the end of the step (see fsi_integrate_step).  Mode 2
    // launches the kernels directly, mode 1 replays the step as a captured
    // HIP graph (measured slower on ROCm 7 for the ~460-launch step), mode 0
    // polls each solve.
    int use_graph = 2;
    int vel_warm = 1;  // predictor warm starts from the stage histories (FSI_STAGE_WARM=0: from u^n)
    int have_history = 0;  // iteration counts of a completed step are known
    int plan_vel[3] = {0, 0, 0};   // planned iterations (mixed precision: inner iterations per cycle)
    int plan_cyc[3] = {0, 0, 0};   // mixed precision: planned refinement cycles
    int last_imax[3] = {8, 8, 8};  // mixed precision: most inner iterations of a cycle in the last solve
    int vel_check_from[3] = {0, 0, 0};  // the fewest iterations of the same history
    int plan_proj[3] = {0, 0, 0};
    // iteration counts of the last kHist completed steps per stage (the plan
    // is their maximum plus a margin, so alternating counts keep one plan)
    static constexpr int kHist = 4;
    int hist_vel[3][kHist] = {};
    int hist_cyc[3][kHist] = {};   // mixed precision: cycles and most inner iterations of a cycle
    int hist_imax[3][kHist] = {};
    int hist_proj[3][kHist] = {};
    int hist_n = 0;
    long long spec_steps = 0, spec_misses = 0;
    // growth and resource path: any growth, diffusion or consumption closure
    // active (TimeIntegrationIMEX.jl:1004-1009) selects the growth step
    int growth = 0;
    GrowthP gp{};
};

struct fsi_state {
    fsi_context* ctx = nullptr;
    double *U = nullptr, *V = nullptr, *phi = nullptr, *res = nullptr;
    double *sxx = nullptr, *syy = nullptr, *sxy = nullptr;
    double res_min = INFINITY;
    long long res_nonfinite = 0;
};

struct fsi_workspace {
    fsi_context* ctx = nullptr;
    double *Us = nullptr, *Vs = nullptr, *phis = nullptr, *sxxs = nullptr, *syys = nullptr, *sxys = nullptr;
    double *Eu[3] = {}, *Ev[3] = {}, *Iu[3] = {}, *Iv[3] = {};
    double *Ep[3] = {}, *Exx[3] = {}, *Eyy[3] = {}, *Exy[3] = {};
    double *ru = nullptr, *rv = nullptr, *r_u[2] = {}, *r_v[2] = {};
    // velocity right-hand sides of stages 2..4, summed by the momentum kernels
    double *rhs_u[3] = {}, *rhs_v[3] = {};
    // phi and sigma of stages 2 and 3, summed by the sigma kernels (stage 4
    // goes to phis, sxxs, syys, sxys, which the state swaps with)
    double *cphi[2] = {}, *csxx[2] = {}, *csyy[2] = {}, *csxy[2] = {};
    double *p_u[2] = {}, *p_v[2] = {}, *q_u[2] = {}, *q_v[2] = {};
    // mixed precision: the FP32 r, p, q of the inner CG (null in FP64 mode)
    float *fr_u[2] = {}, *fr_v[2] = {}, *fp_u[2] = {}, *fp_v[2] = {}, *fq_u[2] = {}, *fq_v[2] = {};
    // `correction` (work.correction, :828-833) is corr[corr_cur].  A step or
    // a projection writes its correction into corr[1 - corr_cur] and flips
    // corr_cur only when it has succeeded, so a failing call leaves the
    // correction as it was (project! throws before copyto!, Projection.jl:279-298).
    double *rhs = nullptr, *corr[2] = {}, *r = nullptr, *z = nullptr, *pp[2] = {}, *w = nullptr;
    int corr_cur = 0;
    // Warm starts of the three stage projections of a step: stage s starts
    // from stage s of the previous step (closer than the previous stage's
    // correction), extrapolated as 2 q^{n-1} - q^{n-2}.  The step solves into
    // cq_next; the three buffers of a stage trade roles only after a
    // successful step, so a failed or redone step starts from the same
    // history.  Only iteration counts depend on it.
    double *cq_cur[3] = {}, *cq_prev[3] = {}, *cq_next[3] = {};
    double* xs = nullptr;  // the iterate of a standalone projection (fsi_project)
    double *Us2 = nullptr, *Vs2 = nullptr;  // the stage-2 velocity (stages 1 and 3 use Us, Vs)
    double *mun_u = nullptr, *mun_v = nullptr;  // M u^n of the step (the reference residual of the predictor)
    // the same for the predictor solves: stage s of the previous two steps
    // velocity stage solutions of the last four steps (newest first) and a
    // fifth buffer that the solve writes into (the projection then writes
    // the projected stage velocity to Us, leaving the solution in place);
    // the five trade roles once a step has succeeded
    double *vel_cur_u[3] = {}, *vel_cur_v[3] = {}, *vel_prev_u[3] = {}, *vel_prev_v[3] = {};
    double *vel_old_u[3] = {}, *vel_old_v[3] = {}, *vel_old2_u[3] = {}, *vel_old2_v[3] = {};
    double *vel_next_u[3] = {}, *vel_next_v[3] = {};
    int stage_warm = 0;         // completed steps whose stage solutions the history buffers hold (0 .. 3)
    std::vector<MgLevel> lv;
    std::vector<double*> owned;
    // captured steps, keyed by everything their kernel arguments depend on:
    // the plan, dt, the history depth (it changes the launch sequence) and
    // every buffer the step reads or writes whose role rotates between steps
    struct StepGraph {
        std::vector<uintptr_t> key;
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> graphs;
    // the discrete adjoint's tape: when set, the step copies its stage-2
    // velocity (the first projected stage, overwritten by the last) here
    double* tape_U1 = nullptr;
    double* tape_V1 = nullptr;
    // growth step: G_s, the resource tendencies E_c, I_c of stages 1..3, the
    // face diffusivities of phi_s, the resource right-hand side, the resource
    // stage value (swapped into the state with the other fields), the CG
    // vectors of the resource solve and its partials
    double *G = nullptr, *Ec[3] = {}, *Ic[3] = {}, *Dx = nullptr, *Dy = nullptr, *crhs = nullptr, *cs = nullptr;
    double *rc_r = nullptr, *rc_p = nullptr, *rc_q = nullptr, *gpart = nullptr;
    const double* proj_target = nullptr;  // the target divergence of the projections in flight (null: G = 0)
};


#define HIPCHECK(c, expr)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail((c), FSI_E_HIP, "HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                               \
    } while (0)

// helpers of fsi_api.hip
fsi_status fail(fsi_context* c, fsi_status code, const char* fmt, ...);
hipError_t dalloc(fsi_context* c, double** p, size_t n);
BcSet to_bcset(const fsi_bc_setup& s);
fsi_status upload2d(fsi_context* c, double* dst, const double* src, int width, int height);
fsi_status download2d(fsi_context* c, double* dst, const double* src, int width, int height);
void status_reset(fsi_context* c);
fsi_status status_push(fsi_context* c);
fsi_status status_pull(fsi_context* c);
fsi_status device_error(fsi_context* c);
// The predictor solve of stage `slot` from the right-hand side (ru, rv) into
// (xu, xv); fixed > 0: exactly that many iterations without read-back
fsi_status velocity_solve(fsi_context* c, fsi_state* s, fsi_workspace* w, double alpha, int slot, int fixed,
                          const double* ru, const double* rv, double* xu, double* xv, bool x0_formed = false);
// Projects (Uin, Vin) into (U, V) (the same arrays for a projection in place)
fsi_status projection(fsi_context* c, fsi_workspace* w, double* Uin, double* Vin, double* U, double* V, int inject,
                      int slot, int fixed);
