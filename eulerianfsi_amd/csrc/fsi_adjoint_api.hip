// fsi_adjoint_api.hip -- C ABI of the discrete adjoint of integrate_step!
// (src/TimeIntegrationIMEX.jl:1152-1256) and of a run of steps with binomial
// checkpointing.
//
// One step, as the device runs it (stages s = 1..3 after the stage-1
// tendencies; G = 0, NoGrowth):
//   phi_s, sigma_s = reset(phi^n + dt sum_q aE[s][q] E_phi,q, sigma^n~ + dt sum_q aE[s][q] E_sigma,q)
//   rhs_s = u^n + dt sum_q (aE[s][q] E_u,q + aI[s][q] I_q)
//   u_s   = P K_s^-1 (rhs_s + alpha_s b),  K_s = I - alpha_s M,  alpha_s = dt aI[s][s]
// with the tendencies of stage q evaluated at (u_q, phi_q, sigma_q) and
// sigma^n~ the safety-reset sigma^n.  Its transpose runs the stages
// backwards: the cotangent of u_s goes through the projection P and the
// predictor solve K_s^-1 -- both symmetric (P = I - G A^+ D with G = -D^T on
// the MAC grid; M is the symmetric viscous block of the free faces), so the
// transposed solves are the forward solves themselves, with zero velocity
// data on the walls: they run on a twin context whose velocity boundary
// values are zero -- and then into u^n and the stage tendencies' cotangents
// by the tableau's coefficients; phi_s, sigma_s likewise (the reset cells
// are constants).  The cotangents of the tendencies of stage q go back to
// (u_q, phi_q, sigma_q) and to G through the transposed linearisation
// (fsi_adjoint.hip).
//
// The linearisation points come from a tape: the step recomputed with its
// stage fields copied out (the stage-2 velocity as the step runs, the rest
// after it).  A run of n steps keeps states at binomial checkpoints (Griewank
// and Walther's revolve schedule) and recomputes the rest; with the velocity
// warm-start history off (x0 = u^n) and the projection solved directly, a
// step is a function of the state alone, so every recomputation reproduces
// the recorded trajectory bit for bit and a checkpoint is the six evolving
// fields (the resource is constant under NoGrowth).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "fsi_internal.hpp"

namespace {

// ---- the binomial checkpointing schedule (revolve) --------------------------
enum ActKind : int32_t { ACT_ADVANCE = 0, ACT_STORE = 1, ACT_RESTORE = 2, ACT_RECORD = 3, ACT_ADJOINT = 4,
                         ACT_FREE = 5 };
struct Act {
    int32_t kind;
    int32_t step;   // ADVANCE: the first step; RECORD / ADJOINT: the step; STORE / RESTORE / FREE: the slot's step
    int32_t slot;
    int32_t count;  // ADVANCE: steps
};

// beta(c, r) = C(c + r, c): the most steps c checkpoints (the one at the
// start included) reverse when no step is advanced more than r times, and
// T(l, c) = r l - beta(c + 1, r - 1) for beta(c, r - 1) < l <= beta(c, r):
// the fewest forward steps that reverse l steps with c checkpoints
// (Griewank, Walther, ACM TOMS 26 (2000), 19-45)
double beta(int c, int r) {
    if (r < 0) return 0.0;
    const int k = std::min(c, r);  // C(c + r, c) = C(c + r, r)
    double b = 1.0;
    for (int q = 1; q <= k; ++q) b = b * (double)(c + r - k + q) / (double)q;
    return b;
}
int repetitions(int l, int c) {
    int r = 0;
    while (beta(c, r) < (double)l) ++r;
    return r;
}
double fewest_steps(int l, int c) {
    if (l <= 1) return 0.0;
    if (c <= 1) return 0.5 * (double)l * (double)(l - 1);
    const int r = repetitions(l, c);
    return (double)r * (double)l - beta(c + 1, r - 1);
}

struct Planner {
    std::vector<Act> acts;
    std::vector<int> slot_step;  // step held by each slot, -1 free
    int cur = 0;                  // step of the working state
    int held = 0, most_held = 0;
    int rep = 0;                  // the largest repetition number used

    int take_slot() {
        for (size_t k = 0; k < slot_step.size(); ++k)
            if (slot_step[k] < 0) return (int)k;
        return -1;
    }
    // bring the working state to step k, from the checkpoint `sa` at step a
    void go(int k, int a, int sa) {
        if (!(cur >= a && cur <= k)) {
            acts.push_back({ACT_RESTORE, a, sa, 0});
            cur = a;
        }
        if (k > cur) acts.push_back({ACT_ADVANCE, cur, -1, k - cur});
        cur = k;
    }
    void record(int k) {
        acts.push_back({ACT_RECORD, k, -1, 1});
        acts.push_back({ACT_ADJOINT, k, -1, 0});
        cur = k + 1;
    }
    // reverse steps [a, b) with the state of step a in slot sa and `f` free slots
    void gen(int a, int b, int sa, int f) {
        const int l = b - a;
        if (l == 1) {
            go(a, a, sa);
            record(a);
            return;
        }
        if (f == 0) {
            for (int k = b - 1; k >= a; --k) {
                go(k, a, sa);
                record(k);
            }
            rep = std::max(rep, l - 1);
            return;
        }
        // the first checkpoint where the forward steps left to reverse both
        // parts (m - a to reach it, then the right part with one checkpoint
        // fewer and the left part again with them all) are fewest
        const int c = f + 1;
        rep = std::max(rep, repetitions(l, c));
        int best = 1;
        double best_cost = INFINITY;
        for (int k = 1; k < l; ++k) {
            const double cost = (double)k + fewest_steps(l - k, c - 1) + fewest_steps(k, c);
            if (cost < best_cost) {
                best_cost = cost;
                best = k;
            }
        }
        const int m = a + best;
        go(m, a, sa);
        const int sm = take_slot();
        slot_step[sm] = m;
        held++;
        most_held = std::max(most_held, held);
        acts.push_back({ACT_STORE, m, sm, 0});
        gen(m, b, sm, f - 1);
        acts.push_back({ACT_FREE, m, sm, 0});
        slot_step[sm] = -1;
        held--;
        gen(a, m, sa, f);
    }
};

std::vector<Act> plan_run(int n, int snapshots, int* most_held, int* rep) {
    Planner P;
    P.slot_step.assign(std::max(1, snapshots), -1);
    P.slot_step[0] = 0;
    P.held = P.most_held = 1;
    P.acts.push_back({ACT_STORE, 0, 0, 0});
    if (n > 0) P.gen(0, n, 0, std::max(1, snapshots) - 1);
    if (most_held) *most_held = P.most_held;
    if (rep) *rep = std::max(1, P.rep);
    return P.acts;
}

}  // namespace

// The stage fields of one step: stages 0..2 (stage 0 = the step's input) and
// phi of stage 3 (19 fields)
struct Tape {
    double *U[3] = {}, *V[3] = {}, *phi[4] = {}, *sxx[3] = {}, *syy[3] = {}, *sxy[3] = {};
    static constexpr int kFields = 19;
    double** slot(int k) {
        double** all[kFields] = {&U[0], &U[1], &U[2], &V[0], &V[1], &V[2], &phi[0], &phi[1], &phi[2], &phi[3],
                                 &sxx[0], &sxx[1], &sxx[2], &syy[0], &syy[1], &syy[2], &sxy[0], &sxy[1], &sxy[2]};
        return all[k];
    }
};

struct fsi_adjoint {
    fsi_context* ctx = nullptr;
    fsi_workspace* ws = nullptr;  // the caller's (the forward part of a run)
    fsi_context* twin = nullptr;  // zero velocity data on the walls: the transposed solves
    fsi_workspace* twin_ws = nullptr;
    fsi_state* twin_zero = nullptr;  // x0 = 0 of its predictor solves
    fsi_state* work = nullptr;       // recomputations of the backward sweep
    fsi_workspace* work_ws = nullptr;
    // the tape of the last recorded step (`own`), or with every tape kept
    // (snapshots = 0) the tapes of all steps of the run (`pool`); `cur` is
    // the tape being written or transposed
    Tape own;
    Tape* cur = &own;
    std::vector<Tape> pool;
    double* pool_mem = nullptr;
    int keep_tapes = 0;
    int taped = 0;
    double tape_dt = 0.0;
    // cotangents of the stage tendencies and of the stage fields (stage 0: the step's input)
    double *bEu[3] = {}, *bEv[3] = {}, *bIu[3] = {}, *bIv[3] = {}, *bEp[3] = {}, *bExx[3] = {}, *bEyy[3] = {},
           *bExy[3] = {};
    double *gU[3] = {}, *gV[3] = {}, *gphi[3] = {}, *gsxx[3] = {}, *gsyy[3] = {}, *gsxy[3] = {};
    double *wU = nullptr, *wV = nullptr;
    // the transposed predictor solutions of the three stages: each solve
    // starts from the one of the step transposed before it
    double *rU[3] = {}, *rV[3] = {};
    double *partials = nullptr, *dG = nullptr, *dG_save = nullptr;
    int planned = 0;           // the transposed solves run planned iteration counts
    int plans = 0;             // plan_vel, plan_proj hold plans
    int plan_vel[3] = {}, plan_proj[3] = {};
    std::vector<double*> owned;
    double bytes = 0.0;
    // the run
    int snapshots = 0;
    std::vector<std::array<double*, 6>> slots;
    std::vector<Act> acts;
    size_t next_act = 0;
    int run_steps = -1;
    double t0 = 0.0, dt = 0.0;
    fsi_adjoint_stats stats{};
};

namespace {

fsi_status adj_alloc(fsi_adjoint* a, double** p) {
    fsi_context* c = a->ctx;
    HIPCHECK(c, dalloc(c, p, c->felems));
    a->owned.push_back(*p);
    a->bytes += (double)c->felems * sizeof(double);
    return FSI_OK;
}

fsi_status copy_field(fsi_context* c, hipStream_t s, double* dst, const double* src) {
    HIPCHECK(c, hipMemcpyAsync(dst, src, c->felems * sizeof(double), hipMemcpyDeviceToDevice, s));
    return FSI_OK;
}

std::array<double*, 6> fields_of(fsi_state* s) { return {s->U, s->V, s->phi, s->sxx, s->syy, s->sxy}; }

// The step with its tape: the input state, the stage-2 velocity (copied as
// the step runs) and the stage-3, stage-4 fields left in the workspace.
fsi_status record_step(fsi_adjoint* a, fsi_state* st, fsi_workspace* w, double t, double dt) {
    fsi_context* c = a->ctx;
    hipStream_t s = c->stream;
    const std::array<double*, 6> in = fields_of(st);
    double* const t0[6] = {a->cur->U[0], a->cur->V[0], a->cur->phi[0], a->cur->sxx[0], a->cur->syy[0], a->cur->sxy[0]};
    for (int k = 0; k < 6; ++k)
        if (fsi_status rc = copy_field(c, s, t0[k], in[k])) return rc;
    w->tape_U1 = a->cur->U[1];
    w->tape_V1 = a->cur->V[1];
    const int graph = c->use_graph;
    if (graph == 1) c->use_graph = 2;  // a captured step would not copy the tape
    fsi_status rc = fsi_integrate_step(c, st, w, t, dt, nullptr, nullptr);
    c->use_graph = graph;
    w->tape_U1 = w->tape_V1 = nullptr;
    a->taped = 0;
    if (rc != FSI_OK) return rc;
    const double* src[] = {w->Us2, w->Vs2, w->cphi[0], w->csxx[0], w->csyy[0], w->csxy[0],
                           w->cphi[1], w->csxx[1], w->csyy[1], w->csxy[1], st->phi};
    double* dst[] = {a->cur->U[2], a->cur->V[2], a->cur->phi[1], a->cur->sxx[1], a->cur->syy[1], a->cur->sxy[1],
                     a->cur->phi[2], a->cur->sxx[2], a->cur->syy[2], a->cur->sxy[2], a->cur->phi[3]};
    for (int k = 0; k < 11; ++k)
        if ((rc = copy_field(c, s, dst[k], src[k])) != FSI_OK) return rc;
    HIPCHECK(c, hipStreamSynchronize(s));
    a->taped = 1;
    a->tape_dt = dt;
    a->stats.recordings++;
    return FSI_OK;
}

AdjStage stage_of(const fsi_adjoint* a, int s) {
    return AdjStage{a->cur->U[s], a->cur->V[s], a->cur->phi[s], a->cur->sxx[s], a->cur->syy[s], a->cur->sxy[s],
                    s == 0 ? a->ctx->has_reset : 0};
}

// The transpose of the recorded step: `bar` holds the cotangent of the step's
// output on entry and that of its input on return; *dG (device) gains the
// shear-modulus part.  With `planned` the transposed solves run the
// iteration counts of the last step's and nothing is read back until the
// end; if one of them has not converged by then, *missed is set, `bar` is
// left as it was, and the caller runs the step again with polling.
fsi_status adjoint_step_once(fsi_adjoint* a, fsi_state* bar, bool planned, bool* missed) {
    fsi_context* c = a->ctx;
    fsi_context* tw = a->twin;
    const Params& p = c->prm;
    hipStream_t S = tw->stream;
    const size_t n = c->felems;
    const double dt = a->tape_dt;
    *missed = false;
    AdjPtrs z{};
    for (int q = 0; q < 3; ++q) {
        double* f[] = {a->bEu[q], a->bEv[q], a->bIu[q], a->bIv[q], a->bEp[q], a->bExx[q], a->bEyy[q], a->bExy[q],
                       a->gU[q], a->gV[q], a->gphi[q], a->gsxx[q], a->gsyy[q], a->gsxy[q]};
        for (double* x : f) z.p[z.n++] = x;
    }
    launch_adj_zero(S, n, z);
    status_reset(tw);
    if (fsi_status rc = status_push(tw)) return rc;
    const std::array<double*, 6> out = fields_of(bar);
    auto bar_of = [&](int q) {
        return AdjBar{a->bEu[q], a->bEv[q], a->bIu[q], a->bIv[q], a->bEp[q], a->bExx[q], a->bEyy[q], a->bExy[q]};
    };
    auto grads_of = [&](int q) {
        return AdjGrads{a->gU[q], a->gV[q], a->gphi[q], a->gsxx[q], a->gsyy[q], a->gsxy[q], nullptr, nullptr};
    };
    fsi_status rc = FSI_OK;
    for (int s = 3; s >= 1; --s) {
        const int k = s - 1;
        // the cotangent of stage s's fields: the output's, or gathered from its tendencies
        std::array<double*, 6> g = out;
        if (s <= 2) {
            launch_adj_tendencies(S, p, c->bpsi, stage_of(a, s), bar_of(s), grads_of(s), a->partials,
                                  p.elastic ? a->dG : nullptr);
            g = {a->gU[s], a->gV[s], a->gphi[s], a->gsxx[s], a->gsyy[s], a->gsxy[s]};
        }
        // u_s = P K_s^-1 (rhs_s + alpha_s b): the wall faces are data
        launch_adj_walls(S, p, g[0], g[1]);
        const int pproj = planned ? a->plan_proj[k] : 0;
        if ((rc = projection(tw, a->twin_ws, g[0], g[1], a->wU, a->wV, 1, k, pproj)) != FSI_OK) return rc;
        if (p.visc) {
            const int pvel = planned ? a->plan_vel[k] : 0;
            if ((rc = velocity_solve(tw, a->twin_zero, a->twin_ws, dt * AI[s][s], k, pvel, a->wU, a->wV, a->rU[k],
                                     a->rV[k], true)) != FSI_OK)
                return rc;
        } else {
            if ((rc = copy_field(c, S, a->rU[k], a->wU)) || (rc = copy_field(c, S, a->rV[k], a->wV))) return rc;
        }
        launch_adj_walls(S, p, a->rU[k], a->rV[k]);  // the Dirichlet rows and the faces outside the compact predictor
        // rhs_s = u^n + dt sum_q (aE E_q + aI I_q)
        AdjCombine cv{};
        cv.src[0] = a->rU[k];
        cv.src[1] = a->rV[k];
        cv.nsrc = 2;
        auto add = [&](AdjCombine& cb, double* dst, double cf, int from) {
            if (cf == 0.0) return;
            cb.dst[cb.ndst] = dst;
            cb.coef[cb.ndst] = cf;
            cb.from[cb.ndst] = from;
            cb.ndst++;
        };
        add(cv, a->gU[0], 1.0, 0);
        add(cv, a->gV[0], 1.0, 1);
        for (int q = 0; q < s; ++q) {
            add(cv, a->bEu[q], dt * AE[s][q], 0);
            add(cv, a->bEv[q], dt * AE[s][q], 1);
            add(cv, a->bIu[q], dt * AI[s][q], 0);
            add(cv, a->bIv[q], dt * AI[s][q], 1);
        }
        launch_adj_combine(S, n, cv);
        // phi_s, sigma_s = reset(phi^n + dt sum_q aE E_phi,q, sigma^n~ + ...): the reset cells are constants
        AdjCombine cc{};
        for (int q = 0; q < 4; ++q) cc.src[q] = g[2 + q];
        cc.nsrc = 4;
        cc.phi = c->has_reset ? a->cur->phi[s] : nullptr;
        cc.cutoff = p.ext_cutoff;
        cc.masked = 2 | 4 | 8;
        double* acc0[4] = {a->gphi[0], a->gsxx[0], a->gsyy[0], a->gsxy[0]};
        for (int q = 0; q < 4; ++q) add(cc, acc0[q], 1.0, q);
        for (int q = 0; q < s; ++q) {
            const double cf = dt * AE[s][q];
            add(cc, a->bEp[q], cf, 0);
            add(cc, a->bExx[q], cf, 1);
            add(cc, a->bEyy[q], cf, 2);
            add(cc, a->bExy[q], cf, 3);
        }
        launch_adj_combine(S, n, cc);
    }
    // the stage-1 tendencies of the input (sigma^n with the safety reset)
    launch_adj_tendencies(S, p, c->bpsi, stage_of(a, 0), bar_of(0), grads_of(0), a->partials,
                          p.elastic ? a->dG : nullptr);
    if (c->has_reset) launch_adj_mask(S, p, a->cur->phi[0], a->gsxx[0], a->gsyy[0], a->gsxy[0]);
    launch_adj_walls(S, p, a->gU[0], a->gV[0]);
    HIPCHECK(c, hipGetLastError());
    if ((rc = status_pull(tw)) != FSI_OK) return rc;
    if (planned) {
        const Status& st = *tw->st_host;
        int mask = 0;
        for (int k = 0; k < 3; ++k) {
            if (p.visc && !st.vel_ok[k]) mask |= 1 << k;
            if (!st.proj_ok[k]) mask |= 8 << k;
        }
        if (mask) {  // those solves get more iterations from now on
            *missed = true;
            a->stats.last_shortfall = mask;
            for (int k = 0; k < 3; ++k) {
                if (mask & (1 << k)) a->plan_vel[k] += 2 + a->plan_vel[k] / 4;
                if (mask & (8 << k)) a->plan_proj[k] += 2 + a->plan_proj[k] / 4;
            }
            return FSI_OK;
        }
    }
    if ((rc = device_error(tw)) != FSI_OK) return fail(c, rc, "adjoint solve: %s", fsi_last_error(tw));
    {
        const Status& st = *tw->st_host;
        // the plans: from the first (polled) step; after a planned step they
        // shrink by one iteration at a time towards what the solves took (a
        // polled rerun after a shortfall teaches nothing: it starts from the
        // shortfall's iterate)
        for (int k = 0; k < 3; ++k) {
            const int v = std::max(1, st.vel_iters[k]) + 1, q = std::max(0, st.proj_iters[k]) + 1;
            if (!a->plans) {
                a->plan_vel[k] = v;
                a->plan_proj[k] = q;
            } else if (planned) {
                a->plan_vel[k] = std::max(v, a->plan_vel[k] - 1);
                a->plan_proj[k] = std::max(q, a->plan_proj[k] - 1);
            }
            a->stats.solve_iterations[k] = st.vel_iters[k];
        }
        a->plans = 1;
    }
    double* const g0[6] = {a->gU[0], a->gV[0], a->gphi[0], a->gsxx[0], a->gsyy[0], a->gsxy[0]};
    for (int k = 0; k < 6; ++k)
        if ((rc = copy_field(c, S, out[k], g0[k])) != FSI_OK) return rc;
    HIPCHECK(c, hipStreamSynchronize(S));
    return FSI_OK;
}

fsi_status adjoint_step(fsi_adjoint* a, fsi_state* bar) {
    fsi_context* c = a->ctx;
    if (!a->taped) return fail(c, FSI_E_ARGUMENT, "fsi_adjoint_step: no recorded step to transpose.");
    HIPCHECK(c, hipStreamSynchronize(c->stream));  // the cotangent and the tape, written on the step's stream
    bool missed = false;
    fsi_status rc = FSI_OK;
    if (a->planned) {
        HIPCHECK(c, hipMemcpyAsync(a->dG_save, a->dG, sizeof(double), hipMemcpyDeviceToDevice, a->twin->stream));
        if ((rc = adjoint_step_once(a, bar, true, &missed)) != FSI_OK) return rc;
        if (missed) {
            a->stats.replanned++;
            HIPCHECK(c, hipMemcpyAsync(a->dG, a->dG_save, sizeof(double), hipMemcpyDeviceToDevice, a->twin->stream));
        }
    }
    if (!a->planned || missed) {
        if ((rc = adjoint_step_once(a, bar, false, &missed)) != FSI_OK) return rc;
        a->planned = 1;
    }
    a->stats.adjoint_steps++;
    return FSI_OK;
}

// the configurations the transposed kernels cover
fsi_status adjoint_supported(fsi_context* c) {
    const fsi_models& m = c->models;
    if (c->grid.periodic_x || c->grid.periodic_y)
        return fail(c, FSI_E_UNSUPPORTED, "The device adjoint covers walled domains only (no periodic side).");
    if (m.phase_method != FSI_PHASE_ACDI || m.acdi_dynamic_gamma)
        return fail(c, FSI_E_UNSUPPORTED, "The device adjoint covers the ACDI phase field with a fixed gamma.");
    if (c->num.precision != FSI_PRECISION_FP64)
        return fail(c, FSI_E_UNSUPPORTED, "The device adjoint differentiates the FP64 step.");
    if (c->growth)
        return fail(c, FSI_E_UNSUPPORTED, "The device adjoint does not cover growth or resource closures.");
    if (c->models.capillarity_kind != FSI_CAPILLARITY_NONE)
        return fail(c, FSI_E_UNSUPPORTED, "The device adjoint does not cover capillary forcing.");
    const fsi_bc_setup* v[2] = {&c->bcs.U, &c->bcs.V};
    for (auto* b : v)
        for (const fsi_bc_side* sd : {&b->left, &b->right, &b->bottom, &b->top})
            if (sd->kind == FSI_BC_PERIODIC)
                return fail(c, FSI_E_UNSUPPORTED, "The device adjoint covers walled domains only.");
    if (c->bcs.U.right.kind != FSI_BC_DIRICHLET || c->bcs.V.top.kind != FSI_BC_DIRICHLET)
        return fail(c, FSI_E_UNSUPPORTED, "The device adjoint needs Dirichlet U-right and V-top walls.");
    return FSI_OK;
}

fsi_status make_twin(fsi_adjoint* a) {
    fsi_context* c = a->ctx;
    fsi_bcs b = c->bcs;
    for (fsi_bc_setup* v : {&b.U, &b.V})
        for (fsi_bc_side* sd : {&v->left, &v->right, &v->bottom, &v->top}) sd->value = 0.0;
    fsi_numerics num = c->num;
    num.velocity_method = FSI_SOLVER_DIRECT;
    num.check_stage1 = 0;
    num.projection_rtol = std::min(num.projection_rtol, 1e-12);
    fsi_status rc = fsi_context_create(&c->grid, &b, &c->models, &num, &a->twin);
    if (rc != FSI_OK) return fail(c, rc, "adjoint twin context: %s", a->twin ? fsi_last_error(a->twin) : "");
    a->twin->vel_warm = 0;
    if ((rc = fsi_workspace_create(a->twin, &a->twin_ws)) != FSI_OK) return fail(c, rc, "adjoint twin workspace");
    if ((rc = fsi_state_create(a->twin, &a->twin_zero)) != FSI_OK) return fail(c, rc, "adjoint twin state");
    HIPCHECK(c, hipStreamSynchronize(a->twin->stream));
    return FSI_OK;
}

}  // namespace

extern "C" fsi_status fsi_adjoint_create(fsi_context* c, fsi_workspace* ws, int32_t snapshots, fsi_adjoint** out) {
    if (!c || !ws || !out || ws->ctx != c || snapshots < 0) return FSI_E_ARGUMENT;
    *out = nullptr;
    if (fsi_status rc = adjoint_supported(c)) return rc;
    size_t free0 = 0, total = 0;
    (void)hipMemGetInfo(&free0, &total);
    fsi_adjoint* a = new fsi_adjoint();
    a->ctx = c;
    a->ws = ws;
    a->snapshots = snapshots;
    a->keep_tapes = snapshots == 0;
    fsi_status rc = make_twin(a);
    if (rc == FSI_OK && (rc = fsi_state_create(c, &a->work)) == FSI_OK) rc = fsi_workspace_create(c, &a->work_ws);
    for (int s = 0; s < 3 && rc == FSI_OK; ++s) {
        double** f[] = {&a->own.U[s], &a->own.V[s], &a->own.phi[s], &a->own.sxx[s], &a->own.syy[s], &a->own.sxy[s],
                        &a->bEu[s], &a->bEv[s], &a->bIu[s], &a->bIv[s], &a->bEp[s], &a->bExx[s], &a->bEyy[s],
                        &a->bExy[s], &a->gU[s], &a->gV[s], &a->gphi[s], &a->gsxx[s], &a->gsyy[s], &a->gsxy[s],
                        &a->rU[s], &a->rV[s]};
        for (double** pp : f)
            if ((rc = adj_alloc(a, pp)) != FSI_OK) break;
    }
    double** extra[] = {&a->own.phi[3], &a->wU, &a->wV};
    for (double** pp : extra)
        if (rc == FSI_OK) rc = adj_alloc(a, pp);
    if (rc == FSI_OK) {
        const size_t np = (size_t)adj_block_count(c->prm) + 1;
        if (hipMalloc((void**)&a->partials, np * sizeof(double)) != hipSuccess ||
            hipMalloc((void**)&a->dG, 2 * sizeof(double)) != hipSuccess)
            rc = fail(c, FSI_E_HIP, "adjoint buffers: out of device memory");
        a->bytes += (double)(np + 2) * sizeof(double);
        a->dG_save = a->dG + 1;
    }
    a->slots.resize(snapshots);
    for (auto& sl : a->slots)
        for (double*& f : sl)
            if (rc == FSI_OK) rc = adj_alloc(a, &f);
    if (rc == FSI_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, FSI_E_HIP, "adjoint create");
    if (rc != FSI_OK) {
        fsi_adjoint_destroy(a);
        return rc;
    }
    size_t free1 = 0;
    (void)hipMemGetInfo(&free1, &total);
    a->stats.device_bytes = free0 > free1 ? (double)(free0 - free1) : a->bytes;
    *out = a;
    return FSI_OK;
}

extern "C" void fsi_adjoint_destroy(fsi_adjoint* a) {
    if (!a) return;
    if (a->ctx) (void)hipStreamSynchronize(a->ctx->stream);
    if (a->twin) (void)hipStreamSynchronize(a->twin->stream);
    for (double* p : a->owned) (void)hipFree(p);
    if (a->partials) (void)hipFree(a->partials);
    if (a->dG) (void)hipFree(a->dG);
    if (a->pool_mem) (void)hipFree(a->pool_mem);
    fsi_workspace_destroy(a->work_ws);
    fsi_state_destroy(a->work);
    fsi_state_destroy(a->twin_zero);
    fsi_workspace_destroy(a->twin_ws);
    fsi_context_destroy(a->twin);
    delete a;
}

extern "C" fsi_status fsi_adjoint_record_step(fsi_adjoint* a, fsi_state* st, double t, double dt) {
    if (!a || !st || st->ctx != a->ctx) return FSI_E_ARGUMENT;
    fsi_context* c = a->ctx;
    a->cur = &a->own;
    const int warm = c->vel_warm;
    c->vel_warm = 0;
    fsi_status rc = record_step(a, st, a->ws, t, dt);
    c->vel_warm = warm;
    return rc;
}

extern "C" fsi_status fsi_adjoint_step(fsi_adjoint* a, fsi_state* bar, double* d_shear_modulus) {
    if (!a || !bar || bar->ctx != a->ctx) return FSI_E_ARGUMENT;
    fsi_context* c = a->ctx;
    HIPCHECK(c, hipMemsetAsync(a->dG, 0, sizeof(double), a->twin->stream));
    fsi_status rc = adjoint_step(a, bar);
    if (rc != FSI_OK) return rc;
    double g = 0.0;
    HIPCHECK(c, hipMemcpyAsync(&g, a->dG, sizeof(double), hipMemcpyDeviceToHost, a->twin->stream));
    HIPCHECK(c, hipStreamSynchronize(a->twin->stream));
    if (d_shear_modulus) *d_shear_modulus = g;
    return FSI_OK;
}

namespace {

// run the actions from next_act up to and including the first RECORD
// (forward) or to the end (backward); `st`, `w`: the working state
fsi_status run_actions(fsi_adjoint* a, fsi_state* st, fsi_workspace* w, fsi_state* bar, bool forward) {
    fsi_context* c = a->ctx;
    hipStream_t s = c->stream;
    while (a->next_act < a->acts.size()) {
        const Act& x = a->acts[a->next_act];
        fsi_status rc = FSI_OK;
        switch (x.kind) {
        case ACT_STORE: {
            const std::array<double*, 6> f = fields_of(st);
            for (int k = 0; k < 6 && rc == FSI_OK; ++k) rc = copy_field(c, s, a->slots[x.slot][k], f[k]);
            break;
        }
        case ACT_RESTORE: {
            const std::array<double*, 6> f = fields_of(st);
            for (int k = 0; k < 6 && rc == FSI_OK; ++k) rc = copy_field(c, s, f[k], a->slots[x.slot][k]);
            if (rc == FSI_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(c, FSI_E_HIP, "restore");
            break;
        }
        case ACT_ADVANCE:
            for (int k = 0; k < x.count && rc == FSI_OK; ++k) {
                rc = fsi_integrate_step(c, st, w, a->t0 + (double)(x.step + k) * a->dt, a->dt, nullptr, nullptr);
                a->stats.advances++;
            }
            break;
        case ACT_RECORD: rc = record_step(a, st, w, a->t0 + (double)x.step * a->dt, a->dt); break;
        case ACT_ADJOINT:
            if (forward) return FSI_OK;
            rc = adjoint_step(a, bar);
            break;
        case ACT_FREE: break;
        }
        if (rc != FSI_OK) return rc;
        a->next_act++;
        if (forward && x.kind == ACT_RECORD) return FSI_OK;
    }
    return FSI_OK;
}

}  // namespace

extern "C" fsi_status fsi_adjoint_forward(fsi_adjoint* a, fsi_state* st, double t0, double dt, int32_t n_steps) {
    if (!a || !st || st->ctx != a->ctx || n_steps < 1) return FSI_E_ARGUMENT;
    fsi_context* c = a->ctx;
    if (!(dt > 0.0) || !std::isfinite(dt) || !std::isfinite(t0))
        return fail(c, FSI_E_ARGUMENT, "dt must be finite and strictly positive.");
    int held = 0, rep = 0;
    a->acts = a->keep_tapes ? std::vector<Act>{} : plan_run(n_steps, a->snapshots, &held, &rep);
    a->next_act = 0;
    a->run_steps = n_steps;
    a->t0 = t0;
    a->dt = dt;
    const double device_bytes = a->stats.device_bytes;
    a->stats = fsi_adjoint_stats{};
    a->stats.device_bytes = device_bytes;
    a->stats.steps = n_steps;
    a->stats.snapshots_used = held;
    a->stats.repetitions = rep;
    for (int k = 0; k < 3; ++k) {  // the transposed solves of the run start from zero
        HIPCHECK(c, hipMemsetAsync(a->rU[k], 0, c->felems * sizeof(double), c->stream));
        HIPCHECK(c, hipMemsetAsync(a->rV[k], 0, c->felems * sizeof(double), c->stream));
    }
    if (a->keep_tapes) {  // every step taped as it runs: nothing is recomputed
        if ((int)a->pool.size() < n_steps) {
            if (a->pool_mem) (void)hipFree(a->pool_mem);
            a->pool_mem = nullptr;
            a->pool.clear();
            const size_t per = (size_t)Tape::kFields * c->felems;
            if (hipMalloc((void**)&a->pool_mem, (size_t)n_steps * per * sizeof(double)) != hipSuccess) {
                a->run_steps = -1;
                return fail(c, FSI_E_HIP, "adjoint: no device memory for %d tapes of %.1f MB (use checkpoints)",
                            n_steps, per * sizeof(double) / 1e6);
            }
            a->pool.resize(n_steps);
            for (int k = 0; k < n_steps; ++k)
                for (int f = 0; f < Tape::kFields; ++f) *a->pool[k].slot(f) = a->pool_mem + ((size_t)k * Tape::kFields + f) * c->felems;
        }
        a->stats.snapshots_used = 0;
        a->stats.repetitions = 0;
        size_t fr = 0, tot = 0;
        (void)hipMemGetInfo(&fr, &tot);
        for (int k = 0; k < n_steps; ++k) {
            a->cur = &a->pool[k];
            if (fsi_status rc = record_step(a, st, a->ws, t0 + (double)k * dt, dt)) {
                a->run_steps = -1;
                return rc;
            }
        }
        a->stats.device_bytes = a->bytes + (double)a->pool.size() * Tape::kFields * c->felems * sizeof(double);
        return FSI_OK;
    }
    // the recomputations start from checkpoints of the six evolving fields;
    // the resource (constant under NoGrowth) is the caller's
    a->cur = &a->own;
    if (fsi_status rc = copy_field(c, c->stream, a->work->res, st->res)) return rc;
    const int warm = c->vel_warm;
    c->vel_warm = 0;
    fsi_status rc = run_actions(a, st, a->ws, nullptr, true);
    c->vel_warm = warm;
    if (rc != FSI_OK) a->run_steps = -1;
    return rc;
}

extern "C" fsi_status fsi_adjoint_backward(fsi_adjoint* a, fsi_state* bar, double* d_shear_modulus) {
    if (!a || !bar || bar->ctx != a->ctx) return FSI_E_ARGUMENT;
    fsi_context* c = a->ctx;
    if (a->run_steps < 1 ||
        (!a->keep_tapes && (a->next_act >= a->acts.size() || a->acts[a->next_act].kind != ACT_ADJOINT)))
        return fail(c, FSI_E_ARGUMENT, "fsi_adjoint_backward: no run recorded by fsi_adjoint_forward.");
    HIPCHECK(c, hipMemsetAsync(a->dG, 0, sizeof(double), a->twin->stream));
    if (a->keep_tapes) {
        fsi_status rc = FSI_OK;
        for (int k = a->run_steps - 1; k >= 0 && rc == FSI_OK; --k) {
            a->cur = &a->pool[k];
            a->taped = 1;
            a->tape_dt = a->dt;
            rc = adjoint_step(a, bar);
        }
        a->run_steps = -1;
        a->taped = 0;
        if (rc != FSI_OK) return rc;
        double g = 0.0;
        HIPCHECK(c, hipMemcpyAsync(&g, a->dG, sizeof(double), hipMemcpyDeviceToHost, a->twin->stream));
        HIPCHECK(c, hipStreamSynchronize(a->twin->stream));
        if (d_shear_modulus) *d_shear_modulus = g;
        return FSI_OK;
    }
    const int warm = c->vel_warm;
    c->vel_warm = 0;
    fsi_status rc = run_actions(a, a->work, a->work_ws, bar, false);
    c->vel_warm = warm;
    a->run_steps = -1;
    if (rc != FSI_OK) return rc;
    double g = 0.0;
    HIPCHECK(c, hipMemcpyAsync(&g, a->dG, sizeof(double), hipMemcpyDeviceToHost, a->twin->stream));
    HIPCHECK(c, hipStreamSynchronize(a->twin->stream));
    if (d_shear_modulus) *d_shear_modulus = g;
    return FSI_OK;
}

extern "C" fsi_status fsi_adjoint_get_stats(const fsi_adjoint* a, fsi_adjoint_stats* out) {
    if (!a || !out) return FSI_E_ARGUMENT;
    *out = a->stats;
    return FSI_OK;
}

// J = |x_c - target|^2 of the phase centroid (run_sugiyama_lid_driven_cavity.jl:496-509)
// and its cotangent with respect to the state: (x_i - x_c) a / M on phi
extern "C" fsi_status fsi_adjoint_centroid_objective(fsi_context* c, const fsi_state* st, double tx, double ty,
                                                     fsi_state* bar, double* J) {
    if (!c || !st || !bar || bar->ctx != c || st->ctx != c) return FSI_E_ARGUMENT;
    double xc = 0.0, yc = 0.0, mass = 0.0;
    fsi_status rc = fsi_phase_centroid(c, st, &xc, &yc, &mass);
    if (rc != FSI_OK) return rc;
    const double area = c->prm.hx * c->prm.hy;
    const double A = 2.0 * (xc - tx) * area / mass, B = 2.0 * (yc - ty) * area / mass;
    for (double* f : {bar->U, bar->V, bar->sxx, bar->syy, bar->sxy})
        HIPCHECK(c, hipMemsetAsync(f, 0, c->felems * sizeof(double), c->stream));
    launch_adj_affine(c->stream, c->prm, bar->phi, A, B, -(A * xc + B * yc));
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (J) *J = (xc - tx) * (xc - tx) + (yc - ty) * (yc - ty);
    return FSI_OK;
}

// Kernel parity: the transposed stage tendencies at the state (sigma with the
// safety reset, as the first stage reads it), cotangents given and returned
// as host arrays; and the point functions' tendencies in FP64.
extern "C" fsi_status fsi_adjoint_tendencies(fsi_context* c, const fsi_state* st, const double* const* bar_in,
                                             double* const* grad_out, double* d_shear_modulus) {
    if (!c || !st || st->ctx != c || !bar_in || !grad_out) return FSI_E_ARGUMENT;
    if (fsi_status rc = adjoint_supported(c)) return rc;
    const int nx = c->grid.nx, ny = c->grid.ny;
    std::vector<double*> buf(15, nullptr);
    fsi_status rc = FSI_OK;
    for (auto& b : buf)
        if (rc == FSI_OK && hipMalloc((void**)&b, c->felems * sizeof(double)) != hipSuccess)
            rc = fail(c, FSI_E_HIP, "out of device memory");
    double* dG = nullptr;
    const size_t np = (size_t)adj_block_count(c->prm) + 1;
    double* partials = nullptr;
    if (rc == FSI_OK && (hipMalloc((void**)&dG, sizeof(double)) != hipSuccess ||
                         hipMalloc((void**)&partials, np * sizeof(double)) != hipSuccess))
        rc = fail(c, FSI_E_HIP, "out of device memory");
    for (size_t k = 0; k < buf.size() && rc == FSI_OK; ++k)
        if (hipMemsetAsync(buf[k], 0, c->felems * sizeof(double), c->stream) != hipSuccess)
            rc = fail(c, FSI_E_HIP, "memset");
    if (rc == FSI_OK && hipMemsetAsync(dG, 0, sizeof(double), c->stream) != hipSuccess) rc = fail(c, FSI_E_HIP, "x");
    // bar: E_u, E_v, I_u, I_v (x-faces / y-faces), E_phi, E_sxx, E_syy, E_sxy (cells)
    const int wid[8] = {nx + 1, nx, nx + 1, nx, nx, nx, nx, nx}, hei[8] = {ny, ny + 1, ny, ny + 1, ny, ny, ny, ny};
    for (int k = 0; k < 8 && rc == FSI_OK; ++k)
        if (bar_in[k]) rc = upload2d(c, buf[k], bar_in[k], wid[k], hei[k]);
    if (rc == FSI_OK) {
        const AdjStage s{st->U, st->V, st->phi, st->sxx, st->syy, st->sxy, c->has_reset};
        const AdjBar b{buf[0], buf[1], buf[2], buf[3], buf[4], buf[5], buf[6], buf[7]};
        const AdjGrads g{buf[8], buf[9], buf[10], buf[11], buf[12], buf[13], nullptr, nullptr};
        launch_adj_tendencies(c->stream, c->prm, c->bpsi, s, b, g, partials, c->prm.elastic ? dG : nullptr);
        if (hipGetLastError() != hipSuccess) rc = fail(c, FSI_E_HIP, "adjoint tendency kernels");
    }
    // out: U (x-faces), V (y-faces), phi, sxx, syy, sxy
    const int ow[6] = {nx + 1, nx, nx, nx, nx, nx}, oh[6] = {ny, ny + 1, ny, ny, ny, ny};
    for (int k = 0; k < 6 && rc == FSI_OK; ++k)
        if (grad_out[k]) rc = download2d(c, grad_out[k], buf[8 + k], ow[k], oh[k]);
    double g = 0.0;
    if (rc == FSI_OK && hipMemcpyAsync(&g, dG, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = fail(c, FSI_E_HIP, "copy");
    if (rc == FSI_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, FSI_E_HIP, "sync");
    if (rc == FSI_OK && d_shear_modulus) *d_shear_modulus = g;
    (void)hipStreamSynchronize(c->stream);
    for (double* b : buf) (void)hipFree(b);
    (void)hipFree(dG);
    (void)hipFree(partials);
    return rc;
}

extern "C" fsi_status fsi_adjoint_primal_tendencies(fsi_context* c, const fsi_state* st, double* const* out) {
    if (!c || !st || st->ctx != c || !out) return FSI_E_ARGUMENT;
    if (fsi_status rc = adjoint_supported(c)) return rc;
    const int nx = c->grid.nx, ny = c->grid.ny;
    std::vector<double*> buf(8, nullptr);
    fsi_status rc = FSI_OK;
    for (auto& b : buf)
        if (rc == FSI_OK && hipMalloc((void**)&b, c->felems * sizeof(double)) != hipSuccess)
            rc = fail(c, FSI_E_HIP, "out of device memory");
    for (size_t k = 0; k < buf.size() && rc == FSI_OK; ++k)
        if (hipMemsetAsync(buf[k], 0, c->felems * sizeof(double), c->stream) != hipSuccess)
            rc = fail(c, FSI_E_HIP, "memset");
    if (rc == FSI_OK) {
        const AdjStage s{st->U, st->V, st->phi, st->sxx, st->syy, st->sxy, c->has_reset};
        const AdjGrads g{buf[0], buf[1], buf[4], buf[5], buf[6], buf[7], buf[2], buf[3]};
        launch_adj_primal(c->stream, c->prm, c->bpsi, s, g);
        if (hipGetLastError() != hipSuccess) rc = fail(c, FSI_E_HIP, "adjoint primal kernel");
    }
    // E_u, E_v, I_u, I_v, E_phi, E_sxx, E_syy, E_sxy
    const int wid[8] = {nx + 1, nx, nx + 1, nx, nx, nx, nx, nx}, hei[8] = {ny, ny + 1, ny, ny + 1, ny, ny, ny, ny};
    for (int k = 0; k < 8 && rc == FSI_OK; ++k)
        if (out[k]) rc = download2d(c, out[k], buf[k], wid[k], hei[k]);
    if (rc == FSI_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, FSI_E_HIP, "sync");
    (void)hipStreamSynchronize(c->stream);
    for (double* b : buf) (void)hipFree(b);
    return rc;
}

// The schedule alone (host only): counts of a run of n steps with the given
// snapshots, and optionally its actions (kind, step, slot, count per entry).
extern "C" fsi_status fsi_adjoint_plan(int32_t n_steps, int32_t snapshots, int64_t* advances, int32_t* most_held,
                                       int32_t* repetitions, int32_t* actions, int64_t max_actions,
                                       int64_t* n_actions) {
    if (n_steps < 1 || snapshots < 1) return FSI_E_ARGUMENT;
    int held = 0, rep = 0;
    const std::vector<Act> acts = plan_run(n_steps, snapshots, &held, &rep);
    int64_t adv = 0;
    for (const Act& x : acts)
        if (x.kind == ACT_ADVANCE) adv += x.count;
    if (advances) *advances = adv;
    if (most_held) *most_held = held;
    if (repetitions) *repetitions = rep;
    if (n_actions) *n_actions = (int64_t)acts.size();
    if (actions)
        for (int64_t k = 0; k < std::min<int64_t>(max_actions, (int64_t)acts.size()); ++k) {
            actions[4 * k + 0] = acts[k].kind;
            actions[4 * k + 1] = acts[k].step;
            actions[4 * k + 2] = acts[k].slot;
            actions[4 * k + 3] = acts[k].count;
        }
    return FSI_OK;
}
