// fsi_api.hip -- C ABI of libfsi_hip.so: context, device-resident state and
// workspace, and the host side of the four-stage IMEX step
// (src/TimeIntegrationIMEX.jl:1152-1256).
//
// One HIP stream per context.  A step enqueues every kernel of the four
// stages and reads back one small Status block at the end (the StepReport,
// the first error raised by any kernel, iteration counts).  The Krylov solves
// run with their convergence flags in device memory: the host enqueues
// iterations in chunks and polls the flag between chunks, sized from the
// previous step's iteration counts.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fsi_kernels.hpp"

using namespace fsi;

#include "fsi_internal.hpp"

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
fsi_status fail(fsi_context* c, fsi_status code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}


thread_local LaunchRecorder* fsi::g_recorder = nullptr;


hipError_t dalloc(fsi_context* c, double** p, size_t n) {
    hipError_t e = hipMalloc((void**)p, n * sizeof(double));
    if (e == hipSuccess) e = hipMemsetAsync(*p, 0, n * sizeof(double), c->stream);
    return e;
}

BcSet to_bcset(const fsi_bc_setup& s) {
    BcSet b;
    b.l = {s.left.kind, s.left.value};
    b.r = {s.right.kind, s.right.value};
    b.b = {s.bottom.kind, s.bottom.value};
    b.t = {s.top.kind, s.top.value};
    return b;
}

fsi_status upload2d(fsi_context* c, double* dst, const double* src, int width, int height) {
    HIPCHECK(c, hipMemcpy2DAsync(dst, (size_t)c->P * sizeof(double), src, (size_t)width * sizeof(double),
                                 (size_t)width * sizeof(double), height, hipMemcpyHostToDevice, c->stream));
    return FSI_OK;
}
fsi_status download2d(fsi_context* c, double* dst, const double* src, int width, int height) {
    HIPCHECK(c, hipMemcpy2DAsync(dst, (size_t)width * sizeof(double), src, (size_t)c->P * sizeof(double),
                                 (size_t)width * sizeof(double), height, hipMemcpyDeviceToHost, c->stream));
    return FSI_OK;
}

static const char* side_name(int k) {
    static const char* n[4] = {"left", "right", "bottom", "top"};
    return n[k];
}

static fsi_status scratch_lookup(fsi_context* c);

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static fsi_status validate(fsi_context* c) {
    const fsi_grid& g = c->grid;
    if (g.nx < 2 || g.ny < 2) return fail(c, FSI_E_ARGUMENT, "Nx and Ny must be at least 2.");
    if (!(g.lx > 0.0) || !(g.ly > 0.0)) return fail(c, FSI_E_ARGUMENT, "Lx and Ly must be positive.");
    const fsi_bc_setup* all[8] = {&c->bcs.U, &c->bcs.V, &c->bcs.correction, &c->bcs.phi,
                                  &c->bcs.resource, &c->bcs.sigma_xx, &c->bcs.sigma_yy, &c->bcs.sigma_xy};
    const char* names[8] = {"U", "V", "correction", "phi", "resource", "sigma_xx", "sigma_yy", "sigma_xy"};
    for (int k = 0; k < 8; ++k) {
        const fsi_bc_side* sd[4] = {&all[k]->left, &all[k]->right, &all[k]->bottom, &all[k]->top};
        for (int q = 0; q < 4; ++q) {
            if (sd[q]->kind < 0 || sd[q]->kind > 2)
                return fail(c, FSI_E_UNSUPPORTED, "%s %s boundary kind %d is not supported on the device.", names[k],
                            side_name(q), sd[q]->kind);
            if (!std::isfinite(sd[q]->value))
                return fail(c, FSI_E_ARGUMENT, "%s %s boundary value must be finite.", names[k], side_name(q));
        }
        int lp = all[k]->left.kind == FSI_BC_PERIODIC, rp = all[k]->right.kind == FSI_BC_PERIODIC;
        int bp = all[k]->bottom.kind == FSI_BC_PERIODIC, tp = all[k]->top.kind == FSI_BC_PERIODIC;
        if (lp != rp || bp != tp)  // BoundaryConditions.jl:74-84
            return fail(c, FSI_E_ARGUMENT, "Periodic boundaries must be specified as an opposing pair (%s).",
                        names[k]);
        if (lp != (g.periodic_x != 0))  // TimeIntegrationIMEX.jl:447-466
            return fail(c, FSI_E_ARGUMENT, "%s boundary conditions do not match MeshMetrics periodicity in x.",
                        names[k]);
        if (bp != (g.periodic_y != 0))
            return fail(c, FSI_E_ARGUMENT, "%s boundary conditions do not match MeshMetrics periodicity in y.",
                        names[k]);
    }
    const fsi_models& m = c->models;
    switch (m.phase_method) {
    case FSI_PHASE_UPWIND: break;
    case FSI_PHASE_CDI:  // ConservativeDiffuseInterfaceFlux, InterfaceRegularization.jl:30-63
        if (!(m.cdi_inverse_peclet >= 0.0) || !std::isfinite(m.cdi_inverse_peclet))
            return fail(c, FSI_E_ARGUMENT, "inverse_peclet must be finite and non-negative.");
        if (!(m.cdi_epsilon > 0.0)) return fail(c, FSI_E_ARGUMENT, "epsilon must be strictly positive.");
        if (!(m.cdi_delta_normal > 0.0)) return fail(c, FSI_E_ARGUMENT, "delta_normal must be strictly positive.");
        break;
    case FSI_PHASE_ACDI:  // PhaseFieldMethods.jl:58-91
        if (!(m.acdi_gamma >= 0.0) || !std::isfinite(m.acdi_gamma))
            return fail(c, FSI_E_ARGUMENT, "gamma must be finite and non-negative.");
        if (!(m.acdi_epsilon > 0.0) || !std::isfinite(m.acdi_epsilon))
            return fail(c, FSI_E_ARGUMENT, "epsilon must be finite and strictly positive.");
        if (!(m.acdi_psi_floor > 0.0 && m.acdi_psi_floor < 0.5))
            return fail(c, FSI_E_ARGUMENT, "psi_floor must lie strictly between 0 and 0.5.");
        if (!(m.acdi_delta_normal > 0.0)) return fail(c, FSI_E_ARGUMENT, "delta_normal must be strictly positive.");
        break;
    case FSI_PHASE_UWCDI:  // PhaseFieldMethods.jl:100-119
        if (!(m.uwcdi_alpha >= 1.0) || !std::isfinite(m.uwcdi_alpha))
            return fail(c, FSI_E_ARGUMENT, "alpha must be finite and >= 1 for the published UWCDI stability range.");
        if (!(m.uwcdi_delta_normal > 0.0)) return fail(c, FSI_E_ARGUMENT, "delta_normal must be strictly positive.");
        if (m.uwcdi_normal_method < FSI_UWCDI_NORMAL_STANDARD || m.uwcdi_normal_method > FSI_UWCDI_NORMAL_SMOOTHED9)
            return fail(c, FSI_E_ARGUMENT, "normal_reconstruction must be :standard, :youngs, or :smoothed9.");
        if (m.uwcdi_normal_method == FSI_UWCDI_NORMAL_SMOOTHED9) {  // _validate_smoothed9_boundaries :347-362
            const fsi_bc_side* ps[4] = {&c->bcs.phi.left, &c->bcs.phi.right, &c->bcs.phi.bottom, &c->bcs.phi.top};
            const char* pn[4] = {"left", "right", "bottom", "top"};
            for (int q = 0; q < 4; ++q) {
                if (ps[q]->kind == FSI_BC_PERIODIC) continue;
                if (ps[q]->kind == FSI_BC_NEUMANN && std::fabs(ps[q]->value) <= 100 * 2.220446049250313e-16) continue;
                return fail(c, FSI_E_ARGUMENT,
                            "smoothed9 normal reconstruction requires homogeneous Neumann or periodic phase "
                            "boundaries; %s does not.", pn[q]);
            }
        }
        // :youngs needs a uniform grid (:285-296): every device grid is uniform
        break;
    default: return fail(c, FSI_E_UNSUPPORTED, "phase method %d is not supported on the device.", m.phase_method);
    }
    switch (m.capillarity_kind) {
    case FSI_CAPILLARITY_NONE: break;
    case FSI_CAPILLARITY_DIFFUSE_AREA:  // Capillarity.jl:21-41
        if (!std::isfinite(m.capillary_coefficient)) return fail(c, FSI_E_ARGUMENT, "coefficient must be finite.");
        if (!(m.capillary_coefficient >= 0.0)) return fail(c, FSI_E_ARGUMENT, "coefficient must be nonnegative.");
        if (!std::isfinite(m.capillary_delta_normal)) return fail(c, FSI_E_ARGUMENT, "delta_normal must be finite.");
        if (!(m.capillary_delta_normal > 0.0)) return fail(c, FSI_E_ARGUMENT, "delta_normal must be positive.");
        if (m.capillary_smoothing_passes < 0) return fail(c, FSI_E_ARGUMENT, "smoothing_passes must be nonnegative.");
        break;
    default: return fail(c, FSI_E_UNSUPPORTED, "capillarity %d is not supported on the device.", m.capillarity_kind);
    }
    if (m.viscosity_kind == FSI_VISC_CONSTANT) {
        if (!(m.viscosity > 0.0) || !std::isfinite(m.viscosity))
            return fail(c, FSI_E_ARGUMENT, "Viscosity must be strictly positive.");
        // _require_supported_viscous_boundaries, Operators.jl:91-107
        const fsi_bc_side* vs[8] = {&c->bcs.U.left, &c->bcs.U.right, &c->bcs.U.bottom, &c->bcs.U.top,
                                    &c->bcs.V.left, &c->bcs.V.right, &c->bcs.V.bottom, &c->bcs.V.top};
        for (int q = 0; q < 8; ++q)
            if (vs[q]->kind == FSI_BC_NEUMANN)
                return fail(c, FSI_E_ARGUMENT,
                            "%s-%s uses Neumann. The current compact viscous operator supports Dirichlet or "
                            "Periodic velocity boundaries.", q < 4 ? "U" : "V", side_name(q & 3));
    } else if (m.viscosity_kind != FSI_VISC_NONE)
        return fail(c, FSI_E_UNSUPPORTED, "viscosity model %d is not supported on the device.", m.viscosity_kind);
    if (m.elastic_kind == FSI_ELASTIC_LINEAR) {
        if (!(m.shear_modulus >= 0.0) || !std::isfinite(m.shear_modulus))
            return fail(c, FSI_E_ARGUMENT, "Elastic-stress coefficient must be finite and non-negative.");
    } else if (m.elastic_kind != FSI_ELASTIC_NONE)
        return fail(c, FSI_E_UNSUPPORTED, "elastic model %d is not supported on the device.", m.elastic_kind);
    if (m.relaxation_kind == FSI_RELAX_MAXWELL) {
        if (!(m.relaxation_rate >= 0.0) || !std::isfinite(m.relaxation_rate))
            return fail(c, FSI_E_ARGUMENT, "Maxwell relaxation rate must be finite and non-negative.");
    } else if (m.relaxation_kind != FSI_RELAX_NONE)
        return fail(c, FSI_E_UNSUPPORTED, "relaxation model %d is not supported on the device.", m.relaxation_kind);
    if (m.body_force_kind == FSI_FORCE_KOLMOGOROV) {
        if (!(m.force_wavenumber > 0.0)) return fail(c, FSI_E_ARGUMENT, "Kolmogorov-force wavenumber must be positive.");
    } else if (m.body_force_kind != FSI_FORCE_NONE && m.body_force_kind != FSI_FORCE_CONSTANT)
        return fail(c, FSI_E_UNSUPPORTED, "body force %d is not supported on the device.", m.body_force_kind);
    const fsi_numerics& n = c->num;
    if (n.extension_kind == FSI_EXT_GEOMETRIC) {  // StructureTensor.jl:106-131
        if (!(n.ext_max_rate >= 0.0) || !std::isfinite(n.ext_max_rate))
            return fail(c, FSI_E_ARGUMENT, "Geometric extension rate must be non-negative.");
        if (!(n.ext_exponent > 0.0) || !std::isfinite(n.ext_exponent))
            return fail(c, FSI_E_ARGUMENT, "Geometric extension exponent must be positive.");
        if (!(n.ext_cutoff >= 0.0 && n.ext_cutoff < 1.0))
            return fail(c, FSI_E_ARGUMENT, "Geometric extension safety cutoff must lie in [0, 1).");
    } else if (n.extension_kind == FSI_EXT_IDENTITY) {
        if (!(n.ext_cutoff >= 0.0 && n.ext_cutoff <= 1.0))
            return fail(c, FSI_E_ARGUMENT, "Structure-extension cutoff must lie in [0, 1].");
    } else if (n.extension_kind != FSI_EXT_NONE)
        return fail(c, FSI_E_UNSUPPORTED, "extension model %d is not supported on the device.", n.extension_kind);
    // growth and resource closures (GrowthModels.jl:55-166, TransportConstitutive.jl:186-295)
    switch (m.growth_source_kind) {
    case FSI_GROWTH_NONE: break;
    case FSI_GROWTH_CONSTANT:
        if (!std::isfinite(m.growth_rate)) return fail(c, FSI_E_ARGUMENT, "Intrinsic growth rate must be finite.");
        break;
    case FSI_GROWTH_SATURATING:
        if (!std::isfinite(m.growth_rate)) return fail(c, FSI_E_ARGUMENT, "Resource-limited growth rate must be finite.");
        if (!std::isfinite(m.growth_half_saturation))
            return fail(c, FSI_E_ARGUMENT, "Growth half-saturation constant must be finite.");
        if (!(m.growth_half_saturation > 0.0))
            return fail(c, FSI_E_ARGUMENT, "Growth half-saturation constant must be strictly positive.");
        break;
    default: return fail(c, FSI_E_UNSUPPORTED, "growth source %d is not supported on the device.", m.growth_source_kind);
    }
    if (m.growth_tensor_kind == FSI_GROWTH_TENSOR_DIRECTIONAL) {
        if (!(std::isfinite(m.growth_direction_x) && std::isfinite(m.growth_direction_y)))
            return fail(c, FSI_E_ARGUMENT, "Directional-growth components must be finite.");
        if (!(std::hypot(m.growth_direction_x, m.growth_direction_y) > 0.0))
            return fail(c, FSI_E_ARGUMENT, "Directional growth requires a non-zero direction vector.");
    } else if (m.growth_tensor_kind != FSI_GROWTH_TENSOR_ISOTROPIC)
        return fail(c, FSI_E_UNSUPPORTED, "growth tensor %d is not supported on the device.", m.growth_tensor_kind);
    switch (m.diffusivity_kind) {
    case FSI_DIFFUSION_NONE: break;
    case FSI_DIFFUSION_CONSTANT:
        if (!std::isfinite(m.diffusivity_fluid)) return fail(c, FSI_E_ARGUMENT, "Diffusivity must be finite.");
        if (!(m.diffusivity_fluid > 0.0)) return fail(c, FSI_E_ARGUMENT, "Diffusivity must be strictly positive.");
        break;
    case FSI_DIFFUSION_HARMONIC:
        if (!std::isfinite(m.diffusivity_fluid)) return fail(c, FSI_E_ARGUMENT, "Fluid diffusivity must be finite.");
        if (!std::isfinite(m.diffusivity_structure))
            return fail(c, FSI_E_ARGUMENT, "Structure diffusivity must be finite.");
        if (!(m.diffusivity_fluid > 0.0)) return fail(c, FSI_E_ARGUMENT, "Fluid diffusivity must be strictly positive.");
        if (!(m.diffusivity_structure > 0.0))
            return fail(c, FSI_E_ARGUMENT, "Structure diffusivity must be strictly positive.");
        break;
    default: return fail(c, FSI_E_UNSUPPORTED, "diffusivity %d is not supported on the device.", m.diffusivity_kind);
    }
    if (m.consumption_kind == FSI_CONSUMPTION_SATURATING) {
        if (!std::isfinite(m.consumption_rate)) return fail(c, FSI_E_ARGUMENT, "Consumption rate must be finite.");
        if (!std::isfinite(m.consumption_half_saturation))
            return fail(c, FSI_E_ARGUMENT, "Consumption half-saturation constant must be finite.");
        if (!(m.consumption_rate >= 0.0)) return fail(c, FSI_E_ARGUMENT, "Consumption rate must be non-negative.");
        if (!(m.consumption_half_saturation > 0.0))
            return fail(c, FSI_E_ARGUMENT, "Consumption half-saturation constant must be strictly positive.");
    } else if (m.consumption_kind != FSI_CONSUMPTION_NONE)
        return fail(c, FSI_E_UNSUPPORTED, "resource consumption %d is not supported on the device.", m.consumption_kind);
    if (n.resource_method != FSI_SOLVER_DIRECT && n.resource_method != FSI_SOLVER_GMRES)
        return fail(c, FSI_E_ARGUMENT, "resource_method must be :direct or :gmres.");
    if (!(n.projection_rtol > 0.0)) return fail(c, FSI_E_ARGUMENT, "projection_reltol must be positive.");
    if (n.projection_maxiter <= 0) return fail(c, FSI_E_ARGUMENT, "projection_maxiter must be positive.");
    if (!(n.velocity_rtol > 0.0)) return fail(c, FSI_E_ARGUMENT, "Linear-solver reltol must be positive.");
    if (n.velocity_maxiter <= 0) return fail(c, FSI_E_ARGUMENT, "Linear-solver maxiter must be positive.");
    if (n.velocity_method != FSI_SOLVER_DIRECT && n.velocity_method != FSI_SOLVER_GMRES)
        return fail(c, FSI_E_ARGUMENT, "velocity_method must be :direct or :gmres.");
    if (!(n.fallback_rtol > 0.0)) return fail(c, FSI_E_ARGUMENT, "fallback_reltol must be positive.");
    if (!(n.stage1_atol >= 0.0) || !(n.stage1_rtol >= 0.0))
        return fail(c, FSI_E_ARGUMENT, "stage-1 constraint tolerances must be non-negative.");
    if (n.precision != FSI_PRECISION_FP64 && n.precision != FSI_PRECISION_MIXED)
        return fail(c, FSI_E_ARGUMENT, "precision must be FSI_PRECISION_FP64 or FSI_PRECISION_MIXED.");
    return FSI_OK;
}

extern "C" fsi_status fsi_context_create(const fsi_grid* grid, const fsi_bcs* bcs, const fsi_models* models,
                                         const fsi_numerics* numerics, fsi_context** out) {
    if (!out) return FSI_E_ARGUMENT;
    *out = nullptr;
    fsi_context* c = new fsi_context();
    *out = c;
    if (!grid || !bcs || !models || !numerics) return fail(c, FSI_E_ARGUMENT, "null configuration pointer");
    c->grid = *grid;
    c->bcs = *bcs;
    c->models = *models;
    c->num = *numerics;
    {  // FSI_STEP_GRAPH=0 keeps every step on the polling path
        const char* e = std::getenv("FSI_STEP_GRAPH");  // 0 polling, 1 planned graph, 2 planned direct
        if (e && (e[0] == '0' || e[0] == '1' || e[0] == '2')) c->use_graph = e[0] - '0';
        e = std::getenv("FSI_STAGE_WARM");
        if (e && e[0] == '0') c->vel_warm = 0;
        e = std::getenv("FSI_SPECTRAL");
        if (e && e[0] == '0') c->use_spectral = 0;
    }
    fsi_status rc = validate(c);
    if (rc != FSI_OK) return rc;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(c, FSI_E_HIP, "no HIP device is available");
    HIPCHECK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    {
        const char* e = std::getenv("FSI_DUAL_STREAM");  // 0: every kernel of a step on one stream
        c->dual = (e && e[0] == '0') ? 0 : 1;
    }
    // the side stream is created by the first step that uses it (enqueue_step):
    // contexts that only step in ensembles never hold one, and their main
    // streams spread over the hardware queues
    HIPCHECK(c, hipGetDevice(&c->device));
    if ((rc = scratch_lookup(c)) != FSI_OK) return rc;

    const int nx = grid->nx, ny = grid->ny;
    c->P = ((nx + 1 + 31) / 32) * 32;
    c->felems = (size_t)(ny + 1) * c->P;
    Params& p = c->prm;
    p.nx = nx;
    p.ny = ny;
    p.P = c->P;
    p.px = grid->periodic_x != 0;
    p.py = grid->periodic_y != 0;
    // generate_grid(:uniform): h = L/N on both the face and the centre grids
    p.hx = grid->lx / nx;
    p.hy = grid->ly / ny;
    p.ihx = 1.0 / p.hx;
    p.ihy = 1.0 / p.hy;
    p.bU = to_bcset(bcs->U);
    p.bV = to_bcset(bcs->V);
    p.bQ = to_bcset(bcs->correction);
    p.bPhi = to_bcset(bcs->phi);
    p.bSxx = to_bcset(bcs->sigma_xx);
    p.bSyy = to_bcset(bcs->sigma_yy);
    p.bSxy = to_bcset(bcs->sigma_xy);
    const fsi_models& m = *models;
    p.phase = m.phase_method;
    p.cdi_G = m.cdi_inverse_peclet;
    p.cdi_eps = m.cdi_epsilon;
    p.cdi_d2 = m.cdi_delta_normal * m.cdi_delta_normal;
    p.acdi_gamma = m.acdi_gamma;
    p.acdi_eps = m.acdi_epsilon;
    p.acdi_floor = m.acdi_psi_floor;
    p.acdi_d2 = m.acdi_delta_normal * m.acdi_delta_normal;
    p.acdi_dyn = m.acdi_dynamic_gamma;
    p.uw_alpha = m.uwcdi_alpha;
    p.uw_d2 = m.uwcdi_delta_normal * m.uwcdi_delta_normal;
    p.uw_normal = p.phase == FSI_PHASE_UWCDI ? m.uwcdi_normal_method : FSI_UWCDI_NORMAL_STANDARD;
    p.cap_coef = m.capillary_coefficient;
    p.cap_d2 = m.capillary_delta_normal * m.capillary_delta_normal;
    p.cap_passes = m.capillary_smoothing_passes;
    p.visc = m.viscosity_kind == FSI_VISC_CONSTANT;
    p.mu = m.viscosity;
    p.elastic = m.elastic_kind == FSI_ELASTIC_LINEAR;
    p.Gmod = m.shear_modulus;
    p.relax = m.relaxation_kind == FSI_RELAX_MAXWELL;
    p.relax_rate = m.relaxation_rate;
    p.ext = numerics->extension_kind;
    p.ext_rate = numerics->ext_max_rate;
    p.ext_exp = numerics->ext_exponent;
    p.ext_cutoff = numerics->ext_cutoff;
    c->has_reset = (p.ext == FSI_EXT_IDENTITY) || (p.ext == FSI_EXT_GEOMETRIC && p.ext_cutoff > 0.0);
    p.force = m.body_force_kind;
    p.ffx = m.force_fx;
    p.ffy = m.force_fy;
    // ACDI psi boundary setup (_psi_boundary_setup, PhaseFieldMethods.jl:493-529)
    c->bpsi = p.bPhi;
    BcSide* ps[4] = {&c->bpsi.l, &c->bpsi.r, &c->bpsi.b, &c->bpsi.t};
    for (int q = 0; q < 4; ++q) {
        if (ps[q]->kind == FSI_BC_NEUMANN) {
            if (p.phase == FSI_PHASE_ACDI && std::fabs(ps[q]->value) > 100 * 2.220446049250313e-16)
                return fail(c, FSI_E_ARGUMENT, "ACDI currently supports only homogeneous Neumann phase gradients.");
            ps[q]->value = 0.0;
        } else if (ps[q]->kind == FSI_BC_DIRICHLET) {
            double q0 = std::min(std::max(ps[q]->value, 0.0), 1.0), e = m.acdi_psi_floor;
            ps[q]->value = m.acdi_epsilon * std::log((q0 + e) / (1.0 - q0 + e));
        }
    }
    const fsi_bc_setup& bq = bcs->correction;
    c->has_null = !(bq.left.kind == FSI_BC_DIRICHLET || bq.right.kind == FSI_BC_DIRICHLET ||
                    bq.bottom.kind == FSI_BC_DIRICHLET || bq.top.kind == FSI_BC_DIRICHLET);

    // constant device data
    std::vector<double> host((size_t)ny, 0.0);
    HIPCHECK(c, dalloc(c, &c->force_row, (size_t)ny));
    if (m.body_force_kind == FSI_FORCE_KOLMOGOROV) {  // TransportConstitutive.jl:385-397
        double y = 0.0;
        for (int j = 0; j < ny; ++j) {
            y += 0.5 * p.hy;
            host[j] = m.force_amplitude * std::sin(m.force_wavenumber * y + m.force_phase);
            y += 0.5 * p.hy;
        }
        HIPCHECK(c, hipMemcpyAsync(c->force_row, host.data(), ny * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    p.force_row = c->force_row;
    if (p.uw_normal == FSI_UWCDI_NORMAL_SMOOTHED9) {
        HIPCHECK(c, dalloc(c, &c->uw_smooth, c->felems));
        p.uw_field = c->uw_smooth;
    }
    if (m.capillarity_kind == FSI_CAPILLARITY_DIFFUSE_AREA) {
        HIPCHECK(c, dalloc(c, &c->cap_buf, 7 * c->felems));
        p.cap_fx = c->cap_buf;
        p.cap_fy = c->cap_buf + c->felems;
        p.cap_work = c->cap_buf + 2 * c->felems;
    }
    p.exnb = grid_blocks(p);
    HIPCHECK(c, dalloc(c, &c->ext, (size_t)EX_COUNT * p.exnb));
    // the growth slots stay zero (neutral) where the no-growth step never writes them
    HIPCHECK(c, hipMemsetAsync(c->ext, 0, (size_t)EX_COUNT * p.exnb * sizeof(double), c->stream));
    {
        const fsi_models& m = c->models;
        GrowthP& g = c->gp;
        c->growth = m.growth_source_kind != FSI_GROWTH_NONE || m.diffusivity_kind != FSI_DIFFUSION_NONE ||
                    m.consumption_kind != FSI_CONSUMPTION_NONE;
        g.src = m.growth_source_kind;
        g.rate = m.growth_rate;
        g.K = m.growth_half_saturation;
        g.kxx = 0.5;  // IsotropicGrowthTensor: G/2 I (GrowthModels.jl:268-289)
        g.kxy = 0.0;
        g.kyy = 0.5;
        if (m.growth_tensor_kind == FSI_GROWTH_TENSOR_DIRECTIONAL) {  // G t (x) t, t normalised (:150-165, :291-318)
            const double mag = std::hypot(m.growth_direction_x, m.growth_direction_y);
            const double tx = m.growth_direction_x / mag, ty = m.growth_direction_y / mag;
            g.kxx = tx * tx;
            g.kxy = tx * ty;
            g.kyy = ty * ty;
        }
        g.diff = m.diffusivity_kind;
        g.Df = m.diffusivity_fluid;
        g.Ds = m.diffusivity_structure;
        g.cons = m.consumption_kind;
        g.crate = m.consumption_rate;
        g.cK = m.consumption_half_saturation;
        g.visc = p.visc;
        g.mu = p.mu;
        g.bC = to_bcset(c->bcs.resource);
    }
    {
        const int nv = std::max(DIAG_STATE_VALUES, std::max(DIAG_TENSOR_VALUES, DIAG_GEOMETRY_VALUES));
        HIPCHECK(c, dalloc(c, &c->diag, (size_t)nv * p.exnb));
        HIPCHECK(c, dalloc(c, &c->diag_out, (size_t)nv));
    }
    p.exbuf = c->ext;
    // b_bc of the projection Laplacian: D * gradient_bc (Operators.jl:361-384)
    {
        std::vector<double> bc((size_t)nx * ny, 0.0);
        auto gx = [&](int i, int j) -> double {  // boundary part of the x-face gradient
            (void)j;
            if (p.px) return 0.0;
            if (i == 0) return bq.left.kind == FSI_BC_DIRICHLET ? -bq.left.value / (p.hx / 2.0) : -bq.left.value;
            if (i == nx) return bq.right.kind == FSI_BC_DIRICHLET ? bq.right.value / (p.hx / 2.0) : bq.right.value;
            return 0.0;
        };
        auto gy = [&](int i, int j) -> double {
            (void)i;
            if (p.py) return 0.0;
            if (j == 0) return bq.bottom.kind == FSI_BC_DIRICHLET ? -bq.bottom.value / (p.hy / 2.0) : -bq.bottom.value;
            if (j == ny) return bq.top.kind == FSI_BC_DIRICHLET ? bq.top.value / (p.hy / 2.0) : bq.top.value;
            return 0.0;
        };
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                bc[(size_t)j * nx + i] = (gx(i + 1, j) - gx(i, j)) / p.hx + (gy(i, j + 1) - gy(i, j)) / p.hy;
        HIPCHECK(c, dalloc(c, &c->lap_bc, c->felems));
        rc = upload2d(c, c->lap_bc, bc.data(), nx, ny);
        if (rc != FSI_OK) return rc;
    }
    HIPCHECK(c, dalloc(c, &c->bvisc_u, c->felems));
    HIPCHECK(c, dalloc(c, &c->bvisc_v, c->felems));
    if (p.visc) launch_visc_boundary(c->stream, p, nullptr, c->bvisc_u, c->bvisc_v);
    // multigrid geometry: halve while both sizes are even, >= 4, and the level
    // is larger than the single-workgroup coarse solve
    {
        MgGeom g{};
        g.nx = nx;
        g.ny = ny;
        g.P = c->P;
        g.px = p.px;
        g.py = p.py;
        g.dl = bq.left.kind == FSI_BC_DIRICHLET;
        g.dr = bq.right.kind == FSI_BC_DIRICHLET;
        g.db = bq.bottom.kind == FSI_BC_DIRICHLET;
        g.dt = bq.top.kind == FSI_BC_DIRICHLET;
        g.ihx2 = 1.0 / (p.hx * p.hx);
        g.ihy2 = 1.0 / (p.hy * p.hy);
        // degree-2 Chebyshev smoothing on [1/3, 2] of the spectrum of D^-1 A
        // (eigenvalues in [0, 2]): weights 1/r, r = 7/6 -+ (5/6) cos(pi/4)
        g.om1 = 1.0 / (7.0 / 6.0 + 5.0 / 6.0 * std::sqrt(0.5));
        g.om2 = 1.0 / (7.0 / 6.0 - 5.0 / 6.0 * std::sqrt(0.5));
        if (const char* e = std::getenv("FSI_MG_SMOOTH")) {  // "w1,w2" overrides (tuning)
            double a = 0.0, b = 0.0;
            if (std::sscanf(e, "%lf,%lf", &a, &b) == 2 && a > 0.0 && b > 0.0) {
                g.om1 = a;
                g.om2 = b;
            }
        }
        g.d0 = g.ihx2 * 2.0 + g.ihy2 * 2.0;
        g.id0 = 1.0 / g.d0;
        c->geo.push_back(g);
        const int co_cells = coarse_max_cells();
        while ((int)c->geo.size() < kMaxLevels && (long)g.nx * g.ny > co_cells && !(g.nx & 1) && !(g.ny & 1) &&
               g.nx >= 8 && g.ny >= 8) {
            g.nx /= 2;
            g.ny /= 2;
            g.P = ((g.nx + 31) / 32) * 32;
            g.ihx2 *= 0.25;
            g.ihy2 *= 0.25;
            g.d0 = g.ihx2 * 2.0 + g.ihy2 * 2.0;
            g.id0 = 1.0 / g.d0;
            c->geo.push_back(g);
        }
        std::vector<double> dense;
        if (coarse_plan(c->geo.back(), co_cells, c->cplan, dense) && c->cplan.dense_n > 0) {
            HIPCHECK(c, dalloc(c, &c->co_dense, dense.size()));
            HIPCHECK(c, hipMemcpyAsync(c->co_dense, dense.data(), dense.size() * sizeof(double),
                                       hipMemcpyHostToDevice, c->stream));
            HIPCHECK(c, hipStreamSynchronize(c->stream));
            c->cplan.dense = c->co_dense;
            std::vector<float> df(dense.begin(), dense.end());
            HIPCHECK(c, hipMalloc((void**)&c->co_dense_f, df.size() * sizeof(float)));
            HIPCHECK(c, hipMemcpyAsync(c->co_dense_f, df.data(), df.size() * sizeof(float), hipMemcpyHostToDevice,
                                       c->stream));
            HIPCHECK(c, hipStreamSynchronize(c->stream));
            c->cplan.dense_f = c->co_dense_f;
        }
        const MgGeom& g0 = c->geo[0];
        const int kx = spec_kind(g0.nx, g0.px, g0.dl, g0.dr), ky = spec_kind(g0.ny, g0.py, g0.db, g0.dt);
        if (c->use_spectral && kx != SPEC_NONE && ky != SPEC_NONE) {
            const std::vector<double> tx = spec_tables(g0.nx, kx, g0.ihx2), ty = spec_tables(g0.ny, ky, g0.ihy2);
            const size_t npart = (size_t)SPEC_PART_ROWS * ((g0.ny + 1) / 2);
            HIPCHECK(c, dalloc(c, &c->spec_tab, tx.size() + ty.size() + npart));
            c->spec_part = c->spec_tab + tx.size() + ty.size();
            HIPCHECK(c, hipMemcpyAsync(c->spec_tab, tx.data(), tx.size() * sizeof(double), hipMemcpyHostToDevice,
                                       c->stream));
            HIPCHECK(c, hipMemcpyAsync(c->spec_tab + tx.size(), ty.data(), ty.size() * sizeof(double),
                                       hipMemcpyHostToDevice, c->stream));
            HIPCHECK(c, hipStreamSynchronize(c->stream));
            auto axis = [](int n, int kind, double* t) {
                SpecAxis a{};
                a.n = n;
                a.r3 = (n & (n - 1)) ? 1 : 0;  // 3 * 2^m
                a.m = 0;
                while (((a.r3 ? 3 : 1) << a.m) < n) ++a.m;
                a.kind = kind;
                a.wfft = reinterpret_cast<const double2*>(t);
                a.wpost = reinterpret_cast<const double2*>(t + n);
                a.mu = t + 3 * n;
                return a;
            };
            c->spec.x = axis(g0.nx, kx, c->spec_tab);
            c->spec.y = axis(g0.ny, ky, c->spec_tab + tx.size());
            c->spec.enabled = spec_prepare(c->spec) ? 1 : 0;
        }
    }
    HIPCHECK(c, dalloc(c, &c->partials, partials_size(p)));
    HIPCHECK(c, hipMalloc((void**)&c->st, sizeof(Status)));
    HIPCHECK(c, hipMemsetAsync(c->st, 0, sizeof(Status), c->stream));
    HIPCHECK(c, hipHostMalloc((void**)&c->st_host, sizeof(Status), hipHostMallocDefault));
    std::memset(c->st_host, 0, sizeof(Status));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipGetLastError());
    return FSI_OK;
}

extern "C" void fsi_context_destroy(fsi_context* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    double* bufs[] = {c->force_row, c->lap_bc, c->bvisc_u,  c->bvisc_v, c->partials, c->co_dense,
                      c->spec_tab,  c->ext,    c->diag,     c->diag_out, c->uw_smooth, c->cap_buf};
    for (double* b : bufs)
        if (b) (void)hipFree(b);
    if (c->co_dense_f) (void)hipFree(c->co_dense_f);
    if (c->batch_dev) (void)hipFree(c->batch_dev);
    if (c->batch_host) (void)hipHostFree(c->batch_host);
    if (c->st) (void)hipFree(c->st);
    if (c->st_host) (void)hipHostFree(c->st_host);
    if (c->side) (void)hipStreamSynchronize(c->side);
    for (int k = 0; k < 3; ++k) {
        if (c->ev_proj[k]) (void)hipEventDestroy(c->ev_proj[k]);
        if (c->ev_cells[k]) (void)hipEventDestroy(c->ev_cells[k]);
        if (c->ev_x0[k]) (void)hipEventDestroy(c->ev_x0[k]);
    }
    if (c->ev_start) (void)hipEventDestroy(c->ev_start);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* fsi_last_error(const fsi_context* c) { return c ? c->err.c_str() : "null context"; }

// ---- driver diagnostics -----------------------------------------------------
static fsi_status diag_fetch(fsi_context* c, double* v, int n) {
    HIPCHECK(c, hipMemcpyAsync(v, c->diag_out, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return FSI_OK;
}
// pass 1 of the diagnostics: sums 0..10 then maxima 11..16 (k_diag_state)
static fsi_status diag_state(fsi_context* c, const fsi_state* s, double density, double* v) {
    launch_diag_state(c->stream, c->prm, s->U, s->V, s->phi, c->grid.lx, c->grid.ly, density, c->diag,
                      c->diag_out);
    return diag_fetch(c, v, DIAG_STATE_VALUES);
}
static double nan_if(double count, double x) { return count > 0.0 ? NAN : x; }

extern "C" fsi_status fsi_velocity_extrema(fsi_context* c, const fsi_state* s, double* max_abs_u, double* max_abs_v) {
    if (!c || !s || s->ctx != c || !max_abs_u || !max_abs_v) return FSI_E_ARGUMENT;
    double v[DIAG_STATE_VALUES];
    fsi_status rc = diag_state(c, s, 1.0, v);
    if (rc != FSI_OK) return rc;
    *max_abs_u = nan_if(v[10], v[11 + 4]);
    *max_abs_v = nan_if(v[10], v[11 + 5]);
    return FSI_OK;
}

// phase_centroid: mass > eps(Float64) || error(...) (:497-498)
static fsi_status centroid_of(fsi_context* c, const double* v, double& xc, double& yc, double& mass) {
    mass = v[0];
    if (!(mass > 2.220446049250313e-16)) return fail(c, FSI_E_SOLVER, "Phase integral is too small to define a centroid.");
    xc = v[1] / mass;
    yc = v[2] / mass;
    return FSI_OK;
}

// ---- the allocation-free centroid ------------------------------------------
// `scratch` holds the per-block partials of k_diag_state (DIAG_STATE_VALUES
// rows of up to FSI_NOALLOC_MAX_TILES values) and, after them, the values of
// its final reduction.  Every block writes its own slot and the final
// reduction writes every value, so nothing needs clearing between calls.  In
// host code `scratch` only names the symbol: kernels get its device address,
// looked up once per device with hipGetSymbolAddress (fsi_context_create).
constexpr size_t kScratchValues = (size_t)DIAG_STATE_VALUES * FSI_NOALLOC_MAX_TILES + DIAG_STATE_VALUES;
static __device__ __align__(256) double scratch[kScratchValues];
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mutex;

static fsi_status scratch_lookup(fsi_context* c) {
    if (c->device < 0 || c->device >= kMaxDevices)
        return fail(c, FSI_E_HIP, "device index %d is beyond the %d the library tracks", c->device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[c->device]) return FSI_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(c, FSI_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[c->device] = static_cast<double*>(ptr);
    return FSI_OK;
}

extern "C" fsi_status fsi_phase_centroid_noalloc(fsi_context* c, const fsi_state* s, double* result, void* stream) {
    if (!c || !s || s->ctx != c || !result) return FSI_E_ARGUMENT;
    if (c->prm.exnb > FSI_NOALLOC_MAX_TILES)
        return fail(c, FSI_E_ARGUMENT,
                    "fsi_phase_centroid_noalloc covers grids of at most %d tiles of 64 x 16 cells; %d x %d has %d.",
                    FSI_NOALLOC_MAX_TILES, c->grid.nx, c->grid.ny, c->prm.exnb);
    double* buf = g_scratch_ptr[c->device];  // set when the context was created
    if (!buf) return fail(c, FSI_E_HIP, "the device scratch of fsi_phase_centroid_noalloc was not looked up");
    hipStream_t hs = stream ? static_cast<hipStream_t>(stream) : c->stream;
    double* reduced = buf + (size_t)DIAG_STATE_VALUES * FSI_NOALLOC_MAX_TILES;
    launch_diag_state(hs, c->prm, s->U, s->V, s->phi, c->grid.lx, c->grid.ly, 1.0, buf, reduced);
    launch_centroid_result(hs, reduced, result);
    HIPCHECK(c, hipGetLastError());
    return FSI_OK;
}

extern "C" fsi_status fsi_phase_centroid(fsi_context* c, const fsi_state* s, double* x_centroid, double* y_centroid,
                                         double* phase_integral) {
    if (!c || !s || s->ctx != c || !x_centroid || !y_centroid || !phase_integral) return FSI_E_ARGUMENT;
    double v[DIAG_STATE_VALUES];
    fsi_status rc = diag_state(c, s, 1.0, v);
    if (rc != FSI_OK) return rc;
    return centroid_of(c, v, *x_centroid, *y_centroid, *phase_integral);
}

// tensor_region_diagnostics (:753-822) from the sums 0..5 and maxima 6..11
// of k_diag_tensor
static void fill_region(const double* v, fsi_tensor_region& r) {
    const double cells = v[0], area = v[1];
    if (cells == 0.0 || area == 0.0) {
        r = fsi_tensor_region{0, 0.0, NAN, NAN, NAN, NAN, NAN, NAN, 0, NAN, NAN};
        return;
    }
    const double nan = v[5];
    r.cells = (int64_t)cells;
    r.area = area;
    r.determinant_min = nan_if(nan, -v[6 + 1]);
    r.determinant_max = nan_if(nan, v[6 + 0]);
    r.max_abs_determinant_error = nan_if(nan, v[6 + 2]);
    r.mean_abs_determinant_error = v[2] / area;
    r.lambda_min = nan_if(nan, -v[6 + 4]);
    r.lambda_max = nan_if(nan, v[6 + 3]);
    r.non_spd_cells = (int64_t)v[4];
    r.max_effective_stress = nan_if(nan, v[6 + 5]);
    r.mean_effective_stress = v[3] / area;
}

extern "C" fsi_status fsi_diagnostics_compute(fsi_context* c, const fsi_state* s, const fsi_diag_params* prm,
                                              fsi_diagnostics* out) {
    if (!c || !s || s->ctx != c || !prm || !out) return FSI_E_ARGUMENT;
    const fsi_diag_params& q = *prm;
    if (q.deformation_mode_max < 0 || q.deformation_mode_max > FSI_DIAG_MAX_MODES)
        return fail(c, FSI_E_ARGUMENT, "deformation_mode_max must lie in 0..%d.", FSI_DIAG_MAX_MODES);
    if (!(0.0 <= q.sigma_safety_cutoff && q.sigma_safety_cutoff < 1.0))
        return fail(c, FSI_E_ARGUMENT, "sigma_safety_cutoff must lie in [0,1).");
    if (!(0.0 < q.sigma_interface_cutoff && q.sigma_interface_cutoff < q.sigma_core_cutoff &&
          q.sigma_core_cutoff <= 1.0))
        return fail(c, FSI_E_ARGUMENT, "Require 0 < sigma_interface_cutoff < sigma_core_cutoff <= 1.");
    if (!(q.sigma_safety_cutoff < q.sigma_interface_cutoff))
        return fail(c, FSI_E_ARGUMENT, "sigma_safety_cutoff must be smaller than sigma_interface_cutoff.");
    *out = fsi_diagnostics{};
    double v1[DIAG_STATE_VALUES];
    fsi_status rc = diag_state(c, s, q.density, v1);
    if (rc != FSI_OK) return rc;
    double xc, yc, mass;
    if ((rc = centroid_of(c, v1, xc, yc, mass)) != FSI_OK) return rc;
    const double* m1 = v1 + 11;
    // phase_bound_diagnostics
    out->phi_min = nan_if(v1[8], -m1[1]);
    out->phi_max = nan_if(v1[8], m1[0]);
    out->lower_violation = std::isnan(out->phi_min) ? NAN : std::fmax(0.0, -out->phi_min);
    out->upper_violation = std::isnan(out->phi_max) ? NAN : std::fmax(0.0, out->phi_max - 1.0);
    out->out_of_bounds_cells = (int64_t)v1[7];
    // flow_diagnostics (max_speed starts from 0.0)
    out->maximum_speed = nan_if(v1[9], std::fmax(0.0, m1[2]));
    out->kinetic_energy = v1[4];
    out->divergence_linf = nan_if(v1[9], m1[3]);
    out->divergence_l2 = std::sqrt(v1[5] / v1[6]);
    out->max_abs_u = nan_if(v1[10], m1[4]);
    out->max_abs_v = nan_if(v1[10], m1[5]);
    // phase_geometry_diagnostics: second pass about the centroid, the modes
    // in launches of diag_modes_per_launch()
    const int nmax = q.deformation_mode_max, per = diag_modes_per_launch();
    double perimeter = 0.0, cxx = 0.0, cyy = 0.0, cxy = 0.0;
    for (int n0 = 0; n0 == 0 || n0 < nmax; n0 += per) {
        double v2[DIAG_GEOMETRY_VALUES];
        launch_diag_geometry(c->stream, c->prm, s->phi, c->grid.lx, c->grid.ly, xc, yc, n0, nmax, c->diag,
                             c->diag_out);
        if ((rc = diag_fetch(c, v2, DIAG_GEOMETRY_VALUES)) != FSI_OK) return rc;
        if (n0 == 0) {
            perimeter = v2[0];
            cxx = v2[1];
            cyy = v2[2];
            cxy = v2[3];
        }
        for (int k = 0; k < per && n0 + k < nmax; ++k) {
            out->Rc[n0 + k] = v2[4 + 2 * k];
            out->Rs[n0 + k] = v2[5 + 2 * k];
        }
    }
    cxx /= mass;
    cyy /= mass;
    cxy /= mass;
    const double gap = std::hypot(cxx - cyy, 2.0 * cxy);
    out->x_centroid = xc;
    out->y_centroid = yc;
    out->phase_integral = mass;
    out->perimeter_proxy = perimeter;
    out->mixed_layer_measure = v1[3];
    out->covariance_major = 0.5 * (cxx + cyy + gap);
    out->covariance_minor = 0.5 * (cxx + cyy - gap);
    out->aspect_ratio = out->covariance_minor > 0.0 ? std::sqrt(out->covariance_major / out->covariance_minor)
                                                    : INFINITY;
    out->orientation_radians = gap > 100.0 * 2.220446049250313e-16 ? 0.5 * std::atan2(2.0 * cxy, cxx - cyy) : 0.0;
    out->R0 = perimeter / (2.0 * M_PI);
    out->nmax = nmax;
    for (int n = 0; n < nmax; ++n) out->amplitudes[n] = std::hypot(out->Rc[n], out->Rs[n]);
    out->effective_interface_width = perimeter > 2.220446049250313e-16 ? v1[3] / perimeter : NAN;
    out->effective_interface_cells = out->effective_interface_width / q.h_interface;
    // tensor_diagnostics_detailed: global, core, interface, tail
    const DiagRegion regions[4] = {
        {1, 0.0, 0.0, 0, 0},
        {0, q.sigma_core_cutoff, INFINITY, 1, 1},
        {0, q.sigma_interface_cutoff, q.sigma_core_cutoff, 0, 0},
        {0, q.sigma_safety_cutoff, q.sigma_interface_cutoff, 0, 1},
    };
    fsi_tensor_region* dst[4] = {&out->global_region, &out->core, &out->interface_region, &out->tail};
    for (int r = 0; r < 4; ++r) {
        double vt[DIAG_TENSOR_VALUES];
        launch_diag_tensor(c->stream, c->prm, s->phi, s->sxx, s->syy, s->sxy, q.elastic_modulus, regions[r],
                           c->diag, c->diag_out);
        if ((rc = diag_fetch(c, vt, DIAG_TENSOR_VALUES)) != FSI_OK) return rc;
        fill_region(vt, *dst[r]);
    }
    return FSI_OK;
}

extern "C" const char* fsi_version(void) {
    return "libfsi_hip 0.1 (gfx950, FP64 IMEX step: CG velocity predictor, MG-PCG projection)";
}

// ---------------------------------------------------------------------------
// state and workspace
// ---------------------------------------------------------------------------
extern "C" fsi_status fsi_state_create(fsi_context* c, fsi_state** out) {
    if (!c || !out) return FSI_E_ARGUMENT;
    fsi_state* s = new fsi_state();
    s->ctx = c;
    *out = s;
    double** f[7] = {&s->U, &s->V, &s->phi, &s->res, &s->sxx, &s->syy, &s->sxy};
    for (auto ptr : f) HIPCHECK(c, dalloc(c, ptr, c->felems));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return FSI_OK;
}

extern "C" void fsi_state_destroy(fsi_state* s) {
    if (!s) return;
    double* f[7] = {s->U, s->V, s->phi, s->res, s->sxx, s->syy, s->sxy};
    for (double* p : f)
        if (p) (void)hipFree(p);
    delete s;
}

extern "C" fsi_status fsi_state_upload(fsi_context* c, fsi_state* s, const double* U, const double* V,
                                       const double* phi, const double* resource, const double* sxx,
                                       const double* syy, const double* sxy) {
    if (!c || !s) return FSI_E_ARGUMENT;
    const int nx = c->grid.nx, ny = c->grid.ny;
    fsi_status rc;
    if (U && (rc = upload2d(c, s->U, U, nx + 1, ny)) != FSI_OK) return rc;
    if (V && (rc = upload2d(c, s->V, V, nx, ny + 1)) != FSI_OK) return rc;
    if (phi && (rc = upload2d(c, s->phi, phi, nx, ny)) != FSI_OK) return rc;
    if (resource) {
        if ((rc = upload2d(c, s->res, resource, nx, ny)) != FSI_OK) return rc;
        // the resource is not transported under NoGrowth: its minimum and
        // finiteness are fixed at upload (StepReport.minimum_resource)
        double mn = INFINITY;
        long long bad = 0;
        for (size_t k = 0; k < (size_t)nx * ny; ++k) {
            if (!std::isfinite(resource[k])) ++bad;
            mn = std::fmin(mn, resource[k]);
        }
        s->res_min = mn;
        s->res_nonfinite = bad;
    }
    if (sxx && (rc = upload2d(c, s->sxx, sxx, nx, ny)) != FSI_OK) return rc;
    if (syy && (rc = upload2d(c, s->syy, syy, nx, ny)) != FSI_OK) return rc;
    if (sxy && (rc = upload2d(c, s->sxy, sxy, nx, ny)) != FSI_OK) return rc;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return FSI_OK;
}

extern "C" fsi_status fsi_state_download(fsi_context* c, const fsi_state* s, double* U, double* V, double* phi,
                                         double* resource, double* sxx, double* syy, double* sxy) {
    if (!c || !s) return FSI_E_ARGUMENT;
    const int nx = c->grid.nx, ny = c->grid.ny;
    fsi_status rc;
    if (U && (rc = download2d(c, U, s->U, nx + 1, ny)) != FSI_OK) return rc;
    if (V && (rc = download2d(c, V, s->V, nx, ny + 1)) != FSI_OK) return rc;
    if (phi && (rc = download2d(c, phi, s->phi, nx, ny)) != FSI_OK) return rc;
    if (resource && (rc = download2d(c, resource, s->res, nx, ny)) != FSI_OK) return rc;
    if (sxx && (rc = download2d(c, sxx, s->sxx, nx, ny)) != FSI_OK) return rc;
    if (syy && (rc = download2d(c, syy, s->syy, nx, ny)) != FSI_OK) return rc;
    if (sxy && (rc = download2d(c, sxy, s->sxy, nx, ny)) != FSI_OK) return rc;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return FSI_OK;
}

extern "C" fsi_status fsi_workspace_create(fsi_context* c, fsi_workspace** out) {
    if (!c || !out) return FSI_E_ARGUMENT;
    fsi_workspace* w = new fsi_workspace();
    w->ctx = c;
    *out = w;
    auto A = [&](double** p, size_t n) -> hipError_t {
        hipError_t e = dalloc(c, p, n);
        if (e == hipSuccess) w->owned.push_back(*p);
        return e;
    };
    const size_t F = c->felems;
    // the six stage fields trade places with the state's fields at the end of
    // every step, so they are freed through their current owner (not `owned`)
    double** stage[] = {&w->Us, &w->Vs, &w->phis, &w->sxxs, &w->syys, &w->sxys};
    for (auto p : stage) HIPCHECK(c, dalloc(c, p, F));
    double** singles[] = {&w->ru,     &w->rv,     &w->r_u[0], &w->r_u[1], &w->r_v[0], &w->r_v[1], &w->p_u[0],
                          &w->p_u[1], &w->p_v[0], &w->p_v[1], &w->q_u[0], &w->q_u[1], &w->q_v[0], &w->q_v[1],
                          &w->rhs,    &w->corr[0], &w->corr[1], &w->r,     &w->z,      &w->pp[0],  &w->pp[1], &w->w,
                          &w->cq_cur[0], &w->cq_cur[1], &w->cq_cur[2], &w->cq_prev[0], &w->cq_prev[1],
                          &w->cq_prev[2], &w->cq_next[0], &w->cq_next[1], &w->cq_next[2], &w->xs,
                          &w->mun_u, &w->mun_v, &w->Us2, &w->Vs2,
                          &w->vel_cur_u[0], &w->vel_cur_u[1], &w->vel_cur_u[2], &w->vel_cur_v[0], &w->vel_cur_v[1],
                          &w->vel_cur_v[2], &w->vel_prev_u[0], &w->vel_prev_u[1], &w->vel_prev_u[2],
                          &w->vel_prev_v[0], &w->vel_prev_v[1], &w->vel_prev_v[2], &w->vel_old_u[0],
                          &w->vel_old_u[1],  &w->vel_old_u[2],  &w->vel_old_v[0],  &w->vel_old_v[1],
                          &w->vel_old_v[2],  &w->vel_old2_u[0], &w->vel_old2_u[1], &w->vel_old2_u[2],
                          &w->vel_old2_v[0], &w->vel_old2_v[1], &w->vel_old2_v[2], &w->vel_next_u[0],
                          &w->vel_next_u[1], &w->vel_next_u[2], &w->vel_next_v[0], &w->vel_next_v[1],
                          &w->vel_next_v[2], &w->rhs_u[0], &w->rhs_u[1], &w->rhs_u[2], &w->rhs_v[0],
                          &w->rhs_v[1], &w->rhs_v[2], &w->cphi[0], &w->cphi[1], &w->csxx[0], &w->csxx[1],
                          &w->csyy[0], &w->csyy[1], &w->csxy[0], &w->csxy[1]};
    for (auto p : singles) HIPCHECK(c, A(p, F));
    for (int s = 0; s < 3; ++s) {
        double** h[8] = {&w->Eu[s], &w->Ev[s], &w->Iu[s], &w->Iv[s], &w->Ep[s], &w->Exx[s], &w->Eyy[s], &w->Exy[s]};
        for (auto p : h) HIPCHECK(c, A(p, F));
    }
    if (c->num.precision == FSI_PRECISION_MIXED) {
        float** fv[12] = {&w->fr_u[0], &w->fr_u[1], &w->fr_v[0], &w->fr_v[1], &w->fp_u[0], &w->fp_u[1],
                          &w->fp_v[0], &w->fp_v[1], &w->fq_u[0], &w->fq_u[1], &w->fq_v[0], &w->fq_v[1]};
        for (auto q : fv) {
            HIPCHECK(c, hipMalloc((void**)q, F * sizeof(float)));
            HIPCHECK(c, hipMemsetAsync(*q, 0, F * sizeof(float), c->stream));
            w->owned.push_back(reinterpret_cast<double*>(*q));
        }
    }
    if (c->growth) {  // the growth step's buffers
        double** gb[] = {&w->G, &w->Ec[0], &w->Ec[1], &w->Ec[2], &w->Ic[0], &w->Ic[1], &w->Ic[2], &w->Dx, &w->Dy,
                         &w->crhs, &w->rc_r, &w->rc_p, &w->rc_q};
        for (auto q : gb) HIPCHECK(c, A(q, F));
        HIPCHECK(c, A(&w->gpart, (size_t)2 * c->prm.exnb));
        HIPCHECK(c, dalloc(c, &w->cs, F));  // trades places with the state's resource (freed through its owner)
    }
    // multigrid levels: level 0 solves into z with b = r; x and r are scratch
    w->lv.resize(c->geo.size());
    for (size_t l = 0; l < c->geo.size(); ++l) {
        const MgGeom& g = c->geo[l];
        MgLevel& L = w->lv[l];
        L.nx = g.nx;
        L.ny = g.ny;
        L.P = g.P;
        const size_t n = (size_t)g.P * (g.ny + 1);
        HIPCHECK(c, A(&L.x, n));
        HIPCHECK(c, A(&L.r, n));
        if (l == 0) {
            L.t = w->z;
            L.b = w->r;
        } else {
            HIPCHECK(c, A(&L.t, n));
            HIPCHECK(c, A(&L.b, n));
        }
        L.xf = L.tf = L.bf = nullptr;
        if (c->num.precision == FSI_PRECISION_MIXED && c->cplan.levels > 0 && c->geo.size() > 1) {
            // FP32 levels of the mixed-precision V-cycle (the finest keeps r and z in FP64)
            float** fl[3] = {&L.xf, &L.tf, &L.bf};
            for (int q = 0; q < (l == 0 ? 1 : 3); ++q) {
                HIPCHECK(c, hipMalloc((void**)fl[q], n * sizeof(float)));
                HIPCHECK(c, hipMemsetAsync(*fl[q], 0, n * sizeof(float), c->stream));
                w->owned.push_back(reinterpret_cast<double*>(*fl[q]));
            }
        }
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return FSI_OK;
}

extern "C" void fsi_workspace_destroy(fsi_workspace* w) {
    if (!w) return;
    if (w->ctx && w->ctx->stream) (void)hipStreamSynchronize(w->ctx->stream);
    for (auto& g : w->graphs) (void)hipGraphExecDestroy(g.exec);
    for (double* p : w->owned) (void)hipFree(p);
    double* stage[] = {w->Us, w->Vs, w->phis, w->sxxs, w->syys, w->sxys, w->cs};
    for (double* p : stage)
        if (p) (void)hipFree(p);
    delete w;
}

extern "C" fsi_status fsi_workspace_download_correction(fsi_context* c, const fsi_workspace* w, double* corr) {
    if (!c || !w || !corr) return FSI_E_ARGUMENT;
    fsi_status rc = download2d(c, corr, w->corr[w->corr_cur], c->grid.nx, c->grid.ny);
    if (rc != FSI_OK) return rc;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return FSI_OK;
}

extern "C" fsi_status fsi_workspace_upload_correction(fsi_context* c, fsi_workspace* w, const double* corr) {
    if (!c || !w || !corr) return FSI_E_ARGUMENT;
    for (size_t k = 0; k < (size_t)c->grid.nx * c->grid.ny; ++k)
        if (!std::isfinite(corr[k])) return fail(c, FSI_E_ARGUMENT, "correction_guess contains non-finite values.");
    fsi_status rc = upload2d(c, w->corr[w->corr_cur], corr, c->grid.nx, c->grid.ny);
    if (rc != FSI_OK) return rc;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return FSI_OK;
}

// ---------------------------------------------------------------------------
// step pieces
// ---------------------------------------------------------------------------
void status_reset(fsi_context* c) {
    Status& s = *c->st_host;
    std::memset(&s, 0, sizeof(Status));
    s.phi_min = INFINITY;
    s.phi_max = -INFINITY;
    s.min_eig = INFINITY;
}

fsi_status status_push(fsi_context* c) {
    HIPCHECK(c, hipMemcpyAsync(c->st, c->st_host, sizeof(Status), hipMemcpyHostToDevice, c->stream));
    // the pinned buffer is rewritten by the next status_pull; wait for the copy
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return FSI_OK;
}
// The step's starting Status, written by a kernel: nothing to wait for.
static fsi_status status_push_reset(fsi_context* c) {
    launch_status_reset(c->stream, c->st);
    HIPCHECK(c, hipGetLastError());
    return FSI_OK;
}
fsi_status status_pull(fsi_context* c) {
    HIPCHECK(c, hipMemcpyAsync(c->st_host, c->st, sizeof(Status), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipGetLastError());
    return FSI_OK;
}

// Poll the `done` flag of one Krylov solve (a small device->host copy).
static fsi_status poll_done(fsi_context* c, const KrylovScalars* dev, int* done, int* iter) {
    int buf[2];
    HIPCHECK(c, hipMemcpyAsync(buf, &dev->done, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    *done = buf[0];
    *iter = buf[1];
    return FSI_OK;
}

// The stopping rules of the predictor solve from LinearSolverOptions: the
// direct target and the fallback get max(10 maxiter, 1000) iterations more
// (CG on this SPD system gains about a factor 3 per iteration at the
// benchmark sizes, so the direct target takes a few tens).
static VelTol vel_tol(const fsi_context* c) {
    VelTol t{};
    t.rtol = c->num.velocity_rtol;
    t.fb_rtol = c->num.fallback_rtol;
    t.maxiter = c->num.velocity_maxiter;
    const long cap = (long)c->num.velocity_maxiter + std::max(10L * c->num.velocity_maxiter, 1000L);
    t.cap = (int)std::min(cap, 1L << 30);
    t.method = c->num.velocity_method;
    t.fallback = c->num.direct_fallback != 0;
    t.icap = 100;
    return t;
}

// Mixed precision: the predictor by FP64 iterative refinement around an FP32
// CG.  k_vel_start forms x0 and r0 = b - A x0 in FP64 (r0 also in FP32);
// each cycle runs the FP32 inner CG on A d = r (x accumulating in FP64) to
// its threshold, then k_vel_start (restart) forms r = b - A x again in FP64
// and k_vel_check applies the stopping rules of LinearSolverOptions to it.
// Planned: `fixed` inner iterations in each of plan_cyc[slot] cycles;
// otherwise every cycle is polled.
static fsi_status poll_vel(fsi_context* c, KrylovScalars* out) {
    HIPCHECK(c, hipMemcpyAsync(out, &c->st->vel, sizeof(KrylovScalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return FSI_OK;
}
static fsi_status velocity_solve_mixed(fsi_context* c, fsi_workspace* w, VelBufs b, double alpha, int slot,
                                       int fixed, const VelTol& vt) {
    const Params& p = c->prm;
    for (int k = 0; k < 2; ++k) {
        b.fr_u[k] = w->fr_u[k];
        b.fr_v[k] = w->fr_v[k];
        b.fp_u[k] = w->fp_u[k];
        b.fp_v[k] = w->fp_v[k];
        b.fq_u[k] = w->fq_u[k];
        b.fq_v[k] = w->fq_v[k];
    }
    launch_vel_init(c->stream, p, b, alpha, c->st, slot, vt);
    if (fixed > 0) {
        const int cycles = std::max(1, c->plan_cyc[slot]);
        for (int cy = 0; cy < cycles; ++cy) {
            launch_vel_iters_f32(c->stream, p, b, alpha, c->st, 0, std::min(fixed, vt.icap + 1), slot, vt);
            launch_vel_refine(c->stream, p, b, alpha, c->st, slot, vt, cy + 1 < cycles);
        }
        return FSI_OK;
    }
    // every cycle takes at least one inner iteration, so cap + 1 cycles reach any rule's limit
    KrylovScalars ks{};
    for (int cy = 0; cy <= vt.cap; ++cy) {
        int m = 0, chunk = c->last_imax[slot] + 1;
        while (m <= vt.icap) {
            const int n = std::min(chunk, vt.icap + 1 - m);
            launch_vel_iters_f32(c->stream, p, b, alpha, c->st, m, n, slot, vt);
            m += n;
            fsi_status rc = poll_vel(c, &ks);
            if (rc != FSI_OK) return rc;
            if (ks.done || ks.idone) break;
            chunk = 4;
        }
        launch_vel_refine(c->stream, p, b, alpha, c->st, slot, vt, 1);
        fsi_status rc = poll_vel(c, &ks);
        if (rc != FSI_OK) return rc;
        if (ks.done) break;
    }
    c->last_vel_iters[slot] = std::max(1, ks.iter);
    c->last_imax[slot] = std::max(1, ks.imax);
    return FSI_OK;
}

// The warm start of stage `slot`: the stage's solutions of the previous
// steps, extrapolated in time once two (2 u^{n-1} - u^{n-2}), three
// (3 (u^{n-1} - u^{n-2}) + u^{n-3}) or four steps are known, instead of u^n.
// Returns false (b untouched) before the first step or with the warm start off.
static bool warm_history(const fsi_context* c, const fsi_workspace* w, int slot, VelBufs& b) {
    if (!(slot >= 0 && c->vel_warm && w->stage_warm >= 1)) return false;
    b.Un = w->vel_cur_u[slot];
    b.Vn = w->vel_cur_v[slot];
    b.Pu = w->vel_prev_u[slot];
    b.Pv = w->vel_prev_v[slot];
    b.Ou = w->vel_old_u[slot];
    b.Ov = w->vel_old_v[slot];
    b.Wu = w->vel_old2_u[slot];
    b.Wv = w->vel_old2_v[slot];
    b.extrap = std::min(w->stage_warm, 4) - 1;
    return true;
}

// Whether the warm start of stage `slot` is formed ahead of its solve, on
// `on` (k_vel_predict into (xu, xv)): with an extrapolation to form, so that
// the start kernel reads one field pair instead of the history.
static bool predict_x0(fsi_context* c, fsi_workspace* w, int slot, double* xu, double* xv, hipStream_t on) {
    VelBufs b{};
    if (!c->prm.visc || !warm_history(c, w, slot, b) || b.extrap < 1) return false;
    b.xu = xu;
    b.xv = xv;
    launch_vel_predict(on, c->prm, b);
    return true;
}

// fixed > 0: enqueue exactly `fixed` iteration kernels without reading the
// residual back (speculative step); otherwise poll until converged.
// The predictor of stage `slot` from the right-hand side (ru, rv) into (xu, xv);
// with `x0_formed` its warm start is already in (xu, xv) (predict_x0).
fsi_status velocity_solve(fsi_context* c, fsi_state* s, fsi_workspace* w, double alpha, int slot,
                                 int fixed, const double* ru, const double* rv, double* xu, double* xv,
                                 bool x0_formed) {
    const Params& p = c->prm;
    VelBufs b{};
    b.Un = s->U;
    b.Vn = s->V;
    b.ru = ru;
    b.rv = rv;
    b.xu = xu;
    b.xv = xv;
    for (int k = 0; k < 2; ++k) {
        b.r_u[k] = w->r_u[k];
        b.r_v[k] = w->r_v[k];
        b.p_u[k] = w->p_u[k];
        b.p_v[k] = w->p_v[k];
        b.q_u[k] = w->q_u[k];
        b.q_v[k] = w->q_v[k];
    }
    b.partials = c->partials;
    if (!p.visc) {
        launch_vel_copy(c->stream, p, b);
        return FSI_OK;
    }
    // warm start: extrapolated by the start kernel, or formed ahead of it
    if (x0_formed) {
        b.Un = xu;
        b.Vn = xv;
        b.x0_in_place = 1;
    } else {
        (void)warm_history(c, w, slot, b);
    }
    const VelTol vt = vel_tol(c);
    // gmres: the stopping rule measures against |b - A u^n| (the state
    // velocity; M u^n from the stage-1 momentum kernel)
    if (vt.method == FSI_SOLVER_GMRES) {
        b.Su = s->U;
        b.Sv = s->V;
        b.Mu = w->mun_u;
        b.Mv = w->mun_v;
    }
    if (c->num.precision == FSI_PRECISION_MIXED) return velocity_solve_mixed(c, w, b, alpha, slot, fixed, vt);
    launch_vel_init(c->stream, p, b, alpha, c->st, slot, vt);
    // the last iteration any rule can reach (kernel m tests iterate m)
    const int last = (vt.method == FSI_SOLVER_DIRECT || vt.fallback) ? vt.cap : vt.maxiter;
    if (fixed > 0) {
        launch_vel_iters(c->stream, p, b, alpha, c->st, 0, std::min(fixed, last + 1), slot, vt,
                         c->vel_check_from[slot]);
        return FSI_OK;
    }
    int m = 0, chunk = c->last_vel_iters[slot] + 1, done = 0, iter = 0;
    while (m <= last) {
        int n = std::min(chunk, last + 1 - m);
        launch_vel_iters(c->stream, p, b, alpha, c->st, m, n, slot, vt, 1 << 30);
        m += n;
        fsi_status rc = poll_done(c, &c->st->vel, &done, &iter);
        if (rc != FSI_OK) return rc;
        if (done) break;
        chunk = 4;
    }
    c->last_vel_iters[slot] = std::max(1, iter);
    return FSI_OK;
}

// Projects (Uin, Vin) into (U, V); the same arrays for a projection in place.
// The PCG iterates in x (initial guess in, solution out) and the mean-free
// correction goes to corr[1 - corr_cur]; nothing else of the workspace's
// projection state is written, so the caller commits (flips corr_cur) only
// on success.  slot >= 0: a stage projection of a step, iterating in that
// stage's cq_next from the extrapolated history; slot -1: a standalone
// projection from the workspace correction.
fsi_status projection(fsi_context* c, fsi_workspace* w, double* Uin, double* Vin, double* U, double* V,
                             int inject, int slot, int fixed) {
    const Params& p = c->prm;
    ProjBufs b{};
    b.U = U;
    b.V = V;
    b.Uin = Uin;
    b.Vin = Vin;
    b.inject = inject;
    b.lap_bc = c->lap_bc;
    b.rhs = w->rhs;
    b.x = slot >= 0 ? w->cq_next[slot] : w->xs;
    // the initial guess (the direct solve writes x itself)
    if (!c->spec.enabled) {
        if (slot >= 0 && w->stage_warm >= 1)
            launch_extrapolate(c->stream, p, b.x, w->cq_cur[slot], w->cq_prev[slot], w->stage_warm >= 2);
        else
            launch_extrapolate(c->stream, p, b.x, w->corr[w->corr_cur], nullptr, 0);
    }
    b.q_out = w->corr[1 - w->corr_cur];
    b.G = w->proj_target;
    b.r = w->r;
    b.z = w->z;
    b.p[0] = w->pp[0];
    b.p[1] = w->pp[1];
    b.w = w->w;
    b.partials = c->partials;
    if (c->spec.enabled) {  // the right-hand side, its sums and the mean of x from the transform passes
        const int np = (p.ny + 1) / 2;
        b.rhs_part = c->spec_part;
        b.rhs_np = np;
        b.q_part = c->spec_part + (size_t)4 * np;
        b.q_np = np;
        launch_spectral_proj(c->stream, c->spec, p, b);
    } else {
        launch_proj_rhs(c->stream, p, b);
    }
    const int ps = slot < 0 ? 0 : slot;
    const int last = c->num.projection_maxiter;
    if (fixed > 0 && std::min(fixed, last + 1) == 1) {
        // planned with no PCG iteration (the direct solve, or an exact warm
        // start): the initial guess is only tested
        launch_proj_tested(c->stream, p, c->geo[0], b, c->st, c->has_null, c->num.projection_rtol, last,
                           c->grid.lx * c->grid.ly, slot);
        return FSI_OK;
    }
    launch_pcg_init(c->stream, p, c->geo[0], b, c->st, c->has_null, c->num.projection_rtol,
                    c->num.projection_maxiter);
    if (fixed > 0) {
        // Of the planned iterations only the convergence test of the last one
        // matters (its update is never tested: converged earlier, it is
        // skipped; not converged, the step is redone), so it is launched alone.
        const int n = std::min(fixed, last + 1);
        if (n > 1)
            launch_pcg_iters(c->stream, p, b, w->lv.data(), c->geo.data(), (int)c->geo.size(), c->cplan, c->st, 0,
                             n - 1, slot);
        launch_pcg_test(c->stream, p, b, c->st, n - 1, slot);
        launch_proj_finish(c->stream, p, c->geo[0], b, c->st, c->has_null, c->grid.lx * c->grid.ly, slot);
        return FSI_OK;
    }
    int m = 0, chunk = c->last_proj_iters[ps] + 1, done = 0, iter = 0;
    while (m <= last) {
        int n = std::min(chunk, last + 1 - m);
        launch_pcg_iters(c->stream, p, b, w->lv.data(), c->geo.data(), (int)c->geo.size(), c->cplan, c->st, m, n,
                         slot);
        m += n;
        fsi_status rc = poll_done(c, &c->st->pcg, &done, &iter);
        if (rc != FSI_OK) return rc;
        if (done) break;
        chunk = 4;
    }
    c->last_proj_iters[ps] = std::max(0, iter);
    launch_proj_finish(c->stream, p, c->geo[0], b, c->st, c->has_null, c->grid.lx * c->grid.ly, slot);
    return FSI_OK;
}

// The failure the reference would raise first: its sequence within a step is
// the tendencies of stage 1, the solves of stage 2, the tendencies of stage
// 2, ... (the stage loop of integrate_step!, TimeIntegrationIMEX.jl:1152-1256).  The solve and projection
// kernels record theirs in `code` (code_at tells the stage), the phase and
// sigma kernels of each stage in cells_err, so the order holds also where
// the latter run beside the next stage's solve.
static ErrSlot first_error(const Status& s) {
    int sa = 4;  // the stage (1..3) of the solve-side failure, 4: none
    for (int k = 0; k < 3; ++k)
        if (s.code_at[k] != DEV_OK) {
            sa = k + 1;
            break;
        }
    if (s.code != DEV_OK && sa == 4) sa = 0;  // outside a step's stages (a standalone projection)
    for (int q = 0; q < 3; ++q)
        if (s.cells_err[q].code != DEV_OK && q < sa) return s.cells_err[q];
    ErrSlot e{};
    e.code = s.code;
    e.v0 = s.v0;
    e.v1 = s.v1;
    return e;
}

fsi_status device_error(fsi_context* c) {
    const ErrSlot s = first_error(*c->st_host);
    switch (s.code) {
    case DEV_OK: return FSI_OK;
    case DEV_NEUMANN_INFLOW:
        return fail(c, FSI_E_ARGUMENT,
                    "A Neumann condition does not define the incoming advected state (face velocity %.6e). "
                    "Prescribe Dirichlet data for inflow.", s.v0);
    case DEV_NONSPD:
        return fail(c, FSI_E_DOMAIN,
                    "GeometricStructureExtension requires an SPD tensor; encountered a non-positive or "
                    "non-finite eigenvalue (sigma_xx = %.6e, sigma_xy = %.6e).", s.v0, s.v1);
    case DEV_INCOMPATIBLE:
        return fail(c, FSI_E_ARGUMENT,
                    "The prescribed divergence constraint is incompatible with the Neumann/periodic correction "
                    "boundaries. The volume integral of the Poisson right-hand side is %.6e, while the "
                    "compatibility tolerance is %.6e.", s.v0, s.v1);
    case DEV_PROJ_RESIDUAL:
        return fail(c, FSI_E_SOLVER,
                    "Projection Poisson solve did not reach the required true-residual accuracy. Relative "
                    "residual = %.6e, tolerance = %.1e.", s.v0, s.v1);
    case DEV_PROJ_CONSTRAINT:
        return fail(c, FSI_E_SOLVER,
                    "Projection did not satisfy the prescribed divergence constraint. Weighted L2 residual = "
                    "%.6e, tolerance = %.6e.", s.v0, s.v1);
    case DEV_NONHOMOG_NEUMANN:
        return fail(c, FSI_E_ARGUMENT,
                    "Only homogeneous Neumann phase boundaries are supported by the ACDI/UWCDI operators.");
    case DEV_DIFFUSIVITY:  // _validate_positive_field, TransportConstitutive.jl:53-59
        return fail(c, FSI_E_ARGUMENT,
                    std::isfinite(s.v0) ? "evaluated diffusivity must be strictly positive everywhere."
                                        : "evaluated diffusivity contains non-finite values.");
    case DEV_RES_SINK:  // resource_sink! TransportConstitutive.jl:331-333
        return fail(c, FSI_E_ARGUMENT, "Evaluated resource sink contains non-finite values.");
    case DEV_RES_NOCONV:  // _solve_linear_system! :636 for the resource stage
        return fail(c, FSI_E_SOLVER, "resource stage GMRES residual is too large: %.6e", s.v0);
    case DEV_VEL_NOCONV:  // _solve_linear_system! :636 (the device's direct target failing included)
        return fail(c, FSI_E_SOLVER, "velocity stage %d GMRES residual is too large: %.6e", (int)s.v1, s.v0);
    default: return fail(c, FSI_E_HIP, "device reported error code %d", s.code);
    }
}

static void fill_projection_result(const fsi_context* c, fsi_projection_result* r) {
    if (!r) return;
    const Status& s = *c->st_host;
    r->linear_relative_residual = s.proj.lin_rel;
    r->constraint_linf = s.proj.linf;
    r->constraint_l2 = s.proj.l2;
    r->compatibility_defect = s.proj.defect;
    r->iterations = s.pcg.iter;
}

static fsi_status project_state(fsi_context* c, fsi_state* s, fsi_workspace* w, fsi_projection_result* res,
                                bool growth = false) {
    if (!c || !s || !w) return FSI_E_ARGUMENT;
    status_reset(c);
    fsi_status rc = status_push(c);
    if (rc != FSI_OK) return rc;
    if (growth) {  // G of the state's phi and resource (growth_source! :1094)
        GrSrc a{s->phi, s->res, s->U, s->V, s->sxx, s->syy, s->sxy, w->G, w->gpart, 1, 0};
        launch_gr_source(c->stream, c->prm, c->gp, a, c->st);
        w->proj_target = w->G;
    }
    rc = projection(c, w, s->U, s->V, s->U, s->V, 0, -1, 0);
    w->proj_target = nullptr;
    if (rc != FSI_OK) return rc;
    if ((rc = status_pull(c)) != FSI_OK) return rc;
    fill_projection_result(c, res);
    rc = device_error(c);
    if (rc == FSI_OK) w->corr_cur = 1 - w->corr_cur;  // copyto!(work.correction, result.correction)
    return rc;
}

extern "C" fsi_status fsi_make_velocity_compatible(fsi_context* c, fsi_state* s, fsi_workspace* w,
                                                   fsi_projection_result* res) {
    // project onto div v = G(phi, c) (:1088-1109); G = 0 under NoGrowth
    return project_state(c, s, w, res, c && w && c->growth && c->gp.src != FSI_GROWTH_NONE);
}

extern "C" fsi_status fsi_project(fsi_context* c, fsi_state* s, fsi_workspace* w, fsi_projection_result* res) {
    return project_state(c, s, w, res);
}

// The tendencies of stage `slot`; with `sums` / `csums` the momentum and
// sigma kernels add their parts to the velocity right-hand sides and to the
// phi, sigma stage values of the later stages instead of storing E_u, E_v,
// I_u, I_v and E_sigma (E_phi is stored for the sigma kernel).
static void tendencies(fsi_context* c, fsi_workspace* w, int slot, const double* U, const double* V,
                       const double* phi, const double* sxx, const double* syy, const double* sxy, int reset,
                       int stage1, const RhsSums* sums = nullptr, const CellSums* csums = nullptr,
                       hipStream_t cells_stream = nullptr) {
    const Params& p = c->prm;
    hipStream_t cs = cells_stream ? cells_stream : c->stream;
    if (sums)
        launch_momentum(c->stream, p, U, V, phi, sxx, syy, sxy, reset, nullptr, nullptr, nullptr, nullptr, sums);
    else
        launch_momentum(c->stream, p, U, V, phi, sxx, syy, sxy, reset, w->Eu[slot], w->Ev[slot], w->Iu[slot],
                        w->Iv[slot]);
    if (csums)
        launch_cells(cs, p, c->bpsi, U, V, phi, sxx, syy, sxy, reset, w->Ep[slot], nullptr, nullptr, nullptr,
                     &c->st->cells_err[slot], stage1, csums);
    else
        launch_cells(c->stream, p, c->bpsi, U, V, phi, sxx, syy, sxy, reset, w->Ep[slot], w->Exx[slot],
                     w->Eyy[slot], w->Exy[slot], &c->st->cells_err[slot], stage1);
}

extern "C" fsi_status fsi_eval_tendencies(fsi_context* c, const fsi_state* s, fsi_workspace* w, double t,
                                          double* E_u, double* E_v, double* I_u, double* I_v, double* E_phi,
                                          double* E_sxx, double* E_syy, double* E_sxy) {
    if (!c || !s || !w) return FSI_E_ARGUMENT;
    (void)t;  // the supported body forces are steady
    status_reset(c);
    fsi_status rc = status_push(c);
    if (rc != FSI_OK) return rc;
    tendencies(c, w, 0, s->U, s->V, s->phi, s->sxx, s->syy, s->sxy, c->has_reset, 0);
    const int nx = c->grid.nx, ny = c->grid.ny;
    if (E_u && (rc = download2d(c, E_u, w->Eu[0], nx, ny)) != FSI_OK) return rc;
    if (E_v && (rc = download2d(c, E_v, w->Ev[0], nx, ny)) != FSI_OK) return rc;
    if (I_u && (rc = download2d(c, I_u, w->Iu[0], nx, ny)) != FSI_OK) return rc;
    if (I_v && (rc = download2d(c, I_v, w->Iv[0], nx, ny)) != FSI_OK) return rc;
    if (E_phi && (rc = download2d(c, E_phi, w->Ep[0], nx, ny)) != FSI_OK) return rc;
    if (E_sxx && (rc = download2d(c, E_sxx, w->Exx[0], nx, ny)) != FSI_OK) return rc;
    if (E_syy && (rc = download2d(c, E_syy, w->Eyy[0], nx, ny)) != FSI_OK) return rc;
    if (E_sxy && (rc = download2d(c, E_sxy, w->Exy[0], nx, ny)) != FSI_OK) return rc;
    if ((rc = status_pull(c)) != FSI_OK) return rc;
    return device_error(c);
}

// uwcdi_strang_step (PhaseFieldMethods.jl:793-850): Euler sweeps of the
// directional right-hand sides, x (dt/2), y (dt), x (dt/2) (first_direction
// 1: y, x, y), through E_phi of stage slots 0..2 of the workspace.
extern "C" fsi_status fsi_uwcdi_strang_step(fsi_context* c, fsi_state* s, fsi_workspace* w, double dt,
                                            int32_t first_direction) {
    if (!c || !s || !w) return FSI_E_ARGUMENT;
    if (c->prm.phase != FSI_PHASE_UWCDI) return fail(c, FSI_E_ARGUMENT, "uwcdi_strang_step needs the UWCDI phase field.");
    if (!(std::isfinite(dt) && dt >= 0.0)) return fail(c, FSI_E_ARGUMENT, "dt must be finite and non-negative.");
    if (first_direction != 0 && first_direction != 1)
        return fail(c, FSI_E_ARGUMENT, "first_direction must be :x or :y; received %d.", (int)first_direction);
    if (dt == 0.0) return FSI_OK;  // every sweep returns phi unchanged (:802)
    status_reset(c);
    fsi_status rc = status_push(c);
    if (rc != FSI_OK) return rc;
    const Params& p = c->prm;
    const int d1 = first_direction == 0 ? 1 : 2, d2 = 3 - d1;
    ErrSlot* err = &c->st->cells_err[0];
    launch_uwcdi_sweep(c->stream, p, s->U, s->V, s->phi, d1, 0.5 * dt, w->Ep[0], w->Ep[1], err);
    launch_uwcdi_sweep(c->stream, p, s->U, s->V, w->Ep[1], d2, dt, w->Ep[0], w->Ep[2], err);
    launch_uwcdi_sweep(c->stream, p, s->U, s->V, w->Ep[2], d1, 0.5 * dt, w->Ep[0], s->phi, err);
    if ((rc = status_pull(c)) != FSI_OK) return rc;
    return device_error(c);
}

// The resource stage solve (_build_resource_stage! :742-774) into w->cs from
// x0 = c_n, polled every few iterations.
static fsi_status resource_solve(fsi_context* c, fsi_state* s, fsi_workspace* w, double alpha) {
    const Params& p = c->prm;
    ResCg a{};
    a.Dx = w->Dx;
    a.Dy = w->Dy;
    a.b = w->crhs;
    a.x0 = s->res;
    a.x = w->cs;
    a.r = w->rc_r;
    a.p = w->rc_p;
    a.q = w->rc_q;
    a.part = w->gpart;
    a.alpha = alpha;
    const int maxiter = c->num.velocity_maxiter;
    const int cap = maxiter + std::max(10 * maxiter, 1000);  // the direct target's budget, as the predictor's
    a.rtol = c->num.velocity_rtol;
    a.fb = c->num.fallback_rtol;
    a.maxiter_d = cap;
    a.method = c->num.resource_method;
    a.fallback = c->num.direct_fallback;
    a.maxiter = maxiter;
    launch_rcg_start(c->stream, p, c->gp, a, c->st);
    int m = 0, chunk = 8, done = 0, iter = 0;
    fsi_status rc = poll_done(c, &c->st->res, &done, &iter);
    while (rc == FSI_OK && !done && m < cap) {
        const int n = std::min(chunk, cap - m);
        launch_rcg_iters(c->stream, p, c->gp, a, c->st, m, n);
        m += n;
        rc = poll_done(c, &c->st->res, &done, &iter);
        chunk = 16;
    }
    return rc;
}

// The step with growth, resource and diffusion closures
// (TimeIntegrationIMEX.jl:1152-1256): per stage phi_s, sigma_s and the
// resource right-hand side, the resource solve (diffusion), G_s and K_g,s,
// the predictor and the projection onto div v = G_s, then the stored
// tendencies of the no-growth kernels with the growth terms added.  Every
// solve is polled (no planned steps); everything runs on the context's stream.
static fsi_status enqueue_growth_step(fsi_context* c, fsi_state* s, fsi_workspace* w, double dt) {
    const Params& p = c->prm;
    const GrowthP& g = c->gp;
    const bool diff = g.diff != FSI_DIFFUSION_NONE;
    hipStream_t S = c->stream;
    fsi_status rc = FSI_OK;
    double* const sphi[3] = {w->cphi[0], w->cphi[1], w->phis};
    double* const ssxx[3] = {w->csxx[0], w->csxx[1], w->sxxs};
    double* const ssyy[3] = {w->csyy[0], w->csyy[1], w->syys};
    double* const ssxy[3] = {w->csxy[0], w->csxy[1], w->sxys};
    auto stage_u = [&](int si) { return si == 2 ? w->Us2 : w->Us; };
    auto stage_v = [&](int si) { return si == 2 ? w->Vs2 : w->Vs; };
    // stage 1: G of the state, the stage-1 checks, the tendencies
    GrSrc src{s->phi, s->res, s->U, s->V, s->sxx, s->syy, s->sxy, w->G, w->gpart, 1, 1};
    launch_gr_source(S, p, g, src, c->st);
    if (diff) launch_gr_diff(S, p, g, s->phi, w->Dx, w->Dy, c->st);
    tendencies(c, w, 0, s->U, s->V, s->phi, s->sxx, s->syy, s->sxy, c->has_reset, 0);
    GrTend t0{s->U,   s->V,      s->phi,    s->sxx,    s->syy,    s->sxy,    w->G,  s->res,
              w->Dx,  w->Dy,     c->has_reset, w->Eu[0], w->Ev[0], w->Ep[0], w->Exx[0], w->Eyy[0],
              w->Exy[0], w->Ec[0], diff ? w->Ic[0] : nullptr};
    launch_gr_tend(S, p, g, t0, c->st);
    for (int si = 1; si < 4; ++si) {
        const int k = si - 1;
        const double alpha = dt * AI[si][si];
        GrStage a{};
        a.phin = s->phi;
        a.sxxn = s->sxx;
        a.syyn = s->syy;
        a.sxyn = s->sxy;
        a.cn = s->res;
        a.Un = s->U;
        a.Vn = s->V;
        for (int q = 0; q < si; ++q) {
            a.Ep[q] = w->Ep[q];
            a.Exx[q] = w->Exx[q];
            a.Eyy[q] = w->Eyy[q];
            a.Exy[q] = w->Exy[q];
            a.Ec[q] = w->Ec[q];
            a.Ic[q] = diff ? w->Ic[q] : nullptr;
            a.Eu[q] = w->Eu[q];
            a.Ev[q] = w->Ev[q];
            a.Iu[q] = w->Iu[q];
            a.Iv[q] = w->Iv[q];
            a.ce[q] = AE[si][q];
            a.ci[q] = AI[si][q];
        }
        a.dt = dt;
        a.alpha = alpha;
        a.bu = p.visc ? c->bvisc_u : nullptr;
        a.bv = p.visc ? c->bvisc_v : nullptr;
        a.phi = sphi[k];
        a.sxx = ssxx[k];
        a.syy = ssyy[k];
        a.sxy = ssxy[k];
        a.crhs = diff ? w->crhs : w->cs;  // without diffusion c_s is the right-hand side itself
        a.ru = w->rhs_u[k];
        a.rv = w->rhs_v[k];
        const bool ref = p.visc && si == 1 && c->num.velocity_method == FSI_SOLVER_GMRES;
        a.mu = ref ? w->mun_u : nullptr;
        a.mv = ref ? w->mun_v : nullptr;
        a.ns = si;
        a.reset_n = c->has_reset;
        a.reset = c->has_reset;
        launch_gr_stage(S, p, a);
        if (diff) {
            launch_gr_diff(S, p, g, sphi[k], w->Dx, w->Dy, c->st);
            if ((rc = resource_solve(c, s, w, alpha)) != FSI_OK) return rc;
        }
        GrSrc sk{sphi[k], w->cs, nullptr, nullptr, nullptr, nullptr, nullptr, w->G, w->gpart, 0, 0};
        launch_gr_source(S, p, g, sk, c->st);
        const bool keep = p.visc;
        double* xu = keep ? w->vel_next_u[k] : stage_u(si);
        double* xv = keep ? w->vel_next_v[k] : stage_v(si);
        if ((rc = velocity_solve(c, s, w, alpha, k, 0, w->rhs_u[k], w->rhs_v[k], xu, xv)) != FSI_OK) return rc;
        w->proj_target = w->G;
        rc = projection(c, w, xu, xv, stage_u(si), stage_v(si), 1, k, 0);
        w->proj_target = nullptr;
        if (rc != FSI_OK) return rc;
        if (si <= 2) {
            tendencies(c, w, si, stage_u(si), stage_v(si), sphi[k], ssxx[k], ssyy[k], ssxy[k], 0, 0);
            GrTend tk{stage_u(si), stage_v(si), sphi[k], ssxx[k], ssyy[k], ssxy[k], w->G, w->cs, w->Dx, w->Dy, 0,
                      w->Eu[si], w->Ev[si], w->Ep[si], w->Exx[si], w->Eyy[si], w->Exy[si], w->Ec[si],
                      diff ? w->Ic[si] : nullptr};
            launch_gr_tend(S, p, g, tk, c->st);
        }
    }
    launch_gr_report(S, p, w->phis, w->sxxs, w->syys, w->sxys, w->cs);
    launch_ext_final(S, p, c->st);
    return FSI_OK;
}

// The stages of one step (TimeIntegrationIMEX.jl:1152-1256) on the stream;
// with `planned` each solve runs the planned number of iterations and the
// launch sequence is fixed (it can be captured into a graph).
static fsi_status enqueue_step(fsi_context* c, fsi_state* s, fsi_workspace* w, double dt, bool planned) {
    if (c->growth) return enqueue_growth_step(c, s, w, dt);
    const Params& p = c->prm;
    const int reset = c->has_reset;
    fsi_status rc = FSI_OK;
    // the velocity right-hand sides (_build_velocity_stage!) are summed by
    // the momentum kernels of stages q = 0..2 (terms aE[s][q] E_q + aI[s][q]
    // I_q for the later stages s, the lifting alpha_s b at s = q + 1)
    auto sums_of = [&](int q) {
        RhsSums R{};
        for (int t = 0; t < 3; ++t) {
            R.ru[t] = w->rhs_u[t];
            R.rv[t] = w->rhs_v[t];
            R.ce[t] = t >= q ? AE[t + 1][q] : 0.0;
            R.ci[t] = t >= q ? AI[t + 1][q] : 0.0;
        }
        R.alpha = dt * AI[q + 1][q + 1];
        R.dt = dt;
        R.bu = p.visc ? c->bvisc_u : nullptr;
        R.bv = p.visc ? c->bvisc_v : nullptr;
        const bool ref = p.visc && q == 0 && c->num.velocity_method == FSI_SOLVER_GMRES;
        R.mu = ref ? w->mun_u : nullptr;
        R.mv = ref ? w->mun_v : nullptr;
        R.q = q;
        return R;
    };
    // phi, sigma of stage s (_build_explicit_stage!) are summed by the sigma
    // kernels of stages q < s; stage 4's values land in phis .. sxys
    double* const sphi[3] = {w->cphi[0], w->cphi[1], w->phis};
    double* const ssxx[3] = {w->csxx[0], w->csxx[1], w->sxxs};
    double* const ssyy[3] = {w->csyy[0], w->csyy[1], w->syys};
    double* const ssxy[3] = {w->csxy[0], w->csxy[1], w->sxys};
    auto cells_of = [&](int q) {
        CellSums C{};
        for (int t = 0; t < 3; ++t) {
            C.phi[t] = sphi[t];
            C.sxx[t] = ssxx[t];
            C.syy[t] = ssyy[t];
            C.sxy[t] = ssxy[t];
            C.cf[t] = t >= q ? dt * AE[t + 1][q] : 0.0;
        }
        C.Ephi = w->Ep[q];
        C.reset = reset;
        C.q = q;
        return C;
    };
    // With the side stream the phase and sigma kernels of stage q overlap the
    // solve of stage q + 1: they need the stage-q velocity (projected before
    // them) and write only E_phi and the phi, sigma stage sums, which the
    // momentum kernel of stage q + 1 reads after the join.  The warm starts
    // of the three predictor solves (from the previous steps' solutions
    // only) are formed there too, ahead of their solves.
    if (c->dual && !c->side && c->use_graph != 1) {
        HIPCHECK(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIPCHECK(c, hipEventCreateWithFlags(&c->ev_start, hipEventDisableTiming));
        for (int k = 0; k < 3; ++k) {
            HIPCHECK(c, hipEventCreateWithFlags(&c->ev_proj[k], hipEventDisableTiming));
            HIPCHECK(c, hipEventCreateWithFlags(&c->ev_cells[k], hipEventDisableTiming));
            HIPCHECK(c, hipEventCreateWithFlags(&c->ev_x0[k], hipEventDisableTiming));
        }
    }
    const bool dual = c->dual && c->side && c->use_graph != 1;
    hipStream_t A = c->stream, B = dual ? c->side : c->stream;
    bool formed[3] = {false, false, false};
    auto x0_ahead = [&](int k) -> fsi_status {
        formed[k] = dual && predict_x0(c, w, k, w->vel_next_u[k], w->vel_next_v[k], B);
        if (formed[k]) HIPCHECK(c, hipEventRecord(c->ev_x0[k], B));
        return FSI_OK;
    };
    // an error return leaves nothing of the step running on the side stream
    auto bail = [&](fsi_status r) {
        if (dual) (void)hipStreamSynchronize(B);
        return r;
    };
    if (dual) {
        HIPCHECK(c, hipEventRecord(c->ev_start, A));  // after the Status reset
        HIPCHECK(c, hipStreamWaitEvent(B, c->ev_start, 0));
        if ((rc = x0_ahead(0)) != FSI_OK) return bail(rc);
    }
    // the stage velocity: stages 1 and 3 in Us, Vs (swapped into the state),
    // stage 2 in Us2, Vs2 (the side stream still reads stage 1's)
    auto stage_u = [&](int si) { return si == 2 ? w->Us2 : w->Us; };
    auto stage_v = [&](int si) { return si == 2 ? w->Vs2 : w->Vs; };
    // stage 1: the state itself (with the safety reset on sigma_n), tendencies
    RhsSums sums = sums_of(0);
    CellSums csums = cells_of(0);
    tendencies(c, w, 0, s->U, s->V, s->phi, s->sxx, s->syy, s->sxy, reset, 1, &sums, &csums, B);
    if (dual) {
        HIPCHECK(c, hipEventRecord(c->ev_cells[0], B));
        for (int k = 1; k < 3; ++k)
            if ((rc = x0_ahead(k)) != FSI_OK) return bail(rc);
    }
    for (int si = 1; si < 4; ++si) {
        // the predictor from the summed right-hand side
        const double alpha = dt * AI[si][si];
        const int k = si - 1;
        // with the warm-start history the predictor is kept (the next step
        // extrapolates from it) and the projection writes the stage velocity
        const bool keep = p.visc;
        double* xu = keep ? w->vel_next_u[k] : stage_u(si);
        double* xv = keep ? w->vel_next_v[k] : stage_v(si);
        if (formed[k]) HIPCHECK(c, hipStreamWaitEvent(A, c->ev_x0[k], 0));
        if ((rc = velocity_solve(c, s, w, alpha, k, planned ? c->plan_vel[k] : 0, w->rhs_u[k], w->rhs_v[k], xu,
                                 xv, formed[k])) != FSI_OK)
            return bail(rc);
        if ((rc = projection(c, w, xu, xv, stage_u(si), stage_v(si), 1, k, planned ? c->plan_proj[k] : 0)) !=
            FSI_OK)
            return bail(rc);
        if (si == 1 && w->tape_U1) {  // the adjoint's tape (nothing else reads these copies)
            HIPCHECK(c, hipMemcpyAsync(w->tape_U1, stage_u(1), c->felems * sizeof(double), hipMemcpyDeviceToDevice, A));
            HIPCHECK(c, hipMemcpyAsync(w->tape_V1, stage_v(1), c->felems * sizeof(double), hipMemcpyDeviceToDevice, A));
        }
        if (si <= 2) {
            if (dual) {
                HIPCHECK(c, hipEventRecord(c->ev_proj[si], A));
                HIPCHECK(c, hipStreamWaitEvent(B, c->ev_proj[si], 0));     // the stage velocity
                HIPCHECK(c, hipStreamWaitEvent(A, c->ev_cells[si - 1], 0));  // phi, sigma of the stage
            }
            sums = sums_of(si);
            csums = cells_of(si);
            const int t = si - 1;
            tendencies(c, w, si, stage_u(si), stage_v(si), sphi[t], ssxx[t], ssyy[t], ssxy[t], 0, 0, &sums, &csums,
                       B);
            if (dual) HIPCHECK(c, hipEventRecord(c->ev_cells[si], B));
        }
    }
    if (dual) HIPCHECK(c, hipStreamWaitEvent(A, c->ev_cells[2], 0));
    launch_ext_final(c->stream, p, c->st);  // the report of the last stage and the stage-1 checks
    return FSI_OK;
}

// Launch the planned step as a graph, capturing it on first use.  A graph
// is valid for the buffers it was captured with: the state/stage pairs, the
// correction and the warm-start histories rotate between steps, so a few
// graphs serve a run at a fixed dt and plan.
static std::vector<uintptr_t> graph_key(const fsi_context* c, const fsi_state* s, const fsi_workspace* w, double dt) {
    std::vector<uintptr_t> k;
    uintptr_t bits = 0;
    std::memcpy(&bits, &dt, sizeof(bits));
    k.push_back(bits);
    k.push_back((uintptr_t)w->stage_warm);
    k.push_back((uintptr_t)c->vel_warm);  // the adjoint's recomputations step without the warm starts
    k.push_back((uintptr_t)w->corr_cur);
    for (int q = 0; q < 3; ++q) {
        k.push_back((uintptr_t)c->plan_vel[q]);
        k.push_back((uintptr_t)c->plan_proj[q]);
        k.push_back((uintptr_t)c->vel_check_from[q]);
    }
    const double* bufs[] = {s->U, s->V, s->phi, s->sxx, s->syy, s->sxy, w->Us, w->Vs, w->phis, w->sxxs, w->syys,
                            w->sxys, w->Us2, w->Vs2};
    for (const double* b : bufs) k.push_back((uintptr_t)b);
    for (int q = 0; q < 3; ++q) {
        const double* h[] = {w->cq_cur[q],     w->cq_prev[q],    w->cq_next[q],    w->vel_cur_u[q],
                             w->vel_cur_v[q],  w->vel_prev_u[q], w->vel_prev_v[q], w->vel_old_u[q],
                             w->vel_old_v[q],  w->vel_old2_u[q], w->vel_old2_v[q], w->vel_next_u[q],
                             w->vel_next_v[q]};
        for (const double* b : h) k.push_back((uintptr_t)b);
    }
    return k;
}

static fsi_status run_step_graph(fsi_context* c, fsi_state* s, fsi_workspace* w, double dt) {
    const std::vector<uintptr_t> key = graph_key(c, s, w, dt);
    fsi_workspace::StepGraph* hit = nullptr;
    for (auto& g : w->graphs)
        if (g.key == key) {
            hit = &g;
            break;
        }
    if (!hit) {
        fsi_workspace::StepGraph g{};
        g.key = key;
        hipGraph_t graph = nullptr;
        HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        fsi_status rc = enqueue_step(c, s, w, dt, true);
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc != FSI_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        HIPCHECK(c, e);
        e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHECK(c, e);
        if (w->graphs.size() >= 16) {
            (void)hipGraphExecDestroy(w->graphs.front().exec);
            w->graphs.erase(w->graphs.begin());
        }
        w->graphs.push_back(g);
        hit = &w->graphs.back();
    }
    HIPCHECK(c, hipGraphLaunch(hit->exec, c->stream));
    return FSI_OK;
}

static fsi_status step_check(fsi_context* c, fsi_state* s, fsi_workspace* w, double t, double dt) {
    if (!c || !s || !w) return FSI_E_ARGUMENT;
    if (s->res_nonfinite) return fail(c, FSI_E_ARGUMENT, "state.resource contains non-finite values.");
    if (!std::isfinite(t)) return fail(c, FSI_E_ARGUMENT, "t must be finite.");
    if (!std::isfinite(dt)) return fail(c, FSI_E_ARGUMENT, "dt must be finite.");
    if (!(dt > 0.0)) return fail(c, FSI_E_ARGUMENT, "dt must be strictly positive.");
    const Params& p = c->prm;
    // _inject_compact_predictor! requirements (:556-578)
    if (!p.px && c->bcs.U.right.kind != FSI_BC_DIRICHLET)
        return fail(c, FSI_E_ARGUMENT, "The compact predictor currently requires a Dirichlet or Periodic U-right boundary.");
    if (!p.py && c->bcs.V.top.kind != FSI_BC_DIRICHLET)
        return fail(c, FSI_E_ARGUMENT, "The compact predictor currently requires a Dirichlet or Periodic V-top boundary.");
    return FSI_OK;
}

struct StepRun {
    bool planned = false;  // the planned step and its status read-back are enqueued
};

// Enqueues the planned step (once the iteration history allows one) and the
// read-back of its Status; nothing is awaited, so the steps of several
// contexts can be in flight together.
// The iteration plans of a planned step from the context's history; false
// when the step has to poll (no history yet, or planning switched off).
static bool plan_step(fsi_context* c) {
    if (c->growth) return false;  // the growth step polls its solves
    if (!(c->use_graph && c->have_history)) return false;
    // planned step: every solve gets its planned iteration count (kernels
    // past convergence return at once), one read-back at the end
    // a solve of k iterations is seen converged by launch k + 1 (kernel m
    // tests the residual of iterate m), so the plan is the maximum count
    // of the last kHist steps plus one; a larger count redoes the step
    static const int kVelMargin = 1, kProjMargin = 1;
    const int nh = std::min(c->hist_n, fsi_context::kHist);
    // Kernels from the fewest iterations of that history on test the
    // previous iterate before loading the fields, so the launch that sees
    // convergence is short.
    for (int k = 0; k < 3; ++k) {
        int mv = c->last_vel_iters[k], mp = c->last_proj_iters[k], lo = mv;
        for (int h = 0; h < nh; ++h) {
            mv = std::max(mv, c->hist_vel[k][h]);
            mp = std::max(mp, c->hist_proj[k][h]);
            lo = std::min(lo, c->hist_vel[k][h]);
        }
        c->plan_vel[k] = mv + kVelMargin;
        c->vel_check_from[k] = lo;
        c->plan_proj[k] = mp + kProjMargin;
        if (c->num.precision == FSI_PRECISION_MIXED) {
            // mixed: the most inner iterations of a cycle plus one, the most cycles
            int mi = c->last_imax[k], mc = 1;
            for (int h = 0; h < nh; ++h) {
                mi = std::max(mi, c->hist_imax[k][h]);
                mc = std::max(mc, c->hist_cyc[k][h]);
            }
            c->plan_vel[k] = mi + kVelMargin;
            c->plan_cyc[k] = mc;
        }
    }
    return true;
}

static fsi_status step_begin(fsi_context* c, fsi_state* s, fsi_workspace* w, double dt, StepRun& run) {
    run.planned = false;
    fsi_status rc = FSI_OK;
    if (!plan_step(c)) return FSI_OK;
    if ((rc = status_push_reset(c)) != FSI_OK) return rc;
    if (c->use_graph == 2) {  // planned counts, launched directly
        if ((rc = enqueue_step(c, s, w, dt, true)) != FSI_OK) return rc;
    } else if ((rc = run_step_graph(c, s, w, dt)) != FSI_OK) {
        return rc;
    }
    HIPCHECK(c, hipMemcpyAsync(c->st_host, c->st, sizeof(Status), hipMemcpyDeviceToHost, c->stream));
    run.planned = true;
    return FSI_OK;
}

// Waits for a planned step (redoing it on the polling path if a solve needed
// more iterations than planned) or runs the polling path, then applies the
// reference's checks and swaps the new state in.
static fsi_status step_finish(fsi_context* c, fsi_state* s, fsi_workspace* w, double dt, const StepRun& run,
                              fsi_step_report* rep, fsi_step_stats* stats) {
    const Params& p = c->prm;
    fsi_status rc = FSI_OK;
    bool done = false, replayed = false;
    if (run.planned) {
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        HIPCHECK(c, hipGetLastError());
        const Status& st = *c->st_host;
        bool ok = true;
        for (int k = 0; k < 3; ++k) ok = ok && (!p.visc || st.vel_ok[k]) && st.proj_ok[k];
        ++c->spec_steps;
        if (ok) {
            done = replayed = true;
        } else {  // a solve needed more iterations than planned: redo with polling
            ++c->spec_misses;
        }
    }
    if (!done) {
        if ((rc = status_push_reset(c)) != FSI_OK) return rc;
        if ((rc = enqueue_step(c, s, w, dt, false)) != FSI_OK) return rc;
        if ((rc = status_pull(c)) != FSI_OK) return rc;
    }
    {
        const Status& st = *c->st_host;
        for (int k = 0; k < 3; ++k) {
            if (!p.visc || st.vel_ok[k]) {
                c->last_vel_iters[k] = std::max(1, st.vel_iters[k]);
                c->last_imax[k] = std::max(1, st.vel_imax[k]);
            }
            if (st.proj_ok[k]) c->last_proj_iters[k] = std::max(0, st.proj_iters[k]);
        }
        bool all = true;
        for (int k = 0; k < 3; ++k) all = all && (!p.visc || st.vel_ok[k]) && st.proj_ok[k];
        c->have_history = all ? 1 : 0;
        if (all) {
            const int h = c->hist_n % fsi_context::kHist;
            for (int k = 0; k < 3; ++k) {
                c->hist_vel[k][h] = c->last_vel_iters[k];
                c->hist_proj[k][h] = c->last_proj_iters[k];
                c->hist_imax[k][h] = c->last_imax[k];
                c->hist_cyc[k][h] = std::max(1, st.vel_cycles[k]);
            }
            ++c->hist_n;
        }
    }
    const Status& st = *c->st_host;
    // error priority follows the reference's order of checks
    if (st.nonfinite_in) {
        return fail(c, FSI_E_ARGUMENT, "the state contains non-finite values.");
    }
    if (c->num.check_stage1) {  // _check_stage_one_constraint! :1116-1127
        double tol = c->num.stage1_atol + c->num.stage1_rtol * std::fmax(st.stage1_gmax, 1.0);
        if (!(st.stage1_linf <= tol)) {
            return fail(c, FSI_E_ARGUMENT,
                        "The velocity entering stage 1 does not satisfy the prescribed divergence constraint. "
                        "Maximum |div(v)-G| = %.6e, tolerance = %.6e. Initialize the state with "
                        "make_velocity_compatible! or provide a compatible velocity.", st.stage1_linf, tol);
        }
    }
    if ((rc = device_error(c)) != FSI_OK) return rc;
    // the fourth stage is the new state (:1232-1238): swap buffers
    std::swap(s->U, w->Us);
    std::swap(s->V, w->Vs);
    std::swap(s->phi, w->phis);
    std::swap(s->sxx, w->sxxs);
    std::swap(s->syy, w->syys);
    std::swap(s->sxy, w->sxys);
    if (c->growth) {  // copyto!(state.resource, work.c_stage) :1235
        std::swap(s->res, w->cs);
        s->res_min = st.res_min;
    }
    // commit the workspace: the step's correction, the stage histories
    w->corr_cur = 1 - w->corr_cur;
    {
        for (int k = 0; k < 3; ++k) {  // (cur, prev, next) <- (next, cur, prev)
            double* q = w->cq_next[k];
            w->cq_next[k] = w->cq_prev[k];
            w->cq_prev[k] = w->cq_cur[k];
            w->cq_cur[k] = q;
        }
        for (int k = 0; k < 3; ++k) {  // (cur, prev, old, old2, next) <- (next, cur, prev, old, old2)
            double* nu = w->vel_next_u[k];
            double* nv = w->vel_next_v[k];
            w->vel_next_u[k] = w->vel_old2_u[k];
            w->vel_next_v[k] = w->vel_old2_v[k];
            w->vel_old2_u[k] = w->vel_old_u[k];
            w->vel_old2_v[k] = w->vel_old_v[k];
            w->vel_old_u[k] = w->vel_prev_u[k];
            w->vel_old_v[k] = w->vel_prev_v[k];
            w->vel_prev_u[k] = w->vel_cur_u[k];
            w->vel_prev_v[k] = w->vel_cur_v[k];
            w->vel_cur_u[k] = nu;
            w->vel_cur_v[k] = nv;
        }
        if (w->stage_warm < 4) ++w->stage_warm;
    }
    if (rep) {
        rep->maximum_projection_constraint_linf = std::fmax(st.stage1_linf, st.max_constraint_linf);
        rep->maximum_growth_trace_error = st.growth_trace;
        rep->minimum_phi = st.phi_min;
        rep->maximum_phi = st.phi_max;
        rep->minimum_resource = s->res_min;
        rep->minimum_sigma_eigenvalue = st.min_eig;
        rep->maximum_abs_sigma_determinant_error = st.max_det_err;
        rep->non_spd_cells = st.nonspd;
    }
    if (stats) {
        for (int k = 0; k < 3; ++k) {
            stats->velocity_iterations[k] = st.vel_iters[k];
            stats->projection_iterations[k] = st.proj_iters[k];
        }
        stats->planned = replayed ? 1 : 0;
        stats->max_abs_u = st.max_abs_u;
        stats->max_abs_v = st.max_abs_v;
        for (int k = 0; k < 3; ++k) {
            stats->velocity_residual[k] = p.visc ? st.vel_res[k] : 0.0;
            stats->velocity_reference_residual[k] = p.visc ? st.vel_ref[k] : 0.0;
            stats->velocity_cycles[k] = p.visc ? st.vel_cycles[k] : 0;
        }
    }
    return FSI_OK;
}

extern "C" fsi_status fsi_integrate_step(fsi_context* c, fsi_state* s, fsi_workspace* w, double t, double dt,
                                         fsi_step_report* rep, fsi_step_stats* stats) {
    fsi_status rc = step_check(c, s, w, t, dt);
    if (rc != FSI_OK) return rc;
    StepRun run;
    if ((rc = step_begin(c, s, w, dt, run)) != FSI_OK) return rc;
    return step_finish(c, s, w, dt, run, rep, stats);
}

// What must agree for two instances' planned steps to launch the same
// kernel sequence: the grid shape and device, the precision, and per stage
// whether the projection only tests the direct solve (a plan of one) or runs
// PCG iterations; in mixed precision also the planned refinement cycles.
// Larger iteration plans are harmless (kernels past convergence return).
static std::vector<long long> batch_key(const fsi_context* c) {
    std::vector<long long> key{c->grid.nx, c->grid.ny, c->device, c->num.precision};
    for (int k = 0; k < 3; ++k) {
        key.push_back(std::min(c->plan_proj[k], c->num.projection_maxiter + 1) == 1);
        if (c->num.precision == FSI_PRECISION_MIXED) key.push_back(c->plan_cyc[k]);
    }
    return key;
}

// A group of instances whose planned steps run as batched launches: the
// launch sequence recorded per instance, and the argument tables (launch j:
// the K tuples `stride[j]` bytes apart from off[j]) staged in the leader's
// pinned buffer.
struct BatchGroup {
    std::vector<int> ids;
    std::vector<LaunchRecorder> rec;
    std::vector<size_t> off, stride;
    size_t total = 0;
};

// Records one planned step of each instance of the group with the plans
// equalised (the largest iteration counts of the group: kernels past a
// solve's convergence return at once, so each instance's result is the one
// its own planned step gives) and stages the argument tables.  False, with
// nothing enqueued, when the sequences differ kernel by kernel or the tables
// cannot be allocated; the instances are then stepped one by one.
static bool record_group(BatchGroup& g, fsi_context* const* ctx, fsi_state* const* st, fsi_workspace* const* w,
                         const double* dt) {
    const int K = (int)g.ids.size();
    fsi_context* lead = ctx[g.ids[0]];
    for (int k = 0; k < 3; ++k) {
        int pv = 0, pp = 0, pc = 0, cf = 1 << 30;
        for (int i : g.ids) {
            pv = std::max(pv, ctx[i]->plan_vel[k]);
            pp = std::max(pp, ctx[i]->plan_proj[k]);
            pc = std::max(pc, ctx[i]->plan_cyc[k]);
            cf = std::min(cf, ctx[i]->vel_check_from[k]);
        }
        for (int i : g.ids) {
            ctx[i]->plan_vel[k] = pv;
            ctx[i]->plan_proj[k] = pp;
            ctx[i]->plan_cyc[k] = pc;
            ctx[i]->vel_check_from[k] = cf;
        }
    }
    g.rec.assign((size_t)K, LaunchRecorder{});
    bool ok = true;
    for (int q = 0; q < K && ok; ++q) {
        fsi_context* c = ctx[g.ids[q]];
        g_recorder = &g.rec[q];
        const int dual = c->dual;
        c->dual = 0;
        ok = status_push_reset(c) == FSI_OK &&
             enqueue_step(c, st[g.ids[q]], w[g.ids[q]], dt[g.ids[q]], true) == FSI_OK;
        c->dual = dual;
        g_recorder = nullptr;
        ok = ok && !g.rec[q].overflow;
    }
    const size_t L = g.rec[0].launches.size();
    for (int q = 1; q < K && ok; ++q) {
        ok = g.rec[q].launches.size() == L;
        for (size_t j = 0; j < L && ok; ++j) {
            const RecordedLaunch &a = g.rec[0].launches[j], &b = g.rec[q].launches[j];
            ok = a.launch == b.launch && a.size == b.size && a.shared == b.shared && a.grid.x == b.grid.x &&
                 a.grid.y == b.grid.y && a.block.x == b.block.x && a.block.y == b.block.y &&
                 a.block.z == b.block.z;
        }
    }
    if (!ok || L == 0) return false;
    g.off.resize(L);
    g.stride.resize(L);
    g.total = 0;
    for (size_t j = 0; j < L; ++j) {
        g.stride[j] = (g.rec[0].launches[j].size + 15) / 16 * 16;
        g.off[j] = g.total;
        g.total = (g.total + (size_t)K * g.stride[j] + 255) / 256 * 256;
    }
    if (g.total > lead->batch_cap) {
        if (lead->batch_dev) (void)hipFree(lead->batch_dev);
        if (lead->batch_host) (void)hipHostFree(lead->batch_host);
        lead->batch_dev = nullptr;
        lead->batch_host = nullptr;
        lead->batch_cap = 0;
        const size_t cap = g.total + g.total / 2;
        if (hipMalloc((void**)&lead->batch_dev, cap) != hipSuccess ||
            hipHostMalloc((void**)&lead->batch_host, cap, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        lead->batch_cap = cap;
    }
    for (size_t j = 0; j < L; ++j)
        for (int q = 0; q < K; ++q)
            std::memcpy(lead->batch_host + g.off[j] + (size_t)q * g.stride[j],
                        g.rec[q].args.data() + g.rec[q].launches[j].offset, g.rec[q].launches[j].size);
    return true;
}

// The group's tables, its launches (every kernel once for the whole group,
// grid.z = instances) and the read-back of each instance's Status, on S.
static bool enqueue_group(const BatchGroup& g, fsi_context* const* ctx, hipStream_t S) {
    const fsi_context* lead = ctx[g.ids[0]];
    const int K = (int)g.ids.size();
    if (hipMemcpyAsync(lead->batch_dev, lead->batch_host, g.total, hipMemcpyHostToDevice, S) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    for (size_t j = 0; j < g.rec[0].launches.size(); ++j) {
        const RecordedLaunch& a = g.rec[0].launches[j];
        a.launch(a.grid, a.block, a.shared, S, lead->batch_dev + g.off[j], K, (int)g.stride[j]);
    }
    for (int i : g.ids) (void)hipMemcpyAsync(ctx[i]->st_host, ctx[i]->st, sizeof(Status), hipMemcpyDeviceToHost, S);
    return true;
}

extern "C" fsi_status fsi_ensemble_step(int32_t n, fsi_context* const* ctx, fsi_state* const* st,
                                        fsi_workspace* const* w, const double* t, const double* dt,
                                        fsi_step_report* reps, fsi_step_stats* stats, int32_t* status) {
    if (n < 0) return FSI_E_ARGUMENT;
    if (n == 0) return FSI_OK;
    if (!ctx || !st || !w || !t || !dt || !status) return FSI_E_ARGUMENT;
    for (int k = 0; k < n; ++k)
        if (!ctx[k]) return FSI_E_ARGUMENT;
    std::vector<StepRun> runs((size_t)n);
    // Instances with a planned step and the same launch sequence (batch_key)
    // form groups stepped by batched launches: every group's launches are
    // enqueued back to back on one stream (the groups one after the other:
    // groups running side by side on several streams gave results that
    // differed run to run in the last bits, which one stream does not), then
    // each instance's step is finished as fsi_integrate_step finishes it.
    // The rest, and any group whose sequences turn out to differ, are stepped
    // one by one by host threads, each instance on its own stream.
    // FSI_ENSEMBLE_BATCH=0 steps every instance on its own.
    const char* eb = std::getenv("FSI_ENSEMBLE_BATCH");
    const bool batch = !(eb && eb[0] == '0');
    std::vector<BatchGroup> groups;
    std::vector<int> single;
    {
        std::vector<char> planned((size_t)n, 0), taken((size_t)n, 0);
        for (int k = 0; k < n; ++k) {
            status[k] = step_check(ctx[k], st[k], w[k], t[k], dt[k]);
            planned[k] = batch && status[k] == FSI_OK && plan_step(ctx[k]);
        }
        for (int k = 0; k < n; ++k) {
            if (status[k] != FSI_OK || taken[k]) continue;
            taken[k] = 1;
            if (!planned[k]) {
                single.push_back(k);
                continue;
            }
            BatchGroup g;
            g.ids.push_back(k);
            const std::vector<long long> key = batch_key(ctx[k]);
            for (int j = k + 1; j < n; ++j)
                if (planned[j] && !taken[j] && status[j] == FSI_OK && batch_key(ctx[j]) == key) {
                    g.ids.push_back(j);
                    taken[j] = 1;
                }
            if (g.ids.size() >= 2) groups.push_back(std::move(g));
            else single.push_back(k);
        }
    }
    auto finish = [&](int k) {
        if (status[k] != FSI_OK) return;
        const int dual = ctx[k]->dual;
        ctx[k]->dual = 0;
        status[k] = step_finish(ctx[k], st[k], w[k], dt[k], runs[k], reps ? &reps[k] : nullptr,
                                stats ? &stats[k] : nullptr);
        ctx[k]->dual = dual;
    };
    // the batched groups
    if (!groups.empty()) {
        const int dev = ctx[groups[0].ids[0]]->device;
        std::vector<char> recorded(groups.size(), 0);
        bool dev_ok = hipSetDevice(dev) == hipSuccess;
        for (size_t q = 0; q < groups.size(); ++q) {
            BatchGroup& g = groups[q];
            recorded[q] = dev_ok && ctx[g.ids[0]]->device == dev && record_group(g, ctx, st, w, dt);
            if (!recorded[q])
                for (int k : g.ids) single.push_back(k);  // one by one, each with its own plan again
        }
        hipStream_t S = nullptr;
        for (size_t q = 0; q < groups.size(); ++q)
            if (recorded[q]) {
                if (!S) S = ctx[groups[q].ids[0]]->stream;
                for (int k : groups[q].ids)  // after whatever their own streams still hold (state uploads)
                    if (ctx[k]->stream != S && hipStreamSynchronize(ctx[k]->stream) != hipSuccess)
                        status[k] = fail(ctx[k], FSI_E_HIP, "hipStreamSynchronize failed");
            }
        std::vector<char> enq(groups.size(), 0);
        for (size_t q = 0; q < groups.size(); ++q)
            if (recorded[q]) {
                enq[q] = enqueue_group(groups[q], ctx, S);
                if (!enq[q])
                    for (int k : groups[q].ids) single.push_back(k);
            }
        const hipError_t e = S ? hipStreamSynchronize(S) : hipSuccess;
        for (size_t q = 0; q < groups.size(); ++q)
            if (enq[q])
                for (int k : groups[q].ids) {
                    if (e != hipSuccess && status[k] == FSI_OK)
                        status[k] = fail(ctx[k], FSI_E_HIP, "batched step: %s", hipGetErrorString(e));
                    runs[k].planned = e == hipSuccess;
                    finish(k);
                }
    }
    // the instances stepped one by one: host threads share the launches;
    // thread q enqueues the planned steps of instances q, q + T, ... (every
    // one in flight before it waits for any), then finishes them.  Each runs
    // on its context's main stream alone: a second stream per instance (the
    // single-step overlap of the phase and sigma kernels with the next
    // solve) would put two queues per instance on the hardware's few
    // (measured: 64 cavities 95 M against 166 M cell updates/s).
    int nt = 8;
    if (const char* e = std::getenv("FSI_ENSEMBLE_THREADS")) nt = std::max(1, std::atoi(e));
    const unsigned hw = std::thread::hardware_concurrency();
    if (hw > 0) nt = std::min(nt, (int)hw);
    nt = std::min(nt, (int)single.size());
    auto work = [&](int q) {
        for (size_t i = (size_t)q; i < single.size(); i += (size_t)nt) {
            const int k = single[i];
            if (status[k] != FSI_OK) continue;
            if (hipSetDevice(ctx[k]->device) != hipSuccess) {
                status[k] = fail(ctx[k], FSI_E_HIP, "hipSetDevice(%d) failed", ctx[k]->device);
                continue;
            }
            const int dual = ctx[k]->dual;
            ctx[k]->dual = 0;
            status[k] = step_begin(ctx[k], st[k], w[k], dt[k], runs[k]);
            ctx[k]->dual = dual;
        }
        for (size_t i = (size_t)q; i < single.size(); i += (size_t)nt) finish(single[i]);
    };
    if (nt == 1) {
        work(0);
    } else if (nt > 1) {
        std::vector<std::thread> pool;
        pool.reserve((size_t)nt);
        for (int q = 0; q < nt; ++q) pool.emplace_back(work, q);
        for (auto& th : pool) th.join();
    }
    fsi_status first = FSI_OK;
    for (int k = 0; k < n; ++k)
        if (status[k] != FSI_OK && first == FSI_OK) first = (fsi_status)status[k];
    return first;
}

// ---------------------------------------------------------------------------
// measurement hook
// ---------------------------------------------------------------------------
extern "C" fsi_status fsi_benchmark_kernel(fsi_context* c, fsi_state* s, fsi_workspace* w, const char* name,
                                           int32_t reps, double* avg_ms, double* bytes) {
    if (!c || !s || !w || !name || reps <= 0) return FSI_E_ARGUMENT;
    const Params& p = c->prm;
    VelBufs vb{};
    vb.Un = s->U;
    vb.Vn = s->V;
    vb.ru = w->ru;
    vb.rv = w->rv;
    vb.xu = w->Us;
    vb.xv = w->Vs;
    for (int k = 0; k < 2; ++k) {
        vb.r_u[k] = w->r_u[k];
        vb.r_v[k] = w->r_v[k];
        vb.p_u[k] = w->p_u[k];
        vb.p_v[k] = w->p_v[k];
        vb.q_u[k] = w->q_u[k];
        vb.q_v[k] = w->q_v[k];
        vb.fr_u[k] = w->fr_u[k];  // mixed precision only (null otherwise)
        vb.fr_v[k] = w->fr_v[k];
        vb.fp_u[k] = w->fp_u[k];
        vb.fp_v[k] = w->fp_v[k];
        vb.fq_u[k] = w->fq_u[k];
        vb.fq_v[k] = w->fq_v[k];
    }
    vb.partials = c->partials;
    ProjBufs pb{};
    pb.U = pb.Uin = w->Us;
    pb.V = pb.Vin = w->Vs;
    pb.lap_bc = c->lap_bc;
    pb.rhs = w->rhs;
    pb.x = w->corr[0];
    pb.q_out = w->corr[1];
    pb.r = w->r;
    pb.z = w->z;
    pb.p[0] = w->pp[0];
    pb.p[1] = w->pp[1];
    pb.w = w->w;
    pb.partials = c->partials;
    // keep every kernel running: done = 0, no tolerance reached, no maxiter
    status_reset(c);
    c->st_host->vel.tol2 = c->st_host->pcg.tol2 = -1.0;
    c->st_host->vel.maxiter = c->st_host->pcg.maxiter = 1 << 30;
    c->st_host->pcg.rzs[0] = c->st_host->pcg.rzs[1] = 1.0;
    fsi_status rc = status_push(c);
    if (rc != FSI_OK) return rc;
    pb.x = w->lv[0].r;  // scratch: the correction warm start is left untouched
    const std::string n(name);
    auto launch = [&]() -> double {
        if (n == "momentum") {
            launch_momentum(c->stream, p, s->U, s->V, s->phi, s->sxx, s->syy, s->sxy, c->has_reset, w->Eu[0],
                            w->Ev[0], w->Iu[0], w->Iv[0]);
            return 10.0 * 8.0 * p.nx * (double)p.ny;
        }
        if (n == "cells") {
            launch_cells(c->stream, p, c->bpsi, s->U, s->V, s->phi, s->sxx, s->syy, s->sxy, c->has_reset,
                         w->Ep[0], w->Exx[0], w->Eyy[0], w->Exy[0], &c->st->cells_err[0], 0);
            return 10.0 * 8.0 * p.nx * (double)p.ny;
        }
        return bench_launch(c->stream, name, p, vb, pb, w->lv.data(), c->geo.data(), (int)c->geo.size(), c->cplan,
                            c->st, &c->spec);
    };
    double b = launch();
    if (b < 0) return fail(c, FSI_E_ARGUMENT, "unknown kernel name %s", name);
    if (c->geo.size() < 2 && (n == "mg_prolong" || n == "mg_restrict"))
        return fail(c, FSI_E_ARGUMENT, "the grid has a single multigrid level");
    hipEvent_t e0, e1;
    HIPCHECK(c, hipEventCreate(&e0));
    HIPCHECK(c, hipEventCreate(&e1));
    HIPCHECK(c, hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) launch();
    HIPCHECK(c, hipEventRecord(e1, c->stream));
    HIPCHECK(c, hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHECK(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHECK(c, hipGetLastError());
    if (avg_ms) *avg_ms = (double)ms / reps;
    if (bytes) *bytes = b;
    return FSI_OK;
}
