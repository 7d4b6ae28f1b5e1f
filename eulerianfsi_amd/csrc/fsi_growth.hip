// fsi_growth.hip -- the growth and resource path of the IMEX step: the
// scalar volumetric source G(phi, c) and growth tensor K_g
// (src/GrowthModels.jl), the resource stage (src/ResourceTransport.jl, the
// diffusion operator of src/Operators.jl:361-384 and its implicit solve,
// src/TimeIntegrationIMEX.jl:742-774) and the growth terms of the stage
// tendencies (TimeIntegrationIMEX.jl:926-1071).
//
// These kernels run beside the tendency kernels of the no-growth step, which
// store E_u, E_v, I_u, I_v, E_phi, E_sigma of a stage; the growth terms are
// then added to the stored values (S v, phi S, the K_g part of the structure
// tensor right-hand side, the divergence of -2 mu K_g), and a stage kernel
// forms the next stage's values.  Every kernel streams cell fields of pitch P
// with one thread per cell (blocks of 64 x 4 threads over 64 x 16 cells, as
// the physics kernels), reading neighbours through the caches: the path is
// HBM-bound at a few passes per field and has no LDS reuse worth a tile.
#include "fsi_kernels.hpp"

namespace fsi {

namespace {

__device__ __forceinline__ int gr_bid() { return blockIdx.x + blockIdx.y * gridDim.x; }
__device__ __forceinline__ int gr_lane() { return threadIdx.x + threadIdx.y * blockDim.x; }

#define GR_FOR(p)                                                               \
    const int i = blockIdx.x * TILE_X + threadIdx.x;                            \
    for (int rr_ = 0; rr_ < ROWS; ++rr_)                                        \
        if (const int j = blockIdx.y * TILE_ROWS + threadIdx.y + rr_ * TILE_Y;  \
            i < (p).nx && j < (p).ny)

dim3 gr_grid(const Params& p) { return dim3((p.nx + TILE_X - 1) / TILE_X, (p.ny + TILE_ROWS - 1) / TILE_ROWS); }
const dim3 kGrBlock(TILE_X, TILE_Y);

// block maxima of NV values (lane 0 holds the result)
template <int NV>
__device__ __forceinline__ void block_maxes(double (&v)[NV]) {
    __shared__ double l[NV][BLOCK / 64];
    const int lane = gr_lane();
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_max(v[k]);
    if ((lane & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) l[k][lane >> 6] = v[k];
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k)
            for (int w = 1; w < BLOCK / 64; ++w) v[k] = fmax(v[k], l[k][w]);
    __syncthreads();
}

// growth_source! GrowthModels.jl:214-256
__device__ __forceinline__ double gr_source(const GrowthP& g, double phi, double c) {
    if (g.src == FSI_GROWTH_CONSTANT) return g.rate * phi;
    if (g.src == FSI_GROWTH_SATURATING) return g.rate * phi * c / (g.K + c);
    return 0.0;
}

// diffusivity_field! TransportConstitutive.jl:241-263 (phi unclamped)
__device__ __forceinline__ double gr_diffusivity(const GrowthP& g, double phi) {
    return g.diff == FSI_DIFFUSION_HARMONIC ? 1.0 / ((1.0 - phi) / g.Df + phi / g.Ds) : g.Df;
}

// L x at cell (i, j) of the diffusion operator (Operators.jl:189-384) with
// homogeneous boundary data: face fluxes D_f (grad x) with the resource BC
// rows (Dirichlet: the wall value at h/2, Neumann: no matrix part, periodic:
// wrapped), then their cell divergence
__device__ __forceinline__ double gr_lap(const Params& p, const GrowthP& g, const double* __restrict__ Dx,
                                         const double* __restrict__ Dy, const double* __restrict__ x, int i, int j,
                                         size_t k) {
    const double xc = x[k];
    double fl, fr, fb, ft;
    if (i > 0) fl = Dx[k] * ((xc - x[k - 1]) * p.ihx);
    else if (p.px) fl = Dx[k] * ((xc - x[k + p.nx - 1]) * p.ihx);
    else fl = g.bC.l.kind == FSI_BC_DIRICHLET ? Dx[k] * (xc * (2.0 * p.ihx)) : 0.0;
    if (i < p.nx - 1) fr = Dx[k + 1] * ((x[k + 1] - xc) * p.ihx);
    else if (p.px) fr = Dx[k + 1] * ((x[k - (p.nx - 1)] - xc) * p.ihx);
    else fr = g.bC.r.kind == FSI_BC_DIRICHLET ? Dx[k + 1] * (-xc * (2.0 * p.ihx)) : 0.0;
    const size_t P = p.P;
    if (j > 0) fb = Dy[k] * ((xc - x[k - P]) * p.ihy);
    else if (p.py) fb = Dy[k] * ((xc - x[k + (size_t)(p.ny - 1) * P]) * p.ihy);
    else fb = g.bC.b.kind == FSI_BC_DIRICHLET ? Dy[k] * (xc * (2.0 * p.ihy)) : 0.0;
    if (j < p.ny - 1) ft = Dy[k + P] * ((x[k + P] - xc) * p.ihy);
    else if (p.py) ft = Dy[k + P] * ((x[k - (size_t)(p.ny - 1) * P] - xc) * p.ihy);
    else ft = g.bC.t.kind == FSI_BC_DIRICHLET ? Dy[k + P] * (-xc * (2.0 * p.ihy)) : 0.0;
    return (fr - fl) * p.ihx + (ft - fb) * p.ihy;
}

// the boundary vector b_L of the diffusion operator at cell (i, j): the
// gradient's boundary data (Dirichlet -/+ g / (h/2), Neumann -/+ g on the
// low / high side) times the face diffusivity, through the divergence
__device__ __forceinline__ double gr_lap_bc(const Params& p, const GrowthP& g, const double* __restrict__ Dx,
                                            const double* __restrict__ Dy, int i, int j, size_t k) {
    double s = 0.0;
    auto lo = [](const BcSide& b, double ih) {
        return b.kind == FSI_BC_DIRICHLET ? -b.value * (2.0 * ih) : (b.kind == FSI_BC_NEUMANN ? -b.value : 0.0);
    };
    auto hi = [](const BcSide& b, double ih) {
        return b.kind == FSI_BC_DIRICHLET ? b.value * (2.0 * ih) : (b.kind == FSI_BC_NEUMANN ? b.value : 0.0);
    };
    if (!p.px && i == 0) s -= Dx[k] * lo(g.bC.l, p.ihx) * p.ihx;
    if (!p.px && i == p.nx - 1) s += Dx[k + 1] * hi(g.bC.r, p.ihx) * p.ihx;
    if (!p.py && j == 0) s -= Dy[k] * lo(g.bC.b, p.ihy) * p.ihy;
    if (!p.py && j == p.ny - 1) s += Dy[k + p.P] * hi(g.bC.t, p.ihy) * p.ihy;
    return s;
}

// ---------------------------------------------------------------------------
// G, K_g trace error, stage-1 checks (_evaluate_growth! :842-865,
// _stage_one_constraint_linf :1111-1114)
// ---------------------------------------------------------------------------
FSI_KERNEL void k_gr_source(Params p, GrowthP g, GrSrc a) {
    double mx[3] = {0.0, 0.0, 0.0};  // trace error, |div v - G|, |G|
    double sm[2] = {0.0, 0.0};       // G^2 dA, non-finite inputs
    GR_FOR(p) {
        const size_t k = at(p, i, j);
        const double ph = a.phi[k], cc = a.c[k];
        const double G = gr_source(g, ph, cc);
        a.G[k] = G;
        mx[0] = fmax(mx[0], fabs(g.kxx * G + g.kyy * G - G));
        sm[0] += G * G;
        if (a.stage1) {
            const double div = (a.U[k + 1] - a.U[k]) * p.ihx + (a.V[k + p.P] - a.V[k]) * p.ihy;
            mx[1] = fmax(mx[1], fabs(div - G));
            mx[2] = fmax(mx[2], fabs(G));
            int bad = !isfinite(ph) + !isfinite(cc) + !isfinite(a.sxx[k]) + !isfinite(a.syy[k]) +
                      !isfinite(a.sxy[k]) + !isfinite(a.U[k]) + !isfinite(a.V[k]);
            if (i == p.nx - 1) bad += !isfinite(a.U[k + 1]);
            if (j == p.ny - 1) bad += !isfinite(a.V[k + p.P]);
            sm[1] += (double)bad;
        }
    }
    sm[0] *= p.hx * p.hy;
    block_maxes<3>(mx);
    __shared__ double lds[2 * 16];
    block_sum<2>(sm, lds);
    if (gr_lane() != 0) return;
    const size_t nb = p.exnb, b = gr_bid();
    double* tr = p.exbuf + (size_t)EX_TRACE * nb + b;
    *tr = a.first ? mx[0] : fmax(*tr, mx[0]);
    if (a.stage1) {
        p.exbuf[(size_t)EX_STAGE1 * nb + b] = mx[1];
        p.exbuf[(size_t)EX_GMAX * nb + b] = mx[2];
        p.exbuf[(size_t)EX_NONFINITE_IN * nb + b] = sm[1];
    }
    a.part[b] = sm[0];
}

// one block: the sum of G^2 dA (the scale of the projection's constraint check)
FSI_KERNEL void k_gr_total(const double* part, int nb, Status* st) {
    double t[1];
    get_totals<1>(part, nb, t);
    if (gr_lane() == 0) st->proj.tgt_l2sq = t[0];
}

// ---------------------------------------------------------------------------
// Stage values (_build_explicit_stage! :716-740, the resource right-hand side
// of _build_resource_stage! :742-751, the velocity right-hand side of
// _build_velocity_stage! :776-800)
// ---------------------------------------------------------------------------
FSI_KERNEL void k_gr_stage(Params p, GrStage a) {
    GR_FOR(p) {
        const size_t k = at(p, i, j);
        const double phin = a.phin[k];
        double ph = phin, xx = a.sxxn[k], yy = a.syyn[k], xy = a.sxyn[k];
        if (a.reset_n && phin < p.ext_cutoff) {  // the safety reset of sigma^n (:686-702)
            xx = 1.0;
            yy = 1.0;
            xy = 0.0;
        }
        double cr = a.crhs ? a.cn[k] : 0.0;
        double ru = a.Un[k], rv = a.Vn[k];
        for (int q = 0; q < a.ns; ++q) {
            const double ae = a.ce[q], ai = a.ci[q];
            if (ae != 0.0) {
                const double f = a.dt * ae;
                ph += f * a.Ep[q][k];
                xx += f * a.Exx[q][k];
                yy += f * a.Eyy[q][k];
                xy += f * a.Exy[q][k];
            }
            if (ae != 0.0 || ai != 0.0) {
                if (a.crhs) cr += a.dt * (ae * a.Ec[q][k] + ai * (a.Ic[q] ? a.Ic[q][k] : 0.0));
                ru += a.dt * (ae * a.Eu[q][k] + ai * a.Iu[q][k]);
                rv += a.dt * (ae * a.Ev[q][k] + ai * a.Iv[q][k]);
            }
        }
        if (a.reset && ph < p.ext_cutoff) {  // apply_structure_extension! StructureTensor.jl:590-613
            xx = 1.0;
            yy = 1.0;
            xy = 0.0;
        }
        if (a.bu) {
            ru += a.alpha * a.bu[k];
            rv += a.alpha * a.bv[k];
        }
        a.phi[k] = ph;
        a.sxx[k] = xx;
        a.syy[k] = yy;
        a.sxy[k] = xy;
        if (a.crhs) a.crhs[k] = cr;
        a.ru[k] = ru;
        a.rv[k] = rv;
        if (a.mu) {
            a.mu[k] = a.Iu[0][k] - a.bu[k];
            a.mv[k] = a.Iv[0][k] - a.bv[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Face diffusivities (interpolate_center_to_x/y_faces, Operators.jl:110-159)
// ---------------------------------------------------------------------------
FSI_KERNEL void k_gr_diff(Params p, GrowthP g, const double* __restrict__ phi, double* __restrict__ Dx,
                          double* __restrict__ Dy, Status* st) {
    GR_FOR(p) {
        const size_t k = at(p, i, j), P = p.P;
        const double ph = phi[k];
        const double D = gr_diffusivity(g, ph);
        if (!isfinite(ph)) raise_error(st, DEV_NONFINITE_STATE, ph, 0.0);
        else if (!(isfinite(D) && D > 0.0)) raise_error(st, DEV_DIFFUSIVITY, D, ph);
        // x-faces i (and nx on the last column)
        const double Dw = i > 0 ? gr_diffusivity(g, phi[k - 1])
                                : (p.px ? gr_diffusivity(g, phi[k + p.nx - 1]) : D);
        Dx[k] = (i > 0 || p.px) ? 0.5 * Dw + 0.5 * D : D;
        if (i == p.nx - 1) Dx[k + 1] = p.px ? 0.5 * D + 0.5 * gr_diffusivity(g, phi[k - (p.nx - 1)]) : D;
        const double Ds = j > 0 ? gr_diffusivity(g, phi[k - P])
                                : (p.py ? gr_diffusivity(g, phi[k + (size_t)(p.ny - 1) * P]) : D);
        Dy[k] = (j > 0 || p.py) ? 0.5 * Ds + 0.5 * D : D;
        if (j == p.ny - 1) Dy[k + P] = p.py ? 0.5 * D + 0.5 * gr_diffusivity(g, phi[k - (size_t)(p.ny - 1) * P]) : D;
    }
}

// ---------------------------------------------------------------------------
// Resource stage solve: CG on (I - alpha L) x = b + alpha b_L from x0 = c_n,
// three kernels per iteration (q = A p with p.q; x, r update with r.r; the
// direction and the stopping rules of _solve_linear_system! :602-641), the
// r.r of iteration m in st->res.rzs[m & 1].
// ---------------------------------------------------------------------------
FSI_KERNEL void k_rcg_start(Params p, GrowthP g, ResCg a) {
    double acc[2] = {0.0, 0.0};  // r.r (= |b - A x0|^2), b.b
    GR_FOR(p) {
        const size_t k = at(p, i, j);
        const double b = a.b[k] + a.alpha * gr_lap_bc(p, g, a.Dx, a.Dy, i, j, k);
        const double x0 = a.x0[k];
        const double r = b - (x0 - a.alpha * gr_lap(p, g, a.Dx, a.Dy, a.x0, i, j, k));
        a.x[k] = x0;
        a.r[k] = r;
        a.p[k] = r;
        acc[0] += r * r;
        acc[1] += b * b;
    }
    put_partials<2>(acc, a.part, (int)p.exnb);
}

FSI_KERNEL void k_rcg_setup(ResCg a, int nb, Status* st) {
    double t[2];
    get_totals<2>(a.part, nb, t);
    if (gr_lane() != 0) return;
    KrylovScalars& s = st->res;
    const double eps = 2.220446049250313e-16;
    const double bscale = fmax(sqrt(t[1]), eps);
    const double fb = a.fb * bscale;
    const double dtol = fmin(1e-14 * bscale, fb);
    s.bnorm2 = t[1];
    s.ref2 = t[0];
    s.fb2 = fb * fb;
    s.dtol2 = dtol * dtol;
    s.sw = a.method == FSI_SOLVER_DIRECT;
    const double tol = s.sw ? dtol : a.rtol * sqrt(t[0]);
    s.tol2 = tol * tol;
    s.cap = s.sw ? (int)a.maxiter_d : a.maxiter;
    s.maxiter = (int)a.maxiter_d;
    s.rzs[0] = t[0];
    s.iter = 0;
    s.done = 0;
    s.rr = t[0];
    if (!(t[0] <= 1.0e300)) {
        s.done = 1;
        raise_error(st, DEV_RES_NOCONV, t[0], 0.0);
    } else if (t[0] <= s.tol2) {
        s.done = 1;
    }
}

FSI_KERNEL void k_rcg_q(Params p, GrowthP g, ResCg a, const Status* st) {
    if (st->res.done) return;
    double acc[1] = {0.0};
    GR_FOR(p) {
        const size_t k = at(p, i, j);
        const double pk = a.p[k];
        const double q = pk - a.alpha * gr_lap(p, g, a.Dx, a.Dy, a.p, i, j, k);
        a.q[k] = q;
        acc[0] += pk * q;
    }
    put_partials<1>(acc, a.part, (int)p.exnb);
}

FSI_KERNEL void k_rcg_x(Params p, ResCg a, const Status* st, int m) {
    if (st->res.done) return;
    double t[1];
    get_totals<1>(a.part, (int)p.exnb, t);
    const double al = st->res.rzs[m & 1] / t[0];
    double acc[1] = {0.0};
    GR_FOR(p) {
        const size_t k = at(p, i, j);
        a.x[k] += al * a.p[k];
        const double r = a.r[k] - al * a.q[k];
        a.r[k] = r;
        acc[0] += r * r;
    }
    put_partials<1>(acc, a.part + p.exnb, (int)p.exnb);
}

FSI_KERNEL void k_rcg_dir(Params p, ResCg a, Status* st, int m) {
    if (st->res.done) return;
    double t[1];
    get_totals<1>(a.part + p.exnb, (int)p.exnb, t);
    const double rr = st->res.rzs[m & 1], rn = t[0];
    const double beta = rn / rr;
    GR_FOR(p) {
        const size_t k = at(p, i, j);
        a.p[k] = a.r[k] + beta * a.p[k];
    }
    if (gr_bid() != 0 || gr_lane() != 0) return;
    // the stopping rules (only block 0 decides; the others read none of it)
    KrylovScalars& s = st->res;
    s.rzs[(m + 1) & 1] = rn;
    s.rr = rn;
    s.iter = m + 1;
    if (!(rn <= 1.0e300)) {
        s.done = 1;
        raise_error(st, DEV_RES_NOCONV, rn, 0.0);
    } else if (rn <= s.tol2) {
        s.done = 1;
    } else if (m + 1 >= s.cap) {
        if (!s.sw && rn <= s.fb2) {
            s.done = 1;  // accepted below fallback_reltol (:629-631)
        } else if (!s.sw && a.fallback) {
            s.sw = 1;  // the direct solve (:632-634)
            s.tol2 = s.dtol2;
            s.cap = s.maxiter;
        } else {
            s.done = 1;
            raise_error(st, DEV_RES_NOCONV, sqrt(rn / fmax(s.bnorm2, 4.930380657631324e-32)), 0.0);
        }
    }
}

// ---------------------------------------------------------------------------
// Growth terms of the stage tendencies and the resource tendencies
// ---------------------------------------------------------------------------
// upwind resource flux through a face with velocity u between the low-side
// state cl and the high-side state ch (ResourceTransport.jl:161-233); on a wall
// the outside state is the boundary's (:84-159)
__device__ __forceinline__ double gr_flux(double u, double cl, double ch) { return u * (u >= 0.0 ? cl : ch); }

__device__ __forceinline__ double gr_wall_state(const BcSide& b, double u, bool low, double interior, Status* st) {
    const bool inflow = low ? u > 0.0 : u < 0.0;
    if (u == 0.0 || !inflow) return interior;
    if (b.kind == FSI_BC_DIRICHLET) return b.value;
    raise_error(st, DEV_NEUMANN_INFLOW, u, 0.0);
    return interior;
}

FSI_KERNEL void k_gr_tend(Params p, GrowthP g, GrTend a, Status* st) {
    GR_FOR(p) {
        const size_t k = at(p, i, j), P = p.P;
        const int nx = p.nx, ny = p.ny;
        const double G = a.G[k], ph = a.phi[k];
        // phi S (PhaseFieldMethods.jl:670, :866; InterfaceTransport.jl:275)
        a.Ep[k] += ph * G;
        // the K_g part of the structure tensor right-hand side (StructureTensor.jl:431-437)
        double xx = a.sxx[k], yy = a.syy[k], xy = a.sxy[k];
        if (a.reset && ph < p.ext_cutoff) {
            xx = 1.0;
            yy = 1.0;
            xy = 0.0;
        }
        const double Kxx = g.kxx * G, Kxy = g.kxy * G, Kyy = g.kyy * G;
        a.Exx[k] += -2.0 * (Kxx * xx + Kxy * xy);
        a.Eyy[k] += -2.0 * (Kxy * xy + Kyy * yy);
        a.Exy[k] += -(Kxy * xx + Kxy * yy + (Kxx + Kyy) * xy);
        // resource: -div(upwind c v) - C(phi, c) (:1011-1026), then L c + b (:1064-1069)
        const double cc = a.c[k];
        double fxl, fxr, fyb, fyt;
        {
            const double ul = a.U[k], ur = a.U[k + 1];
            const double cw = i > 0 ? a.c[k - 1] : (p.px ? a.c[k + nx - 1] : gr_wall_state(g.bC.l, ul, true, cc, st));
            fxl = gr_flux(ul, cw, cc);
            const double ce = i < nx - 1 ? a.c[k + 1]
                                         : (p.px ? a.c[k - (nx - 1)] : gr_wall_state(g.bC.r, ur, false, cc, st));
            fxr = gr_flux(ur, cc, ce);
            const double vb = a.V[k], vt = a.V[k + P];
            const double cs = j > 0 ? a.c[k - P]
                                    : (p.py ? a.c[k + (size_t)(ny - 1) * P] : gr_wall_state(g.bC.b, vb, true, cc, st));
            fyb = gr_flux(vb, cs, cc);
            const double cn = j < ny - 1 ? a.c[k + P]
                                         : (p.py ? a.c[k - (size_t)(ny - 1) * P] : gr_wall_state(g.bC.t, vt, false, cc, st));
            fyt = gr_flux(vt, cc, cn);
        }
        double ec = -((fxr - fxl) * p.ihx + (fyt - fyb) * p.ihy);
        if (g.cons == FSI_CONSUMPTION_SATURATING) {
            const double sink = g.crate * ph * cc / (g.cK + cc);
            if (!isfinite(sink)) raise_error(st, DEV_RES_SINK, sink, cc);
            ec = ec - sink;
        }
        a.Ec[k] = ec;
        if (a.Ic) a.Ic[k] = gr_lap(p, g, a.Dx, a.Dy, a.c, i, j, k) + gr_lap_bc(p, g, a.Dx, a.Dy, i, j, k);

        // momentum: + S v on the face (MomentumTransport.jl:313-334; wall
        // faces keep zero convection) and the divergence of the growth stress
        // -2 mu K_g (:926-942, symmetric_stress_divergence Operators.jl:620-735
        // on the uniform grid: wall derivatives one-sided, wall values of the
        // shear stress linearly extrapolated)
        auto Gc = [&](int ii, int jj) -> double {
            ii = ii < 0 ? ii + nx : (ii >= nx ? ii - nx : ii);
            jj = jj < 0 ? jj + ny : (jj >= ny ? jj - ny : jj);
            return a.G[(size_t)jj * P + ii];
        };
        const double m2 = -2.0 * g.mu;
        // x-face (i, j)
        {
            double e = 0.0;
            if (i > 0 || p.px) e += (0.5 * Gc(i - 1, j) + 0.5 * G) * a.U[k];
            if (g.visc) {
                double dxx;
                if (i > 0 || p.px) dxx = (m2 * (g.kxx * G) - m2 * (g.kxx * Gc(i - 1, j))) * p.ihx;
                else dxx = (m2 * (g.kxx * Gc(1, j)) - m2 * (g.kxx * G)) * p.ihx;
                double dxy = 0.0;
                if (g.kxy != 0.0) {
                    // Txy = -mu (K_xy + K_yx) on the x-face column i, rows jj
                    auto Tf = [&](int jj) -> double {
                        auto T = [&](int ii) { const double q = g.kxy * Gc(ii, jj); return -g.mu * (q + q); };
                        if (i > 0 || p.px) return 0.5 * T(i - 1) + 0.5 * T(i);
                        return 1.5 * T(0) - 0.5 * T(1);
                    };
                    if (p.py) dxy = (Tf(j + 1) - Tf(j - 1)) * (0.5 * p.ihy);
                    else if (j == 0) dxy = (Tf(1) - Tf(0)) * p.ihy;
                    else if (j == ny - 1) dxy = (Tf(ny - 1) - Tf(ny - 2)) * p.ihy;
                    else dxy = (Tf(j + 1) - Tf(j - 1)) * (0.5 * p.ihy);
                }
                e += dxx + dxy;
            }
            a.Eu[k] += e;
        }
        // y-face (i, j)
        {
            double e = 0.0;
            if (j > 0 || p.py) e += (0.5 * Gc(i, j - 1) + 0.5 * G) * a.V[k];
            if (g.visc) {
                double dyy;
                if (j > 0 || p.py) dyy = (m2 * (g.kyy * G) - m2 * (g.kyy * Gc(i, j - 1))) * p.ihy;
                else dyy = (m2 * (g.kyy * Gc(i, 1)) - m2 * (g.kyy * G)) * p.ihy;
                double dxy = 0.0;
                if (g.kxy != 0.0) {
                    auto Tf = [&](int ii) -> double {
                        auto T = [&](int jj) { const double q = g.kxy * Gc(ii, jj); return -g.mu * (q + q); };
                        if (j > 0 || p.py) return 0.5 * T(j - 1) + 0.5 * T(j);
                        return 1.5 * T(0) - 0.5 * T(1);
                    };
                    if (p.px) dxy = (Tf(i + 1) - Tf(i - 1)) * (0.5 * p.ihx);
                    else if (i == 0) dxy = (Tf(1) - Tf(0)) * p.ihx;
                    else if (i == nx - 1) dxy = (Tf(nx - 1) - Tf(nx - 2)) * p.ihx;
                    else dxy = (Tf(i + 1) - Tf(i - 1)) * (0.5 * p.ihx);
                }
                e += dyy + dxy;
            }
            a.Ev[k] += e;
        }
    }
}

// ---------------------------------------------------------------------------
// StepReport partials of the new state (structure_tensor_diagnostics
// StructureTensor.jl:627-662, phi extrema, minimum resource)
// ---------------------------------------------------------------------------
FSI_KERNEL void k_gr_report(Params p, const double* __restrict__ phi, const double* __restrict__ sxx,
                            const double* __restrict__ syy, const double* __restrict__ sxy,
                            const double* __restrict__ c) {
    double mx[5] = {-INFINITY, -INFINITY, -INFINITY, 0.0, -INFINITY};  // -phi_min, phi_max, -eig_min, det err, -c_min
    double sm[2] = {0.0, 0.0};                                         // non-SPD, non-finite
    GR_FOR(p) {
        const size_t k = at(p, i, j);
        const double a = sxx[k], d = syy[k], b = sxy[k], ph = phi[k], cc = c[k];
        const double det = a * d - b * b;
        const double lmin = 0.5 * (a + d - hypot(a - d, 2.0 * b));
        mx[0] = fmax(mx[0], -ph);
        mx[1] = fmax(mx[1], ph);
        mx[2] = fmax(mx[2], -lmin);
        mx[3] = fmax(mx[3], fabs(det - 1.0));
        mx[4] = fmax(mx[4], -cc);
        sm[0] += (isfinite(lmin) && lmin > 0.0) ? 0.0 : 1.0;
        sm[1] += (double)(!isfinite(ph) + !isfinite(a) + !isfinite(d) + !isfinite(b) + !isfinite(cc));
    }
    block_maxes<5>(mx);
    __shared__ double lds[2 * 16];
    block_sum<2>(sm, lds);
    if (gr_lane() != 0) return;
    const size_t nb = p.exnb, b = gr_bid();
    p.exbuf[(size_t)EX_PHI_NEGMIN * nb + b] = mx[0];
    p.exbuf[(size_t)EX_PHI_MAX * nb + b] = mx[1];
    p.exbuf[(size_t)EX_EIG_NEGMIN * nb + b] = mx[2];
    p.exbuf[(size_t)EX_DET_ERR * nb + b] = mx[3];
    p.exbuf[(size_t)EX_RES_NEGMIN * nb + b] = mx[4];
    p.exbuf[(size_t)EX_NONSPD * nb + b] = sm[0];
    p.exbuf[(size_t)EX_NONFINITE * nb + b] = sm[1];
}

}  // namespace

void launch_gr_source(hipStream_t s, const Params& p, const GrowthP& g, const GrSrc& a, Status* st) {
    fsi_launch<&k_gr_source, BLOCK, 0>(gr_grid(p), kGrBlock, 0, s, p, g, a);
    fsi_launch<&k_gr_total, BLOCK, 0>(1, BLOCK, 0, s, (const double*)a.part, p.exnb, st);
}
void launch_gr_stage(hipStream_t s, const Params& p, const GrStage& a) {
    fsi_launch<&k_gr_stage, BLOCK, 0>(gr_grid(p), kGrBlock, 0, s, p, a);
}
void launch_gr_diff(hipStream_t s, const Params& p, const GrowthP& g, const double* phi, double* Dx, double* Dy,
                    Status* st) {
    fsi_launch<&k_gr_diff, BLOCK, 0>(gr_grid(p), kGrBlock, 0, s, p, g, phi, Dx, Dy, st);
}
void launch_rcg_start(hipStream_t s, const Params& p, const GrowthP& g, const ResCg& a, Status* st) {
    fsi_launch<&k_rcg_start, BLOCK, 0>(gr_grid(p), kGrBlock, 0, s, p, g, a);
    fsi_launch<&k_rcg_setup, BLOCK, 0>(1, BLOCK, 0, s, a, p.exnb, st);
}
void launch_rcg_iters(hipStream_t s, const Params& p, const GrowthP& g, const ResCg& a, Status* st, int m0,
                      int count) {
    for (int m = m0; m < m0 + count; ++m) {
        fsi_launch<&k_rcg_q, BLOCK, 0>(gr_grid(p), kGrBlock, 0, s, p, g, a, (const Status*)st);
        fsi_launch<&k_rcg_x, BLOCK, 0>(gr_grid(p), kGrBlock, 0, s, p, a, (const Status*)st, m);
        fsi_launch<&k_rcg_dir, BLOCK, 0>(gr_grid(p), kGrBlock, 0, s, p, a, st, m);
    }
}
void launch_gr_tend(hipStream_t s, const Params& p, const GrowthP& g, const GrTend& a, Status* st) {
    fsi_launch<&k_gr_tend, BLOCK, 0>(gr_grid(p), kGrBlock, 0, s, p, g, a, st);
}
void launch_gr_report(hipStream_t s, const Params& p, const double* phi, const double* sxx, const double* syy,
                      const double* sxy, const double* c) {
    fsi_launch<&k_gr_report, BLOCK, 0>(gr_grid(p), kGrBlock, 0, s, p, phi, sxx, syy, sxy, c);
}

}  // namespace fsi
