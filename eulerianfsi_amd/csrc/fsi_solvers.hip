// fsi_solvers.hip -- the two linear solves of the IMEX step on gfx950.
//
// (1) Velocity predictor: (I - alpha M) uv = rhs with Dirichlet rows
//     (src/TimeIntegrationIMEX.jl:794-812).  M is symmetric negative
//     semidefinite on uniform grids, so after eliminating the Dirichlet rows
//     the system is SPD and is solved by matrix-free CG.  The reference uses
//     UMFPACK or GMRES(100); both return the same unique solution.
// (2) Projection: A q = div(v*) - G - b_bc with the unit-coefficient
//     Laplacian (src/Projection.jl:170-307), solved by CG preconditioned with
//     one geometric-multigrid V-cycle (cell-centred, damped-Jacobi smoothing,
//     bilinear prolongation P, restriction R = P^T/4, rediscretised coarse
//     operators, a single-workgroup LDS solve for the coarsest levels).  The
//     reference uses GMRES(50) + Ruge-Stueben AMG; A q = b has a unique
//     zero-mean solution, so the projected velocity is the same.
//
// All dot products are deterministic: wave64 shuffles, LDS in wave order and
// one partial per block; every block of the consuming kernel sums all the
// partials in the same fixed order, so no kernel synchronises across blocks.
// Every kernel of an iteration first checks the solver's `done` flag, so the
// host can enqueue iterations ahead without reading the residual back.
#include "fsi_kernels.hpp"
#include "fsi_stencils.hpp"

#include <string>

namespace fsi {



static __device__ __forceinline__ int block_id() { return blockIdx.x + blockIdx.y * gridDim.x; }
static __device__ __forceinline__ int lane_id() { return threadIdx.x + threadIdx.y * blockDim.x; }

#define FOR_TILE(p)                                                              \
    const int i = tile_id().x * TILE_X + threadIdx.x;                             \
    for (int rr_ = 0; rr_ < ROWS; ++rr_)                                         \
        if (const int j = tile_id().y * TILE_ROWS + threadIdx.y + rr_ * TILE_Y;  \
            i < (p).nx && j < (p).ny)

// Partial-sum sets (see put_partials / get_totals in fsi_device.hpp): each
// set holds up to 4 quantities x nb blocks.
enum PartialSet : int { PS_VEL0 = 0, PS_VEL1 = 1, PS_VELI = 2, PS_RHS = 3, PS_RR = 4, PS_RZ = 5, PS_PW = 6,
                        PS_MEAN = 7, PS_CHECK = 8, PS_VMAX = 9, PS_VELR = 10, PS_COUNT = 11 };
__device__ __forceinline__ double* pset(double* base, int nb, int set) { return base + (size_t)set * 4 * nb; }

// ===========================================================================
// Velocity predictor: single-reduction CG, one kernel per iteration
// ===========================================================================
// Kernel m forms r_{m+1} = r_m - a_m q_m and p_{m+1} = r_{m+1} + b_{m+1} p_m
// (once per point of a halo-1 LDS tile), q_{m+1} = (I - alpha M) p_{m+1},
// x_{m+1} = x_m + a_m p_m, and the block partials of p.q, r.q, q.q, r.r.  a_m
// and b_{m+1} come from the partials of kernel m-1: a_m = rr_m / pq_m,
// b_{m+1} = rr_{m+1} / rr_m with rr_{m+1} = rr_m - 2 a_m r.q + a_m^2 q.q.
// Convergence is tested on rr_m.  r, p, q are ping-ponged by m; x is updated
// in place.

// LDS tile of a velocity component on halo 1 around the 64 x 16 cells of a
// block, read by global (wrapped) index.
constexpr int VT_W = TILE_X + 2, VT_H = TILE_ROWS + 2;  // 66 x 18
template <class T>
struct VelTileT {
    const T* a;
    int i0, j0, nx, ny;
    __device__ __forceinline__ double operator()(int gi, int gj) const {
        int li = gi - i0 + 1, lj = gj - j0 + 1;
        if (li < 0) li += nx;
        else if (li >= VT_W) li -= nx;
        if (lj < 0) lj += ny;
        else if (lj >= VT_H) lj -= ny;
        return (double)a[lj * VT_W + li];
    }
};
using VelTile = VelTileT<double>;

// The CG vectors r, p, q (ping-pong pairs) in their storage type: FP64, or
// FP32 for the inner solve of the mixed-precision predictor.
template <class T>
struct CgView {
    using Pair = T* const*;
    Pair r_u, r_v, p_u, p_v, q_u, q_v;
};
template <class T>
__device__ __forceinline__ CgView<T> cg_view(const VelBufs& b);
template <>
__device__ __forceinline__ CgView<double> cg_view<double>(const VelBufs& b) {
    return {b.r_u, b.r_v, b.p_u, b.p_v, b.q_u, b.q_v};
}
template <>
__device__ __forceinline__ CgView<float> cg_view<float>(const VelBufs& b) {
    return {b.fr_u, b.fr_v, b.fp_u, b.fp_v, b.fq_u, b.fq_v};
}

// LDS tile of halo H around the 64 x 16 cells of a block, by global index.
// IN: the tile and its halo lie inside the domain, so no index wraps.
template <int H, bool IN = false>
struct HTile {
    static constexpr int W = TILE_X + 2 * H, R = TILE_ROWS + 2 * H;
    const double* a;
    int i0, j0, nx, ny;
    __device__ __forceinline__ double operator()(int gi, int gj) const {
        int li = gi - i0 + H, lj = gj - j0 + H;
        if (!IN) {
            if (li < 0) li += nx;
            else if (li >= W) li -= nx;
            if (lj < 0) lj += ny;
            else if (lj >= R) lj -= ny;
        }
        return a[lj * W + li];
    }
};

// The warm-start extrapolations in time with their roundings fixed (explicit
// fused multiply-adds), so that every kernel forming them gives the same bits:
// 2 x - y, 3 (x - y) + z, 4 (x + z) - 6 y - w.
__device__ __forceinline__ double extrap1(double x, double y) { return __builtin_fma(2.0, x, -y); }
__device__ __forceinline__ double extrap2(double x, double y, double z) { return __builtin_fma(3.0, x - y, z); }
__device__ __forceinline__ double extrap3(double x, double y, double z, double w) {
    return __builtin_fma(4.0, x + z, __builtin_fma(-6.0, y, -w));
}

// x0 = cur, or the warm start extrapolated in time from the last two
// (2 cur - prev), three (3 (cur - prev) + old) or four
// (4 (cur + old) - 6 prev - old2) stage solutions, with the Dirichlet values,
// at the tile points it * BLOCK + tid < NPT, A <= it < B, written to the LDS tile
template <int A, int B, int NPT>
__device__ __forceinline__ void x0_batch(const VelBufs& b, const size_t* idx, const int* fl, int tid, double gul,
                                         double gvb, double* sxu, double* sxv) {
    constexpr int M = B - A;
    double xu[M], xv[M];
#pragma unroll
    for (int t = 0; t < M; ++t) {
        xu[t] = b.Un[idx[A + t]];
        xv[t] = b.Vn[idx[A + t]];
    }
    if (b.extrap == 1) {
        double yu[M], yv[M];
#pragma unroll
        for (int t = 0; t < M; ++t) {
            yu[t] = b.Pu[idx[A + t]];
            yv[t] = b.Pv[idx[A + t]];
        }
#pragma unroll
        for (int t = 0; t < M; ++t) {
            xu[t] = extrap1(xu[t], yu[t]);
            xv[t] = extrap1(xv[t], yv[t]);
        }
    } else if (b.extrap == 2) {
        double yu[M], yv[M], zu[M], zv[M];
#pragma unroll
        for (int t = 0; t < M; ++t) {
            yu[t] = b.Pu[idx[A + t]];
            yv[t] = b.Pv[idx[A + t]];
            zu[t] = b.Ou[idx[A + t]];
            zv[t] = b.Ov[idx[A + t]];
        }
#pragma unroll
        for (int t = 0; t < M; ++t) {
            xu[t] = extrap2(xu[t], yu[t], zu[t]);
            xv[t] = extrap2(xv[t], yv[t], zv[t]);
        }
    } else if (b.extrap == 3) {
        double yu[M], yv[M], zu[M], zv[M], wu[M], wv[M];
#pragma unroll
        for (int t = 0; t < M; ++t) {
            yu[t] = b.Pu[idx[A + t]];
            yv[t] = b.Pv[idx[A + t]];
            zu[t] = b.Ou[idx[A + t]];
            zv[t] = b.Ov[idx[A + t]];
            wu[t] = b.Wu[idx[A + t]];
            wv[t] = b.Wv[idx[A + t]];
        }
#pragma unroll
        for (int t = 0; t < M; ++t) {
            xu[t] = extrap3(xu[t], yu[t], zu[t], wu[t]);
            xv[t] = extrap3(xv[t], yv[t], zv[t], wv[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < M; ++t) {
        const int k = tid + (A + t) * BLOCK;
        if (k >= NPT) break;
        const int f = fl[A + t];
        sxu[k] = (f & 1) ? ((f & 2) ? gul : xu[t]) : 0.0;
        sxv[k] = (f & 1) ? ((f & 4) ? gvb : xv[t]) : 0.0;
    }
}

// Start of a predictor solve: x0 = the warm start with the Dirichlet
// values, r0 = rhs - (I - alpha M) x0 (zero on Dirichlet rows) and the
// partials of |r0|^2, |b|^2 and |b - A u^n|^2.  b and A are the system of
// _impose_row_dirichlet (TimeIntegrationIMEX.jl:649-672): identity rows with
// the Dirichlet value on the right at U-left / V-bottom, the free rows with
// their Dirichlet columns kept, so |b - A u^n| is the initial residual
// IterativeSolvers measures reltol against (x0 = uv_n, :807); with M u^n
// from the stage-1 momentum kernel it is b - u^n + alpha M u^n on the free
// rows.  x0 sits on a halo-1 LDS tile; r0 is formed on the block's own cells
// only.  p0 = r0 is not stored, and q0 = (I - alpha M) r0 with the first
// iteration's partials is k_vel_q0's.
template <bool IN>
__device__ __forceinline__ void vel_start_tile(const Params& p, const VelBufs& b, double alpha, int nb, int i0, int j0,
                                               int tid, double* sxu, double* sxv) {
    using T1 = HTile<1, IN>;
    const double gul = p.bU.l.value, gvb = p.bV.b.value;
    auto wrapped = [&](int ui, int uj, int& gi, int& gj) -> bool {
        gi = ui;
        gj = uj;
        if (IN) return true;
        if (gi < 0 || gi >= p.nx) {
            if (!p.px) return false;
            gi = ((gi % p.nx) + p.nx) % p.nx;
        }
        if (gj < 0 || gj >= p.ny) {
            if (!p.py) return false;
            gj = ((gj % p.ny) + p.ny) % p.ny;
        }
        return true;
    };
    constexpr int N1 = T1::R * T1::W, NI1 = (N1 + BLOCK - 1) / BLOCK;
    size_t idx1[NI1];
    int fl[NI1];  // 1: inside, 2: U-left Dirichlet column, 4: V-bottom Dirichlet row
#pragma unroll
    for (int it = 0; it < NI1; ++it) {
        const int k = min(tid + it * BLOCK, N1 - 1);
        int gi, gj;
        const bool in = wrapped(i0 + k % T1::W - 1, j0 + k / T1::W - 1, gi, gj);
        idx1[it] = in ? at(p, gi, gj) : 0;
        fl[it] = in ? (1 | (dir_u<IN>(p, gi) ? 2 : 0) | (dir_v<IN>(p, gj) ? 4 : 0)) : 0;
    }
    // the own cells' right-hand side first, then x0 in two batches of tile points
    const int i = i0 + threadIdx.x;
    bool own[ROWS];
    size_t kc[ROWS];
    double bu[ROWS], bv[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int j = j0 + threadIdx.y + r * TILE_Y;
        own[r] = i < p.nx && j < p.ny;
        kc[r] = own[r] ? at(p, i, j) : 0;
        bu[r] = b.ru[kc[r]];
        bv[r] = b.rv[kc[r]];
    }
    // |b|^2 and |b - A u^n|^2 from u^n and M u^n on the own cells, before x0
    const bool ref = b.Su != nullptr;
    double acc3[3] = {0.0, 0.0, 0.0};  // |r0|^2, |b|^2, |b - A u^n|^2
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const double nu = ref ? b.Su[kc[r]] : 0.0, nv = ref ? b.Sv[kc[r]] : 0.0;
        const double mu = ref ? b.Mu[kc[r]] : 0.0, mv = ref ? b.Mv[kc[r]] : 0.0;
        if (!own[r]) continue;
        const int j = j0 + threadIdx.y + r * TILE_Y;
        const bool fu = !dir_u<IN>(p, i), fv = !dir_v<IN>(p, j);
        const double cu = fu ? bu[r] : gul, cv = fv ? bv[r] : gvb;
        const double eu = fu ? cu - (nu - alpha * mu) : cu - nu;
        const double ev = fv ? cv - (nv - alpha * mv) : cv - nv;
        acc3[1] += cu * cu;
        acc3[1] += cv * cv;
        if (ref) acc3[2] += eu * eu + ev * ev;
    }
    x0_batch<0, NI1 / 2, N1>(b, idx1, fl, tid, gul, gvb, sxu, sxv);
    x0_batch<NI1 / 2, NI1, N1>(b, idx1, fl, tid, gul, gvb, sxu, sxv);
    __syncthreads();
    const T1 Xu{sxu, i0, j0, p.nx, p.ny}, Xv{sxv, i0, j0, p.nx, p.ny};
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        if (!own[r]) continue;
        const int j = j0 + threadIdx.y + r * TILE_Y;
        const size_t k = kc[r];
        double ru = 0.0, rv = 0.0;
        if (!dir_u<IN>(p, i)) ru = bu[r] - (Xu(i, j) - alpha * visc_u<IN>(p, Xu, Xv, i, j, 0));
        if (!dir_v<IN>(p, j)) rv = bv[r] - (Xv(i, j) - alpha * visc_v<IN>(p, Xu, Xv, i, j, 0));
        if (!b.restart && !b.x0_in_place) {  // a restart (or a formed x0) reads x as it is
            b.xu[k] = Xu(i, j);
            b.xv[k] = Xv(i, j);
        }
        if (b.fr_u[0]) {  // mixed precision: the inner CG starts from the FP32 copy
            b.fr_u[0][k] = (float)ru;
            b.fr_v[0][k] = (float)rv;
        } else {
            b.r_u[0][k] = ru;  // p0 = r0: the first iteration reads it from r
            b.r_v[0][k] = rv;
        }
        acc3[0] += ru * ru + rv * rv;
    }
    put_partials<3>(acc3, pset(b.partials, nb, b.restart ? PS_VELR : PS_VELI), nb);
}

FSI_KERNEL void k_vel_start(Params p, VelBufs b, double alpha, Status* st, int nb) {
    constexpr int N1 = HTile<1>::R * HTile<1>::W;
    __shared__ double sxu[N1], sxv[N1];
    const int tid = threadIdx.x + threadIdx.y * TILE_X;
    const TileId T = tile_id();
    const int i0 = T.x * TILE_X, j0 = T.y * TILE_ROWS;
    if (b.restart) {
        if (st->vel.done) return;  // the solve ended at an earlier check
    } else if (T.x == 0 && T.y == 0 && tid == 0) {
        st->vel.done = 0;  // the iteration kernels follow this one
    }
    // the block's own operator rows are not the first or last face / row and
    // their stencils stay inside the halo-1 tile: no boundary rows, no wraps
    const bool inner = i0 >= 1 && i0 + TILE_X - 1 <= p.nx - 2 && j0 >= 1 && j0 + TILE_ROWS - 1 <= p.ny - 2;
    if (inner) vel_start_tile<true>(p, b, alpha, nb, i0, j0, tid, sxu, sxv);
    else vel_start_tile<false>(p, b, alpha, nb, i0, j0, tid, sxu, sxv);
}

// q0 = (I - alpha M) p0 with p0 = r0 (k_vel_start's) and the partials of the
// first iteration (p.q, r.q, q.q, r.r): r0 on a halo-1 LDS tile, the block's
// own cells and the halo ring as in k_vel_iter.  With `first`, block 0 also
// sets the solve's stopping thresholds from k_vel_start's totals (|b|^2,
// |b - A u^n|^2, and for the FP32 inner CG of the mixed-precision solve the
// threshold of its first cycle).  T: the storage of r, p, q.
template <class T>
FSI_KERNEL void k_vel_q0(Params p, VelBufs b, double alpha, int nb, Status* st, int slot,
                                                  VelTol vt, int first) {
    constexpr bool F32 = sizeof(T) == 4;
    if (!first && st->vel.done) return;  // a later cycle of a solve already converged
    if (first && block_id() == 0) {
        double ti[3];
        get_totals<3>(pset(b.partials, nb, PS_VELI), nb, ti);
        if (lane_id() == 0) {
            const double bscale2 = fmax(ti[1], 4.930380657631324e-32);  // max(|b|, eps)^2
            const double fb2 = vt.fb_rtol * vt.fb_rtol * bscale2;
            KrylovScalars& sk = st->vel;
            sk.bnorm2 = ti[1];
            sk.ref2 = ti[2];
            sk.tol2 = vt.rtol * vt.rtol * ti[2];
            sk.maxiter = vt.method == 0 ? vt.cap : vt.maxiter;
            sk.fb2 = fb2;
            // the direct target, never looser than the acceptance threshold
            sk.dtol2 = fmin(VEL_DIRECT_RTOL * VEL_DIRECT_RTOL * bscale2, fb2);
            sk.iter = 0;
            sk.sw = 0;
            st->vel_iters[slot] = 0;
            if (F32) {  // the inner CG reduces r.r to a quarter of the target, or by 1e-12 at most
                const double target = vt.method == 0 ? sk.dtol2 : sk.tol2;
                sk.itol2 = fmax(0.25 * target, 1e-12 * ti[0]);
                sk.idone = 0;
                sk.ibase = 0;
                sk.cycles = 0;
                sk.imax = 0;
            }
        }
    }
    const CgView<T> cv = cg_view<T>(b);
    __shared__ T su[VT_H * VT_W], sv[VT_H * VT_W];
    const int tid = threadIdx.x + threadIdx.y * TILE_X;
    const TileId Tl = tile_id();
    const int i0 = Tl.x * TILE_X, j0 = Tl.y * TILE_ROWS;
    const T *ru = cv.r_u[0], *rv = cv.r_v[0];
    auto wrapped = [&](int li, int lj) -> long {
        int gi = i0 + li - 1, gj = j0 + lj - 1;
        if (gi < 0 || gi >= p.nx) {
            if (!p.px) return -1;
            gi = ((gi % p.nx) + p.nx) % p.nx;
        }
        if (gj < 0 || gj >= p.ny) {
            if (!p.py) return -1;
            gj = ((gj % p.ny) + p.ny) % p.ny;
        }
        return (long)at(p, gi, gj);
    };
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = i0 + tx;
    size_t kc[ROWS];
    bool own[ROWS], fill[ROWS];
    double cru[ROWS], crv[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int j = j0 + ty + r * TILE_Y;
        own[r] = i < p.nx && j < p.ny;
        const long w = wrapped(tx + 1, ty + r * TILE_Y + 1);
        fill[r] = w >= 0;
        kc[r] = fill[r] ? (size_t)w : 0;
        cru[r] = ru[kc[r]];
        crv[r] = rv[kc[r]];
    }
    constexpr int NRING = 2 * VT_W + 2 * TILE_ROWS;
    if (tid < NRING) {
        int li, lj;
        if (tid < VT_W) { li = tid; lj = 0; }
        else if (tid < 2 * VT_W) { li = tid - VT_W; lj = VT_H - 1; }
        else if (tid < 2 * VT_W + TILE_ROWS) { li = 0; lj = tid - 2 * VT_W + 1; }
        else { li = VT_W - 1; lj = tid - 2 * VT_W - TILE_ROWS + 1; }
        const long w = wrapped(li, lj);
        const size_t kk = w >= 0 ? (size_t)w : 0;
        su[lj * VT_W + li] = w >= 0 ? ru[kk] : T(0);
        sv[lj * VT_W + li] = w >= 0 ? rv[kk] : T(0);
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int slot_c = (ty + r * TILE_Y + 1) * VT_W + tx + 1;
        su[slot_c] = fill[r] ? (T)cru[r] : T(0);
        sv[slot_c] = fill[r] ? (T)crv[r] : T(0);
    }
    __syncthreads();
    const VelTileT<T> tu{su, i0, j0, p.nx, p.ny}, tv{sv, i0, j0, p.nx, p.ny};
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        if (!own[r]) continue;
        const int j = j0 + ty + r * TILE_Y;
        const size_t k = kc[r];
        // q as stored (the FP32 rounding included, so the partials describe the stored vector)
        const double qu = (double)(T)(dir_u(p, i) ? 0.0 : cru[r] - alpha * visc_u(p, tu, tv, i, j, 0));
        const double qv = (double)(T)(dir_v(p, j) ? 0.0 : crv[r] - alpha * visc_v(p, tu, tv, i, j, 0));
        cv.q_u[0][k] = (T)qu;
        cv.q_v[0][k] = (T)qv;
        acc[0] += cru[r] * qu + crv[r] * qv;
        acc[1] += cru[r] * qu + crv[r] * qv;  // r = p on the first iteration
        acc[2] += qu * qu + qv * qv;
        acc[3] += cru[r] * cru[r] + crv[r] * crv[r];
    }
    put_partials<4>(acc, pset(b.partials, nb, PS_VEL0), nb);
}

// One CG iteration m >= 1 with a single reduction phase: the totals of
// iteration m-1 give a = rr/pq and, from rr' = rr - 2a rq + a^2 qq,
// beta = rr'/rr; then r' = r - a q, p' = r' + beta p, x += a p, q' = A p'
// with the partials of p'.q', r'.q', q'.q', r'.r' for iteration m+1.  p' is
// formed once per point on a halo-1 LDS tile and A is applied from there.
// The field loads are normally issued before the totals are summed, so the
// two latencies overlap; with check_first (m >= 1, an iteration at which
// convergence is expected) the totals come first, so the launch that sees
// convergence ends without touching the fields.
//
// T = float: an iteration of the FP32 inner CG of a mixed-precision solve
// (r, p, q stored in FP32, x in FP64, dot products summed in FP64).  Its
// cycle ends at the cycle's threshold itol2 (set by k_vel_q0 or
// k_vel_check), at icap iterations, or where the outer rule needs the FP64
// check (the iteration limit); k_vel_check then decides with the FP64
// residual whether the solve has converged.
template <class T>
FSI_KERNEL void k_vel_iter(Params p, VelBufs b, double alpha, Status* st, int nb, int m,
                                                    int slot, VelTol vt, int check_first) {
    constexpr bool F32 = sizeof(T) == 4;
    if (st->vel.done || (F32 && st->vel.idone)) return;
    // The test of iterate m, a function of the totals, m and the rule's
    // constants alone, so every block decides the same.  Returns true when
    // the solve ends here.  :direct: the direct target within `cap`
    // iterations.  :gmres: reltol * |b - A u^n| within `maxiter`; where
    // GMRES stops (the rule met or the limit reached) the iterate is kept
    // below fallback_reltol (:629-631), else with direct_fallback the solve
    // goes on to the direct target within `cap` iterations in all (the
    // sparse direct solve of :632-634; `sw` marks it for the later
    // launches), else the solver error (:636).
    auto test = [&](double rr, double tol2, double fb2, double dtol2) -> bool {
        bool end = false, ok = false;
        if (!(rr <= 1.0e300)) {  // NaN or Inf: no rule can be met
            end = true;
        } else if (vt.method == 0 || m > vt.maxiter || st->vel.sw) {
            ok = rr <= dtol2;
            end = ok || m >= vt.cap;
        } else if (rr <= tol2 || m == vt.maxiter) {
            ok = rr <= fb2;
            end = ok || !vt.fallback;
            // every block takes the same decision whether or not it sees sw
            if (!end && block_id() == 0 && lane_id() == 0) st->vel.sw = 1;
        }
        if (end && block_id() == 0 && lane_id() == 0) {
            st->vel.done = 1;
            st->vel.rr = rr;
            st->vel.iter = m;
            st->vel_iters[slot] = m;
            const double bs2 = fmax(st->vel.bnorm2, 4.930380657631324e-32);
            st->vel_res[slot] = sqrt(rr / bs2);
            st->vel_ref[slot] = sqrt(st->vel.ref2 / bs2);
            if (ok) st->vel_ok[slot] = 1;
            else raise_error(st, DEV_VEL_NOCONV, sqrt(rr / fmax(st->vel.bnorm2, 4.930380657631324e-32)),
                             (double)(slot + 2));
        }
        return end;
    };
    // the inner rule of a refinement cycle (FP32)
    auto test_inner = [&](double rr) -> bool {
        const KrylovScalars& sk = st->vel;
        const int total = sk.ibase + m;
        const bool bad = !(rr <= 1.0e300);
        const bool limit =
            vt.method == 1 && !sk.sw && total <= vt.maxiter ? total >= vt.maxiter : total >= vt.cap;
        const bool end = bad || rr <= sk.itol2 || m >= vt.icap || limit;
        if (end && block_id() == 0 && lane_id() == 0) {
            st->vel.idone = 1;
            st->vel_iters[slot] = total;
            if (bad) {  // nothing to refine: the solve ends with the error
                st->vel.done = 1;
                raise_error(st, DEV_VEL_NOCONV, NAN, (double)(slot + 2));
            }
        }
        return end;
    };
    const CgView<T> cv = cg_view<T>(b);
    double tot[4];
    if (check_first) {
        get_totals<4>(pset(b.partials, nb, (m & 1) ? PS_VEL1 : PS_VEL0), nb, tot);
        const KrylovScalars& sk = st->vel;
        if (F32 ? test_inner(tot[3]) : test(tot[3], sk.tol2, sk.fb2, sk.dtol2)) return;
    }
    const int o = m & 1, n = (m + 1) & 1;
    // p0 = r0 is not stored: the first iteration reads p from r
    const T *ru = cv.r_u[o], *rv = cv.r_v[o], *qu0 = cv.q_u[o], *qv0 = cv.q_v[o];
    const T *pu0 = m == 0 ? ru : cv.p_u[o], *pv0 = m == 0 ? rv : cv.p_v[o];
    __shared__ T su[VT_H * VT_W], sv[VT_H * VT_W];  // p' on halo 1
    const int tid = threadIdx.x + threadIdx.y * TILE_X;
    const TileId Tl = tile_id();
    const int i0 = Tl.x * TILE_X, j0 = Tl.y * TILE_ROWS;
    // wrapped global index of tile point (li, lj) of the halo-1 tile, or -1
    // outside a non-periodic wall
    auto wrapped = [&](int li, int lj) -> long {
        int gi = i0 + li - 1, gj = j0 + lj - 1;
        if (gi < 0 || gi >= p.nx) {
            if (!p.px) return -1;
            gi = ((gi % p.nx) + p.nx) % p.nx;
        }
        if (gj < 0 || gj >= p.ny) {
            if (!p.py) return -1;
            gj = ((gj % p.ny) + p.ny) % p.ny;
        }
        return (long)at(p, gi, gj);
    };
    // (1) the block's own cells (column tx, rows ty + 4r) and the halo ring
    // (2 rows of VT_W, 2 columns of TILE_ROWS) in registers: every load is
    // issued before the partials of the previous iteration are summed, so the
    // two latencies overlap.  Tile slots of cells past the domain edge get the
    // wrapped values (periodic) or 0.
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = i0 + tx;
    size_t kc[ROWS];
    bool own[ROWS], fill[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int j = j0 + ty + r * TILE_Y;
        own[r] = i < p.nx && j < p.ny;
        const long w = wrapped(tx + 1, ty + r * TILE_Y + 1);
        fill[r] = w >= 0;
        kc[r] = fill[r] ? (size_t)w : 0;
    }
    T cru[ROWS], crv[ROWS], cqu[ROWS], cqv[ROWS], cpu[ROWS], cpv[ROWS];  // in the storage type
    double cxu[ROWS], cxv[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        cru[r] = ru[kc[r]];
        crv[r] = rv[kc[r]];
        cqu[r] = qu0[kc[r]];
        cqv[r] = qv0[kc[r]];
        cpu[r] = pu0[kc[r]];
        cpv[r] = pv0[kc[r]];
        const size_t kx = own[r] ? kc[r] : 0;
        cxu[r] = b.xu[kx];
        cxv[r] = b.xv[kx];
    }
    constexpr int NRING = 2 * VT_W + 2 * TILE_ROWS;
    int hslot = -1;
    bool hin = false;
    T hru = 0, hrv = 0, hqu = 0, hqv = 0, hpu = 0, hpv = 0;
    if (tid < NRING) {
        int li, lj;
        if (tid < VT_W) { li = tid; lj = 0; }
        else if (tid < 2 * VT_W) { li = tid - VT_W; lj = VT_H - 1; }
        else if (tid < 2 * VT_W + TILE_ROWS) { li = 0; lj = tid - 2 * VT_W + 1; }
        else { li = VT_W - 1; lj = tid - 2 * VT_W - TILE_ROWS + 1; }
        hslot = lj * VT_W + li;
        const long w = wrapped(li, lj);
        hin = w >= 0;
        const size_t kk = hin ? (size_t)w : 0;
        hru = ru[kk];
        hrv = rv[kk];
        hqu = qu0[kk];
        hqv = qv0[kk];
        hpu = pu0[kk];
        hpv = pv0[kk];
    }
    // (one partial in flight at a time: the field loads hold the registers)
    if (!check_first) get_totals<4, 1>(pset(b.partials, nb, (m & 1) ? PS_VEL1 : PS_VEL0), nb, tot);
    const double pq = tot[0], rq = tot[1], qq = tot[2], rr = tot[3];
    if (!check_first) {
        const KrylovScalars& sk = st->vel;
        if (F32 ? test_inner(rr) : test(rr, sk.tol2, sk.fb2, sk.dtol2)) return;
    }
    const double a = rr / pq;
    const double beta = fmax(rr - 2.0 * a * rq + a * a * qq, 0.0) / rr;
    double hp_u = 0.0, hp_v = 0.0;
    if (hin) {
        hp_u = ((double)hru - a * (double)hqu) + beta * (double)hpu;
        hp_v = ((double)hrv - a * (double)hqv) + beta * (double)hpv;
    }
    // r', p' as stored (FP32 rounding included): the partials and the
    // stencil then describe the stored vectors
    double r1u[ROWS], r1v[ROWS], p1u[ROWS], p1v[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        r1u[r] = (double)(T)((double)cru[r] - a * (double)cqu[r]);
        r1v[r] = (double)(T)((double)crv[r] - a * (double)cqv[r]);
        p1u[r] = (double)(T)(r1u[r] + beta * (double)cpu[r]);
        p1v[r] = (double)(T)(r1v[r] + beta * (double)cpv[r]);
        const int slot_c = (ty + r * TILE_Y + 1) * VT_W + tx + 1;
        su[slot_c] = fill[r] ? (T)p1u[r] : T(0);
        sv[slot_c] = fill[r] ? (T)p1v[r] : T(0);
        if (own[r]) {
            const size_t k = kc[r];
            b.xu[k] = cxu[r] + a * (double)cpu[r];
            b.xv[k] = cxv[r] + a * (double)cpv[r];
            cv.r_u[n][k] = (T)r1u[r];
            cv.r_v[n][k] = (T)r1v[r];
            cv.p_u[n][k] = (T)p1u[r];
            cv.p_v[n][k] = (T)p1v[r];
        }
    }
    if (hslot >= 0) {
        su[hslot] = (T)hp_u;
        sv[hslot] = (T)hp_v;
    }
    __syncthreads();
    // (3) q' = (I - alpha M) p' from the tile, and the four partials
    const VelTileT<T> tu{su, i0, j0, p.nx, p.ny}, tv{sv, i0, j0, p.nx, p.ny};
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        if (!own[r]) continue;
        const int j = j0 + ty + r * TILE_Y;
        const size_t k = kc[r];
        const double q1u = (double)(T)(dir_u(p, i) ? 0.0 : p1u[r] - alpha * visc_u(p, tu, tv, i, j, 0));
        const double q1v = (double)(T)(dir_v(p, j) ? 0.0 : p1v[r] - alpha * visc_v(p, tu, tv, i, j, 0));
        cv.q_u[n][k] = (T)q1u;
        cv.q_v[n][k] = (T)q1v;
        acc[0] += p1u[r] * q1u + p1v[r] * q1v;
        acc[1] += r1u[r] * q1u + r1v[r] * q1v;
        acc[2] += q1u * q1u + q1v * q1v;
        acc[3] += r1u[r] * r1u[r] + r1v[r] * r1v[r];
    }
    put_partials<4>(acc, pset(b.partials, nb, n ? PS_VEL1 : PS_VEL0), nb);
    if (block_id() == 0 && lane_id() == 0) st->vel_iters[slot] = (F32 ? st->vel.ibase : 0) + m + 1;
}

// The FP64 check at the end of a refinement cycle of the mixed-precision
// solve, on the residual r = b - A x that k_vel_start (restart) formed from
// the FP64 x: the stopping rules of k_vel_iter's `test` applied to the true
// residual after `total` inner iterations.  When the solve goes on, the next
// cycle's inner threshold is set from it.  One block.
FSI_KERNEL void k_vel_check(VelBufs b, Status* st, int nb, int slot, VelTol vt) {
    if (st->vel.done) return;
    double tot[3];
    get_totals<3>(pset(b.partials, nb, PS_VELR), nb, tot);
    if (lane_id() != 0) return;
    KrylovScalars& sk = st->vel;
    const double rr = tot[0];
    const int total = st->vel_iters[slot];
    sk.cycles += 1;
    sk.imax = max(sk.imax, total - sk.ibase);
    bool end = false, ok = false;
    const bool fallback_phase = vt.method == 0 || total > vt.maxiter || sk.sw;
    if (!(rr <= 1.0e300)) {
        end = true;
    } else if (fallback_phase) {
        ok = rr <= sk.dtol2;
        end = ok || total >= vt.cap;
    } else if (rr <= sk.tol2 || total >= vt.maxiter) {
        ok = rr <= sk.fb2;
        end = ok || !vt.fallback;
        if (!end) sk.sw = 1;
    }
    st->vel_cycles[slot] = sk.cycles;
    st->vel_imax[slot] = sk.imax;
    if (end) {
        sk.done = 1;
        sk.rr = rr;
        sk.iter = total;
        const double bs2 = fmax(sk.bnorm2, 4.930380657631324e-32);
        st->vel_res[slot] = sqrt(rr / bs2);
        st->vel_ref[slot] = sqrt(sk.ref2 / bs2);
        if (ok) st->vel_ok[slot] = 1;
        else raise_error(st, DEV_VEL_NOCONV, sqrt(rr / bs2), (double)(slot + 2));
    } else {
        const double target = (fallback_phase || sk.sw) ? sk.dtol2 : sk.tol2;
        sk.itol2 = fmax(0.25 * target, 1e-12 * rr);
        sk.ibase = total;
        sk.idone = 0;
    }
}

// The warm start of a predictor solve formed ahead of it (on the side
// stream, beside the previous stage): x0_batch's extrapolation of the
// stage's solutions of the last steps, with the Dirichlet values, on the
// compact cells, written into the solve's iterate.  The start kernel then
// reads it as it is (x0_in_place) instead of the four history fields.
FSI_KERNEL void k_vel_predict(Params p, VelBufs b) {
    const double gul = p.bU.l.value, gvb = p.bV.b.value;
    FOR_TILE(p) {
        const size_t k = at(p, i, j);
        double xu = b.Un[k], xv = b.Vn[k];
        if (b.extrap == 1) {
            xu = extrap1(xu, b.Pu[k]);
            xv = extrap1(xv, b.Pv[k]);
        } else if (b.extrap == 2) {
            xu = extrap2(xu, b.Pu[k], b.Ou[k]);
            xv = extrap2(xv, b.Pv[k], b.Ov[k]);
        } else if (b.extrap == 3) {
            xu = extrap3(xu, b.Pu[k], b.Ou[k], b.Wu[k]);
            xv = extrap3(xv, b.Pv[k], b.Ov[k], b.Wv[k]);
        }
        b.xu[k] = dir_u(p, i) ? gul : xu;
        b.xv[k] = dir_v(p, j) ? gvb : xv;
    }
}
// NoViscosity: the predictor is the stage sum with Dirichlet values (:794-796).
FSI_KERNEL void k_vel_copy(Params p, VelBufs b) {
    FOR_TILE(p) {
        const size_t k = at(p, i, j);
        b.xu[k] = dir_u(p, i) ? p.bU.l.value : b.ru[k];
        b.xv[k] = dir_v(p, j) ? p.bV.b.value : b.rv[k];
    }
}

static dim3 grid2(const Params& p) { return dim3((p.nx + TILE_X - 1) / TILE_X, (p.ny + TILE_ROWS - 1) / TILE_ROWS); }
int grid_blocks(const Params& p) {
    dim3 g = grid2(p);
    return (int)(g.x * g.y);
}
size_t partials_size(const Params& p) { return (size_t)PS_COUNT * 4 * (size_t)grid_blocks(p); }

// ===========================================================================
// Projection: rhs, PCG, multigrid V-cycle, finish
// ===========================================================================

// -A x at a cell of a uniform level (A = unit Laplacian with the correction
// BCs, homogeneous data).  Dirichlet sides add 2/h^2 to the diagonal
// (gradient (q1 - g)/(h/2)), Neumann sides add nothing, periodic sides wrap.
template <class FX>
__device__ __forceinline__ double neg_lap(const MgGeom& g, const FX& x, int i, int j, double* diag) {
    double c = x(i, j), sx = 0.0, sy = 0.0, cx = 0.0, cy = 0.0;
    if (i > 0) { sx += x(i - 1, j); cx += 1.0; }
    else if (g.px) { sx += x(g.nx - 1, j); cx += 1.0; }
    else if (g.dl) cx += 2.0;
    if (i < g.nx - 1) { sx += x(i + 1, j); cx += 1.0; }
    else if (g.px) { sx += x(0, j); cx += 1.0; }
    else if (g.dr) cx += 2.0;
    if (j > 0) { sy += x(i, j - 1); cy += 1.0; }
    else if (g.py) { sy += x(i, g.ny - 1); cy += 1.0; }
    else if (g.db) cy += 2.0;
    if (j < g.ny - 1) { sy += x(i, j + 1); cy += 1.0; }
    else if (g.py) { sy += x(i, 0); cy += 1.0; }
    else if (g.dt) cy += 2.0;
    if (diag) *diag = g.ihx2 * cx + g.ihy2 * cy;
    return g.ihx2 * (cx * c - sx) + g.ihy2 * (cy * c - sy);
}

struct GAcc {
    const double* a;
    int P;
    __device__ __forceinline__ double operator()(int i, int j) const { return a[(size_t)j * P + i]; }
};
// neg_lap's value from the cell and its four neighbours (periodic ones
// wrapped) already in registers; the same operations in the same order
__device__ __forceinline__ double neg_lap5(const MgGeom& g, int i, int j, double c, double w, double e, double s,
                                           double n) {
    double sx = 0.0, sy = 0.0, cx = 0.0, cy = 0.0;
    if (i > 0 || g.px) { sx += w; cx += 1.0; }
    else if (g.dl) cx += 2.0;
    if (i < g.nx - 1 || g.px) { sx += e; cx += 1.0; }
    else if (g.dr) cx += 2.0;
    if (j > 0 || g.py) { sy += s; cy += 1.0; }
    else if (g.db) cy += 2.0;
    if (j < g.ny - 1 || g.py) { sy += n; cy += 1.0; }
    else if (g.dt) cy += 2.0;
    return g.ihx2 * (cx * c - sx) + g.ihy2 * (cy * c - sy);
}

// 1-D bilinear prolongation of fine index f from the coarse grid: parent
// weight 3/4, nearer neighbour 1/4; outside a non-periodic wall the ghost is
// the mirrored parent (+, Neumann) or its negative (-, Dirichlet).
__device__ __forceinline__ void prol_1d(int f, int nf, int per, int dlo, int dhi, int& I0, double& w0, int& I1,
                                        double& w1) {
    const int nc = nf >> 1, parent = f >> 1;
    int other = (f & 1) ? parent + 1 : parent - 1;
    I0 = parent;
    w0 = 0.75;
    I1 = other;
    w1 = 0.25;
    if (other < 0 || other >= nc) {
        if (per) I1 = (other + nc) % nc;
        else {
            w0 += ((other < 0 ? dlo : dhi) ? -0.25 : 0.25);
            I1 = parent;
            w1 = 0.0;
        }
    }
}
__device__ __forceinline__ double pweight(int f, int I, int nf, int per, int dlo, int dhi) {
    int I0, I1;
    double w0, w1;
    prol_1d(f, nf, per, dlo, dhi, I0, w0, I1, w1);
    double w = 0.0;
    if (I0 == I) w += w0;
    if (I1 == I) w += w1;
    return w;
}

// rhs = (div(v*) - G) - b_bc, with the injected boundary faces of the
// predictor (_inject_compact_predictor! :543-579) written first.
FSI_KERNEL void k_proj_rhs(Params p, ProjBufs b, int nb) {
    const double area = p.hx * p.hy;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // every load of the thread's rows first (index 0 stands in outside)
    const int i = tile_id().x * TILE_X + threadIdx.x;
    const int jb = tile_id().y * TILE_ROWS + threadIdx.y;
    bool in[ROWS];
    size_t kk[ROWS];
    double uL[ROWS], uR[ROWS], vB[ROWS], vT[ROWS], lb[ROWS], gk[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int j = jb + r * TILE_Y;
        in[r] = i < p.nx && j < p.ny;
        const size_t k = in[r] ? at(p, i, j) : 0;
        kk[r] = k;
        uL[r] = b.Uin[k];
        uR[r] = b.Uin[k + 1];
        vB[r] = b.Vin[k];
        vT[r] = b.Vin[k + p.P];
        lb[r] = b.lap_bc[k];
        gk[r] = b.G ? b.G[k] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        if (!in[r]) continue;
        const int j = jb + r * TILE_Y;
        const size_t k = kk[r];
        // _inject_compact_predictor!: the far faces take the boundary data
        if (b.inject && i == p.nx - 1) {
            uR[r] = p.px ? b.Uin[at(p, 0, j)] : p.bU.r.value;
            b.Uin[k + 1] = uR[r];
        }
        if (b.inject && j == p.ny - 1) {
            vT[r] = p.py ? b.Vin[at(p, i, 0)] : p.bV.t.value;
            b.Vin[k + p.P] = vT[r];
        }
        double div = (uR[r] - uL[r]) * p.ihx + (vT[r] - vB[r]) * p.ihy;
        double rhs = (div - gk[r]) - lb[r];
        b.rhs[k] = rhs;
        acc[0] += rhs * area;
        acc[1] += fabs(rhs) * area;
        acc[2] += rhs * rhs;
        acc[3] += div * div * area;
    }
    put_partials<4>(acc, pset(b.partials, nb, PS_RHS), nb);
}

// The projection scalars from the right-hand side totals: compatibility
// check (src/Projection.jl:206-221) and solver tolerance.  Returns the mean
// of the right-hand side (removed from b when A is singular).
__device__ __forceinline__ double proj_scalars_of(const Params& p, const double (&tot)[4], Status* st,
                                                  int has_null, double rtol, int maxiter) {
    const double area = p.hx * p.hy;
    const double n = (double)p.nx * (double)p.ny;
    const double mean = has_null ? (tot[0] / area) / n : 0.0;
    if (block_id() == 0 && lane_id() == 0) {
        ProjScalars& q = st->proj;
        q.defect = has_null ? tot[0] : 0.0;
        q.scale = tot[1];
        q.rhs2 = tot[2];
        q.pdiv_l2sq = tot[3];
        q.mean_rhs = mean;
        q.linf = 0.0;
        if (has_null) {
            double tol = 1e-12 + 1e-10 * fmax(tot[1], 2.220446049250313e-16);
            if (!(fabs(tot[0]) <= tol)) raise_error(st, DEV_INCOMPATIBLE, tot[0], tol);
        }
        KrylovScalars& s = st->pcg;
        s.bnorm2 = fmax(tot[2] - n * mean * mean, 0.0);
        s.tol2 = rtol * rtol * s.bnorm2;
        s.iter = 0;
        s.maxiter = maxiter;
        s.done = 0;
    }
    return mean;
}
__device__ __forceinline__ double proj_scalars(const Params& p, const ProjBufs& b, Status* st, int nb, int has_null,
                                               double rtol, int maxiter) {
    double tot[4];
    if (b.rhs_part) get_totals<4>(b.rhs_part, b.rhs_np, tot);  // from the first transform pass
    else get_totals<4>(pset(b.partials, nb, PS_RHS), nb, tot);
    return proj_scalars_of(p, tot, st, has_null, rtol, maxiter);
}

// the projection scalars and r = b - (-A) x with b = -(rhs - mean)
FSI_KERNEL void k_pcg_init(Params p, MgGeom g, ProjBufs b, Status* st, int nb, int has_null,
                                                    double rtol, int maxiter) {
    const double mean = proj_scalars(p, b, st, nb, has_null, rtol, maxiter);
    GAcc X{b.x, p.P};
    double acc[1] = {0.0};
    FOR_TILE(p) {
        const size_t k = at(p, i, j);
        double r = -(b.rhs[k] - mean) - neg_lap(g, X, i, j, nullptr);
        b.r[k] = r;
        acc[0] += r * r;
    }
    put_partials<1>(acc, pset(b.partials, nb, PS_RR), nb);
}

// A solution that is only tested (no PCG iteration follows it: a planned
// projection whose initial guess is expected to meet the tolerance, such as
// the direct solve), in one pass over x: the projection scalars, and per
// block the partials of r.r (r = b - (-A) x, the convergence test of
// iteration 0), of the true residual |A x - rhs|^2 (:261-264; x and
// x - mean(x) have the same image under A) and of the area-weighted sum of x
// (its mean, :255-259).  r is not stored.
FSI_KERNEL void k_proj_direct(Params p, MgGeom g, ProjBufs b, Status* st, int nb,
                                                       int has_null, double rtol, int maxiter) {
    // the block's own cells loaded before the right-hand side totals are
    // summed, so the two latencies overlap
    const int i = tile_id().x * TILE_X + threadIdx.x;
    double xr[ROWS], rh[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int j = tile_id().y * TILE_ROWS + threadIdx.y + r * TILE_Y;
        const size_t k = (i < p.nx && j < p.ny) ? at(p, i, j) : 0;
        xr[r] = b.x[k];
        rh[r] = b.rhs[k];
    }
    const double mean = proj_scalars(p, b, st, nb, has_null, rtol, maxiter);
    const double area = p.hx * p.hy;
    GAcc X{b.x, p.P};
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int j = tile_id().y * TILE_ROWS + threadIdx.y + r * TILE_Y;
        if (i >= p.nx || j >= p.ny) continue;
        const double nl = neg_lap(g, X, i, j, nullptr);
        const double rr = -(rh[r] - mean) - nl;
        const double res = -nl - rh[r];
        acc[0] += rr * rr;
        acc[1] += res * res;
        acc[2] += xr[r] * area;
    }
    __shared__ double lds[3 * 16];
    block_sum<3>(acc, lds);
    if (lane_id() == 0) {
        const int bid = block_id();
        pset(b.partials, nb, PS_RR)[bid] = acc[0];
        pset(b.partials, nb, PS_CHECK)[bid] = acc[1];
        pset(b.partials, nb, PS_MEAN)[bid] = acc[2];
    }
}

// ---- multigrid level kernels (every one returns at once once PCG is done)
FSI_KERNEL void k_mg_jacobi0(MgGeom g, const double* b, double* x, double omega,
                                                      const int* done) {
    if (*done) return;
    FOR_TILE(g) {
        double d;
        neg_lap(g, GAcc{b, g.P}, i, j, &d);
        x[(size_t)j * g.P + i] = omega * b[(size_t)j * g.P + i] / d;
    }
}
FSI_KERNEL void k_mg_jacobi(MgGeom g, const double* xin, const double* b, double* xout,
                                                     double omega, const int* done) {
    if (*done) return;
    FOR_TILE(g) {
        double d;
        double ax = neg_lap(g, GAcc{xin, g.P}, i, j, &d);
        const size_t k = (size_t)j * g.P + i;
        xout[k] = xin[k] + omega * (b[k] - ax) / d;
    }
}
// single-level hierarchy: the convergence test the finest down-kernel
// would do, and the r.z partials the finest up-kernel would write
FSI_KERNEL void k_pcg_check(Status* st, const double* rr, int nb, int m, int slot) {
    if (st->pcg.done) return;
    double tot[1];
    get_totals<1>(rr, nb, tot);
    if (threadIdx.x == 0 && (tot[0] <= st->pcg.tol2 || m >= st->pcg.maxiter || !(tot[0] <= 1.0e300))) {
        st->pcg.done = 1;
        st->pcg.iter = m;
        st->pcg.rr = tot[0];
        if (slot >= 0) st->proj_iters[slot] = m;
        if (slot >= 0 && tot[0] <= st->pcg.tol2) st->proj_ok[slot] = 1;
    }
}
FSI_KERNEL void k_rz(Params p, ProjBufs b, double* rz, int nb, const int* done) {
    if (*done) return;
    double acc[1] = {0.0};
    FOR_TILE(p) {
        const size_t k = at(p, i, j);
        acc[0] += b.r[k] * b.z[k];
    }
    put_partials<1>(acc, rz, nb);
}

// p_m = z + beta p_{m-1} (beta = rz_m / rz_{m-1}), w = (-A) p_m, partial p.w
FSI_KERNEL void k_pcg_dir(Params p, MgGeom g, ProjBufs b, Status* st, int nb, int m) {
    if (st->pcg.done) return;
    double tot[1];
    get_totals<1>(pset(b.partials, nb, PS_RZ), nb, tot);
    const double rz = tot[0];
    const double beta = m == 0 ? 0.0 : rz / st->pcg.rzs[(m - 1) & 1];
    if (block_id() == 0 && lane_id() == 0) st->pcg.rzs[m & 1] = rz;
    const double* po = b.p[(m + 1) & 1];
    double* pn = b.p[m & 1];
    // p_m on a halo-1 LDS tile, each point formed once (loads batched)
    __shared__ double sp[VT_H * VT_W];
    const int tid = threadIdx.x + threadIdx.y * TILE_X;
    const int i0 = tile_id().x * TILE_X, j0 = tile_id().y * TILE_ROWS;
    constexpr int NPT = VT_H * VT_W, NIT = (NPT + BLOCK - 1) / BLOCK;
    size_t idx[NIT];
    bool in[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int k = min(tid + it * BLOCK, NPT - 1);
        int gi = i0 + k % VT_W - 1, gj = j0 + k / VT_W - 1;
        bool ok = true;
        if (gi < 0 || gi >= p.nx) {
            if (p.px) gi = ((gi % p.nx) + p.nx) % p.nx;
            else ok = false;
        }
        if (gj < 0 || gj >= p.ny) {
            if (p.py) gj = ((gj % p.ny) + p.ny) % p.ny;
            else ok = false;
        }
        in[it] = ok;
        idx[it] = ok ? at(p, gi, gj) : 0;
    }
    double vz[NIT], vo[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        vz[it] = b.z[idx[it]];
        vo[it] = m == 0 ? 0.0 : po[idx[it]];
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int k = tid + it * BLOCK;
        if (k < NPT) sp[k] = in[it] ? (m == 0 ? vz[it] : vz[it] + beta * vo[it]) : 0.0;
    }
    __syncthreads();
    const VelTile P{sp, i0, j0, p.nx, p.ny};
    double acc[1] = {0.0};
    FOR_TILE(p) {
        const size_t k = at(p, i, j);
        const double pv = sp[(j - j0 + 1) * VT_W + (i - i0 + 1)];
        const double w = neg_lap(g, P, i, j, nullptr);
        pn[k] = pv;
        b.w[k] = w;
        acc[0] += pv * w;
    }
    put_partials<1>(acc, pset(b.partials, nb, PS_PW), nb);
}
// x += a p_m; r -= a w (a = rz_m / p.w); partial r.r
FSI_KERNEL void k_pcg_upd(Params p, ProjBufs b, Status* st, int nb, int m, int slot) {
    if (st->pcg.done) return;
    double tot[1];
    get_totals<1>(pset(b.partials, nb, PS_PW), nb, tot);
    const double a = st->pcg.rzs[m & 1] / tot[0];
    const double* pn = b.p[m & 1];
    double acc[1] = {0.0};
    FOR_TILE(p) {
        const size_t k = at(p, i, j);
        b.x[k] += a * pn[k];
        double r = b.r[k] - a * b.w[k];
        b.r[k] = r;
        acc[0] += r * r;
    }
    put_partials<1>(acc, pset(b.partials, nb, PS_RR), nb);
    if (block_id() == 0 && lane_id() == 0 && slot >= 0) st->proj_iters[slot] = m + 1;
}

// mean of q (area weighted, src/Projection.jl:255-259)
FSI_KERNEL void k_proj_mean(Params p, ProjBufs b, int nb) {
    double acc[1] = {0.0};
    const double area = p.hx * p.hy;
    FOR_TILE(p) { acc[0] += b.x[at(p, i, j)] * area; }
    put_partials<1>(acc, pset(b.partials, nb, PS_MEAN), nb);
}

// The face gradients of q with their boundary rows (scalar_gradient,
// Operators.jl:189-356), any face: x-face (i, j), i = 0..nx; y-face (i, j),
// j = 0..ny.
__device__ __forceinline__ double grad_face_x(const Params& p, const GAcc& Q, int i, int j) {
    if (p.px) {
        int k = (i == p.nx) ? 0 : i;
        int km = k == 0 ? p.nx - 1 : k - 1;
        return p.ihx * Q(k, j) - p.ihx * Q(km, j);
    } else if (i == 0) {
        if (p.bQ.l.kind == FSI_BC_DIRICHLET) {
            double dist = p.hx / 2.0;
            return (1.0 / dist) * Q(0, j) + (-p.bQ.l.value / dist);
        }
        return -p.bQ.l.value;
    } else if (i == p.nx) {
        if (p.bQ.r.kind == FSI_BC_DIRICHLET) {
            double dist = p.hx / 2.0;
            return (-1.0 / dist) * Q(p.nx - 1, j) + p.bQ.r.value / dist;
        }
        return p.bQ.r.value;
    }
    return p.ihx * Q(i, j) - p.ihx * Q(i - 1, j);
}
__device__ __forceinline__ double grad_face_y(const Params& p, const GAcc& Q, int i, int j) {
    if (p.py) {
        int k = (j == p.ny) ? 0 : j;
        int km = k == 0 ? p.ny - 1 : k - 1;
        return p.ihy * Q(i, k) - p.ihy * Q(i, km);
    } else if (j == 0) {
        if (p.bQ.b.kind == FSI_BC_DIRICHLET) {
            double dist = p.hy / 2.0;
            return (1.0 / dist) * Q(i, 0) + (-p.bQ.b.value / dist);
        }
        return -p.bQ.b.value;
    } else if (j == p.ny) {
        if (p.bQ.t.kind == FSI_BC_DIRICHLET) {
            double dist = p.hy / 2.0;
            return (-1.0 / dist) * Q(i, p.ny - 1) + p.bQ.t.value / dist;
        }
        return p.bQ.t.value;
    }
    return p.ihy * Q(i, j) - p.ihy * Q(i, j - 1);
}

// v -= grad q on every face (src/Projection.jl:266-273), q with its BC data;
// read from Uin/Vin, written to U/V (the same arrays for a projection in
// place).  Launched on the cell grid: the thread of cell (i, j) updates the
// x-face (i, j) and the y-face (i, j), the last column and row also the far
// faces x = nx and y = ny; every load of the thread's rows is issued first.
FSI_KERNEL void k_proj_update(Params p, ProjBufs b) {
    GAcc Q{b.x, p.P};
    const int i = tile_id().x * TILE_X + threadIdx.x;
    const int jb = tile_id().y * TILE_ROWS + threadIdx.y;
    bool in[ROWS];
    size_t kk[ROWS];
    double qc[ROWS], qw[ROWS], qs[ROWS], ui[ROWS], vi[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int j = jb + r * TILE_Y;
        in[r] = i < p.nx && j < p.ny;
        const size_t k = in[r] ? at(p, i, j) : 0;
        kk[r] = k;
        qc[r] = b.x[k];
        qw[r] = b.x[in[r] && i > 0 ? k - 1 : k];
        qs[r] = b.x[in[r] && j > 0 ? k - p.P : k];
        ui[r] = b.Uin[k];
        vi[r] = b.Vin[k];
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        if (!in[r]) continue;
        const int j = jb + r * TILE_Y;
        const size_t k = kk[r];
        const double gx = i > 0 ? p.ihx * qc[r] - p.ihx * qw[r] : grad_face_x(p, Q, 0, j);
        const double gy = j > 0 ? p.ihy * qc[r] - p.ihy * qs[r] : grad_face_y(p, Q, i, 0);
        b.U[k] = ui[r] - gx;
        b.V[k] = vi[r] - gy;
        if (i == p.nx - 1) b.U[k + 1] = b.Uin[k + 1] - grad_face_x(p, Q, p.nx, j);
        if (j == p.ny - 1) b.V[k + p.P] = b.Vin[k + p.P] - grad_face_y(p, Q, i, p.ny);
    }
}

// q_out = q - mean; true residual |A q - rhs| (:261-264); constraint div(v) - G
// in L-inf and area-weighted L2 (:275-297)
template <int RESID>
FSI_KERNEL void k_proj_check(Params p, MgGeom g, ProjBufs b, Status* st, int nb,
                                                      int has_null, double lxly) {
    double tm[1];
    get_totals<1>(pset(b.partials, nb, PS_MEAN), nb, tm);
    const double mean = has_null ? tm[0] / lxly : 0.0;
    const double area = p.hx * p.hy;
    auto Q = [&](int a, int c) { return b.x[at(p, a, c)] - mean; };
    // sums: true residual, constraint, NaN faces; max |U| over every x-face,
    // max |V| over every y-face and their NaN count are the velocity extrema
    // of the driver's timestep bound, for free here
    double acc[3] = {0.0, 0.0, 0.0};
    double linf = 0.0;
    double umax = 0.0, vmax = 0.0;
    // every load of the thread's rows first (index 0 stands in outside)
    const int i = tile_id().x * TILE_X + threadIdx.x;
    const int jb = tile_id().y * TILE_ROWS + threadIdx.y;
    bool in[ROWS];
    size_t kk[ROWS];
    double xq[ROWS], uL[ROWS], uR[ROWS], vB[ROWS], vT[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int j = jb + r * TILE_Y;
        in[r] = i < p.nx && j < p.ny;
        kk[r] = in[r] ? at(p, i, j) : 0;
        const size_t k = kk[r];
        xq[r] = b.x[k];
        uL[r] = b.U[k];
        uR[r] = b.U[k + 1];
        vB[r] = b.V[k];
        vT[r] = b.V[k + p.P];
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        if (!in[r]) continue;
        const int j = jb + r * TILE_Y;
        const size_t k = kk[r];
        b.q_out[k] = xq[r] - mean;
        if (RESID) {
            double aq = -neg_lap(g, Q, i, j, nullptr);
            double res = aq - b.rhs[k];
            acc[0] += res * res;
        }
        double div = (uR[r] - uL[r]) * p.ihx + (vT[r] - vB[r]) * p.ihy;
        double c = div - (b.G ? b.G[k] : 0.0);
        linf = fmax(linf, fabs(c));
        if (!isfinite(c)) linf = INFINITY;
        acc[1] += c * c * area;
        // the far faces x = nx and y = ny belong to the last column / row
        const bool ue = i == p.nx - 1, ve = j == p.ny - 1;
        umax = fmax(umax, fmax(fabs(uL[r]), ue ? fabs(uR[r]) : 0.0));
        vmax = fmax(vmax, fmax(fabs(vB[r]), ve ? fabs(vT[r]) : 0.0));
        acc[2] += ((isnan(uL[r]) ? 1.0 : 0.0) + (isnan(vB[r]) ? 1.0 : 0.0)) +
                  ((ue && isnan(uR[r]) ? 1.0 : 0.0) + (ve && isnan(vT[r]) ? 1.0 : 0.0));
    }
    __shared__ double lds[3 * 16];
    block_sum<3>(acc, lds);
    // block maxima: the constraint as a third partial of PS_CHECK, the face
    // velocities as PS_VMAX with the NaN count (maxima taken by k_proj_final)
    __shared__ double lmax[3][BLOCK / 64];
    linf = wave_max(linf);
    umax = wave_max(umax);
    vmax = wave_max(vmax);
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    if ((lane & 63) == 0) {
        lmax[0][lane >> 6] = linf;
        lmax[1][lane >> 6] = umax;
        lmax[2][lane >> 6] = vmax;
    }
    __syncthreads();
    if (lane == 0) {
        double m[3];
        for (int q = 0; q < 3; ++q) {
            m[q] = lmax[q][0];
            for (int w = 1; w < BLOCK / 64; ++w) m[q] = fmax(m[q], lmax[q][w]);
        }
        const int bid = blockIdx.x + blockIdx.y * gridDim.x;
        double* dst = pset(b.partials, nb, PS_CHECK);
        if (RESID) dst[bid] = acc[0];  // else slot 0 (true residual) comes from k_proj_direct
        dst[(size_t)nb + bid] = acc[1];
        dst[2 * (size_t)nb + bid] = m[0];
        double* vm = pset(b.partials, nb, PS_VMAX);
        vm[bid] = m[1];
        vm[(size_t)nb + bid] = m[2];
        vm[2 * (size_t)nb + bid] = acc[2];
    }
}

// The tested projection's update and check in one pass: v = v* - grad q on
// the thread's own faces (and the far faces of the last column and row), and
// the checks of k_proj_check<0> (q_out = q - mean, constraint, velocity
// extrema) from the updated faces of each cell, the right and top ones
// recomputed exactly as the neighbouring thread computes them.  DIRECT: also
// k_proj_direct's test of the solution (the projection scalars, r.r of PCG
// iteration 0 and the true residual), with the mean of q from the partials
// of the last transform pass; otherwise the test and the mean are
// k_proj_direct's.
template <bool DIRECT>
FSI_KERNEL void k_proj_upcheck(Params p, MgGeom g, ProjBufs b, Status* st, int nb,
                                                        int has_null, double rtol, int maxiter, double lxly) {
    const double area = p.hx * p.hy;
    GAcc Q{b.x, p.P};
    const int i = tile_id().x * TILE_X + threadIdx.x;
    const int jb = tile_id().y * TILE_ROWS + threadIdx.y;
    bool in[ROWS];
    size_t kk[ROWS];
    double qc[ROWS], qw[ROWS], qs[ROWS], qe[ROWS], qn[ROWS], uL[ROWS], uR[ROWS], vB[ROWS], vT[ROWS], rh[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int j = jb + r * TILE_Y;
        in[r] = i < p.nx && j < p.ny;
        const size_t k = in[r] ? at(p, i, j) : 0;
        kk[r] = k;
        // the neighbours inside the grid; across a periodic side the wrapped
        // one (only the residual uses those; the boundary faces' gradients
        // are grad_face_x/y's)
        const size_t kw = i > 0 ? k - 1 : (DIRECT && p.px ? k + p.nx - 1 : k);
        const size_t ke = i < p.nx - 1 ? k + 1 : (DIRECT && p.px ? k - (p.nx - 1) : k);
        const size_t ks = j > 0 ? k - p.P : (DIRECT && p.py ? k + (size_t)(p.ny - 1) * p.P : k);
        const size_t kn = j < p.ny - 1 ? k + p.P : (DIRECT && p.py ? k - (size_t)(p.ny - 1) * p.P : k);
        qc[r] = b.x[k];
        qw[r] = b.x[in[r] ? kw : k];
        qs[r] = b.x[in[r] ? ks : k];
        qe[r] = b.x[in[r] ? ke : k];
        qn[r] = b.x[in[r] ? kn : k];
        uL[r] = b.Uin[k];
        uR[r] = b.Uin[k + 1];
        vB[r] = b.Vin[k];
        vT[r] = b.Vin[k + p.P];
        rh[r] = DIRECT ? b.rhs[k] : 0.0;
    }
    double mean;
    // true residual (DIRECT), constraint, NaN faces, r.r of PCG iteration 0 (DIRECT)
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (DIRECT) {
        double nl[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
            nl[r] = neg_lap5(g, i, jb + r * TILE_Y, qc[r], qw[r], qe[r], qs[r], qn[r]);
        // the right-hand side totals and the sum of x in one pass (q_part
        // follows rhs_part's four rows)
        double tot[5];
        get_totals<5, 1>(b.rhs_part, b.rhs_np, tot);  // a few partials per thread: one in flight
        const double trhs[4] = {tot[0], tot[1], tot[2], tot[3]};
        const double mrhs = proj_scalars_of(p, trhs, st, has_null, rtol, maxiter);
        mean = has_null ? tot[4] / lxly : 0.0;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            if (!in[r]) continue;
            const double rr = -(rh[r] - mrhs) - nl[r];
            const double res = -nl[r] - rh[r];
            acc[3] += rr * rr;
            acc[0] += res * res;
        }
    } else {
        double tm[1];
        get_totals<1>(pset(b.partials, nb, PS_MEAN), nb, tm);
        mean = has_null ? tm[0] / lxly : 0.0;
    }
    double linf = 0.0, umax = 0.0, vmax = 0.0;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        if (!in[r]) continue;
        const int j = jb + r * TILE_Y;
        const size_t k = kk[r];
        const bool ue = i == p.nx - 1, ve = j == p.ny - 1;
        const double gxL = i > 0 ? p.ihx * qc[r] - p.ihx * qw[r] : grad_face_x(p, Q, 0, j);
        const double gyB = j > 0 ? p.ihy * qc[r] - p.ihy * qs[r] : grad_face_y(p, Q, i, 0);
        const double gxR = !ue ? p.ihx * qe[r] - p.ihx * qc[r] : grad_face_x(p, Q, p.nx, j);
        const double gyT = !ve ? p.ihy * qn[r] - p.ihy * qc[r] : grad_face_y(p, Q, i, p.ny);
        const double u0 = uL[r] - gxL, u1 = uR[r] - gxR, v0 = vB[r] - gyB, v1 = vT[r] - gyT;
        b.U[k] = u0;
        b.V[k] = v0;
        if (ue) b.U[k + 1] = u1;
        if (ve) b.V[k + p.P] = v1;
        b.q_out[k] = qc[r] - mean;
        double div = (u1 - u0) * p.ihx + (v1 - v0) * p.ihy;
        double c = div - (b.G ? b.G[k] : 0.0);
        linf = fmax(linf, fabs(c));
        if (!isfinite(c)) linf = INFINITY;
        acc[1] += c * c * area;
        umax = fmax(umax, fmax(fabs(u0), ue ? fabs(u1) : 0.0));
        vmax = fmax(vmax, fmax(fabs(v0), ve ? fabs(v1) : 0.0));
        acc[2] += ((isnan(u0) ? 1.0 : 0.0) + (isnan(v0) ? 1.0 : 0.0)) +
                  ((ue && isnan(u1) ? 1.0 : 0.0) + (ve && isnan(v1) ? 1.0 : 0.0));
    }
    __shared__ double lds[4 * 16];
    block_sum<4>(acc, lds);
    __shared__ double lmax[3][BLOCK / 64];
    linf = wave_max(linf);
    umax = wave_max(umax);
    vmax = wave_max(vmax);
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    if ((lane & 63) == 0) {
        lmax[0][lane >> 6] = linf;
        lmax[1][lane >> 6] = umax;
        lmax[2][lane >> 6] = vmax;
    }
    __syncthreads();
    if (lane == 0) {
        double m[3];
        for (int q = 0; q < 3; ++q) {
            m[q] = lmax[q][0];
            for (int w = 1; w < BLOCK / 64; ++w) m[q] = fmax(m[q], lmax[q][w]);
        }
        const int bid = blockIdx.x + blockIdx.y * gridDim.x;
        double* dst = pset(b.partials, nb, PS_CHECK);
        if (DIRECT) {
            pset(b.partials, nb, PS_RR)[bid] = acc[3];
            dst[bid] = acc[0];
        }
        dst[(size_t)nb + bid] = acc[1];
        dst[2 * (size_t)nb + bid] = m[0];
        double* vm = pset(b.partials, nb, PS_VMAX);
        vm[bid] = m[1];
        vm[(size_t)nb + bid] = m[2];
        vm[2 * (size_t)nb + bid] = acc[2];
    }
}

// the checks of project! (:279-298), one block
FSI_KERNEL void k_proj_final(ProjBufs b, Status* st, int nb, double lxly, int pcg_m,
                                                      int slot) {
    // one pass over the partials: the r.r of PCG iteration pcg_m (its
    // convergence test, as k_pcg_check makes it), the true-residual and
    // constraint sums (each in the order of get_totals) and the constraint
    // maximum
    const double* rrp = pset(b.partials, nb, PS_RR);
    const double* chk = pset(b.partials, nb, PS_CHECK);
    const double* vm = pset(b.partials, nb, PS_VMAX);
    const int lane = lane_id();
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    double linf = 0.0, umax = 0.0, vmax = 0.0;
    constexpr int UNR = 4;  // partials of UNR blocks in flight at once
    for (int k0 = lane; k0 < nb; k0 += UNR * BLOCK) {
        double v[UNR][7];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int k = k0 + u * BLOCK;
            const bool in = k < nb;
            v[u][0] = in && pcg_m >= 0 ? rrp[k] : 0.0;
            v[u][1] = in ? chk[k] : 0.0;
            v[u][2] = in ? chk[(size_t)nb + k] : 0.0;
            v[u][3] = in ? vm[2 * (size_t)nb + k] : 0.0;
            v[u][4] = in ? chk[2 * (size_t)nb + k] : 0.0;
            v[u][5] = in ? vm[k] : 0.0;
            v[u][6] = in ? vm[(size_t)nb + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            acc[0] += v[u][0];
            acc[1] += v[u][1];
            acc[2] += v[u][2];
            acc[3] += v[u][3];
            linf = fmax(linf, v[u][4]);
            umax = fmax(umax, v[u][5]);
            vmax = fmax(vmax, v[u][6]);
        }
    }
    __shared__ double lds[4 * 16];
    __shared__ double lmax[3][BLOCK / 64];
    block_sum<4>(acc, lds);
    linf = wave_max(linf);
    umax = wave_max(umax);
    vmax = wave_max(vmax);
    if ((lane & 63) == 0) {
        lmax[0][lane >> 6] = linf;
        lmax[1][lane >> 6] = umax;
        lmax[2][lane >> 6] = vmax;
    }
    __syncthreads();
    if (lane != 0) return;
    for (int w = 0; w < BLOCK / 64; ++w) {
        umax = fmax(umax, lmax[1][w]);
        vmax = fmax(vmax, lmax[2][w]);
    }
    st->max_abs_u = acc[3] > 0.0 ? NAN : umax;  // Julia's maximum propagates NaN
    st->max_abs_v = acc[3] > 0.0 ? NAN : vmax;
    if (pcg_m >= 0 && !st->pcg.done && (acc[0] <= st->pcg.tol2 || pcg_m >= st->pcg.maxiter || !(acc[0] <= 1.0e300))) {
        st->pcg.done = 1;
        st->pcg.iter = pcg_m;
        st->pcg.rr = acc[0];
        if (slot >= 0) st->proj_iters[slot] = pcg_m;
        if (slot >= 0 && acc[0] <= st->pcg.tol2) st->proj_ok[slot] = 1;
    }
    for (int w = 1; w < BLOCK / 64; ++w) linf = fmax(linf, lmax[0][w]);
    linf = fmax(linf, lmax[0][0]);
    ProjScalars& q = st->proj;
    q.linf = linf;
    q.resid2 = acc[1];
    q.l2sq = acc[2];
    q.lin_rel = sqrt(acc[1]) / fmax(sqrt(q.rhs2), 2.220446049250313e-16);
    q.l2 = sqrt(acc[2] / lxly);
    st->max_constraint_linf = fmax(st->max_constraint_linf, q.linf);
    if (!(q.lin_rel <= 1e-8)) raise_error(st, DEV_PROJ_RESIDUAL, q.lin_rel, 1e-8);
    double cs = fmax(fmax(sqrt(q.pdiv_l2sq / lxly), b.G ? sqrt(q.tgt_l2sq / lxly) : 0.0), 1.0);
    double ctol = 1e-12 + 1e-8 * cs;
    if (!(q.l2 <= ctol)) raise_error(st, DEV_PROJ_CONSTRAINT, q.l2, ctol);
    if (slot >= 0) st->code_at[slot] = st->code;  // the first failure up to this stage's projection
}

// Plain copy, the calibration kernel of the traffic measurement.
FSI_KERNEL void k_copy(Params p, const double* __restrict__ a, double* __restrict__ b) {
    FOR_TILE(p) { b[at(p, i, j)] = a[at(p, i, j)]; }
}

// ===========================================================================
// launchers
// ===========================================================================
static dim3 grid_of(int nx, int ny) { return dim3((nx + TILE_X - 1) / TILE_X, (ny + TILE_ROWS - 1) / TILE_ROWS); }
static const dim3 kBlock(TILE_X, TILE_Y);

void launch_vel_init(hipStream_t s, const Params& p, const VelBufs& b, double alpha, Status* st, int slot,
                     const VelTol& tol) {
    fsi_launch<&k_vel_start, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b, alpha, st, grid_blocks(p));
    if (b.fr_u[0])
        fsi_launch<&k_vel_q0<float>, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b, alpha, grid_blocks(p), st, slot, tol, 1);
    else
        fsi_launch<&k_vel_q0<double>, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b, alpha, grid_blocks(p), st, slot, tol, 1);
}
void launch_vel_iters_f32(hipStream_t s, const Params& p, const VelBufs& b, double alpha, Status* st, int m0,
                          int count, int slot, const VelTol& tol) {
    const int nb = grid_blocks(p);
    for (int c = 0; c < count; ++c)
        fsi_launch<&k_vel_iter<float>, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b, alpha, st, nb, m0 + c, slot, tol, 0);
}
void launch_vel_refine(hipStream_t s, const Params& p, const VelBufs& b, double alpha, Status* st, int slot,
                       const VelTol& tol, int next) {
    const int nb = grid_blocks(p);
    VelBufs rb = b;
    rb.restart = 1;
    rb.Un = b.xu;  // x as it is (its Dirichlet values already in place)
    rb.Vn = b.xv;
    rb.extrap = 0;
    rb.Su = rb.Sv = rb.Mu = rb.Mv = nullptr;
    fsi_launch<&k_vel_start, BLOCK, 0>(grid2(p), kBlock, 0, s, p, rb, alpha, st, nb);
    fsi_launch<&k_vel_check, BLOCK, 0>(1, BLOCK, 0, s, rb, st, nb, slot, tol);
    if (next) fsi_launch<&k_vel_q0<float>, BLOCK, 0>(grid2(p), kBlock, 0, s, p, rb, alpha, nb, st, slot, tol, 0);
}
void launch_vel_iters(hipStream_t s, const Params& p, const VelBufs& b, double alpha, Status* st, int m0, int count,
                      int slot, const VelTol& tol, int check_from) {
    const int nb = grid_blocks(p);
    for (int c = 0; c < count; ++c) {
        const int m = m0 + c;
        fsi_launch<&k_vel_iter<double>, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b, alpha, st, nb, m, slot, tol,
                                                       m >= 1 && m >= check_from ? 1 : 0);
    }
}
void launch_vel_predict(hipStream_t s, const Params& p, const VelBufs& b) {
    fsi_launch<&k_vel_predict, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b);
}
void launch_vel_copy(hipStream_t s, const Params& p, const VelBufs& b) {
    fsi_launch<&k_vel_copy, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b);
}
void launch_pcg_check(hipStream_t s, Status* st, const double* rr, int nb, int m, int slot) {
    fsi_launch<&k_pcg_check, 1024, 0>(1, BLOCK, 0, s, st, rr, nb, m, slot);
}
void launch_pcg_test(hipStream_t s, const Params& p, const ProjBufs& b, Status* st, int m, int slot) {
    const int nb = grid_blocks(p);
    fsi_launch<&k_pcg_check, 1024, 0>(1, BLOCK, 0, s, st, b.partials + (size_t)PS_RR * 4 * nb, nb, m, slot);
}
void launch_rz(hipStream_t s, const Params& p, const ProjBufs& b, double* rz, int nb, const int* done) {
    fsi_launch<&k_rz, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b, rz, nb, done);
}

// Initial guess of a projection: x = 2 cur - prev (extrap, linear in time)
// or x = cur.
FSI_KERNEL void k_extrapolate(Params p, double* __restrict__ x, const double* __restrict__ cur,
                                                       const double* __restrict__ prev, int extrap) {
    FOR_TILE(p) {
        const size_t k = at(p, i, j);
        x[k] = extrap ? 2.0 * cur[k] - prev[k] : cur[k];
    }
}
void launch_extrapolate(hipStream_t s, const Params& p, double* x, const double* cur, const double* prev,
                        int extrap) {
    fsi_launch<&k_extrapolate, BLOCK, 0>(grid2(p), kBlock, 0, s, p, x, cur, prev, extrap);
}

void launch_proj_rhs(hipStream_t s, const Params& p, const ProjBufs& b) {
    fsi_launch<&k_proj_rhs, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b, grid_blocks(p));
}
void launch_pcg_init(hipStream_t s, const Params& p, const MgGeom& g0, const ProjBufs& b, Status* st, int has_null,
                     double rtol, int maxiter) {
    fsi_launch<&k_pcg_init, BLOCK, 0>(grid2(p), kBlock, 0, s, p, g0, b, st, grid_blocks(p), has_null, rtol, maxiter);
}

static void coarse_solve(hipStream_t s, const MgLevel* lv, const MgGeom* geo, int nlev, const CoarsePlan& cp,
                         const int* done) {
    const int l = nlev - 1;
    const MgGeom& g = geo[l];
    if (cp.levels > 0 && lv[0].xf) {  // mixed precision: FP32 coarse levels
        launch_mg_coarse(s, g, cp, lv[l].bf, lv[l].tf, done);
    } else if (cp.levels > 0) {
        launch_mg_coarse(s, g, cp, lv[l].b, lv[l].t, done);
    } else {  // odd-sized large coarsest level: a fixed number of damped-Jacobi sweeps
        dim3 gr = grid_of(g.nx, g.ny);
        const double omega = 0.8;  // damped-Jacobi sweeps of the fallback
        fsi_launch<&k_mg_jacobi0, BLOCK, 0>(gr, kBlock, 0, s, g, lv[l].b, lv[l].x, omega, done);
        const int sweeps = 2 * (g.nx > g.ny ? g.nx : g.ny);
        for (int it = 0; it < sweeps; ++it) {
            fsi_launch<&k_mg_jacobi, BLOCK, 0>(gr, kBlock, 0, s, g, lv[l].x, lv[l].b, lv[l].t, omega, done);
            fsi_launch<&k_mg_jacobi, BLOCK, 0>(gr, kBlock, 0, s, g, lv[l].t, lv[l].b, lv[l].x, omega, done);
        }
        fsi_launch<&k_mg_jacobi, BLOCK, 0>(gr, kBlock, 0, s, g, lv[l].x, lv[l].b, lv[l].t, omega, done);
    }
}

void launch_pcg_iters(hipStream_t s, const Params& p, const ProjBufs& b, const MgLevel* lv, const MgGeom* geo,
                      int nlev, const CoarsePlan& cp, Status* st, int m0, int count, int slot) {
    const int nb = grid_blocks(p);
    const int* done = &st->pcg.done;
    double* rr = b.partials + (size_t)PS_RR * 4 * nb;
    double* rz = b.partials + (size_t)PS_RZ * 4 * nb;
    for (int c = 0; c < count; ++c) {
        const int m = m0 + c;
        if (nlev > 1 && lv[0].xf) {  // mixed precision: FP32 V-cycle, FP64 r in and z out
            launch_mg_down(s, geo[0], geo[1], lv[0].b, lv[0].xf, lv[1].bf, done, st, rr, nb, m, slot);
            for (int l = 1; l < nlev - 1; ++l)
                launch_mg_down(s, geo[l], geo[l + 1], lv[l].bf, lv[l].xf, lv[l + 1].bf, done, nullptr, rr, nb, m,
                               slot);
            coarse_solve(s, lv, geo, nlev, cp, done);
            for (int l = nlev - 2; l >= 1; --l)
                launch_mg_up(s, geo[l], geo[l + 1], lv[l].xf, lv[l].bf, lv[l + 1].tf, lv[l].tf, done, nullptr, nb);
            launch_mg_up(s, geo[0], geo[1], lv[0].xf, lv[0].b, lv[1].tf, lv[0].t, done, rz, nb);
        } else if (nlev > 1) {
            for (int l = 0; l < nlev - 1; ++l)
                launch_mg_down(s, geo[l], geo[l + 1], lv[l].b, lv[l].x, lv[l + 1].b, done, l == 0 ? st : nullptr,
                               rr, nb, m, slot);
            coarse_solve(s, lv, geo, nlev, cp, done);
            for (int l = nlev - 2; l >= 0; --l)
                launch_mg_up(s, geo[l], geo[l + 1], lv[l].x, lv[l].b, lv[l + 1].t, lv[l].t, done,
                             l == 0 ? rz : nullptr, nb);
        } else {
            launch_pcg_check(s, st, rr, nb, m, slot);
            coarse_solve(s, lv, geo, nlev, cp, done);
            launch_rz(s, p, b, rz, nb, done);
        }
        fsi_launch<&k_pcg_dir, BLOCK, 0>(grid2(p), kBlock, 0, s, p, geo[0], b, st, nb, m);
        fsi_launch<&k_pcg_upd, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b, st, nb, m, slot);
    }
}

void launch_proj_finish(hipStream_t s, const Params& p, const MgGeom& g0, const ProjBufs& b, Status* st,
                        int has_null, double lxly, int slot) {
    const int nb = grid_blocks(p);
    fsi_launch<&k_proj_mean, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b, nb);
    fsi_launch<&k_proj_update, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b);
    fsi_launch<&k_proj_check<1>, BLOCK, 0>(grid2(p), kBlock, 0, s, p, g0, b, st, nb, has_null, lxly);
    fsi_launch<&k_proj_final, BLOCK, 0>(1, BLOCK, 0, s, b, st, nb, lxly, -1, slot);
}
void launch_proj_tested(hipStream_t s, const Params& p, const MgGeom& g0, const ProjBufs& b, Status* st,
                        int has_null, double rtol, int maxiter, double lxly, int slot) {
    const int nb = grid_blocks(p);
    const bool apart = b.Uin != b.U && b.Vin != b.V;  // the fused pass reads the neighbours' unprojected faces
    if (apart && b.rhs_part && b.q_part == b.rhs_part + (size_t)4 * b.rhs_np && b.q_np == b.rhs_np) {
        // the direct solve's own partials: test, update and checks in one pass
        fsi_launch<&k_proj_upcheck<true>, BLOCK, 0>(grid2(p), kBlock, 0, s, p, g0, b, st, nb, has_null, rtol, maxiter, lxly);
        fsi_launch<&k_proj_final, BLOCK, 0>(1, BLOCK, 0, s, b, st, nb, lxly, 0, slot);
        return;
    }
    fsi_launch<&k_proj_direct, BLOCK, 0>(grid2(p), kBlock, 0, s, p, g0, b, st, nb, has_null, rtol, maxiter);
    if (apart) {
        fsi_launch<&k_proj_upcheck<false>, BLOCK, 0>(grid2(p), kBlock, 0, s, p, g0, b, st, nb, has_null, rtol, maxiter, lxly);
    } else {
        fsi_launch<&k_proj_update, BLOCK, 0>(grid2(p), kBlock, 0, s, p, b);
        fsi_launch<&k_proj_check<0>, BLOCK, 0>(grid2(p), kBlock, 0, s, p, g0, b, st, nb, has_null, lxly);
    }
    fsi_launch<&k_proj_final, BLOCK, 0>(1, BLOCK, 0, s, b, st, nb, lxly, 0, slot);
}

// One launch of a named solver kernel with the buffers of a live solve, for
// timing.  Returns the algorithmic HBM bytes of one launch (fields read +
// written, 8 bytes per cell), or a negative value for an unknown name.
double bench_launch(hipStream_t s, const char* name, const Params& p, const VelBufs& vb, const ProjBufs& pb,
                    const MgLevel* lv, const MgGeom* geo, int nlev, const CoarsePlan& cp, Status* st,
                    const SpecPlan* sp) {
    const int nb = grid_blocks(p);
    const double cell = 8.0 * (double)p.nx * (double)p.ny;
    const std::string n(name);
    if (n == "copy") {
        fsi_launch<&k_copy, BLOCK, 0>(grid2(p), kBlock, 0, s, p, pb.rhs, pb.w);
        return 2 * cell;
    }
    if (n == "vel_iter") {
        fsi_launch<&k_vel_iter<double>, BLOCK, 0>(grid2(p), kBlock, 0, s, p, vb, 1e-3, st, nb, 2, 0,
                                                       VelTol{-1.0, 0.0, 1 << 30, 1 << 30, 1, 0, 1 << 30}, 0);
        return 16 * cell;  // r, p, q, x (u, v) read; r', p', q', x (u, v) written
    }
    if (n == "vel_iter_f32") {
        if (!vb.fr_u[0]) return -1.0;  // FP64 mode: no FP32 vectors
        fsi_launch<&k_vel_iter<float>, BLOCK, 0>(grid2(p), kBlock, 0, s, p, vb, 1e-3, st, nb, 2, 0,
                                                      VelTol{-1.0, 0.0, 1 << 30, 1 << 30, 1, 0, 1 << 30}, 0);
        return 10 * cell;  // r, p, q (FP32) and x (FP64), u and v, read and written
    }
    if (n == "pcg_dir") {
        fsi_launch<&k_pcg_dir, BLOCK, 0>(grid2(p), kBlock, 0, s, p, geo[0], pb, st, nb, 2);
        return 4 * cell;  // z, p_old read; p_new, w written
    }
    if (n == "pcg_upd") {
        fsi_launch<&k_pcg_upd, BLOCK, 0>(grid2(p), kBlock, 0, s, p, pb, st, nb, 2, 0);
        return 6 * cell;  // x, p, r, w read; x, r written
    }
    if (n == "mg_down0" && nlev > 1 && lv[0].xf) {
        launch_mg_down(s, geo[0], geo[1], lv[0].b, lv[0].xf, lv[1].bf, &st->pcg.done, nullptr, nullptr, nb, 2, 0);
        return 1.625 * cell;  // b read (FP64); x2 and b_c written (FP32)
    }
    if (n == "mg_up0" && nlev > 1 && lv[0].xf) {
        launch_mg_up(s, geo[0], geo[1], lv[0].xf, lv[0].b, lv[1].tf, lv[0].t, &st->pcg.done, nullptr, nb);
        return 2.625 * cell;  // x2, e_c (FP32) and b (FP64) read; x5 written (FP64)
    }
    if (n == "mg_down0" && nlev > 1) {
        launch_mg_down(s, geo[0], geo[1], lv[0].b, lv[0].x, lv[1].b, &st->pcg.done, nullptr, nullptr, nb, 2, 0);
        return 2.25 * cell;  // b read; x2 and b_c written
    }
    if (n == "mg_up0" && nlev > 1) {
        launch_mg_up(s, geo[0], geo[1], lv[0].x, lv[0].b, lv[1].t, lv[0].t, &st->pcg.done, nullptr, nb);
        return 3.25 * cell;  // x2, b, e_c read; x5 written
    }
    if (n == "mg_coarse" && cp.levels > 0) {
        const MgGeom& g = geo[nlev - 1];
        if (lv[0].xf) {
            launch_mg_coarse(s, g, cp, lv[nlev - 1].bf, lv[nlev - 1].tf, &st->pcg.done);
            return 8.0 * g.nx * g.ny;
        }
        launch_mg_coarse(s, g, cp, lv[nlev - 1].b, lv[nlev - 1].t, &st->pcg.done);
        return 16.0 * g.nx * g.ny;  // b read, t written
    }
    if (n == "spec_solve" && sp && sp->enabled) {  // direct solve: three passes, each reading and writing once
        launch_spectral_solve(s, *sp, p.P, pb.rhs, pb.w);
        return 6.0 * cell;
    }
    if (n == "pcg_iter") {  // one whole PCG iteration: V-cycle, direction, update
        launch_pcg_iters(s, p, pb, lv, geo, nlev, cp, st, 2, 1, -1);
        return 0.0;
    }
    if (n == "mg_jacobi") {
        fsi_launch<&k_mg_jacobi, BLOCK, 0>(grid_of(geo[0].nx, geo[0].ny), kBlock, 0, s, geo[0], lv[0].x, lv[0].b, lv[0].t, 0.8,
                                                                    &st->pcg.done);
        return 3 * cell;  // x, b read; x' written
    }
    return -1.0;
}

}  // namespace fsi
