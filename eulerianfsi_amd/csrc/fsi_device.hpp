// fsi_device.hpp -- device-side data layout and shared helpers of libfsi_hip.
//
// HBM layout: every field is stored row-major in y with x contiguous (the
// reference's column-major (i,j) order), padded to a pitch P that is a
// multiple of 32 doubles (256 B), so that each grid row starts on a 256-byte
// boundary and a 64-lane wavefront reads 512 contiguous bytes.  All fields of
// a grid share one allocation size (ny+1)*P so x-faces (nx+1 columns), y-faces
// (ny+1 rows) and cells use the same indexing at(i,j) = j*P + i.
//
// The compact velocity vector of the implicit solve, uv = [U(1:nx,1:ny);
// V(1:nx,1:ny)] (src/TimeIntegrationIMEX.jl:529-541), is stored in place in
// the U / V arrays: the compact u view ignores column nx, the compact v view
// ignores row ny.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fsi_api.h"
#include "fsi_launch.hpp"

namespace fsi {

constexpr int TILE_X = 64;         // threads along x (one wavefront row)
constexpr int TILE_Y = 4;          // thread rows per block
constexpr int ROWS = 4;            // grid rows processed per thread
constexpr int BLOCK = TILE_X * TILE_Y;
constexpr int TILE_ROWS = TILE_Y * ROWS;  // 16 grid rows per block

struct BcSide {
    int kind;      // fsi_bc_kind
    double value;
};
struct BcSet {
    BcSide l, r, b, t;  // left, right, bottom, top
};

// Parameters passed by value to every physics kernel.
struct Params {
    int nx, ny, P;
    int px, py;
    double hx, hy, ihx, ihy;
    BcSet bU, bV, bQ, bPhi, bSxx, bSyy, bSxy;
    // phase field
    int phase;
    double cdi_G, cdi_eps, cdi_d2;
    double acdi_gamma, acdi_eps, acdi_floor, acdi_d2;
    int acdi_dyn;
    double uw_alpha, uw_d2;
    int uw_normal;      // fsi_uwcdi_normal
    double* uw_field;   // the 9-point-filtered phi of a smoothed9 stage (cells), or null
    // DiffuseAreaCapillarity: the face force read by the momentum kernel
    // (null: no capillarity), its work cells and coefficients
    const double* cap_fx;
    const double* cap_fy;
    double* cap_work;  // 5 fields: two filter buffers, n_x, n_y, kappa
    double cap_coef, cap_d2;
    int cap_passes;
    // closures
    int visc;
    double mu;
    int elastic;
    double Gmod;
    int relax;
    double relax_rate;
    int ext;
    double ext_rate, ext_exp, ext_cutoff;
    int force;
    double ffx, ffy;
    const double* force_row;  // KolmogorovBodyForce per-row value (ny), or null
    // Per-block extrema and counts of the report and checks (EX_* slots of
    // exnb values each, one per block of the cell grid); reduced in a fixed
    // order by their consumer, so no global atomics are needed.
    double* exbuf;
    int exnb;
};
enum ExtSlot : int {
    EX_PHI_NEGMIN = 0,  // -min phi (max-reduced)
    EX_PHI_MAX = 1,
    EX_EIG_NEGMIN = 2,  // -min eigenvalue of sigma
    EX_DET_ERR = 3,     // max |det sigma - 1|
    EX_NONSPD = 4,      // counts (sum-reduced)
    EX_NONFINITE = 5,
    EX_STAGE1 = 6,      // max |div v - G| of the input velocity
    EX_NONFINITE_IN = 7,
    EX_SPEED = 8,       // max |u_c| (dynamic ACDI gamma)
    EX_GMAX = 9,        // max |G| at stage 1 (growth step)
    EX_TRACE = 10,      // max |tr K_g - G| over the stages (growth step)
    EX_RES_NEGMIN = 11, // -min of the new resource (growth step)
    EX_COUNT = 12
};

// Error codes raised from kernels into Status::code (first error wins).
enum DevError : int {
    DEV_OK = 0,
    DEV_NEUMANN_INFLOW = 1,      // ArgumentError: Neumann BC on an inflow face
    DEV_NONSPD = 2,              // DomainError: geometric extension of non-SPD sigma
    DEV_NONFINITE_STATE = 3,     // ArgumentError: non-finite state entry
    DEV_STAGE1 = 4,              // ArgumentError: stage-1 divergence check
    DEV_INCOMPATIBLE = 5,        // ArgumentError: projection compatibility
    DEV_PROJ_RESIDUAL = 6,       // ErrorException: projection true residual
    DEV_PROJ_CONSTRAINT = 7,     // ErrorException: projection constraint
    DEV_NONHOMOG_NEUMANN = 8,    // ArgumentError: non-homogeneous Neumann phase BC
    DEV_VEL_NOCONV = 9,          // ErrorException: velocity predictor solve not converged
    DEV_DIFFUSIVITY = 10,        // ArgumentError: evaluated diffusivity not finite and positive
    DEV_RES_SINK = 11,           // ArgumentError: evaluated resource sink not finite
    DEV_RES_NOCONV = 12          // ErrorException: resource stage solve not converged
};

// Scalars of one Krylov solve, resident in device memory.
struct KrylovScalars {
    double rr, bnorm2, tol2;
    double rzs[2];  // r.z of the last two PCG iterations (by parity)
    int done, iter, maxiter;
    int sw;  // velocity: 1 once GMRES stopped above fallback_reltol and the solve went on to the direct target
    // velocity predictor only: the stopping rules of _solve_linear_system!
    // (src/TimeIntegrationIMEX.jl:602-641)
    double fb2;    // accepted at the iteration limit when r.r <= fb2 (fallback_reltol)
    double dtol2;  // the direct-solve target of :direct and of the fallback
    double ref2;   // |b - A u^n|^2 (:gmres)
    // mixed precision: the FP32 inner CG of the current refinement cycle
    double itol2;  // its stopping threshold on r.r
    int idone;     // 1 once the cycle's inner CG has ended (the FP64 check decides what follows)
    int ibase;     // inner iterations of the finished cycles
    int cycles;    // refinement cycles checked so far
    int imax;      // most inner iterations of one cycle
    int cap;       // resource solve: the iteration limit of its current phase
};

// Stopping rules of the velocity predictor solve, passed to its kernels.
//   method 0 (:direct): ||r|| <= DIRECT_RTOL max(||b||, eps), at most `cap` iterations
//   method 1 (:gmres):  ||r|| <= rtol ||b - A u^n|| (IterativeSolvers' reltol
//                       rule, x0 = uv_n), at most `maxiter`; then accepted when
//                       ||r|| <= fb_rtol max(||b||, eps), else (fallback) the
//                       direct target up to `cap` iterations in all, else an error
// (the direct target is min(DIRECT_RTOL, fb_rtol) max(||b||, eps))
struct VelTol {
    double rtol, fb_rtol;
    int maxiter, cap;
    int method, fallback;
    int icap;  // mixed precision: inner iterations of one refinement cycle at most
};
constexpr double VEL_DIRECT_RTOL = 1e-14;

// Projection scalars (src/Projection.jl:170-307).
struct ProjScalars {
    double defect, scale, mean_rhs, rhs2, pdiv_l2sq, mean_q;
    double resid2, linf, l2sq;
    double lin_rel, l2;
    double tgt_l2sq;  // sum of G^2 dA of the target divergence (growth step; scales the constraint check)
};

// A first-failure record (a DevError code and two values describing it).
struct ErrSlot {
    int code;
    int pad0;
    double v0, v1;
};

// Everything the host reads back once per step.
struct Status {
    int code;        // DevError of the first failure
    int pad0;
    double v0, v1;   // values describing the first failure
    // StepReport accumulators (src/TimeIntegrationIMEX.jl:318-327)
    double max_constraint_linf;
    double phi_min, phi_max;
    double min_eig, max_det_err;
    long long nonspd, nonfinite;
    long long nonfinite_in;  // non-finite entries of the input state (stage 1)
    double stage1_linf;      // max |div(v) - G| of the input velocity (stage 1)
    double gamma_umax;       // dynamic ACDI gamma: max |u_c|
    double max_abs_u, max_abs_v;  // face-velocity extrema after the last projection
    int vel_iters[3], proj_iters[3];
    int vel_ok[3], proj_ok[3];  // 1 once the solve of that stage has converged
    // per predictor solve: final |r| / max(|b|, eps) and |b - A u^n| / max(|b|, eps)
    double vel_res[3], vel_ref[3];
    int vel_cycles[3], vel_imax[3];  // mixed precision: refinement cycles, most inner iterations of one
    KrylovScalars vel;
    KrylovScalars pcg;
    ProjScalars proj;
    KrylovScalars res;    // the resource stage solve (growth step)
    double stage1_gmax;   // max |G| at stage 1 (the scale of the stage-1 check)
    double growth_trace;  // maximum_growth_trace_error of the step
    double res_min;       // minimum_resource of the new state (growth step)
    // The phase and sigma kernels of stage q report their failures here (they
    // may run beside the solve of stage q + 1); code_at[s] is `code` after the
    // projection of stage s + 1, so the host can order the two as the
    // reference's sequence of operations does.
    ErrSlot cells_err[3];
    int code_at[3];
    int pad9;
};

__device__ __forceinline__ size_t at(const Params& p, int i, int j) { return (size_t)j * p.P + i; }

// XCD-aware tile order.  The hardware deals consecutive workgroups round-robin
// to the 8 XCDs, each with its own L2; renumbering the tiles so that every XCD
// works a contiguous run of them (row-major) keeps the halo rows its tiles
// share in that XCD's L2.  A bijection whenever the block count is a
// multiple of 8, otherwise the identity.
constexpr int kXcds = 8;
struct TileId {
    int x, y;
};
__device__ __forceinline__ TileId tile_id() {
    const int nbx = gridDim.x, nb = gridDim.x * gridDim.y;
    const int b = blockIdx.x + blockIdx.y * nbx;
    const int t = (nb % kXcds == 0) ? (b % kXcds) * (nb / kXcds) + b / kXcds : b;
    return {t % nbx, t / nbx};
}

__device__ __forceinline__ void raise_error(Status* st, int code, double v0, double v1) {
    if (atomicCAS(&st->code, 0, code) == 0) {
        st->v0 = v0;
        st->v1 = v1;
    }
}
__device__ __forceinline__ void raise_error(ErrSlot* e, int code, double v0, double v1) {
    if (atomicCAS(&e->code, 0, code) == 0) {
        e->v0 = v0;
        e->v1 = v1;
    }
}

// ---- deterministic block reduction (wave64 shuffles, then LDS in wave order)
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /* >= NV*BLOCK/64 */) {
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    const int nthreads = blockDim.x * blockDim.y;
#pragma unroll
    for (int k = 0; k < NV; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
    const int wave = lane >> 6, nwaves = (nthreads + 63) >> 6;
    if ((lane & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[k * 16 + wave] = v[k];
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int w = 0; w < nwaves; ++w) s += lds[k * 16 + w];
            v[k] = s;
        }
    }
    __syncthreads();
}

// Block-wide max / min (order independent, hence deterministic).
__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
    return v;
}

// Cross-block reductions never synchronise blocks inside a kernel: a kernel
// writes one partial per block (put_partials), and every block of the next
// kernel sums all partials in the same fixed order (get_totals), so every
// block derives bitwise-identical totals and the result is deterministic.
// The kernel boundary orders the writes before the reads.
template <int NV>
__device__ __forceinline__ void put_partials(double (&v)[NV], double* dst, int nb) {
    __shared__ double lds[NV * 16];
    block_sum<NV>(v, lds);
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    const int bid = blockIdx.x + blockIdx.y * gridDim.x;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) dst[(size_t)k * nb + bid] = v[k];
}

template <int NV, int UNR = 4>
__device__ __forceinline__ void get_totals(const double* src, int nb, double (&out)[NV]) {
    __shared__ double lds[NV * 16];
    __shared__ double res[NV];
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    const int nthreads = blockDim.x * blockDim.y;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    // UNR partials per value in flight at once (one memory latency per UNR
    // instead of per partial; UNR * NV doubles of registers); the additions
    // keep the order b = lane, lane + nthreads, ... (adding 0.0 past nb
    // leaves the sum unchanged)
    for (int b0 = lane; b0 < nb; b0 += UNR * nthreads) {
        double v[UNR][NV];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int b = b0 + u * nthreads;
#pragma unroll
            for (int k = 0; k < NV; ++k) v[u][k] = b < nb ? src[(size_t)k * nb + b] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u)
#pragma unroll
            for (int k = 0; k < NV; ++k) acc[k] += v[u][k];
    }
    block_sum<NV>(acc, lds);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) res[k] = acc[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = res[k];
}

}  // namespace fsi
