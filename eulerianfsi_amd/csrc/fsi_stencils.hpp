// fsi_stencils.hpp -- stencils shared by the tendency kernels and the
// implicit velocity solve: the viscous operator M of build_viscous_operator
// (src/Operators.jl:929-1083) applied matrix-free to a compact velocity.
//
// The operator is written against accessors u(i,j), v(i,j) so that the same
// code applies it to a stored vector, to the zero vector (boundary part
// only), or to a vector formed on the fly (p = r + beta*p_old in the CG).
// IN = true: the caller guarantees 1 <= i <= nx-2 and 1 <= j <= ny-2 for the
// point and its stencil, so every boundary branch takes its interior side and
// compiles out; the arithmetic is unchanged.
#pragma once

#include "fsi_device.hpp"

namespace fsi {

struct PtrAcc {
    const double* a;
    int P;
    __device__ __forceinline__ double operator()(int i, int j) const { return a[(size_t)j * P + i]; }
};
struct ZeroAcc {
    __device__ __forceinline__ double operator()(int, int) const { return 0.0; }
};

// tau_xy at compact corner (x_u[i], y_v[jc]): 2 mu * (0.5 du/dy + 0.5 dv/dx);
// the du/dy part vanishes on a non-periodic bottom row and the dv/dx part on a
// non-periodic left column (build_velocity_gradient_operators :893-913).
template <bool IN = false, class FU, class FV>
__device__ __forceinline__ double visc_txy(const Params& p, const FU& u, const FV& v, int i, int jc) {
    double duy = 0.0, dvx = 0.0;
    if (IN || jc > 0) duy = -p.ihy * u(i, jc - 1) + p.ihy * u(i, jc);
    else if (p.py) duy = -p.ihy * u(i, p.ny - 1) + p.ihy * u(i, 0);
    if (IN || i > 0) dvx = -p.ihx * v(i - 1, jc) + p.ihx * v(i, jc);
    else if (p.px) dvx = -p.ihx * v(p.nx - 1, jc) + p.ihx * v(0, jc);
    return (0.5 * duy + 0.5 * dvx) * (2.0 * p.mu);
}
// tau_xx at cell (k,j): 2 mu du/dx, with the U-right datum as E_bc (:870-877)
template <bool IN = false, class FU>
__device__ __forceinline__ double visc_txx(const Params& p, const FU& u, int k, int j, int wb) {
    double a = -p.ihx * u(k, j);
    if (IN || k < p.nx - 1) a += p.ihx * u(k + 1, j);
    else if (p.px) a += p.ihx * u(0, j);
    else if (wb) a += p.bU.r.value * p.ihx;
    return a * (2.0 * p.mu);
}
template <bool IN = false, class FV>
__device__ __forceinline__ double visc_tyy(const Params& p, const FV& v, int i, int k, int wb) {
    double b = -p.ihy * v(i, k);
    if (IN || k < p.ny - 1) b += p.ihy * v(i, k + 1);
    else if (p.py) b += p.ihy * v(i, 0);
    else if (wb) b += p.bV.t.value * p.ihy;
    return b * (2.0 * p.mu);
}
// (M uv + wb*b)_u at compact x-face (i,j): Dstress rows :1000-1034 plus the
// tangential Dirichlet wall lifting :1036-1058.
template <bool IN = false, class FU, class FV>
__device__ __forceinline__ double visc_u(const Params& p, const FU& u, const FV& v, int i, int j, int wb) {
    double fu = p.ihx * visc_txx<IN>(p, u, i, j, wb);
    if (IN || i > 0) fu += -p.ihx * visc_txx<IN>(p, u, i - 1, j, wb);
    else if (p.px) fu += -p.ihx * visc_txx<IN>(p, u, p.nx - 1, j, wb);
    fu += -p.ihy * visc_txy<IN>(p, u, v, i, j);
    if (IN || j < p.ny - 1) fu += p.ihy * visc_txy<IN>(p, u, v, i, j + 1);
    else if (p.py) fu += p.ihy * visc_txy<IN>(p, u, v, i, 0);
    if (!IN && !p.py) {
        if (j == 0 && p.bU.b.kind == FSI_BC_DIRICHLET) {
            double w = p.mu / (p.hy * (p.hy / 2.0));
            fu += -w * u(i, 0) + (wb ? w * p.bU.b.value : 0.0);
        }
        if (j == p.ny - 1 && p.bU.t.kind == FSI_BC_DIRICHLET) {
            double w = p.mu / (p.hy * (p.hy / 2.0));
            fu += -w * u(i, j) + (wb ? w * p.bU.t.value : 0.0);
        }
    }
    return fu;
}
// (M uv + wb*b)_v at compact y-face (i,j), lifting :1060-1080.
template <bool IN = false, class FU, class FV>
__device__ __forceinline__ double visc_v(const Params& p, const FU& u, const FV& v, int i, int j, int wb) {
    double fv = -p.ihx * visc_txy<IN>(p, u, v, i, j);
    if (IN || i < p.nx - 1) fv += p.ihx * visc_txy<IN>(p, u, v, i + 1, j);
    else if (p.px) fv += p.ihx * visc_txy<IN>(p, u, v, 0, j);
    fv += p.ihy * visc_tyy<IN>(p, v, i, j, wb);
    if (IN || j > 0) fv += -p.ihy * visc_tyy<IN>(p, v, i, j - 1, wb);
    else if (p.py) fv += -p.ihy * visc_tyy<IN>(p, v, i, p.ny - 1, wb);
    if (!IN && !p.px) {
        if (i == 0 && p.bV.l.kind == FSI_BC_DIRICHLET) {
            double w = p.mu / (p.hx * (p.hx / 2.0));
            fv += -w * v(0, j) + (wb ? w * p.bV.l.value : 0.0);
        }
        if (i == p.nx - 1 && p.bV.r.kind == FSI_BC_DIRICHLET) {
            double w = p.mu / (p.hx * (p.hx / 2.0));
            fv += -w * v(i, j) + (wb ? w * p.bV.r.value : 0.0);
        }
    }
    return fv;
}

// Dirichlet rows of the predictor solve: U-left column and V-bottom row
// (velocity_dirichlet_data, Operators.jl:1085-1110).
template <bool IN = false>
__device__ __forceinline__ bool dir_u(const Params& p, int i) {
    return !IN && !p.px && i == 0 && p.bU.l.kind == FSI_BC_DIRICHLET;
}
template <bool IN = false>
__device__ __forceinline__ bool dir_v(const Params& p, int j) {
    return !IN && !p.py && j == 0 && p.bV.b.kind == FSI_BC_DIRICHLET;
}

}  // namespace fsi
