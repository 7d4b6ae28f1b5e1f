// fsi_adjoint.hip -- kernels of the discrete adjoint of the IMEX step
// (integrate_step!, src/TimeIntegrationIMEX.jl:1152-1256): the transposed
// linearisation of the stage tendencies (_evaluate_stage_tendencies!,
// :944-1071) at a recorded stage, and the elementwise pieces of the
// transposed stage combinations.
//
// The tendencies are local stencils.  Their Jacobian-transpose products are
// formed by gathering: one thread per input point p (a cell, an x-face or a
// y-face) evaluates every tendency whose stencil holds p, in forward-mode
// dual numbers seeded at p, and sums (adjoint of the tendency) x
// (derivative).  No atomics, and every sum has a fixed order, so the
// results are deterministic.
//
// The point functions restate the forward kernels' formulas
// (fsi_physics.hip, fsi_stencils.hpp; the reference's material_convection
// MomentumTransport.jl:252-338, symmetric_stress_divergence
// Operators.jl:620-735, build_viscous_operator :929-1083, the ACDI fluxes
// PhaseFieldMethods.jl:479-671, structure_tensor_rhs
// StructureTensor.jl:267-551) generically in the value type: with T = double
// they reproduce the forward tendencies (k_adj_primal, a parity check), and
// in dual numbers every comparison looks at the value, so the upwind side,
// the psi clamp, the Hermite limit and the safety reset are frozen as the
// forward took them at the recorded stage.  Walls only (no periodic side);
// the host rejects other configurations.
#include "fsi_kernels.hpp"

namespace fsi {
namespace {

// ---------------------------------------------------------------------------
// forward-mode dual numbers
// ---------------------------------------------------------------------------
template <int K>
struct Dual {
    double v;
    double d[K];
    __device__ __forceinline__ Dual() {}
    __device__ __forceinline__ Dual(double x) : v(x) {
#pragma unroll
        for (int k = 0; k < K; ++k) d[k] = 0.0;
    }
};
template <int K>
__device__ __forceinline__ Dual<K> chain(const Dual<K>& a, double f, double s) {
    Dual<K> r;
    r.v = f;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = s * a.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(Dual<K> a, const Dual<K>& b) {
    a.v += b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) a.d[k] += b.d[k];
    return a;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(Dual<K> a, const Dual<K>& b) {
    a.v -= b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) a.d[k] -= b.d[k];
    return a;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v / b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(Dual<K> a, double b) {
    a.v += b;
    return a;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(double b, Dual<K> a) {
    a.v += b;
    return a;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(Dual<K> a, double b) {
    a.v -= b;
    return a;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(double b, const Dual<K>& a) {
    return chain(a, b - a.v, -1.0);
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a) {
    return chain(a, -a.v, -1.0);
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, double b) {
    return chain(a, a.v * b, b);
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(double b, const Dual<K>& a) {
    return chain(a, b * a.v, b);
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, double b) {
    Dual<K> r;
    r.v = a.v / b;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] / b;
    return r;
}

__device__ __forceinline__ double val(double x) { return x; }
template <int K>
__device__ __forceinline__ double val(const Dual<K>& x) {
    return x.v;
}
__device__ __forceinline__ double g_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double g_log(double x) { return log(x); }
__device__ __forceinline__ double g_tanh(double x) { return tanh(x); }
__device__ __forceinline__ double g_hypot(double a, double b) { return hypot(a, b); }
__device__ __forceinline__ double g_pow(double x, double e) { return pow(x, e); }
template <int K>
__device__ __forceinline__ Dual<K> g_sqrt(const Dual<K>& a) {
    const double r = sqrt(a.v);
    return chain(a, r, 0.5 / r);
}
template <int K>
__device__ __forceinline__ Dual<K> g_log(const Dual<K>& a) {
    return chain(a, log(a.v), 1.0 / a.v);
}
template <int K>
__device__ __forceinline__ Dual<K> g_tanh(const Dual<K>& a) {
    const double t = tanh(a.v);
    return chain(a, t, 1.0 - t * t);
}
template <int K>
__device__ __forceinline__ Dual<K> g_hypot(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = hypot(a.v, b.v);
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = r.v > 0.0 ? (a.v * a.d[k] + b.v * b.d[k]) / r.v : 0.0;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> g_pow(const Dual<K>& a, double e) {
    const double pw = pow(a.v, e);
    return chain(a, pw, (e == 0.0 || a.v == 0.0) ? 0.0 : e * pow(a.v, e - 1.0));
}
// the value with derivative 1 in component c (none past the type's components)
template <class T>
struct Seed;
template <>
struct Seed<double> {
    __device__ __forceinline__ static double at(double x, int) { return x; }
};
template <int K>
struct Seed<Dual<K>> {
    __device__ __forceinline__ static Dual<K> at(double x, int c) {
        Dual<K> r(x);
        if (c >= 0 && c < K) r.d[c] = 1.0;
        return r;
    }
};
template <class T>
__device__ __forceinline__ T seeded(double x, int c) {
    return Seed<T>::at(x, c);
}

// ---------------------------------------------------------------------------
// accessors: a field read at (i, j), the value seeded at one point
// ---------------------------------------------------------------------------
template <class T>
struct Acc {
    const double* a;
    int P;
    int si, sj, comp;  // the seeded point and derivative component (si < 0: none)
    __device__ __forceinline__ T operator()(int i, int j) const {
        const double x = a[(size_t)j * P + i];
        return (i == si && j == sj) ? seeded<T>(x, comp) : T(x);
    }
};
template <class T>
struct SymT {
    T xx, yy, xy;
};
// sigma of the stage, with the safety reset on load where `reset`
// (apply_structure_extension!, StructureTensor.jl:590-613): reset cells are
// constants of the linearisation
template <class T>
struct SigAcc {
    Acc<T> xx, yy, xy;
    const double* phi;
    int P, reset;
    double cutoff;
    __device__ __forceinline__ SymT<T> operator()(int i, int j) const {
        if (reset && phi[(size_t)j * P + i] < cutoff) return {T(1.0), T(1.0), T(0.0)};
        return {xx(i, j), yy(i, j), xy(i, j)};
    }
};
// the elastic stress phi G (sigma - I), one component
// (StructuralConstitutive.jl:100-113)
template <class T, int C>
struct StressAcc {
    const SigAcc<T>* s;
    const Acc<T>* phi;
    double G;
    __device__ __forceinline__ T operator()(int i, int j) const {
        const SymT<T> v = (*s)(i, j);
        const T ph = (*phi)(i, j);
        if (C == 0) return ph * (G * (v.xx - 1.0));
        if (C == 1) return ph * (G * (v.yy - 1.0));
        return ph * (G * v.xy);
    }
};

// ---------------------------------------------------------------------------
// point functions (walls only)
// ---------------------------------------------------------------------------
// corner values of U and V (MomentumTransport.jl:161-219)
template <class T, class FU>
__device__ __forceinline__ T a_u_corner(const Params& p, const FU& U, int i, int jc) {
    if (jc == 0)
        return p.bU.b.kind == FSI_BC_DIRICHLET ? T(p.bU.b.value) : U(i, 0) + p.bU.b.value * (p.hy / 2.0);
    if (jc == p.ny)
        return p.bU.t.kind == FSI_BC_DIRICHLET ? T(p.bU.t.value) : U(i, p.ny - 1) + p.bU.t.value * (p.hy / 2.0);
    return 0.5 * U(i, jc - 1) + 0.5 * U(i, jc);
}
template <class T, class FV>
__device__ __forceinline__ T a_v_corner(const Params& p, const FV& V, int i, int jc) {
    if (i == 0)
        return p.bV.l.kind == FSI_BC_DIRICHLET ? T(p.bV.l.value) : V(0, jc) + p.bV.l.value * (p.hx / 2.0);
    if (i == p.nx)
        return p.bV.r.kind == FSI_BC_DIRICHLET ? T(p.bV.r.value) : V(p.nx - 1, jc) + p.bV.r.value * (p.hx / 2.0);
    return 0.5 * V(i - 1, jc) + 0.5 * V(i, jc);
}
// convection at x-face (i, j), zero on the wall-normal row i = 0
// (MomentumTransport.jl:252-338)
template <class T, class FU, class FV>
__device__ __forceinline__ T a_conv_u(const Params& p, const FU& U, const FV& V, int i, int j) {
    if (i == 0) return T(0.0);
    auto u2c = [&](int k) __attribute__((always_inline)) -> T {
        const T uc = 0.5 * U(k, j) + 0.5 * U(k + 1, j);
        return uc * uc;
    };
    auto uv = [&](int jc) __attribute__((always_inline)) -> T { return a_u_corner<T>(p, U, i, jc) * a_v_corner<T>(p, V, i, jc); };
    return (u2c(i) - u2c(i - 1)) * p.ihx + (uv(j + 1) - uv(j)) * p.ihy;
}
template <class T, class FU, class FV>
__device__ __forceinline__ T a_conv_v(const Params& p, const FU& U, const FV& V, int i, int j) {
    if (j == 0) return T(0.0);
    auto v2c = [&](int k) __attribute__((always_inline)) -> T {
        const T vc = 0.5 * V(i, k) + 0.5 * V(i, k + 1);
        return vc * vc;
    };
    auto uv = [&](int ic) __attribute__((always_inline)) -> T { return a_u_corner<T>(p, U, ic, j) * a_v_corner<T>(p, V, ic, j); };
    return (uv(i + 1) - uv(i)) * p.ihx + (v2c(j) - v2c(j - 1)) * p.ihy;
}
// symmetric_stress_divergence at x-face (i, j) / y-face (i, j) with the
// linear wall extrapolation (Operators.jl:620-842)
template <class T, class FT, class FX>
__device__ __forceinline__ T a_elastic_fx(const Params& p, const FT& Txx, const FX& Txy, int i, int j) {
    const int ny = p.ny;
    T fx;
    if (i == 0) {
        const T lw = Txx(0, j) - (Txx(1, j) - Txx(0, j)) * 0.5;
        fx = (Txx(0, j) - lw) * (2.0 * p.ihx);
    } else {
        fx = (Txx(i, j) - Txx(i - 1, j)) * p.ihx;
    }
    auto txy = [&](int jj) __attribute__((always_inline)) -> T {
        if (i == 0) return Txy(0, jj) - (Txy(1, jj) - Txy(0, jj)) * 0.5;
        return 0.5 * Txy(i - 1, jj) + 0.5 * Txy(i, jj);
    };
    if (j == 0) {
        const T t0 = txy(0), t1 = txy(1);
        const T bw = t0 - (t1 - t0) * 0.5;
        fx = fx + (t0 - bw) * (2.0 * p.ihy);
    } else if (j == ny - 1) {
        const T t0 = txy(ny - 1), t1 = txy(ny - 2);
        const T tw = t0 + (t0 - t1) * 0.5;
        fx = fx + (tw - t0) * (2.0 * p.ihy);
    } else {
        fx = fx + (txy(j + 1) - txy(j - 1)) * (0.5 * p.ihy);
    }
    return fx;
}
template <class T, class FT, class FX>
__device__ __forceinline__ T a_elastic_fy(const Params& p, const FT& Tyy, const FX& Txy, int i, int j) {
    const int nx = p.nx;
    T fy;
    if (j == 0) {
        const T bw = Tyy(i, 0) - (Tyy(i, 1) - Tyy(i, 0)) * 0.5;
        fy = (Tyy(i, 0) - bw) * (2.0 * p.ihy);
    } else {
        fy = (Tyy(i, j) - Tyy(i, j - 1)) * p.ihy;
    }
    auto txy = [&](int ii) __attribute__((always_inline)) -> T {
        if (j == 0) return Txy(ii, 0) - (Txy(ii, 1) - Txy(ii, 0)) * 0.5;
        return 0.5 * Txy(ii, j - 1) + 0.5 * Txy(ii, j);
    };
    if (i == 0) {
        const T t0 = txy(0), t1 = txy(1);
        const T lw = t0 - (t1 - t0) * 0.5;
        fy = fy + (t0 - lw) * (2.0 * p.ihx);
    } else if (i == nx - 1) {
        const T t0 = txy(nx - 1), t1 = txy(nx - 2);
        const T rw = t0 + (t0 - t1) * 0.5;
        fy = fy + (rw - t0) * (2.0 * p.ihx);
    } else {
        fy = fy + (txy(i + 1) - txy(i - 1)) * (0.5 * p.ihx);
    }
    return fy;
}
// the viscous operator M (+ wb * its wall lifting) on compact faces
// (Operators.jl:848-1083, constant viscosity)
template <class T, class FU, class FV>
__device__ __forceinline__ T a_visc_txy(const Params& p, const FU& u, const FV& v, int i, int jc) {
    T duy(0.0), dvx(0.0);
    if (jc > 0) duy = -p.ihy * u(i, jc - 1) + p.ihy * u(i, jc);
    if (i > 0) dvx = -p.ihx * v(i - 1, jc) + p.ihx * v(i, jc);
    return (0.5 * duy + 0.5 * dvx) * (2.0 * p.mu);
}
template <class T, class FU>
__device__ __forceinline__ T a_visc_txx(const Params& p, const FU& u, int k, int j, int wb) {
    T a = -p.ihx * u(k, j);
    if (k < p.nx - 1) a = a + p.ihx * u(k + 1, j);
    else if (wb) a = a + p.bU.r.value * p.ihx;
    return a * (2.0 * p.mu);
}
template <class T, class FV>
__device__ __forceinline__ T a_visc_tyy(const Params& p, const FV& v, int i, int k, int wb) {
    T b = -p.ihy * v(i, k);
    if (k < p.ny - 1) b = b + p.ihy * v(i, k + 1);
    else if (wb) b = b + p.bV.t.value * p.ihy;
    return b * (2.0 * p.mu);
}
template <class T, class FU, class FV>
__device__ __forceinline__ T a_visc_u(const Params& p, const FU& u, const FV& v, int i, int j, int wb) {
    T fu = p.ihx * a_visc_txx<T>(p, u, i, j, wb);
    if (i > 0) fu = fu + -p.ihx * a_visc_txx<T>(p, u, i - 1, j, wb);
    fu = fu + -p.ihy * a_visc_txy<T>(p, u, v, i, j);
    if (j < p.ny - 1) fu = fu + p.ihy * a_visc_txy<T>(p, u, v, i, j + 1);
    if (j == 0 && p.bU.b.kind == FSI_BC_DIRICHLET) {
        const double w = p.mu / (p.hy * (p.hy / 2.0));
        fu = fu + -w * u(i, 0) + (wb ? w * p.bU.b.value : 0.0);
    }
    if (j == p.ny - 1 && p.bU.t.kind == FSI_BC_DIRICHLET) {
        const double w = p.mu / (p.hy * (p.hy / 2.0));
        fu = fu + -w * u(i, j) + (wb ? w * p.bU.t.value : 0.0);
    }
    return fu;
}
template <class T, class FU, class FV>
__device__ __forceinline__ T a_visc_v(const Params& p, const FU& u, const FV& v, int i, int j, int wb) {
    T fv = -p.ihx * a_visc_txy<T>(p, u, v, i, j);
    if (i < p.nx - 1) fv = fv + p.ihx * a_visc_txy<T>(p, u, v, i + 1, j);
    fv = fv + p.ihy * a_visc_tyy<T>(p, v, i, j, wb);
    if (j > 0) fv = fv + -p.ihy * a_visc_tyy<T>(p, v, i, j - 1, wb);
    if (i == 0 && p.bV.l.kind == FSI_BC_DIRICHLET) {
        const double w = p.mu / (p.hx * (p.hx / 2.0));
        fv = fv + -w * v(0, j) + (wb ? w * p.bV.l.value : 0.0);
    }
    if (i == p.nx - 1 && p.bV.r.kind == FSI_BC_DIRICHLET) {
        const double w = p.mu / (p.hx * (p.hx / 2.0));
        fv = fv + -w * v(i, j) + (wb ? w * p.bV.r.value : 0.0);
    }
    return fv;
}

// ---- ACDI (PhaseFieldMethods.jl:479-671, Operators.jl:189-356)
template <class T>
__device__ __forceinline__ T a_clamp01(const T& x) {
    return val(x) < 0.0 ? T(0.0) : (val(x) > 1.0 ? T(1.0) : x);
}
template <class T>
__device__ __forceinline__ T a_psi(const Params& p, const T& ph) {
    const T q = a_clamp01(ph);
    const double e = p.acdi_floor;
    return p.acdi_eps * g_log((q + e) / (1.0 - q + e));
}
// scalar_gradient on x-face k / y-face k with the BC rows of `bc`
template <class T, class F>
__device__ __forceinline__ T a_grad_x(const Params& p, const BcSide& l, const BcSide& r, const F& f, int k, int j) {
    if (k == 0) {
        if (l.kind == FSI_BC_DIRICHLET) {
            const double dist = p.hx / 2.0;
            return (1.0 / dist) * f(k, j) + (-l.value / dist);  // k = 0
        }
        return T(-l.value);
    }
    if (k == p.nx) {
        if (r.kind == FSI_BC_DIRICHLET) {
            const double dist = p.hx / 2.0;
            return (-1.0 / dist) * f(k - 1, j) + r.value / dist;  // k = nx
        }
        return T(r.value);
    }
    return p.ihx * f(k, j) - p.ihx * f(k - 1, j);
}
template <class T, class F>
__device__ __forceinline__ T a_grad_y(const Params& p, const BcSide& b, const BcSide& t, const F& f, int i, int k) {
    if (k == 0) {
        if (b.kind == FSI_BC_DIRICHLET) {
            const double dist = p.hy / 2.0;
            return (1.0 / dist) * f(i, k) + (-b.value / dist);  // k = 0
        }
        return T(-b.value);
    }
    if (k == p.ny) {
        if (t.kind == FSI_BC_DIRICHLET) {
            const double dist = p.hy / 2.0;
            return (-1.0 / dist) * f(i, k - 1) + t.value / dist;  // k = ny
        }
        return T(t.value);
    }
    return p.ihy * f(i, k) - p.ihy * f(i, k - 1);
}
// _phase_boundary_value (PhaseFieldMethods.jl:174-196)
template <class T>
__device__ __forceinline__ T a_phase_bvalue(const BcSide& b, const T& interior) {
    return b.kind == FSI_BC_DIRICHLET ? T(b.value) : interior;
}
// E_phi of the ACDI model at cell (i, j); psi is formed once on the diamond
// |di| + |dj| <= 2 (the cells the normals of the five cells around (i, j)
// read) and the normals once per cell
template <class T, class FU, class FV, class FP>
__device__ __forceinline__ T a_acdi(const Params& p, const BcSet& bpsi, const FU& U, const FV& V, const FP& phi,
                                    int i, int j) {
    T ps[5][5];
#pragma unroll
    for (int dj = -2; dj <= 2; ++dj)
#pragma unroll
        for (int di = -2; di <= 2; ++di) {
            const int m = i + di, n = j + dj;
            const bool use = (di < 0 ? -di : di) + (dj < 0 ? -dj : dj) <= 2;
            ps[dj + 2][di + 2] = (use && m >= 0 && m < p.nx && n >= 0 && n < p.ny) ? a_psi(p, phi(m, n)) : T(0.0);
        }
    auto PS = [&](int m, int n) __attribute__((always_inline)) -> T { return ps[n - j + 2][m - i + 2]; };
    // the normal at cell (i + di, j + dj), |di| + |dj| <= 1 (_center_normals :233-244)
    auto normal = [&](int di, int dj, T& nxv, T& nyv) __attribute__((always_inline)) {
        const int m = i + di, n = j + dj;
        const T gx = 0.5 * a_grad_x<T>(p, bpsi.l, bpsi.r, PS, m, n) + 0.5 * a_grad_x<T>(p, bpsi.l, bpsi.r, PS, m + 1, n);
        const T gy = 0.5 * a_grad_y<T>(p, bpsi.b, bpsi.t, PS, m, n) + 0.5 * a_grad_y<T>(p, bpsi.b, bpsi.t, PS, m, n + 1);
        const T mag = g_sqrt(gx * gx + gy * gy + p.acdi_d2);
        nxv = gx / mag;
        nyv = gy / mag;
    };
    T n0c, n1c, n0w(0.0), n1w(0.0), n0e(0.0), n1e(0.0), n0s(0.0), n1s(0.0), n0n(0.0), n1n(0.0);
    normal(0, 0, n0c, n1c);
    if (i > 0) normal(-1, 0, n0w, n1w);
    if (i < p.nx - 1) normal(1, 0, n0e, n1e);
    if (j > 0) normal(0, -1, n0s, n1s);
    if (j < p.ny - 1) normal(0, 1, n0n, n1n);
    const double eps = p.acdi_eps, gam = p.acdi_gamma;
    // convective fluxes (_acdi_convective_fluxes :531-574)
    auto fx = [&](int k) __attribute__((always_inline)) -> T {
        if (k == 0) return U(0, j) * a_phase_bvalue(p.bPhi.l, phi(0, j));
        if (k == p.nx) return U(p.nx, j) * a_phase_bvalue(p.bPhi.r, phi(p.nx - 1, j));
        return U(k, j) * (0.5 * phi(k - 1, j) + 0.5 * phi(k, j));
    };
    auto fy = [&](int k) __attribute__((always_inline)) -> T {
        if (k == 0) return V(i, 0) * a_phase_bvalue(p.bPhi.b, phi(i, 0));
        if (k == p.ny) return V(i, p.ny) * a_phase_bvalue(p.bPhi.t, phi(i, p.ny - 1));
        return V(i, k) * (0.5 * phi(i, k - 1) + 0.5 * phi(i, k));
    };
    // regularisation fluxes (_acdi_regularization_fluxes :598-650), zero on walls
    auto ax = [&](int k, const T& nl, const T& nr) __attribute__((always_inline)) -> T {
        if (k == 0 || k == p.nx) return T(0.0);
        const T t = g_tanh((0.5 * PS(k - 1, j) + 0.5 * PS(k, j)) / (2.0 * eps));
        const T amp = 0.25 * (1.0 - t * t);
        return gam * (eps * a_grad_x<T>(p, p.bPhi.l, p.bPhi.r, phi, k, j) - amp * (0.5 * nl + 0.5 * nr));
    };
    auto ay = [&](int k, const T& nb, const T& nt) __attribute__((always_inline)) -> T {
        if (k == 0 || k == p.ny) return T(0.0);
        const T t = g_tanh((0.5 * PS(i, k - 1) + 0.5 * PS(i, k)) / (2.0 * eps));
        const T amp = 0.25 * (1.0 - t * t);
        return gam * (eps * a_grad_y<T>(p, p.bPhi.b, p.bPhi.t, phi, i, k) - amp * (0.5 * nb + 0.5 * nt));
    };
    const T dF = (fx(i + 1) - fx(i)) * p.ihx + (fy(j + 1) - fy(j)) * p.ihy;
    const T dA = (ax(i + 1, n0c, n0e) - ax(i, n0w, n0c)) * p.ihx + (ay(j + 1, n1c, n1n) - ay(j, n1s, n1c)) * p.ihy;
    return -dF + dA;
}

// ---- structure tensor (StructureTensor.jl:267-551, Operators.jl:431-616)
template <class T>
__device__ __forceinline__ SymT<T> a_upwind3(const T& u, const SymT<T>& lo, const SymT<T>& hi) {
    const SymT<T>& q = val(u) >= 0.0 ? lo : hi;
    return {u * q.xx, u * q.yy, u * q.xy};
}
template <class T>
__device__ __forceinline__ T a_wall_state(const BcSide& b, const T& vel, bool low_side, const T& interior) {
    const double v = val(vel);
    const bool inflow = low_side ? v > 0.0 : v < 0.0;
    if (v == 0.0 || !inflow) return interior;
    return b.kind == FSI_BC_DIRICHLET ? T(b.value) : interior;
}
template <class T>
__device__ __forceinline__ SymT<T> a_wall3(const Params& p, int side, const T& u, bool low_side, const SymT<T>& c) {
    const BcSide& bx = side == 0 ? p.bSxx.l : side == 1 ? p.bSxx.r : side == 2 ? p.bSxx.b : p.bSxx.t;
    const BcSide& by = side == 0 ? p.bSyy.l : side == 1 ? p.bSyy.r : side == 2 ? p.bSyy.b : p.bSyy.t;
    const BcSide& bz = side == 0 ? p.bSxy.l : side == 1 ? p.bSxy.r : side == 2 ? p.bSxy.b : p.bSxy.t;
    return {a_wall_state(bx, u, low_side, c.xx), a_wall_state(by, u, low_side, c.yy),
            a_wall_state(bz, u, low_side, c.xy)};
}
template <class T>
__device__ __forceinline__ T a_pow(const T& x, double e) {
    if (e == 4.0) {
        const T x2 = x * x;
        return x2 * x2;
    }
    if (e == 2.0) return x * x;
    if (e == 1.0) return x;
    return g_pow(x, e);
}
// sigma log sigma by the 2 x 2 spectral closed form (:482-515)
template <class T>
__device__ __forceinline__ bool a_sigma_log(const T& a, const T& d, const T& b, T& ox, T& oy, T& oxy) {
    const T m = 0.5 * (a + d), delta = 0.5 * (a - d);
    const T r = g_hypot(delta, b);
    const T lhi = m + r, llo = m - r;
    if (!(isfinite(val(llo)) && isfinite(val(lhi)) && val(llo) > 0.0)) return false;
    const double scale = fmax(1.0, fabs(val(m)));
    T alpha, beta;
    if (val(r) <= 1.0e-12 * scale) {
        alpha = g_log(m) + 1.0;
        beta = -m;
    } else {
        const T ghi = lhi * g_log(lhi), glo = llo * g_log(llo);
        const T den = lhi - llo;
        alpha = (ghi - glo) / den;
        beta = (lhi * glo - llo * ghi) / den;
    }
    ox = alpha * a + beta;
    oy = alpha * d + beta;
    oxy = alpha * b;
    return true;
}
// The transport part of E_sigma at cell (i, j): upwind advection and
// stretching (structure_tensor_rhs :377-463 without its pointwise terms)
template <class T, class FS, class FU, class FV>
__device__ __forceinline__ void a_sigma_transport(const Params& p, const FS& S, const FU& U, const FV& V, int i,
                                                  int j, T& oxx, T& oyy, T& oxy) {
    const SymT<T> C = S(i, j);
    SymT<T> fl, fr, fb, ft;
    {
        const T ul = U(i, j);
        fl = i > 0 ? a_upwind3(ul, S(i - 1, j), C) : a_upwind3(ul, a_wall3(p, 0, ul, true, C), C);
        const T ur = U(i + 1, j);
        fr = i < p.nx - 1 ? a_upwind3(ur, C, S(i + 1, j)) : a_upwind3(ur, C, a_wall3(p, 1, ur, false, C));
    }
    {
        const T vb = V(i, j);
        fb = j > 0 ? a_upwind3(vb, S(i, j - 1), C) : a_upwind3(vb, a_wall3(p, 2, vb, true, C), C);
        const T vt = V(i, j + 1);
        ft = j < p.ny - 1 ? a_upwind3(vt, C, S(i, j + 1)) : a_upwind3(vt, C, a_wall3(p, 3, vt, false, C));
    }
    const T axx = -((fr.xx - fl.xx) * p.ihx + (ft.xx - fb.xx) * p.ihy);
    const T ayy = -((fr.yy - fl.yy) * p.ihx + (ft.yy - fb.yy) * p.ihy);
    const T axy = -((fr.xy - fl.xy) * p.ihx + (ft.xy - fb.xy) * p.ihy);
    const T uL = U(i, j), uR = U(i + 1, j), vB = V(i, j), vT = V(i, j + 1);
    const T dv = (uR - uL) * p.ihx + (vT - vB) * p.ihy;
    const T ux = (uR - uL) * p.ihx, vy = (vT - vB) * p.ihy;
    auto Uc = [&](int jj) __attribute__((always_inline)) -> T { return 0.5 * U(i, jj) + 0.5 * U(i + 1, jj); };
    auto Vc = [&](int ii) __attribute__((always_inline)) -> T { return 0.5 * V(ii, j) + 0.5 * V(ii, j + 1); };
    T uy, vx;
    if (j == 0) uy = p.bU.b.kind == FSI_BC_DIRICHLET ? (Uc(0) - p.bU.b.value) * (2.0 * p.ihy) : T(-p.bU.b.value);
    else if (j == p.ny - 1)
        uy = p.bU.t.kind == FSI_BC_DIRICHLET ? (p.bU.t.value - Uc(p.ny - 1)) * (2.0 * p.ihy) : T(p.bU.t.value);
    else uy = (Uc(j + 1) - Uc(j - 1)) * (0.5 * p.ihy);
    if (i == 0) vx = p.bV.l.kind == FSI_BC_DIRICHLET ? (Vc(0) - p.bV.l.value) * (2.0 * p.ihx) : T(-p.bV.l.value);
    else if (i == p.nx - 1)
        vx = p.bV.r.kind == FSI_BC_DIRICHLET ? (p.bV.r.value - Vc(p.nx - 1)) * (2.0 * p.ihx) : T(p.bV.r.value);
    else vx = (Vc(i + 1) - Vc(i - 1)) * (0.5 * p.ihx);
    const T sa = C.xx, sd = C.yy, sb = C.xy;
    const T dxx = 2.0 * (ux * sa + uy * sb);
    const T dyy = 2.0 * (vx * sb + vy * sd);
    const T dxy = vx * sa + uy * sd + (ux + vy) * sb;
    oxx = (axx + sa * dv) + dxx;
    oyy = (ayy + sd * dv) + dyy;
    oxy = (axy + sb * dv) + dxy;
}
// The pointwise part of E_sigma: relaxation and the geometric extension of
// sigma at the cell, with phi at the cell
template <class T>
__device__ __forceinline__ void a_sigma_local(const Params& p, const SymT<T>& C, const T& phic, T& oxx, T& oyy,
                                              T& oxy) {
    const T sa = C.xx, sd = C.yy, sb = C.xy;
    oxx = T(0.0);
    oyy = T(0.0);
    oxy = T(0.0);
    if (p.relax) {
        oxx = -p.relax_rate * (sa - 1.0);
        oyy = -p.relax_rate * (sd - 1.0);
        oxy = -p.relax_rate * sb;
    }
    if (p.ext == FSI_EXT_GEOMETRIC) {
        const T kappa = p.ext_rate * a_pow(1.0 - a_clamp01(phic), p.ext_exp);
        if (val(kappa) != 0.0) {
            T gx, gy, gxy;
            if (a_sigma_log(sa, sd, sb, gx, gy, gxy)) {
                oxx = oxx + -kappa * gx;
                oyy = oyy + -kappa * gy;
                oxy = oxy + -kappa * gxy;
            }
        }
    }
}
// E_sigma at cell (i, j): both parts
template <class T, class FS, class FU, class FV>
__device__ __forceinline__ void a_sigma_rhs(const Params& p, const FS& S, const FU& U, const FV& V, const T& phic,
                                            int i, int j, T& oxx, T& oyy, T& oxy) {
    T txx, tyy, txy, lxx, lyy, lxy;
    a_sigma_transport<T>(p, S, U, V, i, j, txx, tyy, txy);
    a_sigma_local<T>(p, S(i, j), phic, lxx, lyy, lxy);
    oxx = txx + lxx;
    oyy = tyy + lyy;
    oxy = txy + lxy;
}

__device__ __forceinline__ double rd(const double* a, size_t k) { return a ? a[k] : 0.0; }

}  // namespace

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Tendencies of the stage by the point functions in FP64 (the same
// quantities as the forward kernels: E_u, E_v, I_u = M u + b, I_v, E_phi,
// E_sigma) -- the parity check of the functions the adjoint differentiates.
__global__ void __launch_bounds__(BLOCK) k_adj_primal(Params p, BcSet bpsi, AdjStage s, AdjGrads out) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (i >= p.nx || j >= p.ny) return;
    using T = double;
    const Acc<T> U{s.U, p.P, -1, -1, 0}, V{s.V, p.P, -1, -1, 0}, ph{s.phi, p.P, -1, -1, 0};
    const SigAcc<T> S{{s.sxx, p.P, -1, -1, 0}, {s.syy, p.P, -1, -1, 0}, {s.sxy, p.P, -1, -1, 0}, s.phi, p.P,
                      s.reset, p.ext_cutoff};
    const StressAcc<T, 0> Txx{&S, &ph, p.Gmod};
    const StressAcc<T, 1> Tyy{&S, &ph, p.Gmod};
    const StressAcc<T, 2> Txy{&S, &ph, p.Gmod};
    const size_t k = at(p, i, j);
    double eu = -a_conv_u<T>(p, U, V, i, j), ev = -a_conv_v<T>(p, U, V, i, j);
    if (p.force == FSI_FORCE_CONSTANT) {
        eu += p.ffx;
        ev += p.ffy;
    } else if (p.force == FSI_FORCE_KOLMOGOROV) {
        eu += p.force_row[j];
    }
    if (p.elastic) {
        eu += a_elastic_fx<T>(p, Txx, Txy, i, j);
        ev += a_elastic_fy<T>(p, Tyy, Txy, i, j);
    }
    out.U[k] = eu;
    out.V[k] = ev;
    out.aux_u[k] = p.visc ? a_visc_u<T>(p, U, V, i, j, 1) : 0.0;
    out.aux_v[k] = p.visc ? a_visc_v<T>(p, U, V, i, j, 1) : 0.0;
    out.phi[k] = a_acdi<T>(p, bpsi, U, V, ph, i, j);
    double oxx, oyy, oxy;
    a_sigma_rhs<T>(p, S, U, V, ph(i, j), i, j, oxx, oyy, oxy);
    out.sxx[k] = oxx;
    out.syy[k] = oyy;
    out.sxy[k] = oxy;
}

// Cells: the derivatives with respect to phi, sigma_xx, sigma_yy, sigma_xy
// at cell (i, j): the transport part of E_sigma on the five-point stencil
// (sigma only, three dual components), its pointwise part (relaxation and
// extension) at the cell (four), and the elastic part of E_u and E_v on the
// faces around the cell (four); E_phi is gathered by k_adj_phase.
__global__ void __launch_bounds__(BLOCK) k_adj_cells(Params p, AdjStage s, AdjBar b, AdjGrads out) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (i >= p.nx || j >= p.ny) return;
    double g[4] = {0.0, 0.0, 0.0, 0.0};  // phi, sigma_xx, sigma_yy, sigma_xy
    const size_t k = at(p, i, j);
    if (b.Exx) {
        using T = Dual<3>;
        const Acc<T> U{s.U, p.P, -1, -1, 0}, V{s.V, p.P, -1, -1, 0};
        const SigAcc<T> S{{s.sxx, p.P, i, j, 0}, {s.syy, p.P, i, j, 1}, {s.sxy, p.P, i, j, 2}, s.phi, p.P, s.reset,
                          p.ext_cutoff};
        const int ci[5] = {i, i - 1, i + 1, i, i}, cj[5] = {j, j, j, j - 1, j + 1};
#pragma unroll 1
        for (int q = 0; q < 5; ++q) {
            if (ci[q] < 0 || ci[q] >= p.nx || cj[q] < 0 || cj[q] >= p.ny) continue;
            const size_t kc = at(p, ci[q], cj[q]);
            const double wx = b.Exx[kc], wy = b.Eyy[kc], wz = b.Exy[kc];
            if (wx == 0.0 && wy == 0.0 && wz == 0.0) continue;
            T oxx, oyy, oxy;
            a_sigma_transport<T>(p, S, U, V, ci[q], cj[q], oxx, oyy, oxy);
#pragma unroll
            for (int c = 0; c < 3; ++c) g[1 + c] += wx * oxx.d[c] + wy * oyy.d[c] + wz * oxy.d[c];
        }
        const double wx = b.Exx[k], wy = b.Eyy[k], wz = b.Exy[k];
        if (wx != 0.0 || wy != 0.0 || wz != 0.0) {
            using T4 = Dual<4>;
            const Acc<T4> ph{s.phi, p.P, i, j, 0};
            const SigAcc<T4> S4{{s.sxx, p.P, i, j, 1}, {s.syy, p.P, i, j, 2}, {s.sxy, p.P, i, j, 3}, s.phi, p.P,
                                s.reset, p.ext_cutoff};
            T4 oxx, oyy, oxy;
            a_sigma_local<T4>(p, S4(i, j), ph(i, j), oxx, oyy, oxy);
#pragma unroll
            for (int c = 0; c < 4; ++c) g[c] += wx * oxx.d[c] + wy * oyy.d[c] + wz * oxy.d[c];
        }
    }
    // the elastic force on the faces around the cell (x-face (fi, fj) reads
    // the stress of cells fi - 1, fi (and 1 on the wall face 0) in rows
    // fj - 1 .. fj + 1 (0, 1 and ny - 2, ny - 1 on the wall rows); the y-faces
    // likewise): faces fi in [i - 1, i + 1], fj in [j - 1, j + 1]
    if (p.elastic && b.Eu) {
        using T = Dual<4>;
        const Acc<T> ph{s.phi, p.P, i, j, 0};
        const SigAcc<T> S{{s.sxx, p.P, i, j, 1}, {s.syy, p.P, i, j, 2}, {s.sxy, p.P, i, j, 3}, s.phi, p.P, s.reset,
                          p.ext_cutoff};
        const StressAcc<T, 0> Txx{&S, &ph, p.Gmod};
        const StressAcc<T, 1> Tyy{&S, &ph, p.Gmod};
        const StressAcc<T, 2> Txy{&S, &ph, p.Gmod};
#pragma unroll 1
        for (int fj = max(j - 1, 0); fj <= min(j + 1, p.ny - 1); ++fj)
#pragma unroll 1
            for (int fi = max(i - 1, 0); fi <= min(i + 1, p.nx - 1); ++fi) {
                const size_t kf = at(p, fi, fj);
                const double wu = b.Eu[kf], wv = b.Ev[kf];
                if (wu != 0.0) {
                    const T f = a_elastic_fx<T>(p, Txx, Txy, fi, fj);
#pragma unroll
                    for (int c = 0; c < 4; ++c) g[c] += wu * f.d[c];
                }
                if (wv != 0.0) {
                    const T f = a_elastic_fy<T>(p, Tyy, Txy, fi, fj);
#pragma unroll
                    for (int c = 0; c < 4; ++c) g[c] += wv * f.d[c];
                }
            }
    }
    out.phi[k] += g[0];
    out.sxx[k] += g[1];
    out.syy[k] += g[2];
    out.sxy[k] += g[3];
}

// Cells: the derivative of E_phi (ACDI) with respect to phi at p = (i, j).
// phi_p enters the convective fluxes of its four faces, the face gradients
// and the psi amplitude of its four faces, and through psi_p the normals of
// the five cells around it, hence the regularisation fluxes of those five
// cells' faces.  Flux F on a face with left (lower) cell L and right (upper)
// cell R adds -F/h to E_phi(L) and +F/h to E_phi(R) (the regularisation flux
// A the opposite: E_phi = -div(convective) + div(A)), so
//   dE^T bar = sum over those faces of (bar(R) - bar(L)) dF/h + (bar(L) - bar(R)) dA/h.
// Only the derivative along phi_p is formed (a closed form per face from
// psi on the cells within two of p and the five normals with their
// derivatives), so the sum costs one psi per cell of that diamond.
__global__ void __launch_bounds__(BLOCK) k_adj_phase(Params p, BcSet bpsi, AdjStage s, AdjBar b, AdjGrads out) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (i >= p.nx || j >= p.ny || !b.Ephi) return;
    const double* phi = s.phi;
    auto PH = [&](int m, int n) __attribute__((always_inline)) { return phi[at(p, m, n)]; };
    auto W = [&](int m, int n) __attribute__((always_inline)) -> double {  // bar E_phi, zero outside the cells
        return (m >= 0 && m < p.nx && n >= 0 && n < p.ny) ? b.Ephi[at(p, m, n)] : 0.0;
    };
    // psi on the diamond |di| + |dj| <= 2 around p and its derivative at p
    double ps[5][5];
#pragma unroll
    for (int dj = -2; dj <= 2; ++dj)
#pragma unroll
        for (int di = -2; di <= 2; ++di) {
            const int m = i + di, n = j + dj;
            const bool use = (di < 0 ? -di : di) + (dj < 0 ? -dj : dj) <= 2;
            ps[dj + 2][di + 2] = (use && m >= 0 && m < p.nx && n >= 0 && n < p.ny) ? a_psi(p, PH(m, n)) : 0.0;
        }
    const double ph = PH(i, j), e = p.acdi_floor;
    const double dpsi = (ph < 0.0 || ph > 1.0) ? 0.0 : p.acdi_eps * (1.0 / (ph + e) + 1.0 / (1.0 - ph + e));
    using D = Dual<1>;
    auto PS = [&](int m, int n) __attribute__((always_inline)) -> D {  // psi with its derivative along phi_p
        D r(ps[n - j + 2][m - i + 2]);
        if (m == i && n == j) r.d[0] = dpsi;
        return r;
    };
    // the five normals (value and derivative), index 0 centre, 1 west, 2 east, 3 south, 4 north
    D nxv[5], nyv[5];
    const int oi[5] = {0, -1, 1, 0, 0}, oj[5] = {0, 0, 0, -1, 1};
    bool has[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const int m = i + oi[q], n = j + oj[q];
        has[q] = m >= 0 && m < p.nx && n >= 0 && n < p.ny;
        nxv[q] = D(0.0);
        nyv[q] = D(0.0);
        if (!has[q]) continue;
        const D gx = 0.5 * a_grad_x<D>(p, bpsi.l, bpsi.r, PS, m, n) + 0.5 * a_grad_x<D>(p, bpsi.l, bpsi.r, PS, m + 1, n);
        const D gy = 0.5 * a_grad_y<D>(p, bpsi.b, bpsi.t, PS, m, n) + 0.5 * a_grad_y<D>(p, bpsi.b, bpsi.t, PS, m, n + 1);
        const D mag = g_sqrt(gx * gx + gy * gy + p.acdi_d2);
        nxv[q] = gx / mag;
        nyv[q] = gy / mag;
    }
    // the index (0..4) among the five of the cell at offset (di, dj) from p,
    // or -1 (a compile-time value at every call: the offsets are constants)
    auto five = [&](int di, int dj) __attribute__((always_inline)) -> int {
        return (di == 0 && dj == 0) ? 0 : (dj == 0 && di == -1) ? 1 : (dj == 0 && di == 1) ? 2
             : (di == 0 && dj == -1) ? 3 : (di == 0 && dj == 1) ? 4 : -1;
    };
    const double eps = p.acdi_eps, gam = p.acdi_gamma;
    double g = 0.0;
    // one interior face between the cells at offsets L and R from p (x-faces:
    // comp 0, y-faces: comp 1); the callers pass compile-time offsets
    auto face = [&](int ldi, int ldj, int rdi, int rdj, int comp, double ih) __attribute__((always_inline)) {
        const int lq = five(ldi, ldj), rq = five(rdi, rdj);
        const bool lin = lq >= 0 && has[lq], rin = rq >= 0 && has[rq];
        const bool adj = (ldi == 0 && ldj == 0) || (rdi == 0 && rdj == 0);
        const double psf = 0.5 * ps[ldj + 2][ldi + 2] + 0.5 * ps[rdj + 2][rdi + 2];
        const double t = tanh(psf / (2.0 * eps));
        const double amp = 0.25 * (1.0 - t * t);
        double dA = 0.0;
        if (adj) {
            const double dgrad = (rdi == 0 && rdj == 0) ? ih : -ih;  // ih (phi_R - phi_L)
            const double damp = -0.25 * t * (1.0 - t * t) / eps * (0.5 * dpsi);  // d amp / d phi_p
            const double nl = lin ? (comp == 0 ? nxv[lq].v : nyv[lq].v) : 0.0;
            const double nr = rin ? (comp == 0 ? nxv[rq].v : nyv[rq].v) : 0.0;
            dA += eps * dgrad - damp * (0.5 * nl + 0.5 * nr);
        }
        double dn = 0.0;
        if (lin) dn += 0.5 * (comp == 0 ? nxv[lq].d[0] : nyv[lq].d[0]);
        if (rin) dn += 0.5 * (comp == 0 ? nxv[rq].d[0] : nyv[rq].d[0]);
        dA -= amp * dn;
        g += gam * dA * (W(i + ldi, j + ldj) - W(i + rdi, j + rdj)) * ih;
    };
    // x-faces of the five cells: row j faces i-1 .. i+2, rows j-1 and j+1 faces i, i+1
    // (face k lies between cells k - 1 and k)
#pragma unroll
    for (int t = -1; t <= 2; ++t)
        if (i + t >= 1 && i + t <= p.nx - 1) face(t - 1, 0, t, 0, 0, p.ihx);
#pragma unroll
    for (int r = -1; r <= 1; r += 2)
#pragma unroll
        for (int t = 0; t <= 1; ++t)
            if (j + r >= 0 && j + r < p.ny && i + t >= 1 && i + t <= p.nx - 1) face(t - 1, r, t, r, 0, p.ihx);
    // y-faces: column i faces j-1 .. j+2, columns i-1 and i+1 faces j, j+1
#pragma unroll
    for (int t = -1; t <= 2; ++t)
        if (j + t >= 1 && j + t <= p.ny - 1) face(0, t - 1, 0, t, 1, p.ihy);
#pragma unroll
    for (int c = -1; c <= 1; c += 2)
#pragma unroll
        for (int t = 0; t <= 1; ++t)
            if (i + c >= 0 && i + c < p.nx && j + t >= 1 && j + t <= p.ny - 1) face(c, t - 1, c, t, 1, p.ihy);
    // convective fluxes of the cell's four faces: F = u phi_f
    {
        const double wp = W(i, j);
        // x-face i (left of p; p is its right cell) and i + 1 (right of p)
        const double uL = s.U[at(p, i, j)], uR = s.U[at(p, i + 1, j)];
        const double dl = i == 0 ? (p.bPhi.l.kind == FSI_BC_DIRICHLET ? 0.0 : 1.0) : 0.5;
        const double dr = i == p.nx - 1 ? (p.bPhi.r.kind == FSI_BC_DIRICHLET ? 0.0 : 1.0) : 0.5;
        g += (wp - W(i - 1, j)) * p.ihx * uL * dl;  // (bar(R) - bar(L)) dF / h, R = p
        g += (W(i + 1, j) - wp) * p.ihx * uR * dr;  // L = p
        const double vB = s.V[at(p, i, j)], vT = s.V[at(p, i, j + 1)];
        const double db = j == 0 ? (p.bPhi.b.kind == FSI_BC_DIRICHLET ? 0.0 : 1.0) : 0.5;
        const double dt = j == p.ny - 1 ? (p.bPhi.t.kind == FSI_BC_DIRICHLET ? 0.0 : 1.0) : 0.5;
        g += (wp - W(i, j - 1)) * p.ihy * vB * db;
        g += (W(i, j + 1) - wp) * p.ihy * vT * dt;
    }
    out.phi[at(p, i, j)] += g;
}

// x-faces: the derivative with respect to U at face (i, j), 0 <= i <= nx, of
// the convection, viscous and sigma-transport tendencies around it, and of
// the ACDI convective flux u phi_f of the face itself (closed form).
__global__ void __launch_bounds__(BLOCK) k_adj_uface(Params p, BcSet bpsi, AdjStage s, AdjBar b, AdjGrads out) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (i > p.nx || j >= p.ny) return;
    using T = Dual<1>;
    const Acc<T> U{s.U, p.P, i, j, 0}, V{s.V, p.P, -1, -1, 0};
    const SigAcc<T> S{{s.sxx, p.P, -1, -1, 0}, {s.syy, p.P, -1, -1, 0}, {s.sxy, p.P, -1, -1, 0}, s.phi, p.P,
                      s.reset, p.ext_cutoff};
    double g = 0.0;
#pragma unroll 1
    for (int fj = max(j - 1, 0); fj <= min(j + 1, p.ny - 1); ++fj)
#pragma unroll 1
        for (int fi = max(i - 1, 0); fi <= min(i + 1, p.nx - 1); ++fi) {
            const size_t kf = at(p, fi, fj);
            const double weu = rd(b.Eu, kf), wev = rd(b.Ev, kf), wiu = rd(b.Iu, kf), wiv = rd(b.Iv, kf);
            if (weu != 0.0) g -= weu * a_conv_u<T>(p, U, V, fi, fj).d[0];
            if (wev != 0.0) g -= wev * a_conv_v<T>(p, U, V, fi, fj).d[0];
            if (p.visc && wiu != 0.0) g += wiu * a_visc_u<T>(p, U, V, fi, fj, 0).d[0];
            if (p.visc && wiv != 0.0) g += wiv * a_visc_v<T>(p, U, V, fi, fj, 0).d[0];
        }
    if (b.Exx) {
#pragma unroll 1
        for (int cj = max(j - 1, 0); cj <= min(j + 1, p.ny - 1); ++cj)
#pragma unroll 1
            for (int ci = max(i - 1, 0); ci <= min(i, p.nx - 1); ++ci) {
                const size_t kc = at(p, ci, cj);
                const double wx = b.Exx[kc], wy = b.Eyy[kc], wz = b.Exy[kc];
                if (wx == 0.0 && wy == 0.0 && wz == 0.0) continue;
                T oxx, oyy, oxy;
                a_sigma_transport<T>(p, S, U, V, ci, cj, oxx, oyy, oxy);
                g += wx * oxx.d[0] + wy * oyy.d[0] + wz * oxy.d[0];
            }
    }
    if (b.Ephi && p.phase == FSI_PHASE_ACDI) {  // F = U phi_f enters E_phi(L) as -F/h, E_phi(R) as +F/h
        const double* phi = s.phi;
        double pf;
        if (i == 0) pf = p.bPhi.l.kind == FSI_BC_DIRICHLET ? p.bPhi.l.value : phi[at(p, 0, j)];
        else if (i == p.nx) pf = p.bPhi.r.kind == FSI_BC_DIRICHLET ? p.bPhi.r.value : phi[at(p, p.nx - 1, j)];
        else pf = 0.5 * phi[at(p, i - 1, j)] + 0.5 * phi[at(p, i, j)];
        const double wl = i > 0 ? b.Ephi[at(p, i - 1, j)] : 0.0, wr = i < p.nx ? b.Ephi[at(p, i, j)] : 0.0;
        g += (wr - wl) * p.ihx * pf;
    }
    out.U[at(p, i, j)] += g;
}

// y-faces: the same with respect to V at face (i, j), 0 <= j <= ny.
__global__ void __launch_bounds__(BLOCK) k_adj_vface(Params p, BcSet bpsi, AdjStage s, AdjBar b, AdjGrads out) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (i >= p.nx || j > p.ny) return;
    using T = Dual<1>;
    const Acc<T> U{s.U, p.P, -1, -1, 0}, V{s.V, p.P, i, j, 0};
    const SigAcc<T> S{{s.sxx, p.P, -1, -1, 0}, {s.syy, p.P, -1, -1, 0}, {s.sxy, p.P, -1, -1, 0}, s.phi, p.P,
                      s.reset, p.ext_cutoff};
    double g = 0.0;
#pragma unroll 1
    for (int fj = max(j - 1, 0); fj <= min(j + 1, p.ny - 1); ++fj)
#pragma unroll 1
        for (int fi = max(i - 1, 0); fi <= min(i + 1, p.nx - 1); ++fi) {
            const size_t kf = at(p, fi, fj);
            const double weu = rd(b.Eu, kf), wev = rd(b.Ev, kf), wiu = rd(b.Iu, kf), wiv = rd(b.Iv, kf);
            if (weu != 0.0) g -= weu * a_conv_u<T>(p, U, V, fi, fj).d[0];
            if (wev != 0.0) g -= wev * a_conv_v<T>(p, U, V, fi, fj).d[0];
            if (p.visc && wiu != 0.0) g += wiu * a_visc_u<T>(p, U, V, fi, fj, 0).d[0];
            if (p.visc && wiv != 0.0) g += wiv * a_visc_v<T>(p, U, V, fi, fj, 0).d[0];
        }
    if (b.Exx) {
#pragma unroll 1
        for (int cj = max(j - 1, 0); cj <= min(j, p.ny - 1); ++cj)
#pragma unroll 1
            for (int ci = max(i - 1, 0); ci <= min(i + 1, p.nx - 1); ++ci) {
                const size_t kc = at(p, ci, cj);
                const double wx = b.Exx[kc], wy = b.Eyy[kc], wz = b.Exy[kc];
                if (wx == 0.0 && wy == 0.0 && wz == 0.0) continue;
                T oxx, oyy, oxy;
                a_sigma_transport<T>(p, S, U, V, ci, cj, oxx, oyy, oxy);
                g += wx * oxx.d[0] + wy * oyy.d[0] + wz * oxy.d[0];
            }
    }
    if (b.Ephi && p.phase == FSI_PHASE_ACDI) {
        const double* phi = s.phi;
        double pf;
        if (j == 0) pf = p.bPhi.b.kind == FSI_BC_DIRICHLET ? p.bPhi.b.value : phi[at(p, i, 0)];
        else if (j == p.ny) pf = p.bPhi.t.kind == FSI_BC_DIRICHLET ? p.bPhi.t.value : phi[at(p, i, p.ny - 1)];
        else pf = 0.5 * phi[at(p, i, j - 1)] + 0.5 * phi[at(p, i, j)];
        const double wb = j > 0 ? b.Ephi[at(p, i, j - 1)] : 0.0, wt = j < p.ny ? b.Ephi[at(p, i, j)] : 0.0;
        g += (wt - wb) * p.ihy * pf;
    }
    out.V[at(p, i, j)] += g;
}

// The shear-modulus derivative: sum over the compact faces of the E_u, E_v
// adjoints times the elastic force of phi (sigma - I) (the force is linear
// in G); one partial per block, summed by k_adj_shear_final in block order.
__global__ void __launch_bounds__(BLOCK) k_adj_shear(Params p, AdjStage s, AdjBar b, double* partials) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    double acc[1] = {0.0};
    if (i < p.nx && j < p.ny && b.Eu) {
        using T = double;
        const Acc<T> ph{s.phi, p.P, -1, -1, 0};
        const SigAcc<T> S{{s.sxx, p.P, -1, -1, 0}, {s.syy, p.P, -1, -1, 0}, {s.sxy, p.P, -1, -1, 0}, s.phi, p.P,
                          s.reset, p.ext_cutoff};
        const StressAcc<T, 0> Txx{&S, &ph, 1.0};
        const StressAcc<T, 1> Tyy{&S, &ph, 1.0};
        const StressAcc<T, 2> Txy{&S, &ph, 1.0};
        const size_t k = at(p, i, j);
        if (b.Eu[k] != 0.0) acc[0] += b.Eu[k] * a_elastic_fx<T>(p, Txx, Txy, i, j);
        if (b.Ev[k] != 0.0) acc[0] += b.Ev[k] * a_elastic_fy<T>(p, Tyy, Txy, i, j);
    }
    __shared__ double lds[16];
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) partials[blockIdx.x + blockIdx.y * gridDim.x] = acc[0];
}
__global__ void __launch_bounds__(BLOCK) k_adj_shear_final(const double* partials, int nb, double* dG) {
    double t[1];
    get_totals<1>(partials, nb, t);
    if (threadIdx.x == 0 && threadIdx.y == 0) *dG += t[0];
}

// zero n values of each of the arrays (one array per block row)
__global__ void __launch_bounds__(256) k_adj_zero(size_t n, AdjPtrs z) {
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n && blockIdx.y < (unsigned)z.n) z.p[blockIdx.y][k] = 0.0;
}
// dst[d] += coef[d] * src[from[d]] over n values, the sources read once; a
// source with `masked` set counts only where phi >= cutoff (the safety
// reset's constants)
__global__ void __launch_bounds__(256) k_adj_combine(size_t n, AdjCombine c) {
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    double v[4];
    const bool keep = !c.phi || !(c.phi[k] < c.cutoff);
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = q < c.nsrc ? ((c.masked & (1 << q)) && !keep ? 0.0 : c.src[q][k]) : 0.0;
    for (int d = 0; d < c.ndst; ++d) c.dst[d][k] += c.coef[d] * v[c.from[d]];
}
// y += a x over n values
__global__ void __launch_bounds__(256) k_adj_axpy(size_t n, double* __restrict__ y, double a,
                                                  const double* __restrict__ x) {
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n) y[k] += a * x[k];
}
// sigma adjoints of the cells the safety reset made constants (phi < cutoff)
__global__ void __launch_bounds__(BLOCK) k_adj_mask(Params p, const double* phi, double* gxx, double* gyy,
                                                    double* gxy) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (i >= p.nx || j >= p.ny) return;
    const size_t k = at(p, i, j);
    if (phi[k] < p.ext_cutoff) {
        gxx[k] = 0.0;
        gyy[k] = 0.0;
        gxy[k] = 0.0;
    }
}
// adjoints of the wall faces (the velocity data of the walls: constants)
__global__ void __launch_bounds__(BLOCK) k_adj_walls(Params p, double* U, double* V) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (j < p.ny && (i == 0 || i == p.nx)) U[at(p, i, j)] = 0.0;
    if (i < p.nx && (j == 0 || j == p.ny)) V[at(p, i, j)] = 0.0;
}
// lambda_phi = (A x_i + B y_j + C) at the cells, zero elsewhere
__global__ void __launch_bounds__(BLOCK) k_adj_affine(Params p, double* phi, double A, double B, double C) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (i >= p.nx || j >= p.ny) return;
    const double x = p.hx * ((double)i + 0.5), y = p.hy * ((double)j + 0.5);
    phi[at(p, i, j)] = A * x + B * y + C;
}

static dim3 pgrid(const Params& p) { return dim3((p.nx + 1 + TILE_X - 1) / TILE_X, (p.ny + 1 + TILE_Y - 1) / TILE_Y); }
static const dim3 kAdjBlock(TILE_X, TILE_Y);

int adj_block_count(const Params& p) {
    const dim3 g = pgrid(p);
    return (int)(g.x * g.y);
}
void launch_adj_primal(hipStream_t st, const Params& p, const BcSet& bpsi, const AdjStage& s, const AdjGrads& out) {
    k_adj_primal<<<pgrid(p), kAdjBlock, 0, st>>>(p, bpsi, s, out);
}
void launch_adj_tendencies(hipStream_t st, const Params& p, const BcSet& bpsi, const AdjStage& s, const AdjBar& b,
                           const AdjGrads& out, double* partials, double* dG) {
    const dim3 g = pgrid(p);
    k_adj_cells<<<g, kAdjBlock, 0, st>>>(p, s, b, out);
    if (p.phase == FSI_PHASE_ACDI) k_adj_phase<<<g, kAdjBlock, 0, st>>>(p, bpsi, s, b, out);
    k_adj_uface<<<g, kAdjBlock, 0, st>>>(p, bpsi, s, b, out);
    k_adj_vface<<<g, kAdjBlock, 0, st>>>(p, bpsi, s, b, out);
    if (p.elastic && dG) {
        k_adj_shear<<<g, kAdjBlock, 0, st>>>(p, s, b, partials);
        k_adj_shear_final<<<1, kAdjBlock, 0, st>>>(partials, (int)(g.x * g.y), dG);
    }
}
void launch_adj_zero(hipStream_t st, size_t n, const AdjPtrs& z) {
    if (z.n > 0) k_adj_zero<<<dim3((unsigned)((n + 255) / 256), (unsigned)z.n), 256, 0, st>>>(n, z);
}
void launch_adj_combine(hipStream_t st, size_t n, const AdjCombine& c) {
    if (c.ndst > 0) k_adj_combine<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, c);
}
void launch_adj_axpy(hipStream_t st, size_t n, double* y, double a, const double* x) {
    if (a == 0.0) return;
    k_adj_axpy<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, y, a, x);
}
void launch_adj_mask(hipStream_t st, const Params& p, const double* phi, double* gxx, double* gyy, double* gxy) {
    k_adj_mask<<<pgrid(p), kAdjBlock, 0, st>>>(p, phi, gxx, gyy, gxy);
}
void launch_adj_walls(hipStream_t st, const Params& p, double* U, double* V) {
    k_adj_walls<<<pgrid(p), kAdjBlock, 0, st>>>(p, U, V);
}
void launch_adj_affine(hipStream_t st, const Params& p, double* phi, double A, double B, double C) {
    k_adj_affine<<<pgrid(p), kAdjBlock, 0, st>>>(p, phi, A, B, C);
}

}  // namespace fsi
