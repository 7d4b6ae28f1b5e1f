// fsi_physics.hip -- explicit stage tendencies and stage combinations of the
// four-stage IMEX step on gfx950.
//
// Every kernel is an HBM-bound stencil: one thread per output point, 64
// threads along x (one wavefront reads one 512-byte row segment), 4 rows per
// thread.  Neighbour values come from the same or adjacent cache lines and
// are served by L1/L2.  No MFMA: nothing here is a dense contraction.
#include "fsi_kernels.hpp"
#include "fsi_stencils.hpp"

namespace fsi {

// ---------------------------------------------------------------------------
// Small device helpers shared by the stencils.
// ---------------------------------------------------------------------------
struct Stage {
    const double* U;
    const double* V;
    const double* phi;
    const double* sxx;
    const double* syy;
    const double* sxy;
    int reset;  // sigma reads apply the safety reset (phi < cutoff -> identity)
};

__device__ __forceinline__ double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

__device__ __forceinline__ void load_sigma(const Params& p, const Stage& s, int i, int j,
                                           double& a, double& d, double& b) {
    size_t k = at(p, i, j);
    a = s.sxx[k];
    d = s.syy[k];
    b = s.sxy[k];
    if (s.reset && s.phi[k] < p.ext_cutoff) {  // apply_structure_extension! StructureTensor.jl:590-613
        a = 1.0;
        d = 1.0;
        b = 0.0;
    }
}

// ---- LDS tiles of the momentum kernel ------------------------------------
// A halo-1 tile around the 64 x 16 faces/cells of a block, read by global
// (wrapped) index: tile position (li, lj) holds global (i0 + li - 1,
// j0 + lj - 1), wrapped across a periodic side.
constexpr int MT_W = TILE_X + 2, MT_H = TILE_ROWS + 2;  // 66 x 18
// IN: the tile and its halo lie inside the domain (no wrap).
template <bool IN = false>
struct HaloTile {
    const double* a;
    int i0, j0, nx, ny;
    __device__ __forceinline__ double operator()(int gi, int gj) const {
        int li = gi - i0 + 1, lj = gj - j0 + 1;
        if (!IN) {
            if (li < 0) li += nx;
            else if (li >= MT_W) li -= nx;
            if (lj < 0) lj += ny;
            else if (lj >= MT_H) lj -= ny;
        }
        return a[lj * MT_W + li];
    }
};

// ---- material convection pieces (MomentumTransport.jl:141-219) -----------
// IN = true: every index lies at least one cell inside the walls, so the wall
// and wrap branches compile out (same arithmetic on interior points).
template <bool IN = false, class FU>
__device__ __forceinline__ double u2c(const Params& p, const FU& U, int k, int j) {
    double ur = (!IN && p.px && k == p.nx - 1) ? U(0, j) : U(k + 1, j);
    double uc = 0.5 * U(k, j) + 0.5 * ur;
    return uc * uc;
}
template <bool IN = false, class FV>
__device__ __forceinline__ double v2c(const Params& p, const FV& V, int i, int k) {
    double vt = (!IN && p.py && k == p.ny - 1) ? V(i, 0) : V(i, k + 1);
    double vc = 0.5 * V(i, k) + 0.5 * vt;
    return vc * vc;
}
template <bool IN = false, class FU>
__device__ __forceinline__ double u_corner(const Params& p, const FU& U, int i, int jc) {
    if (IN) return 0.5 * U(i, jc - 1) + 0.5 * U(i, jc);
    int ii = (p.px && i == p.nx) ? 0 : i;
    if (p.py) {
        if (jc == 0 || jc == p.ny) return 0.5 * U(ii, p.ny - 1) + 0.5 * U(ii, 0);
        return 0.5 * U(ii, jc - 1) + 0.5 * U(ii, jc);
    }
    if (jc == 0) {
        if (p.bU.b.kind == FSI_BC_DIRICHLET) return p.bU.b.value;
        return U(ii, 0) + p.bU.b.value * (p.hy / 2.0);
    }
    if (jc == p.ny) {
        if (p.bU.t.kind == FSI_BC_DIRICHLET) return p.bU.t.value;
        return U(ii, p.ny - 1) + p.bU.t.value * (p.hy / 2.0);
    }
    return 0.5 * U(ii, jc - 1) + 0.5 * U(ii, jc);
}
template <bool IN = false, class FV>
__device__ __forceinline__ double v_corner(const Params& p, const FV& V, int i, int jc) {
    if (IN) return 0.5 * V(i - 1, jc) + 0.5 * V(i, jc);
    int jj = (p.py && jc == p.ny) ? 0 : jc;
    if (p.px) {
        if (i == 0 || i == p.nx) return 0.5 * V(p.nx - 1, jj) + 0.5 * V(0, jj);
        return 0.5 * V(i - 1, jj) + 0.5 * V(i, jj);
    }
    if (i == 0) {
        if (p.bV.l.kind == FSI_BC_DIRICHLET) return p.bV.l.value;
        return V(0, jj) + p.bV.l.value * (p.hx / 2.0);
    }
    if (i == p.nx) {
        if (p.bV.r.kind == FSI_BC_DIRICHLET) return p.bV.r.value;
        return V(p.nx - 1, jj) + p.bV.r.value * (p.hx / 2.0);
    }
    return 0.5 * V(i - 1, jj) + 0.5 * V(i, jj);
}
template <bool IN = false, class FU, class FV>
__device__ __forceinline__ double uv_corner(const Params& p, const FU& U, const FV& V, int i, int jc) {
    return u_corner<IN>(p, U, i, jc) * v_corner<IN>(p, V, i, jc);
}

// ---- linear wall extrapolation (Operators.jl:738-842) ---------------------
template <class F>
__device__ __forceinline__ double lex_lo(F f) {  // value at the low wall (half a cell out) from f(0), f(1)
    return f(0) - (f(1) - f(0)) * 0.5;
}

// ---------------------------------------------------------------------------
// Elastic force: symmetric_stress_divergence (Operators.jl:620-735) of
// T = phi G (sigma - I) (StructuralConstitutive.jl:100-113,
// TimeIntegrationIMEX.jl:911-913), from cell-centred accessors Txx, Tyy, Txy.
// ---------------------------------------------------------------------------
template <bool IN = false, class FT>
__device__ __forceinline__ double elastic_fx(const Params& p, const FT& Txx, const FT& Txy, int i, int j) {
    // x-part at x-face i (Operators.jl:642-683)
    if (IN) {
        double fx = (Txx(i, j) - Txx(i - 1, j)) * p.ihx;
        const double tp = 0.5 * Txy(i - 1, j + 1) + 0.5 * Txy(i, j + 1);
        const double tm = 0.5 * Txy(i - 1, j - 1) + 0.5 * Txy(i, j - 1);
        return fx + (tp - tm) * (0.5 * p.ihy);
    }
    const int nx = p.nx, ny = p.ny;
    double fx;
    if (p.px && i == 0) fx = (Txx(0, j) - Txx(nx - 1, j)) * p.ihx;
    else if (i == 0) {
        auto f = [&](int k) { return Txx(k, j); };
        double lw = lex_lo(f);
        fx = (Txx(0, j) - lw) * (2.0 * p.ihx);
    } else fx = (Txx(i, j) - Txx(i - 1, j)) * p.ihx;
    // Txy on this x-face column, any row jj
    auto txy_x = [&](int jj) -> double {
        if (p.px && i == 0) return 0.5 * Txy(nx - 1, jj) + 0.5 * Txy(0, jj);
        if (i == 0) {
            auto f = [&](int k) { return Txy(k, jj); };
            return lex_lo(f);
        }
        return 0.5 * Txy(i - 1, jj) + 0.5 * Txy(i, jj);
    };
    if (p.py) {
        int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
        fx += (txy_x(jp) - txy_x(jm)) * (0.5 * p.ihy);
    } else if (j == 0) {
        double t0 = txy_x(0), t1 = txy_x(1);
        double bw = t0 - (t1 - t0) * 0.5;
        fx += (t0 - bw) * (2.0 * p.ihy);
    } else if (j == ny - 1) {
        double t0 = txy_x(ny - 1), t1 = txy_x(ny - 2);
        double tw = t0 + (t0 - t1) * 0.5;
        fx += (tw - t0) * (2.0 * p.ihy);
    } else {
        fx += (txy_x(j + 1) - txy_x(j - 1)) * (0.5 * p.ihy);
    }
    return fx;
}

template <bool IN = false, class FT>
__device__ __forceinline__ double elastic_fy(const Params& p, const FT& Tyy, const FT& Txy, int i, int j) {
    if (IN) {
        double fy = (Tyy(i, j) - Tyy(i, j - 1)) * p.ihy;
        const double tp = 0.5 * Txy(i + 1, j - 1) + 0.5 * Txy(i + 1, j);
        const double tm = 0.5 * Txy(i - 1, j - 1) + 0.5 * Txy(i - 1, j);
        return fy + (tp - tm) * (0.5 * p.ihx);
    }
    const int nx = p.nx, ny = p.ny;
    double fy;
    if (p.py && j == 0) fy = (Tyy(i, 0) - Tyy(i, ny - 1)) * p.ihy;
    else if (j == 0) {
        auto f = [&](int k) { return Tyy(i, k); };
        double bw = lex_lo(f);
        fy = (Tyy(i, 0) - bw) * (2.0 * p.ihy);
    } else fy = (Tyy(i, j) - Tyy(i, j - 1)) * p.ihy;
    auto txy_y = [&](int ii) -> double {
        if (p.py && j == 0) return 0.5 * Txy(ii, ny - 1) + 0.5 * Txy(ii, 0);
        if (j == 0) {
            auto f = [&](int k) { return Txy(ii, k); };
            return lex_lo(f);
        }
        return 0.5 * Txy(ii, j - 1) + 0.5 * Txy(ii, j);
    };
    if (p.px) {
        int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
        fy += (txy_y(ip) - txy_y(im)) * (0.5 * p.ihx);
    } else if (i == 0) {
        double t0 = txy_y(0), t1 = txy_y(1);
        double lw = t0 - (t1 - t0) * 0.5;
        fy += (t0 - lw) * (2.0 * p.ihx);
    } else if (i == nx - 1) {
        double t0 = txy_y(nx - 1), t1 = txy_y(nx - 2);
        double rw = t0 + (t0 - t1) * 0.5;
        fy += (rw - t0) * (2.0 * p.ihx);
    } else {
        fy += (txy_y(i + 1) - txy_y(i - 1)) * (0.5 * p.ihx);
    }
    return fy;
}

// ---------------------------------------------------------------------------
// k_momentum: E_u, E_v (explicit momentum, TimeIntegrationIMEX.jl:962-988) and
// I_u, I_v (viscous operator, :1057-1061) on the compact faces.  U, V and
// the three components of phi G (sigma - I) (safety reset applied) sit on
// halo-1 LDS tiles, each value formed once; the face formulas read them
// from there.
// ---------------------------------------------------------------------------
template <bool IN>
__device__ __forceinline__ void momentum_tile(const Params& p, const Stage& s, double* __restrict__ Eu,
                                              double* __restrict__ Ev, double* __restrict__ Iu,
                                              double* __restrict__ Iv, const RhsSums& R, int i0, int j0,
                                              double* su, double* sv, double* sxx, double* syy, double* sxy) {
    const int tid = threadIdx.x + threadIdx.y * TILE_X;
    const bool elastic = p.elastic != 0;
    constexpr int NPT = MT_H * MT_W, NIT = (NPT + BLOCK - 1) / BLOCK;
    // indices of every tile point for U (faces 0..nx), V (rows 0..ny) and
    // the cells; then all loads back to back
    size_t ku[NIT], kv[NIT], kc[NIT];
    bool okc[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int k = min(tid + it * BLOCK, NPT - 1);
        const int ui = i0 + k % MT_W - 1, uj = j0 + k / MT_W - 1;
        auto wrap = [](int u, int n) { return u < 0 ? u + n : (u >= n ? u - n : u); };
        // x index for faces (U) and cells; y index for rows (V) and cells
        const int xf = p.px ? wrap(ui, p.nx) : ui, xc = p.px ? wrap(ui, p.nx) : ui;
        const int yr = p.py ? wrap(uj, p.ny) : uj, yc = p.py ? wrap(uj, p.ny) : uj;
        const bool xu_ok = xf >= 0 && xf <= (p.px ? p.nx - 1 : p.nx);
        const bool yv_ok = yr >= 0 && yr <= (p.py ? p.ny - 1 : p.ny);
        const bool xc_ok = xc >= 0 && xc < p.nx, yc_ok = yc >= 0 && yc < p.ny;
        ku[it] = (xu_ok && yc_ok) ? at(p, xf, yc) : 0;
        kv[it] = (xc_ok && yv_ok) ? at(p, xc, yr) : 0;
        okc[it] = xc_ok && yc_ok;
        kc[it] = okc[it] ? at(p, xc, yc) : 0;
    }
    double vu[NIT], vvv[NIT], vph[NIT], va[NIT], vd[NIT], vb[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        vu[it] = s.U[ku[it]];
        vvv[it] = s.V[kv[it]];
        if (elastic) {
            vph[it] = s.phi[kc[it]];
            va[it] = s.sxx[kc[it]];
            vd[it] = s.syy[kc[it]];
            vb[it] = s.sxy[kc[it]];
        }
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int k = tid + it * BLOCK;
        if (k >= NPT) break;
        su[k] = vu[it];
        sv[k] = vvv[it];
        if (elastic) {
            double a = va[it], d = vd[it], b = vb[it];
            if (s.reset && vph[it] < p.ext_cutoff) {  // apply_structure_extension! StructureTensor.jl:590-613
                a = 1.0;
                d = 1.0;
                b = 0.0;
            }
            const double ph = okc[it] ? vph[it] : 0.0;
            sxx[k] = ph * (p.Gmod * (a - 1.0));
            syy[k] = ph * (p.Gmod * (d - 1.0));
            sxy[k] = ph * (p.Gmod * b);
        }
    }
    __syncthreads();
    using HT = HaloTile<IN>;
    const HT U{su, i0, j0, p.nx, p.ny}, V{sv, i0, j0, p.nx, p.ny};
    const HT Txx{sxx, i0, j0, p.nx, p.ny}, Tyy{syy, i0, j0, p.nx, p.ny}, Txy{sxy, i0, j0, p.nx, p.ny};
    const int i = i0 + threadIdx.x;
    if (i >= p.nx) return;
#pragma unroll 1
    for (int r = 0; r < ROWS; ++r) {
        const int j = j0 + threadIdx.y + r * TILE_Y;
        if (j >= p.ny) return;
        const size_t k = at(p, i, j);
        // the running sums this row updates (stages q+1.., at most two read
        // back for q >= 1) and the lifting b_visc, loaded before the face
        // formulas so that their latency overlaps them
        double su_old[2] = {0.0, 0.0}, sv_old[2] = {0.0, 0.0}, lbu = 0.0, lbv = 0.0;
        if (R.q >= 1) {
#pragma unroll
            for (int t = 0; t < 2; ++t)
                if (R.q + t < 3) {
                    su_old[t] = R.ru[R.q + t][k];
                    sv_old[t] = R.rv[R.q + t][k];
                }
        }
        if (R.q >= 0 && R.bu) {
            lbu = R.bu[k];
            lbv = R.bv[k];
        }
        // ---- x-momentum at x-face (i, j)
        double cu;
        if (!IN && !p.px && i == 0) cu = 0.0;  // wall-normal row (MomentumTransport.jl:322-327)
        else {
            int im = (!IN && i == 0) ? p.nx - 1 : i - 1;
            cu = (u2c<IN>(p, U, i, j) - u2c<IN>(p, U, im, j)) * p.ihx +
                 (uv_corner<IN>(p, U, V, i, j + 1) - uv_corner<IN>(p, U, V, i, j)) * p.ihy;
        }
        double eu = -cu;
        if (p.force == FSI_FORCE_CONSTANT) eu += p.ffx;
        else if (p.force == FSI_FORCE_KOLMOGOROV) eu += p.force_row[j];
        if (p.cap_fx) eu += p.cap_fx[k];  // capillary_force, :980-982
        if (elastic) eu += elastic_fx<IN>(p, Txx, Txy, i, j);
        // ---- y-momentum at y-face (i, j)
        double cv;
        if (!IN && !p.py && j == 0) cv = 0.0;
        else {
            int jm = (!IN && j == 0) ? p.ny - 1 : j - 1;
            cv = (uv_corner<IN>(p, U, V, i + 1, j) - uv_corner<IN>(p, U, V, i, j)) * p.ihx +
                 (v2c<IN>(p, V, i, j) - v2c<IN>(p, V, i, jm)) * p.ihy;
        }
        double ev = -cv;
        if (p.force == FSI_FORCE_CONSTANT) ev += p.ffy;
        if (p.cap_fy) ev += p.cap_fy[k];
        if (elastic) ev += elastic_fy<IN>(p, Tyy, Txy, i, j);
        if (Eu) {
            Eu[k] = eu;
            Ev[k] = ev;
        }
        double iu = 0.0, iv = 0.0;
        if (Iu || R.q >= 0) {
            iu = p.visc ? visc_u<IN>(p, U, V, i, j, 1) : 0.0;
            iv = p.visc ? visc_v<IN>(p, U, V, i, j, 1) : 0.0;
        }
        if (Iu) {
            Iu[k] = iu;
            Iv[k] = iv;
        }
        if (R.q == 0 && R.mu) {  // M u^n, for the reference residual of the predictor solves
            R.mu[k] = iu - lbu;
            R.mv[k] = iv - lbv;
        }
        if (R.q >= 0) {  // the right-hand sides of the later stages
            for (int t = R.q; t < 3; ++t) {
                double ru = R.q == 0 ? U(i, j) : su_old[(t - R.q) & 1];
                double rv = R.q == 0 ? V(i, j) : sv_old[(t - R.q) & 1];
                if (R.ce[t] != 0.0 || R.ci[t] != 0.0) {
                    ru += R.dt * (R.ce[t] * eu + R.ci[t] * iu);
                    rv += R.dt * (R.ce[t] * ev + R.ci[t] * iv);
                }
                if (t == R.q && R.bu) {
                    ru += R.alpha * lbu;
                    rv += R.alpha * lbv;
                }
                R.ru[t][k] = ru;
                R.rv[t][k] = rv;
            }
        }
    }
}

// (47.5 KB of LDS: at most 3 waves per SIMD, so registers up to that occupancy)
FSI_KERNEL void k_momentum(Params p, Stage s, double* __restrict__ Eu, double* __restrict__ Ev, double* __restrict__ Iu,
           double* __restrict__ Iv, RhsSums R) {
    __shared__ double su[MT_H * MT_W], sv[MT_H * MT_W];
    __shared__ double sxx[MT_H * MT_W], syy[MT_H * MT_W], sxy[MT_H * MT_W];
    const TileId T = tile_id();
    const int i0 = T.x * TILE_X, j0 = T.y * TILE_ROWS;
    // every face of the tile at least one cell inside the walls: no wall
    // row, no wrap, no boundary lifting
    const bool inner = i0 >= 1 && i0 + TILE_X - 1 <= p.nx - 2 && j0 >= 1 && j0 + TILE_ROWS - 1 <= p.ny - 2;
    if (inner) momentum_tile<true>(p, s, Eu, Ev, Iu, Iv, R, i0, j0, su, sv, sxx, syy, sxy);
    else momentum_tile<false>(p, s, Eu, Ev, Iu, Iv, R, i0, j0, su, sv, sxx, syy, sxy);
}

// ---------------------------------------------------------------------------
// Phase-field and structure-tensor right-hand sides (cell centred).
// ---------------------------------------------------------------------------

// scalar_gradient on x-face k (Operators.jl:189-356) of the cell function f.
// IN = true: 1 <= k <= nx-1, an interior face (no wall row, no wrap).
// Both products are rounded before the difference, as the reference's
// sparse product forms them: a fused multiply-add would turn the exact zero
// of two equal neighbours into the product's rounding error, which the
// regularised normals g / sqrt(g^2 + delta^2) (delta down to 1e-14) amplify
// to O(1) in flat regions of phi.
template <bool IN = false, class F>
__device__ __forceinline__ double grad_xface(const Params& p, const BcSet& bc, F f, int k, int j) {
#pragma clang fp contract(off)
    if (IN) return p.ihx * f(k, j) - p.ihx * f(k - 1, j);
    if (p.px) {
        if (k == 0 || k == p.nx) return p.ihx * f(0, j) - p.ihx * f(p.nx - 1, j);
    } else if (k == 0) {
        if (bc.l.kind == FSI_BC_DIRICHLET) {
            double dist = p.hx / 2.0;
            return (1.0 / dist) * f(0, j) + (-bc.l.value / dist);
        }
        return -bc.l.value;
    } else if (k == p.nx) {
        if (bc.r.kind == FSI_BC_DIRICHLET) {
            double dist = p.hx / 2.0;
            return (-1.0 / dist) * f(p.nx - 1, j) + bc.r.value / dist;
        }
        return bc.r.value;
    }
    return p.ihx * f(k, j) - p.ihx * f(k - 1, j);
}
template <bool IN = false, class F>
__device__ __forceinline__ double grad_yface(const Params& p, const BcSet& bc, F f, int i, int k) {
#pragma clang fp contract(off)
    if (IN) return p.ihy * f(i, k) - p.ihy * f(i, k - 1);
    if (p.py) {
        if (k == 0 || k == p.ny) return p.ihy * f(i, 0) - p.ihy * f(i, p.ny - 1);
    } else if (k == 0) {
        if (bc.b.kind == FSI_BC_DIRICHLET) {
            double dist = p.hy / 2.0;
            return (1.0 / dist) * f(i, 0) + (-bc.b.value / dist);
        }
        return -bc.b.value;
    } else if (k == p.ny) {
        if (bc.t.kind == FSI_BC_DIRICHLET) {
            double dist = p.hy / 2.0;
            return (-1.0 / dist) * f(i, p.ny - 1) + bc.t.value / dist;
        }
        return bc.t.value;
    }
    return p.ihy * f(i, k) - p.ihy * f(i, k - 1);
}
// centre gradient = face gradients averaged (PhaseFieldMethods.jl:219-231)
template <bool IN = false, class F>
__device__ __forceinline__ void grad_center(const Params& p, const BcSet& bc, F f, int i, int j,
                                            double& gx, double& gy) {
    gx = 0.5 * grad_xface<IN>(p, bc, f, i, j) + 0.5 * grad_xface<IN>(p, bc, f, i + 1, j);
    gy = 0.5 * grad_yface<IN>(p, bc, f, i, j) + 0.5 * grad_yface<IN>(p, bc, f, i, j + 1);
}
// interpolate_center_to_x_faces (Operators.jl:110-134) of f at face k
template <class F>
__device__ __forceinline__ double c2fx(const Params& p, F f, int k, int j) {
    if (p.px && (k == 0 || k == p.nx)) return 0.5 * f(p.nx - 1, j) + 0.5 * f(0, j);
    if (k == 0) return f(0, j);
    if (k == p.nx) return f(p.nx - 1, j);
    return 0.5 * f(k - 1, j) + 0.5 * f(k, j);
}
template <class F>
__device__ __forceinline__ double c2fy(const Params& p, F f, int i, int k) {
    if (p.py && (k == 0 || k == p.ny)) return 0.5 * f(i, p.ny - 1) + 0.5 * f(i, 0);
    if (k == 0) return f(i, 0);
    if (k == p.ny) return f(i, p.ny - 1);
    return 0.5 * f(i, k - 1) + 0.5 * f(i, k);
}

// Incoming state at a wall face for first-order upwinding
// (InterfaceTransport.jl:112-159, StructureTensor.jl:195-265).
__device__ __forceinline__ double upwind_state(const BcSide& b, double vel, bool low_side, double interior,
                                               ErrSlot* st) {
    bool inflow = low_side ? vel > 0.0 : vel < 0.0;
    if (vel == 0.0 || !inflow) return interior;
    if (b.kind == FSI_BC_DIRICHLET) return b.value;
    raise_error(st, DEV_NEUMANN_INFLOW, vel, 0.0);
    return interior;
}

// -div(upwind v f) at cell (i,j) (InterfaceTransport.jl:161-258).
template <class F>
__device__ __forceinline__ double upwind_rhs(const Params& p, const double* U, const double* V, const BcSet& bc, F f,
                             int i, int j, ErrSlot* st) {
    auto ax = [&](int k) -> double {
        if (p.px) {
            int kk = (k == p.nx) ? 0 : k;
            double u = U[at(p, kk, j)];
            double L = f(kk == 0 ? p.nx - 1 : kk - 1, j), R = f(kk, j);
            return u * (u >= 0.0 ? L : R);
        }
        if (k == 0) {
            double u = U[at(p, 0, j)];
            double bs = upwind_state(bc.l, u, true, f(0, j), st);
            return u * (u >= 0.0 ? bs : f(0, j));
        }
        if (k == p.nx) {
            double u = U[at(p, p.nx, j)];
            double bs = upwind_state(bc.r, u, false, f(p.nx - 1, j), st);
            return u * (u >= 0.0 ? f(p.nx - 1, j) : bs);
        }
        double u = U[at(p, k, j)];
        return u * (u >= 0.0 ? f(k - 1, j) : f(k, j));
    };
    auto ay = [&](int k) -> double {
        if (p.py) {
            int kk = (k == p.ny) ? 0 : k;
            double v = V[at(p, i, kk)];
            double B = f(i, kk == 0 ? p.ny - 1 : kk - 1), T = f(i, kk);
            return v * (v >= 0.0 ? B : T);
        }
        if (k == 0) {
            double v = V[at(p, i, 0)];
            double bs = upwind_state(bc.b, v, true, f(i, 0), st);
            return v * (v >= 0.0 ? bs : f(i, 0));
        }
        if (k == p.ny) {
            double v = V[at(p, i, p.ny)];
            double bs = upwind_state(bc.t, v, false, f(i, p.ny - 1), st);
            return v * (v >= 0.0 ? f(i, p.ny - 1) : bs);
        }
        double v = V[at(p, i, k)];
        return v * (v >= 0.0 ? f(i, k - 1) : f(i, k));
    };
    return -((ax(i + 1) - ax(i)) * p.ihx + (ay(j + 1) - ay(j)) * p.ihy);
}

// _phase_boundary_value (PhaseFieldMethods.jl:174-196)
__device__ __forceinline__ double phase_bvalue(const BcSide& b, double interior, ErrSlot* st) {
    if (b.kind == FSI_BC_DIRICHLET) return b.value;
    if (fabs(b.value) > 100 * 2.220446049250313e-16) raise_error(st, DEV_NONHOMOG_NEUMANN, b.value, 0.0);
    return interior;
}

// ---- ACDI psi (PhaseFieldMethods.jl:479-483)
__device__ __forceinline__ double psi_value(const Params& p, double ph) {
    double q = clamp01(ph), e = p.acdi_floor;
    return p.acdi_eps * log((q + e) / (1.0 - q + e));
}

// ---- structure tensor (StructureTensor.jl:377-551, Operators.jl:431-586)
// x^e; small non-negative integer exponents (the geometric extension's
// default is 4) by repeated squaring, others by pow
__device__ __forceinline__ double pow_fast(double x, double e) {
    if (e == 4.0) {
        const double x2 = x * x;
        return x2 * x2;
    }
    if (e == 2.0) return x * x;
    if (e == 1.0) return x;
    return pow(x, e);
}

__device__ __forceinline__ bool sigma_log_product(double a, double d, double b, double& ox, double& oy,
                                                  double& oxy) {
    double m = 0.5 * (a + d), delta = 0.5 * (a - d);
    double r = hypot(delta, b);
    double lhi = m + r, llo = m - r;
    if (!(isfinite(llo) && isfinite(lhi) && llo > 0.0)) return false;
    double scale = fmax(1.0, fabs(m)), alpha, beta;
    if (r <= 1.0e-12 * scale) {  // Hermite limit: lhi log lhi, llo log llo are not needed
        alpha = log(m) + 1.0;
        beta = -m;
    } else {
        double ghi = lhi * log(lhi), glo = llo * log(llo);
        double den = lhi - llo;
        alpha = (ghi - glo) / den;
        beta = (lhi * glo - llo * ghi) / den;
    }
    ox = alpha * a + beta;
    oy = alpha * d + beta;
    oxy = alpha * b;
    return true;
}

struct Sym3 {
    double xx, yy, xy;
};
__device__ __forceinline__ Sym3 load_sym3(const Params& p, const Stage& s, int i, int j) {
    Sym3 v;
    load_sigma(p, s, i, j, v.xx, v.yy, v.xy);
    return v;
}
// upwind flux u * state at a face for the three components
__device__ __forceinline__ Sym3 upwind_flux3(double u, const Sym3& lo, const Sym3& hi) {
    const Sym3& q = u >= 0.0 ? lo : hi;
    return {u * q.xx, u * q.yy, u * q.xy};
}
__device__ __forceinline__ const BcSide& side_of(const BcSet& b, int side) {
    return side == 0 ? b.l : side == 1 ? b.r : side == 2 ? b.b : b.t;
}
// incoming state at a wall face, per component (upwind_state)
__device__ __forceinline__ Sym3 wall_state3(const Params& p, int side, double u, bool low_side, const Sym3& c,
                                            ErrSlot* st) {
    return {upwind_state(side_of(p.bSxx, side), u, low_side, c.xx, st),
            upwind_state(side_of(p.bSyy, side), u, low_side, c.yy, st),
            upwind_state(side_of(p.bSxy, side), u, low_side, c.xy, st)};
}

// Structure-tensor right-hand side at cell (i, j): upwind transport of the
// three components (InterfaceTransport.jl:161-258 per component, the 5-point
// stencil loaded once), stretching, relaxation and geometric extension.
// S(i, j): sigma after the safety reset; U, V: face velocities; phic: phi
// at the cell.
// IN = true: 1 <= i <= nx-2 and 1 <= j <= ny-2, so the wall states, wraps
// and one-sided wall gradients compile out (same arithmetic).
template <bool IN = false, class FS, class FU, class FV>
__device__ __forceinline__ void sigma_rhs(const Params& p, const FS& S, const FU& U, const FV& V, double phic, int i,
                                          int j, ErrSlot* st, double& oxx, double& oyy, double& oxy) {
    const Sym3 C = S(i, j);
    Sym3 fl, fr, fb, ft;
    {  // x faces i (left) and i+1 (right); periodic face nx is face 0
        const double ul = U(i, j);
        if (IN || p.px || i > 0) fl = upwind_flux3(ul, S((IN || i > 0) ? i - 1 : p.nx - 1, j), C);
        else fl = upwind_flux3(ul, wall_state3(p, 0, ul, true, C, st), C);
        if (IN || i < p.nx - 1) {
            fr = upwind_flux3(U(i + 1, j), C, S(i + 1, j));
        } else if (p.px) {
            fr = upwind_flux3(U(0, j), C, S(0, j));
        } else {
            const double ur = U(p.nx, j);
            fr = upwind_flux3(ur, C, wall_state3(p, 1, ur, false, C, st));
        }
    }
    {  // y faces j (bottom) and j+1 (top)
        const double vb = V(i, j);
        if (IN || p.py || j > 0) fb = upwind_flux3(vb, S(i, (IN || j > 0) ? j - 1 : p.ny - 1), C);
        else fb = upwind_flux3(vb, wall_state3(p, 2, vb, true, C, st), C);
        if (IN || j < p.ny - 1) {
            ft = upwind_flux3(V(i, j + 1), C, S(i, j + 1));
        } else if (p.py) {
            ft = upwind_flux3(V(i, 0), C, S(i, 0));
        } else {
            const double vt = V(i, p.ny);
            ft = upwind_flux3(vt, C, wall_state3(p, 3, vt, false, C, st));
        }
    }
    const double axx = -((fr.xx - fl.xx) * p.ihx + (ft.xx - fb.xx) * p.ihy);
    const double ayy = -((fr.yy - fl.yy) * p.ihx + (ft.yy - fb.yy) * p.ihy);
    const double axy = -((fr.xy - fl.xy) * p.ihx + (ft.xy - fb.xy) * p.ihy);
    double uL = U(i, j), uR = U(i + 1, j), vB = V(i, j), vT = V(i, j + 1);
    double dv = (uR - uL) * p.ihx + (vT - vB) * p.ihy;
    double ux = (uR - uL) * p.ihx, vy = (vT - vB) * p.ihy;
    auto Uc = [&](int jj) { return 0.5 * U(i, jj) + 0.5 * U(i + 1, jj); };
    auto Vc = [&](int ii) { return 0.5 * V(ii, j) + 0.5 * V(ii, j + 1); };
    double uy, vx;
    if (IN) {
        uy = (Uc(j + 1) - Uc(j - 1)) * (0.5 * p.ihy);
    } else if (p.py) {
        int jm = j == 0 ? p.ny - 1 : j - 1, jp = j == p.ny - 1 ? 0 : j + 1;
        uy = (Uc(jp) - Uc(jm)) * (0.5 * p.ihy);
    } else if (j == 0) {
        uy = p.bU.b.kind == FSI_BC_DIRICHLET ? (Uc(0) - p.bU.b.value) * (2.0 * p.ihy) : -p.bU.b.value;
    } else if (j == p.ny - 1) {
        uy = p.bU.t.kind == FSI_BC_DIRICHLET ? (p.bU.t.value - Uc(p.ny - 1)) * (2.0 * p.ihy) : p.bU.t.value;
    } else {
        uy = (Uc(j + 1) - Uc(j - 1)) * (0.5 * p.ihy);
    }
    if (IN) {
        vx = (Vc(i + 1) - Vc(i - 1)) * (0.5 * p.ihx);
    } else if (p.px) {
        int im = i == 0 ? p.nx - 1 : i - 1, ip = i == p.nx - 1 ? 0 : i + 1;
        vx = (Vc(ip) - Vc(im)) * (0.5 * p.ihx);
    } else if (i == 0) {
        vx = p.bV.l.kind == FSI_BC_DIRICHLET ? (Vc(0) - p.bV.l.value) * (2.0 * p.ihx) : -p.bV.l.value;
    } else if (i == p.nx - 1) {
        vx = p.bV.r.kind == FSI_BC_DIRICHLET ? (p.bV.r.value - Vc(p.nx - 1)) * (2.0 * p.ihx) : p.bV.r.value;
    } else {
        vx = (Vc(i + 1) - Vc(i - 1)) * (0.5 * p.ihx);
    }
    const double sa = C.xx, sd = C.yy, sb = C.xy;
    double A11 = ux - 0.0, A12 = uy - 0.0, A21 = vx - 0.0, A22 = vy - 0.0;
    double dxx = 2.0 * (A11 * sa + A12 * sb);
    double dyy = 2.0 * (A21 * sb + A22 * sd);
    double dxy = A21 * sa + A12 * sd + (A11 + A22) * sb;
    double rxx = 0.0, ryy = 0.0, rxy = 0.0;
    if (p.relax) {
        rxx = -p.relax_rate * (sa - 1.0);
        ryy = -p.relax_rate * (sd - 1.0);
        rxy = -p.relax_rate * sb;
    }
    double exx = 0.0, eyy = 0.0, exy = 0.0;
    if (p.ext == FSI_EXT_GEOMETRIC) {
        double kappa = p.ext_rate * pow_fast(1.0 - clamp01(phic), p.ext_exp);
        if (kappa != 0.0) {
            double gx, gy, gxy;
            if (sigma_log_product(sa, sd, sb, gx, gy, gxy)) {
                exx = -kappa * gx;
                eyy = -kappa * gy;
                exy = -kappa * gxy;
            } else {
                raise_error(st, DEV_NONSPD, sa, sb);
            }
        }
    }
    oxx = (axx + sa * dv) + dxx + rxx + exx;
    oyy = (ayy + sd * dv) + dyy + ryy + eyy;
    oxy = (axy + sb * dv) + dxy + rxy + exy;
}

// ---- phase right-hand sides on tile accessors ------------------------------
// f: phi, fpsi: psi, N0/N1: cell normals (CDI: sharpening vector); all take
// global (wrapped) indices.
template <class F, class NX, class NY>
__device__ __forceinline__ double phase_cdi_t(const Params& p, const Stage& s, const F& f, const NX& N0, const NY& N1, int i, int j,
                              ErrSlot* st, int with_flux) {
    double adv = upwind_rhs(p, s.U, s.V, p.bPhi, f, i, j, st);
    if (!with_flux) return adv - 0.0;
    const double ell = 0.5 * p.cdi_eps;
    auto jx = [&](int k) -> double {
        if (!p.px && (k == 0 || k == p.nx)) return 0.0;
        return p.cdi_G * (c2fx(p, N0, k, j) - ell * grad_xface(p, p.bPhi, f, k, j));
    };
    auto jy = [&](int k) -> double {
        if (!p.py && (k == 0 || k == p.ny)) return 0.0;
        return p.cdi_G * (c2fy(p, N1, i, k) - ell * grad_yface(p, p.bPhi, f, i, k));
    };
    double divj = (jx(i + 1) - jx(i)) * p.ihx + (jy(j + 1) - jy(j)) * p.ihy;
    return adv - divj;
}

template <class F, class FP, class NX, class NY>
__device__ __forceinline__ double phase_acdi_t(const Params& p, const Stage& s, const F& f, const FP& psi, const NX& N0,
                               const NY& N1, int i, int j, double gamma, ErrSlot* st) {
    auto fx = [&](int k) -> double {
        if (p.px) {
            int kk = (k == p.nx) ? 0 : k;
            double pf = kk == 0 ? 0.5 * f(p.nx - 1, j) + 0.5 * f(0, j) : 0.5 * f(kk - 1, j) + 0.5 * f(kk, j);
            return s.U[at(p, kk, j)] * pf;
        }
        if (k == 0) return s.U[at(p, 0, j)] * phase_bvalue(p.bPhi.l, f(0, j), st);
        if (k == p.nx) return s.U[at(p, p.nx, j)] * phase_bvalue(p.bPhi.r, f(p.nx - 1, j), st);
        return s.U[at(p, k, j)] * (0.5 * f(k - 1, j) + 0.5 * f(k, j));
    };
    auto fy = [&](int k) -> double {
        if (p.py) {
            int kk = (k == p.ny) ? 0 : k;
            double pf = kk == 0 ? 0.5 * f(i, p.ny - 1) + 0.5 * f(i, 0) : 0.5 * f(i, kk - 1) + 0.5 * f(i, kk);
            return s.V[at(p, i, kk)] * pf;
        }
        if (k == 0) return s.V[at(p, i, 0)] * phase_bvalue(p.bPhi.b, f(i, 0), st);
        if (k == p.ny) return s.V[at(p, i, p.ny)] * phase_bvalue(p.bPhi.t, f(i, p.ny - 1), st);
        return s.V[at(p, i, k)] * (0.5 * f(i, k - 1) + 0.5 * f(i, k));
    };
    const double eps = p.acdi_eps;
    auto Ax = [&](int k) -> double {
        if (!p.px && (k == 0 || k == p.nx)) return 0.0;
        double t = tanh(c2fx(p, psi, k, j) / (2.0 * eps));
        double amp = 0.25 * (1.0 - t * t);
        return gamma * (eps * grad_xface(p, p.bPhi, f, k, j) - amp * c2fx(p, N0, k, j));
    };
    auto Ay = [&](int k) -> double {
        if (!p.py && (k == 0 || k == p.ny)) return 0.0;
        double t = tanh(c2fy(p, psi, i, k) / (2.0 * eps));
        double amp = 0.25 * (1.0 - t * t);
        return gamma * (eps * grad_yface(p, p.bPhi, f, i, k) - amp * c2fy(p, N1, i, k));
    };
    double dF = (fx(i + 1) - fx(i)) * p.ihx + (fy(j + 1) - fy(j)) * p.ihy;
    double dA = (Ax(i + 1) - Ax(i)) * p.ihx + (Ay(j + 1) - Ay(j)) * p.ihy;
    return -dF + dA;
}

template <class F, class NX, class NY>
__device__ __forceinline__ double phase_uwcdi_t(const Params& p, const Stage& s, const F& f, const NX& N0, const NY& N1, int i,
                                int j, ErrSlot* st, int dir) {
    auto fx = [&](int k) -> double {
        int kk = (p.px && k == p.nx) ? 0 : k;
        double u = s.U[at(p, kk, j)], pu;
        if (!p.px && kk == 0) {
            double bnd = phase_bvalue(p.bPhi.l, f(0, j), st);
            pu = u >= 0.0 ? bnd : f(0, j);
        } else if (!p.px && kk == p.nx) {
            double bnd = phase_bvalue(p.bPhi.r, f(p.nx - 1, j), st);
            pu = u >= 0.0 ? f(p.nx - 1, j) : bnd;
        } else {
            int iL = kk == 0 ? p.nx - 1 : kk - 1, iR = kk;
            pu = u >= 0.0 ? f(iL, j) : f(iR, j);
        }
        double nf = c2fx(p, N0, kk, j);
        double ny_on = c2fx(p, N1, kk, j);
        nf /= sqrt(nf * nf + ny_on * ny_on + p.uw_d2);
        double ax = p.uw_alpha * fmax(1.0, p.hy / p.hx);
        double beta = fabs(u) / ax;
        return pu * (u + beta * nf * (1.0 - pu));
    };
    auto fy = [&](int k) -> double {
        int kk = (p.py && k == p.ny) ? 0 : k;
        double v = s.V[at(p, i, kk)], pu;
        if (!p.py && kk == 0) {
            double bnd = phase_bvalue(p.bPhi.b, f(i, 0), st);
            pu = v >= 0.0 ? bnd : f(i, 0);
        } else if (!p.py && kk == p.ny) {
            double bnd = phase_bvalue(p.bPhi.t, f(i, p.ny - 1), st);
            pu = v >= 0.0 ? f(i, p.ny - 1) : bnd;
        } else {
            int jB = kk == 0 ? p.ny - 1 : kk - 1, jT = kk;
            pu = v >= 0.0 ? f(i, jB) : f(i, jT);
        }
        double nf = c2fy(p, N1, i, kk);
        double nx_on = c2fy(p, N0, i, kk);
        nf /= sqrt(nx_on * nx_on + nf * nf + p.uw_d2);
        double ay = p.uw_alpha * fmax(1.0, p.hx / p.hy);
        double beta = fabs(v) / ay;
        return pu * (v + beta * nf * (1.0 - pu));
    };
    // dir 1 / 2: -divergence of (fx, 0) / (0, fy) (uwcdi_directional_rhs :768-791)
    if (dir == 1) return -((fx(i + 1) - fx(i)) * p.ihx + 0.0);
    if (dir == 2) return -(0.0 + (fy(j + 1) - fy(j)) * p.ihy);
    double dF = (fx(i + 1) - fx(i)) * p.ihx + (fy(j + 1) - fy(j)) * p.ihy;
    return -dF;
}

// Max |u_c| for dynamic ACDI gamma (PhaseFieldMethods.jl:576-596).
FSI_KERNEL void k_max_speed(Params p, const double* U, const double* V, ErrSlot* st) {
    __shared__ double lds[16];
    const int i = tile_id().x * TILE_X + threadIdx.x;
    double m = 0.0;
    if (i < p.nx)
        for (int r = 0; r < ROWS; ++r) {
            const int j = tile_id().y * TILE_ROWS + threadIdx.y + r * TILE_Y;
            if (j >= p.ny) break;
            double uc = 0.5 * U[at(p, i, j)] + 0.5 * U[at(p, i + 1, j)];
            double vc = 0.5 * V[at(p, i, j)] + 0.5 * V[at(p, i, j + 1)];
            m = fmax(m, hypot(uc, vc));
        }
    m = wave_max(m);
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    if ((lane & 63) == 0) lds[lane >> 6] = m;
    __syncthreads();
    if (lane == 0) {
        double mm = lds[0];
        for (int w = 1; w < BLOCK / 64; ++w) mm = fmax(mm, lds[w]);
        p.exbuf[(size_t)EX_SPEED * p.exnb + blockIdx.x + blockIdx.y * gridDim.x] = mm;
    }
}

// ---------------------------------------------------------------------------
// k_cells: E_phi (phase_field_rhs) and E_sigma (structure_tensor_rhs) of a
// stage; on stage 1 also the finiteness validation of the state
// (TimeIntegrationIMEX.jl:495-523) and the stage-1 divergence L-infinity
// (:1111-1114).
//
// LDS tile of the 64 x 16 output cells: phi (and, for ACDI, psi) on a halo of
// 2, then the cell-centred normals (CDI: the sharpening vector) on a halo of
// 1, each computed once per cell; the face fluxes are formed from the tile.
// Across a periodic side the halo holds the wrapped values; outside a
// non-periodic wall the halo is never read (walls enter through the BC
// branches of the stencils, exactly as in the reference).
// ---------------------------------------------------------------------------
constexpr int CH2 = 2, CW2 = TILE_X + 2 * CH2, CR2 = TILE_ROWS + 2 * CH2;  // 68 x 20
constexpr int CW1 = TILE_X + 2, CR1 = TILE_ROWS + 2;                      // 66 x 18

template <bool IN = false>
struct TileAcc {  // accessor of an LDS tile by global (wrapped) index; IN: no wrap
    const double* a;
    int i0, j0, H, W, R, nx, ny;
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

// ACDI right-hand side (phase_field_rhs, PhaseFieldMethods.jl:652-671), formed
// face by face: every x- and y-face flux -f + A of the tile (convective flux
// f, regularisation flux A) is evaluated once into LDS and each cell takes
// the divergence.  With q = clamp(phi), r = (q + e)/(1 - q + e), psi = eps log r
// and the interior face value psi_f = (psi_a + psi_b)/2 of interpolate_center_to_*,
// the amplitude 1/4 (1 - tanh^2(psi_f / 2 eps)) (:603-614) equals s/(1 + s)^2
// with s = sqrt(r_a) sqrt(r_b), so no tanh is evaluated.  The face
// velocities are loaded with the phi tile.
// s/(1 + s)^2 = 1/4 (1 - tanh^2(y)) with s = e^{2y}; zero where tanh(y) rounds
// to +-1 in double (|y| > 55 ln2 / 2, i.e. s outside [2^-55, 2^55]), as the
// tanh form is, so regions of phi = 0 or 1 receive no flux.
__device__ __forceinline__ double acdi_amp(double sv) {
    if (!(sv >= 0x1p-55 && sv <= 0x1p55)) return 0.0;
    return sv / ((1.0 + sv) * (1.0 + sv));
}
constexpr int AF_XW = TILE_X + 1, AF_YH = TILE_ROWS + 1;  // 65 x-faces per row, 17 rows of y-faces
constexpr int AF_NX = TILE_ROWS * AF_XW, AF_NY = AF_YH * TILE_X;
template <bool IN>
__device__ __forceinline__ void acdi_tile(const Params& p, const Stage& s, const BcSet& bpsi,
                                          double* __restrict__ Ephi, ErrSlot* st, int i0, int j0, double* s_phi,
                                          double* s_psi, double* s_sr, double* s_n0, double* s_n1) {
    static_assert(AF_NX <= CR2 * CW2 && AF_NY <= CR2 * CW2, "face fluxes fit the psi and sqrt(r) tiles");
    // psi and sqrt(r); once the face fluxes are formed, the x- and y-face fluxes
    // (50 KB of LDS in all: three blocks per CU)
    double* s_fx = s_psi;
    double* s_fy = s_sr;
    const int tid = threadIdx.x + threadIdx.y * TILE_X;
    // 1. phi on the halo-2 tile and the tile's face velocities: indices, then
    // every load back to back
    constexpr int NPT = CR2 * CW2, NIT = (NPT + BLOCK - 1) / BLOCK;
    constexpr int NUX = (AF_NX + BLOCK - 1) / BLOCK, NVY = (AF_NY + BLOCK - 1) / BLOCK;
    size_t idx[NIT], iu[NUX], iv[NVY];
    bool inb[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int k = min(tid + it * BLOCK, NPT - 1);
        int gi = i0 + k % CW2 - CH2, gj = j0 + k / CW2 - CH2;
        bool in = true;
        if (!IN && (gi < 0 || gi >= p.nx)) {
            if (p.px) gi = ((gi % p.nx) + p.nx) % p.nx;
            else in = false;
        }
        if (!IN && (gj < 0 || gj >= p.ny)) {
            if (p.py) gj = ((gj % p.ny) + p.ny) % p.ny;
            else in = false;
        }
        inb[it] = in;
        idx[it] = in ? at(p, gi, gj) : 0;
    }
#pragma unroll
    for (int it = 0; it < NUX; ++it) {  // x-face (i0 + fi, j0 + fj); face nx of a periodic x is face 0
        const int q = min(tid + it * BLOCK, AF_NX - 1);
        const int k = i0 + q % AF_XW, j = j0 + q / AF_XW;
        const bool ok = k <= p.nx && j < p.ny;
        iu[it] = ok ? at(p, (!IN && p.px && k == p.nx) ? 0 : k, j) : 0;
    }
#pragma unroll
    for (int it = 0; it < NVY; ++it) {  // y-face (i0 + fi, j0 + fj)
        const int q = min(tid + it * BLOCK, AF_NY - 1);
        const int i = i0 + q % TILE_X, k = j0 + q / TILE_X;
        const bool ok = i < p.nx && k <= p.ny;
        iv[it] = ok ? at(p, i, (!IN && p.py && k == p.ny) ? 0 : k) : 0;
    }
    double vals[NIT], uf[NUX], vf[NVY];
#pragma unroll
    for (int it = 0; it < NIT; ++it) vals[it] = s.phi[idx[it]];
#pragma unroll
    for (int it = 0; it < NUX; ++it) uf[it] = s.U[iu[it]];
#pragma unroll
    for (int it = 0; it < NVY; ++it) vf[it] = s.V[iv[it]];
    const double eps = p.acdi_eps, floor_ = p.acdi_floor;
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int k = tid + it * BLOCK;
        if (k >= NPT) break;
        double v = 0.0, ps = 0.0, sr = 0.0;
        if (inb[it]) {
            v = vals[it];
            const double q = clamp01(v);
            const double r = (q + floor_) / (1.0 - q + floor_);
            ps = eps * log(r);
            sr = sqrt(r);
        }
        s_phi[k] = v;
        s_psi[k] = ps;
        s_sr[k] = sr;
    }
    double gamma = p.acdi_gamma;
    if (p.acdi_dyn) {  // max |u_c| over the block maxima of k_max_speed
        __shared__ double smax[BLOCK / 64];
        double m = 0.0;
        for (int b = tid; b < p.exnb; b += BLOCK) m = fmax(m, p.exbuf[(size_t)EX_SPEED * p.exnb + b]);
        m = wave_max(m);
        if ((tid & 63) == 0) smax[tid >> 6] = m;
        __syncthreads();
        m = smax[0];
        for (int w = 1; w < BLOCK / 64; ++w) m = fmax(m, smax[w]);
        gamma = p.acdi_gamma * m;
    }
    __syncthreads();
    using TA = TileAcc<IN>;
    const TA f{s_phi, i0, j0, CH2, CW2, CR2, p.nx, p.ny};
    const TA fpsi{s_psi, i0, j0, CH2, CW2, CR2, p.nx, p.ny};
    const TA fsr{s_sr, i0, j0, CH2, CW2, CR2, p.nx, p.ny};
    // 2. normals of psi (_center_normals, :233-244) on the halo-1 tile
    for (int k = tid; k < CR1 * CW1; k += BLOCK) {
        int gi = i0 + k % CW1 - 1, gj = j0 + k / CW1 - 1;
        bool in = true;
        if (!IN && (gi < 0 || gi >= p.nx)) {
            if (p.px) gi = ((gi % p.nx) + p.nx) % p.nx;
            else in = false;
        }
        if (!IN && (gj < 0 || gj >= p.ny)) {
            if (p.py) gj = ((gj % p.ny) + p.ny) % p.ny;
            else in = false;
        }
        double n0 = 0.0, n1 = 0.0;
        if (in) {
            double gx, gy;
            grad_center<IN>(p, bpsi, fpsi, gi, gj, gx, gy);
            const double imag = 1.0 / sqrt(gx * gx + gy * gy + p.acdi_d2);
            n0 = gx * imag;
            n1 = gy * imag;
        }
        s_n0[k] = n0;
        s_n1[k] = n1;
    }
    __syncthreads();
    const TA N0{s_n0, i0, j0, 1, CW1, CR1, p.nx, p.ny};
    const TA N1{s_n1, i0, j0, 1, CW1, CR1, p.nx, p.ny};
    // 3. face fluxes -f + A, once per face (psi is no longer needed: its tile
    // takes the x-face fluxes)
    double fxv[NUX], fyv[NVY];
#pragma unroll
    for (int it = 0; it < NUX; ++it) {
        const int q = tid + it * BLOCK;
        const int k = i0 + q % AF_XW, j = j0 + q / AF_XW;
        double F = 0.0;
        if (q < AF_NX && k <= p.nx && j < p.ny) {
            const double u = uf[it];
            if (!IN && !p.px && k == 0) {  // walls: boundary data in f, no regularisation flux (:637-648)
                F = -u * phase_bvalue(p.bPhi.l, f(0, j), st);
            } else if (!IN && !p.px && k == p.nx) {
                F = -u * phase_bvalue(p.bPhi.r, f(p.nx - 1, j), st);
            } else {
                const int ka = (!IN && k == 0) ? p.nx - 1 : k - 1, kb = (!IN && k == p.nx) ? 0 : k;
                const double pa = f(ka, j), pb = f(kb, j);
                const double amp = acdi_amp(fsr(ka, j) * fsr(kb, j));
                const double nf = 0.5 * N0(ka, j) + 0.5 * N0(kb, j);
                const double A = gamma * (eps * (p.ihx * pb - p.ihx * pa) - amp * nf);
                F = A - u * (0.5 * pa + 0.5 * pb);
            }
        }
        fxv[it] = F;
    }
#pragma unroll
    for (int it = 0; it < NVY; ++it) {
        const int q = tid + it * BLOCK;
        const int i = i0 + q % TILE_X, k = j0 + q / TILE_X;
        double F = 0.0;
        if (q < AF_NY && i < p.nx && k <= p.ny) {
            const double v = vf[it];
            if (!IN && !p.py && k == 0) {
                F = -v * phase_bvalue(p.bPhi.b, f(i, 0), st);
            } else if (!IN && !p.py && k == p.ny) {
                F = -v * phase_bvalue(p.bPhi.t, f(i, p.ny - 1), st);
            } else {
                const int ka = (!IN && k == 0) ? p.ny - 1 : k - 1, kb = (!IN && k == p.ny) ? 0 : k;
                const double pa = f(i, ka), pb = f(i, kb);
                const double amp = acdi_amp(fsr(i, ka) * fsr(i, kb));
                const double nf = 0.5 * N1(i, ka) + 0.5 * N1(i, kb);
                const double A = gamma * (eps * (p.ihy * pb - p.ihy * pa) - amp * nf);
                F = A - v * (0.5 * pa + 0.5 * pb);
            }
        }
        fyv[it] = F;
    }
    __syncthreads();  // the psi and sqrt(r) tiles are free now
#pragma unroll
    for (int it = 0; it < NUX; ++it) {
        const int q = tid + it * BLOCK;
        if (q < AF_NX) s_fx[q] = fxv[it];
    }
#pragma unroll
    for (int it = 0; it < NVY; ++it) {
        const int q = tid + it * BLOCK;
        if (q < AF_NY) s_fy[q] = fyv[it];
    }
    __syncthreads();
    // 4. E_phi = div of the face fluxes
    const int i = i0 + threadIdx.x;
    if (i >= p.nx) return;
    for (int r = 0; r < ROWS; ++r) {
        const int lj = threadIdx.y + r * TILE_Y, j = j0 + lj;
        if (j >= p.ny) break;
        const double* fx = s_fx + lj * AF_XW + threadIdx.x;
        const double* fy = s_fy + lj * TILE_X + threadIdx.x;
        Ephi[at(p, i, j)] = (fx[1] - fx[0]) * p.ihx + (fy[TILE_X] - fy[0]) * p.ihy;
    }
}

FSI_KERNEL void k_phase_acdi(Params p, Stage s, BcSet bpsi, double* __restrict__ Ephi,
                                                      ErrSlot* st) {
    __shared__ double s_phi[CR2 * CW2];
    __shared__ double s_psi[CR2 * CW2];
    __shared__ double s_sr[CR2 * CW2];  // sqrt(r)
    __shared__ double s_n0[CR1 * CW1];
    __shared__ double s_n1[CR1 * CW1];
    const int i0 = tile_id().x * TILE_X, j0 = tile_id().y * TILE_ROWS;
    // the halo-2 tile inside the domain: no wrap, every face and normal interior
    const bool inner = i0 >= CH2 && i0 + TILE_X + CH2 - 1 <= p.nx - 1 && j0 >= CH2 &&
                       j0 + TILE_ROWS + CH2 - 1 <= p.ny - 1;
    if (inner) acdi_tile<true>(p, s, bpsi, Ephi, st, i0, j0, s_phi, s_psi, s_sr, s_n0, s_n1);
    else acdi_tile<false>(p, s, bpsi, Ephi, st, i0, j0, s_phi, s_psi, s_sr, s_n0, s_n1);
}

// Phase-field right-hand side, one instantiation per phase method.
template <int PH>
FSI_KERNEL void k_phase(Params p, Stage s, BcSet bpsi, double* __restrict__ Ephi,
                                                 ErrSlot* st, int dir) {
    constexpr bool kPsi = PH == FSI_PHASE_ACDI;
    constexpr bool kNormals = PH != FSI_PHASE_UPWIND;
    // UWCDI :smoothed9 takes its normals from the filtered phi, held in the psi tile
    constexpr bool kSecond = kPsi || PH == FSI_PHASE_UWCDI;
    const bool smooth = PH == FSI_PHASE_UWCDI && p.uw_field != nullptr;
    __shared__ double s_phi[CR2 * CW2];
    __shared__ double s_psi[kSecond ? CR2 * CW2 : 1];
    __shared__ double s_n0[kNormals ? CR1 * CW1 : 1];
    __shared__ double s_n1[kNormals ? CR1 * CW1 : 1];
    const int tid = threadIdx.x + threadIdx.y * TILE_X;
    const int i0 = tile_id().x * TILE_X, j0 = tile_id().y * TILE_ROWS;
    // 1. phi (and psi) on the halo-2 tile: indices, then all loads back to back
    {
        constexpr int NPT = CR2 * CW2, NIT = (NPT + BLOCK - 1) / BLOCK;
        size_t idx[NIT];
        bool inb[NIT];
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int k = min(tid + it * BLOCK, NPT - 1);
            int gi = i0 + k % CW2 - CH2, gj = j0 + k / CW2 - CH2;
            bool in = true;
            if (gi < 0 || gi >= p.nx) {
                if (p.px) gi = ((gi % p.nx) + p.nx) % p.nx;
                else in = false;
            }
            if (gj < 0 || gj >= p.ny) {
                if (p.py) gj = ((gj % p.ny) + p.ny) % p.ny;
                else in = false;
            }
            inb[it] = in;
            idx[it] = in ? at(p, gi, gj) : 0;
        }
        double vals[NIT], sm[NIT];
#pragma unroll
        for (int it = 0; it < NIT; ++it) vals[it] = s.phi[idx[it]];
        if (smooth) {
#pragma unroll
            for (int it = 0; it < NIT; ++it) sm[it] = p.uw_field[idx[it]];
        }
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int k = tid + it * BLOCK;
            if (k >= NPT) break;
            const double v = inb[it] ? vals[it] : 0.0;
            s_phi[k] = v;
            if (kPsi) s_psi[k] = inb[it] ? psi_value(p, v) : 0.0;
            if (smooth) s_psi[k] = inb[it] ? sm[it] : 0.0;
        }
    }
    __syncthreads();
    const TileAcc<> f{s_phi, i0, j0, CH2, CW2, CR2, p.nx, p.ny};
    const TileAcc<> fpsi{s_psi, i0, j0, CH2, CW2, CR2, p.nx, p.ny};
    // 2. cell-centred normals / sharpening vector on the halo-1 tile
    if (kNormals) {
        for (int k = tid; k < CR1 * CW1; k += BLOCK) {
            int li = k % CW1, lj = k / CW1;
            int gi = i0 + li - 1, gj = j0 + lj - 1;
            bool in = true;
            if (gi < 0 || gi >= p.nx) {
                if (p.px) gi = ((gi % p.nx) + p.nx) % p.nx;
                else in = false;
            }
            if (gj < 0 || gj >= p.ny) {
                if (p.py) gj = ((gj % p.ny) + p.ny) % p.ny;
                else in = false;
            }
            double n0 = 0.0, n1 = 0.0;
            if (in) {
                double gx, gy;
                if (PH == FSI_PHASE_ACDI) {  // _center_normals of psi (:233-244)
                    grad_center(p, bpsi, fpsi, gi, gj, gx, gy);
                    double mag = sqrt(gx * gx + gy * gy + p.acdi_d2);
                    n0 = gx / mag;
                    n1 = gy / mag;
                } else if (PH == FSI_PHASE_UWCDI) {
                    // _center_normals of phi, or of the filtered phi (:smoothed9, :385-390)
                    if (smooth) grad_center(p, p.bPhi, fpsi, gi, gj, gx, gy);
                    else grad_center(p, p.bPhi, f, gi, gj, gx, gy);
                    // :youngs (:280-333): the 3x3 stencil on every cell off the non-periodic walls
                    if (p.uw_normal == FSI_UWCDI_NORMAL_YOUNGS && (p.px || (gi > 0 && gi < p.nx - 1)) &&
                        (p.py || (gj > 0 && gj < p.ny - 1))) {
                        const int im = gi - 1, ip = gi + 1, jm = gj - 1, jp = gj + 1;
                        gx = (f(ip, jm) + 2.0 * f(ip, gj) + f(ip, jp) - f(im, jm) - 2.0 * f(im, gj) - f(im, jp)) /
                             (8.0 * p.hx);
                        gy = (f(im, jp) + 2.0 * f(gi, jp) + f(ip, jp) - f(im, jm) - 2.0 * f(gi, jm) - f(ip, jm)) /
                             (8.0 * p.hy);
                    }
                    double mag = sqrt(gx * gx + gy * gy + p.uw_d2);
                    n0 = gx / mag;
                    n1 = gy / mag;
                } else {  // CDI _sharpening_vector (InterfaceRegularization.jl:123-145)
                    grad_center(p, p.bPhi, f, gi, gj, gx, gy);
                    double mag = sqrt(gx * gx + gy * gy + p.cdi_d2);
                    double ph = f(gi, gj);
                    double amp = ph * (1.0 - ph);
                    n0 = amp * gx / mag;
                    n1 = amp * gy / mag;
                }
            }
            s_n0[k] = n0;
            s_n1[k] = n1;
        }
        __syncthreads();
    }
    const TileAcc<> N0{s_n0, i0, j0, 1, CW1, CR1, p.nx, p.ny};
    const TileAcc<> N1{s_n1, i0, j0, 1, CW1, CR1, p.nx, p.ny};
    double gamma = p.acdi_gamma;
    if (PH == FSI_PHASE_ACDI && p.acdi_dyn) {  // max |u_c| over the block maxima of k_max_speed
        __shared__ double smax[BLOCK / 64];
        double m = 0.0;
        for (int b = tid; b < p.exnb; b += BLOCK) m = fmax(m, p.exbuf[(size_t)EX_SPEED * p.exnb + b]);
        m = wave_max(m);
        if ((tid & 63) == 0) smax[tid >> 6] = m;
        __syncthreads();
        m = smax[0];
        for (int w = 1; w < BLOCK / 64; ++w) m = fmax(m, smax[w]);
        gamma = p.acdi_gamma * m;
    }
    const int i = i0 + threadIdx.x;
    if (i >= p.nx) return;
    for (int r = 0; r < ROWS; ++r) {
        const int j = j0 + threadIdx.y + r * TILE_Y;
        if (j >= p.ny) break;
        double e;
        if (PH == FSI_PHASE_ACDI) e = phase_acdi_t(p, s, f, fpsi, N0, N1, i, j, gamma, st);
        else if (PH == FSI_PHASE_UWCDI) e = phase_uwcdi_t(p, s, f, N0, N1, i, j, st, dir);
        else e = phase_cdi_t(p, s, f, N0, N1, i, j, st, PH == FSI_PHASE_CDI);
        Ephi[at(p, i, j)] = e;
    }
}

// Structure-tensor right-hand side, plus on stage 1 the divergence
// L-infinity and the finiteness check of the stage-1 state.  sigma (safety
// reset applied), U and V sit on halo-1 LDS tiles, each value loaded once.
template <bool IN = false>
struct SymTile {
    const double *xx, *yy, *xy;
    int i0, j0, nx, ny;
    __device__ __forceinline__ Sym3 operator()(int gi, int gj) const {
        int li = gi - i0 + 1, lj = gj - j0 + 1;
        if (!IN) {
            if (li < 0) li += nx;
            else if (li >= MT_W) li -= nx;
            if (lj < 0) lj += ny;
            else if (lj >= MT_H) lj -= ny;
        }
        const int k = lj * MT_W + li;
        return {xx[k], yy[k], xy[k]};
    }
};

template <bool IN>
__device__ __forceinline__ void sigma_tile(const Params& p, const Stage& s, double* __restrict__ Exx,
                                           double* __restrict__ Eyy, double* __restrict__ Exy, ErrSlot* st, int stage1,
                                           const CellSums& C, int i0, int j0, double* su, double* sv, double* sxx,
                                           double* syy, double* sxy) {
    __shared__ double lds[16];
    __shared__ long long lds_n[16];
    __shared__ double lds_r[4][16];
    __shared__ long long lds_rn[2][16];
    const int tid = threadIdx.x + threadIdx.y * TILE_X;
    constexpr int NPT = MT_H * MT_W, NIT = (NPT + BLOCK - 1) / BLOCK;
    size_t ku[NIT], kv[NIT], kc[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int k = min(tid + it * BLOCK, NPT - 1);
        const int ui = i0 + k % MT_W - 1, uj = j0 + k / MT_W - 1;
        auto wrap = [](int u, int n) { return u < 0 ? u + n : (u >= n ? u - n : u); };
        const int x = p.px ? wrap(ui, p.nx) : ui, y = p.py ? wrap(uj, p.ny) : uj;
        const bool xu_ok = x >= 0 && x <= (p.px ? p.nx - 1 : p.nx);  // faces 0..nx
        const bool yv_ok = y >= 0 && y <= (p.py ? p.ny - 1 : p.ny);  // rows 0..ny
        const bool xc_ok = x >= 0 && x < p.nx, yc_ok = y >= 0 && y < p.ny;
        ku[it] = (xu_ok && yc_ok) ? at(p, x, y) : 0;
        kv[it] = (xc_ok && yv_ok) ? at(p, x, y) : 0;
        kc[it] = (xc_ok && yc_ok) ? at(p, x, y) : 0;
    }
    double vu[NIT], vvv[NIT], vph[NIT], va[NIT], vd[NIT], vb[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        vu[it] = s.U[ku[it]];
        vvv[it] = s.V[kv[it]];
        vph[it] = s.phi[kc[it]];
        va[it] = s.sxx[kc[it]];
        vd[it] = s.syy[kc[it]];
        vb[it] = s.sxy[kc[it]];
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int k = tid + it * BLOCK;
        if (k >= NPT) break;
        su[k] = vu[it];
        sv[k] = vvv[it];
        double a = va[it], d = vd[it], b = vb[it];
        if (s.reset && vph[it] < p.ext_cutoff) {  // apply_structure_extension! StructureTensor.jl:590-613
            a = 1.0;
            d = 1.0;
            b = 0.0;
        }
        sxx[k] = a;
        syy[k] = d;
        sxy[k] = b;
    }
    __syncthreads();
    const HaloTile<IN> U{su, i0, j0, p.nx, p.ny}, V{sv, i0, j0, p.nx, p.ny};
    const SymTile<IN> S{sxx, syy, sxy, i0, j0, p.nx, p.ny};
    double dmax = 0.0;
    long long nonfinite = 0;
    // StepReport partials of the final stage (structure_tensor_diagnostics
    // StructureTensor.jl:627-662), when this launch completes stage 3
    const bool report = C.q == 2;
    double pmin = INFINITY, pmax = -INFINITY, emin = INFINITY, demax = 0.0;
    long long nonspd = 0, nonfin = 0;
    const int i = i0 + threadIdx.x;
    if (i < p.nx)
#pragma unroll 1
        for (int r = 0; r < ROWS; ++r) {
            const int j = j0 + threadIdx.y + r * TILE_Y;
            if (j >= p.ny) break;
            const size_t k = at(p, i, j);
            // E_phi and the sums this cell updates (at most two read back for
            // q >= 1), loaded before the right-hand side so that their
            // latency overlaps it
            double ep = 0.0, old[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
            if (C.q >= 0) {
                ep = C.Ephi[k];
                if (C.q >= 1) {
#pragma unroll
                    for (int t = 0; t < 2; ++t)
                        if (C.q + t < 3) {
                            old[t][0] = C.phi[C.q + t][k];
                            old[t][1] = C.sxx[C.q + t][k];
                            old[t][2] = C.syy[C.q + t][k];
                            old[t][3] = C.sxy[C.q + t][k];
                        }
                }
            }
            double oxx, oyy, oxy;
            sigma_rhs<IN>(p, S, U, V, s.phi[k], i, j, st, oxx, oyy, oxy);
            if (Exx) {
                Exx[k] = oxx;
                Eyy[k] = oyy;
                Exy[k] = oxy;
            }
            if (C.q >= 0) {  // the stage values of the later stages
                double b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;
                if (C.q == 0) {  // phi^n and sigma^n with the safety reset (the tile's values)
                    const Sym3 b = S(i, j);
                    b0 = s.phi[k];
                    b1 = b.xx;
                    b2 = b.yy;
                    b3 = b.xy;
                }
                for (int t = C.q; t < 3; ++t) {
                    const int u = (t - C.q) & 1;
                    double P = C.q == 0 ? b0 : old[u][0], A = C.q == 0 ? b1 : old[u][1];
                    double D = C.q == 0 ? b2 : old[u][2], B = C.q == 0 ? b3 : old[u][3];
                    const double c = C.cf[t];
                    if (c != 0.0) {
                        P += c * ep;
                        A += c * oxx;
                        D += c * oyy;
                        B += c * oxy;
                    }
                    if (t == C.q && C.reset && P < p.ext_cutoff) {
                        A = 1.0;
                        D = 1.0;
                        B = 0.0;
                    }
                    C.phi[t][k] = P;
                    C.sxx[t][k] = A;
                    C.syy[t][k] = D;
                    C.sxy[t][k] = B;
                    if (report && t == 2) {
                        double det = A * D - B * B;
                        double lmin = 0.5 * (A + D - hypot(A - D, 2.0 * B));
                        demax = fmax(demax, fabs(det - 1.0));
                        emin = fmin(emin, lmin);
                        nonspd += (isfinite(lmin) && lmin > 0.0) ? 0 : 1;
                        pmin = fmin(pmin, P);
                        pmax = fmax(pmax, P);
                        nonfin += (isfinite(P) && isfinite(A) && isfinite(D) && isfinite(B)) ? 0 : 1;
                    }
                }
            }
            if (stage1) {
                double uL = s.U[k], uR = s.U[at(p, i + 1, j)], vB = s.V[k], vT = s.V[at(p, i, j + 1)];
                double dv = (uR - uL) * p.ihx + (vT - vB) * p.ihy;
                dmax = fmax(dmax, fabs(dv - 0.0));
                bool ok = isfinite(uL) && isfinite(vB) && isfinite(s.phi[k]) && isfinite(s.sxx[k]) &&
                          isfinite(s.syy[k]) && isfinite(s.sxy[k]);
                if (i == p.nx - 1) ok = ok && isfinite(uR);
                if (j == p.ny - 1) ok = ok && isfinite(vT);
                nonfinite += ok ? 0 : 1;
                if (!isfinite(dv)) dmax = INFINITY;
            }
        }
    if (report) {
        pmin = wave_min(pmin);
        pmax = wave_max(pmax);
        emin = wave_min(emin);
        demax = wave_max(demax);
        for (int off = 32; off > 0; off >>= 1) {
            nonspd += __shfl_down(nonspd, off, 64);
            nonfin += __shfl_down(nonfin, off, 64);
        }
        if ((tid & 63) == 0) {
            const int w = tid >> 6;
            lds_r[0][w] = pmin;
            lds_r[1][w] = pmax;
            lds_r[2][w] = emin;
            lds_r[3][w] = demax;
            lds_rn[0][w] = nonspd;
            lds_rn[1][w] = nonfin;
        }
        __syncthreads();
        if (tid == 0) {
            double v0 = lds_r[0][0], v1 = lds_r[1][0], v2 = lds_r[2][0], v3 = lds_r[3][0];
            long long n0 = lds_rn[0][0], n1 = lds_rn[1][0];
            for (int w = 1; w < BLOCK / 64; ++w) {
                v0 = fmin(v0, lds_r[0][w]);
                v1 = fmax(v1, lds_r[1][w]);
                v2 = fmin(v2, lds_r[2][w]);
                v3 = fmax(v3, lds_r[3][w]);
                n0 += lds_rn[0][w];
                n1 += lds_rn[1][w];
            }
            const size_t bid = blockIdx.x + blockIdx.y * gridDim.x, nb = p.exnb;
            p.exbuf[EX_PHI_NEGMIN * nb + bid] = -v0;
            p.exbuf[EX_PHI_MAX * nb + bid] = v1;
            p.exbuf[EX_EIG_NEGMIN * nb + bid] = -v2;
            p.exbuf[EX_DET_ERR * nb + bid] = v3;
            p.exbuf[EX_NONSPD * nb + bid] = (double)n0;
            p.exbuf[EX_NONFINITE * nb + bid] = (double)n1;
        }
    }
    if (!stage1) return;
    dmax = wave_max(dmax);
    for (int off = 32; off > 0; off >>= 1) nonfinite += __shfl_down(nonfinite, off, 64);
    if ((tid & 63) == 0) {
        lds[tid >> 6] = dmax;
        lds_n[tid >> 6] = nonfinite;
    }
    __syncthreads();
    if (tid == 0) {
        double mm = lds[0];
        long long nn = lds_n[0];
        for (int w = 1; w < BLOCK / 64; ++w) {
            mm = fmax(mm, lds[w]);
            nn += lds_n[w];
        }
        const size_t bid = blockIdx.x + blockIdx.y * gridDim.x;
        p.exbuf[(size_t)EX_STAGE1 * p.exnb + bid] = mm;
        p.exbuf[(size_t)EX_NONFINITE_IN * p.exnb + bid] = (double)nn;
    }
}

FSI_KERNEL void k_sigma(Params p, Stage s, double* __restrict__ Exx,
                                                 double* __restrict__ Eyy, double* __restrict__ Exy, ErrSlot* st,
                                                 int stage1, CellSums C) {
    __shared__ double su[MT_H * MT_W], sv[MT_H * MT_W];
    __shared__ double sxx[MT_H * MT_W], syy[MT_H * MT_W], sxy[MT_H * MT_W];
    const TileId T = tile_id();
    const int i0 = T.x * TILE_X, j0 = T.y * TILE_ROWS;
    // every cell of the tile at least one cell inside the walls
    const bool inner = i0 >= 1 && i0 + TILE_X - 1 <= p.nx - 2 && j0 >= 1 && j0 + TILE_ROWS - 1 <= p.ny - 2;
    if (inner) sigma_tile<true>(p, s, Exx, Eyy, Exy, st, stage1, C, i0, j0, su, sv, sxx, syy, sxy);
    else sigma_tile<false>(p, s, Exx, Eyy, Exy, st, stage1, C, i0, j0, su, sv, sxx, syy, sxy);
}

// b_visc = M*0 + b (the affine part of the viscous operator), computed once.
FSI_KERNEL void k_visc_boundary(Params p, const double* zero, double* bu, double* bv) {
    const int i = tile_id().x * TILE_X + threadIdx.x;
    if (i >= p.nx) return;
    for (int r = 0; r < ROWS; ++r) {
        const int j = tile_id().y * TILE_ROWS + threadIdx.y + r * TILE_Y;
        if (j >= p.ny) return;
        bu[at(p, i, j)] = visc_u(p, ZeroAcc{}, ZeroAcc{}, i, j, 1);
        bv[at(p, i, j)] = visc_v(p, ZeroAcc{}, ZeroAcc{}, i, j, 1);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static dim3 grid2d(const Params& p) {
    return dim3((p.nx + TILE_X - 1) / TILE_X, (p.ny + TILE_ROWS - 1) / TILE_ROWS);
}

// One block: maxima and sums of the per-block partials in a fixed order.
FSI_KERNEL void k_ext_final(Params p, Status* st) {
    // maxima: -phi_min, phi_max, -eig_min, det_err, stage1, and of the growth
    // step max |G| at stage 1, the trace error, -min resource; sums: counts
    constexpr int NM = 8, NS = 3;
    const int mslot[NM] = {EX_PHI_NEGMIN, EX_PHI_MAX, EX_EIG_NEGMIN, EX_DET_ERR, EX_STAGE1, EX_GMAX, EX_TRACE,
                           EX_RES_NEGMIN};
    const int sslot[NS] = {EX_NONSPD, EX_NONFINITE, EX_NONFINITE_IN};
    __shared__ double lm[NM][BLOCK / 64], ls[NS][BLOCK / 64];
    const int tid = threadIdx.x;
    double m[NM], sm[NS];
#pragma unroll
    for (int q = 0; q < NM; ++q) m[q] = -INFINITY;
#pragma unroll
    for (int q = 0; q < NS; ++q) sm[q] = 0.0;
    const size_t nb = p.exnb;
    constexpr int UNR = 4;  // partials of UNR blocks in flight at once
    for (size_t b0 = tid; b0 < nb; b0 += UNR * BLOCK) {
        double vm[UNR][NM], vs[UNR][NS];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const size_t b = b0 + (size_t)u * BLOCK;
#pragma unroll
            for (int q = 0; q < NM; ++q) vm[u][q] = b < nb ? p.exbuf[mslot[q] * nb + b] : -INFINITY;
#pragma unroll
            for (int q = 0; q < NS; ++q) vs[u][q] = b < nb ? p.exbuf[sslot[q] * nb + b] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
#pragma unroll
            for (int q = 0; q < NM; ++q) m[q] = fmax(m[q], vm[u][q]);
#pragma unroll
            for (int q = 0; q < NS; ++q) sm[q] += vs[u][q];
        }
    }
#pragma unroll
    for (int q = 0; q < NM; ++q) m[q] = wave_max(m[q]);
#pragma unroll
    for (int q = 0; q < NS; ++q)
        for (int off = 32; off > 0; off >>= 1) sm[q] += __shfl_down(sm[q], off, 64);
    if ((tid & 63) == 0) {
#pragma unroll
        for (int q = 0; q < NM; ++q) lm[q][tid >> 6] = m[q];
#pragma unroll
        for (int q = 0; q < NS; ++q) ls[q][tid >> 6] = sm[q];
    }
    __syncthreads();
    if (tid != 0) return;
    for (int w = 1; w < BLOCK / 64; ++w) {
#pragma unroll
        for (int q = 0; q < NM; ++q) lm[q][0] = fmax(lm[q][0], lm[q][w]);
#pragma unroll
        for (int q = 0; q < NS; ++q) ls[q][0] += ls[q][w];
    }
    st->phi_min = -lm[0][0];
    st->phi_max = lm[1][0];
    st->min_eig = -lm[2][0];
    st->max_det_err = fmax(lm[3][0], 0.0);
    st->stage1_linf = fmax(lm[4][0], 0.0);
    st->stage1_gmax = fmax(lm[5][0], 0.0);
    st->growth_trace = fmax(lm[6][0], 0.0);
    st->res_min = -lm[7][0];
    st->nonspd = (long long)ls[0][0];
    st->nonfinite = (long long)ls[1][0];
    st->nonfinite_in = (long long)ls[2][0];
}

// ---------------------------------------------------------------------------
// Point kernels of the UWCDI :smoothed9 normals, the directional Strang sweep
// and DiffuseAreaCapillarity: one thread per cell (faces: `faces` = 1), blocks
// of 64 x 4.
// ---------------------------------------------------------------------------
static inline dim3 point_grid(const Params& p, int faces) {
    return dim3((unsigned)((p.nx + faces + TILE_X - 1) / TILE_X), (unsigned)((p.ny + faces + TILE_Y - 1) / TILE_Y));
}
__device__ __forceinline__ int clamp_or_wrap(int k, int n, int periodic) {
    if (periodic) return k < 0 ? k + n : (k >= n ? k - n : k);
    return k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
}

// The 9-point filter of _smoothed9_field (PhaseFieldMethods.jl:364-383) and
// _smooth_phase_for_capillarity (Capillarity.jl:81-111): the index is
// clamped at a wall and wrapped across a periodic side.
FSI_KERNEL void k_smooth9(Params p, const double* __restrict__ f, double* __restrict__ out) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (i >= p.nx || j >= p.ny) return;
    const int im = clamp_or_wrap(i - 1, p.nx, p.px), ip = clamp_or_wrap(i + 1, p.nx, p.px);
    const int jm = clamp_or_wrap(j - 1, p.ny, p.py), jp = clamp_or_wrap(j + 1, p.ny, p.py);
    auto F = [&](int a, int b) { return f[at(p, a, b)]; };
    out[at(p, i, j)] = 0.25 * F(i, j) + 0.125 * (F(ip, j) + F(im, j) + F(i, jp) + F(i, jm)) +
                       0.0625 * (F(ip, jp) + F(im, jp) + F(ip, jm) + F(im, jm));
}

// phi_out = phi + h rhs on the cells (_uwcdi_directional_euler_sweep :793-810)
FSI_KERNEL void k_cell_axpy(Params p, const double* __restrict__ phi, double h,
                                                     const double* __restrict__ rhs, double* __restrict__ out) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (i >= p.nx || j >= p.ny) return;
    const size_t k = at(p, i, j);
    out[k] = phi[k] + h * rhs[k];
}

// Regularised centre normal of the geometry field (Capillarity.jl:59-79):
// the face gradients with the phase boundary conditions averaged to the centre.
FSI_KERNEL void k_cap_normals(Params p, const double* __restrict__ geo,
                                                       double* __restrict__ nxc, double* __restrict__ nyc) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (i >= p.nx || j >= p.ny) return;
    auto G = [=](int a, int b) { return geo[at(p, a, b)]; };
    double gx, gy;
    grad_center(p, p.bPhi, G, i, j, gx, gy);
    const double den = sqrt(gx * gx + gy * gy + p.cap_d2);
    nxc[at(p, i, j)] = gx / den;
    nyc[at(p, i, j)] = gy / den;
}

// kappa = -div(n) with n carried to the faces by interpolate_center_to_*_faces (Capillarity.jl:168-173)
FSI_KERNEL void k_cap_kappa(Params p, const double* __restrict__ nxc,
                                                     const double* __restrict__ nyc, double* __restrict__ kap) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (i >= p.nx || j >= p.ny) return;
    auto NX = [=](int a, int b) { return nxc[at(p, a, b)]; };
    auto NY = [=](int a, int b) { return nyc[at(p, a, b)]; };
    kap[at(p, i, j)] = -((c2fx(p, NX, i + 1, j) - c2fx(p, NX, i, j)) / p.hx +
                         (c2fy(p, NY, i, j + 1) - c2fy(p, NY, i, j)) / p.hy);
}

// f = coefficient kappa_face grad(phi) on every x- and y-face (Capillarity.jl:175-186)
FSI_KERNEL void k_cap_force(Params p, const double* __restrict__ phi,
                                                     const double* __restrict__ kap, double* __restrict__ fx,
                                                     double* __restrict__ fy) {
    const int i = blockIdx.x * TILE_X + threadIdx.x, j = blockIdx.y * TILE_Y + threadIdx.y;
    if (i > p.nx || j > p.ny) return;
    auto PH = [=](int a, int b) { return phi[at(p, a, b)]; };
    auto K = [=](int a, int b) { return kap[at(p, a, b)]; };
    if (j < p.ny) fx[at(p, i, j)] = p.cap_coef * c2fx(p, K, i, j) * grad_xface(p, p.bPhi, PH, i, j);
    if (i < p.nx) fy[at(p, i, j)] = p.cap_coef * c2fy(p, K, i, j) * grad_yface(p, p.bPhi, PH, i, j);
}

void launch_ext_final(hipStream_t s, const Params& p, Status* st) { fsi_launch<&k_ext_final, BLOCK, 0>(1, BLOCK, 0, s, p, st); }

// capillary_force (Capillarity.jl:134-187) of the stage phi into p.cap_fx, p.cap_fy
static void launch_capillary(hipStream_t s, const Params& p, const double* phi) {
    const size_t F = (size_t)(p.ny + 1) * p.P;
    double* w = p.cap_work;
    const double* geo = phi;
    for (int k = 0; k < p.cap_passes; ++k) {
        double* out = w + (size_t)(k & 1) * F;
        fsi_launch<&k_smooth9, BLOCK, 0>(point_grid(p, 0), dim3(TILE_X, TILE_Y), 0, s, p, geo, out);
        geo = out;
    }
    fsi_launch<&k_cap_normals, BLOCK, 0>(point_grid(p, 0), dim3(TILE_X, TILE_Y), 0, s, p, geo, w + 2 * F, w + 3 * F);
    fsi_launch<&k_cap_kappa, BLOCK, 0>(point_grid(p, 0), dim3(TILE_X, TILE_Y), 0, s, p, w + 2 * F, w + 3 * F, w + 4 * F);
    fsi_launch<&k_cap_force, BLOCK, 0>(point_grid(p, 1), dim3(TILE_X, TILE_Y), 0, s, p, phi, w + 4 * F, const_cast<double*>(p.cap_fx),
                                                                   const_cast<double*>(p.cap_fy));
}

void launch_momentum(hipStream_t s, const Params& p, const double* U, const double* V, const double* phi,
                     const double* sxx, const double* syy, const double* sxy, int reset, double* Eu, double* Ev,
                     double* Iu, double* Iv, const RhsSums* sums) {
    if (p.cap_fx) launch_capillary(s, p, phi);
    Stage st{U, V, phi, sxx, syy, sxy, reset};
    RhsSums R{};
    R.q = -1;
    if (sums) R = *sums;
    fsi_launch<&k_momentum, BLOCK, 3>(grid2d(p), dim3(TILE_X, TILE_Y), 0, s, p, st, Eu, Ev, Iu, Iv, R);
}

void launch_cells(hipStream_t s, const Params& p, const BcSet& bpsi, const double* U, const double* V,
                  const double* phi, const double* sxx, const double* syy, const double* sxy, int reset,
                  double* Ephi, double* Exx, double* Eyy, double* Exy, ErrSlot* st, int stage1,
                  const CellSums* sums) {
    Stage sg{U, V, phi, sxx, syy, sxy, reset};
    if (p.phase == FSI_PHASE_ACDI && p.acdi_dyn)
        fsi_launch<&k_max_speed, BLOCK, 0>(grid2d(p), dim3(TILE_X, TILE_Y), 0, s, p, U, V, st);
    const dim3 g = grid2d(p), blk(TILE_X, TILE_Y);
    switch (p.phase) {
        case FSI_PHASE_ACDI: fsi_launch<&k_phase_acdi, BLOCK, 0>(g, blk, 0, s, p, sg, bpsi, Ephi, st); break;
        case FSI_PHASE_UWCDI:
            if (p.uw_field) fsi_launch<&k_smooth9, BLOCK, 0>(point_grid(p, 0), blk, 0, s, p, phi, p.uw_field);
            fsi_launch<&k_phase<FSI_PHASE_UWCDI>, BLOCK, 0>(g, blk, 0, s, p, sg, bpsi, Ephi, st, 0);
            break;
        case FSI_PHASE_CDI: fsi_launch<&k_phase<FSI_PHASE_CDI>, BLOCK, 0>(g, blk, 0, s, p, sg, bpsi, Ephi, st, 0); break;
        default: fsi_launch<&k_phase<FSI_PHASE_UPWIND>, BLOCK, 0>(g, blk, 0, s, p, sg, bpsi, Ephi, st, 0); break;
    }
    CellSums C{};
    C.q = -1;
    if (sums) C = *sums;
    fsi_launch<&k_sigma, BLOCK, 0>(g, blk, 0, s, p, sg, Exx, Eyy, Exy, st, stage1, C);
}

void launch_uwcdi_sweep(hipStream_t s, const Params& p, const double* U, const double* V, const double* phi, int dir,
                        double h, double* rhs, double* phi_out, ErrSlot* err) {
    Stage sg{U, V, phi, nullptr, nullptr, nullptr, 0};
    const dim3 blk(TILE_X, TILE_Y);
    if (p.uw_field) fsi_launch<&k_smooth9, BLOCK, 0>(point_grid(p, 0), blk, 0, s, p, phi, p.uw_field);
    fsi_launch<&k_phase<FSI_PHASE_UWCDI>, BLOCK, 0>(grid2d(p), blk, 0, s, p, sg, p.bPhi, rhs, err, dir);
    fsi_launch<&k_cell_axpy, BLOCK, 0>(point_grid(p, 0), blk, 0, s, p, phi, h, rhs, phi_out);
}

void launch_visc_boundary(hipStream_t s, const Params& p, const double* zero, double* bu, double* bv) {
    fsi_launch<&k_visc_boundary, BLOCK, 0>(grid2d(p), dim3(TILE_X, TILE_Y), 0, s, p, zero, bu, bv);
}

// ===========================================================================
// Driver diagnostics (scripts/benchmarks/run_sugiyama_lid_driven_cavity.jl
// :483-934, BenchmarkUtilities.jl:62-88).  Each kernel writes per block NS
// sums and NM maxima (a minimum is stored negated) as rows of nb values;
// k_diag_final reduces them in a fixed order, so the results are
// deterministic.  Sums that count cells are exact in double up to 2^53.
// Julia's minimum/maximum propagate NaN while fmax skips it, so every kernel
// also counts the NaN values its extrema see; the host reports NaN for those
// extrema when the count is not zero.
// ===========================================================================
template <int NS, int NM>
__device__ __forceinline__ void diag_put(double (&sm)[NS], double (&mx)[NM], double* buf, int nb) {
    __shared__ double ls[NS * 16];
    __shared__ double lm[NM][BLOCK / 64];
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    const int bid = blockIdx.x + blockIdx.y * gridDim.x;
    block_sum<NS>(sm, ls);
#pragma unroll
    for (int q = 0; q < NM; ++q) mx[q] = wave_max(mx[q]);
    if ((lane & 63) == 0)
#pragma unroll
        for (int q = 0; q < NM; ++q) lm[q][lane >> 6] = mx[q];
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NS; ++q) buf[(size_t)q * nb + bid] = sm[q];
#pragma unroll
        for (int q = 0; q < NM; ++q) {
            double v = lm[q][0];
            for (int w = 1; w < BLOCK / 64; ++w) v = fmax(v, lm[q][w]);
            buf[(size_t)(NS + q) * nb + bid] = v;
        }
    }
}

// x_p, y_p of a uniform grid (MeshGeneration.jl:215: L * ((i - 0.5) / N), 1-based)
__device__ __forceinline__ double centre_x(double lx, int nx, int i) { return lx * (((double)i + 0.5) / (double)nx); }

// pass 1: phase integral and first moments (phase_centroid :496-509), mixed
// layer (:656), bounds (phase_bound_diagnostics :709-722), flow
// (flow_diagnostics :724-751, with the MAC divergence), velocity extrema
// (comparison_timestep_bounds :862-866).
// sums: 0 sum phi a, 1 sum x phi a, 2 sum y phi a, 3 sum a 4 phi~ (1 - phi~),
//       4 kinetic energy, 5 sum a div^2, 6 sum a, 7 cells outside [0, 1],
//       NaN counts: 8 phi, 9 speed or divergence, 10 U or V
// maxima: 0 phi max, 1 -phi min, 2 speed, 3 |div|, 4 |U|, 5 |V|
constexpr int DG1_S = 11, DG1_M = 6;
FSI_KERNEL void k_diag_state(Params p, const double* U, const double* V, const double* phi,
                                                      double lx, double ly, double density, double* buf) {
    double sm[DG1_S], mx[DG1_M];
#pragma unroll
    for (int q = 0; q < DG1_S; ++q) sm[q] = 0.0;
#pragma unroll
    for (int q = 0; q < DG1_M; ++q) mx[q] = -INFINITY;
    const double area = p.hx * p.hy;
    const TileId T = tile_id();
    const int i = T.x * TILE_X + threadIdx.x;
    for (int r = 0; r < ROWS; ++r) {
        const int j = T.y * TILE_ROWS + threadIdx.y + r * TILE_Y;
        if (i >= p.nx || j >= p.ny) break;
        const double ph = phi[at(p, i, j)];
        const double pb = clamp01(ph);
        sm[0] += ph * area;
        const double w = ph * area;
        sm[1] += centre_x(lx, p.nx, i) * w;
        sm[2] += centre_x(ly, p.ny, j) * w;
        sm[3] += area * 4.0 * pb * (1.0 - pb);
        const double uL = U[at(p, i, j)], uR = U[at(p, i + 1, j)], vB = V[at(p, i, j)], vT = V[at(p, i, j + 1)];
        const double uc = 0.5 * (uL + uR), vc = 0.5 * (vB + vT);
        const double s2 = uc * uc + vc * vc;
        sm[4] += 0.5 * density * s2 * area;
        const double div = (uR - uL) / p.hx + (vT - vB) / p.hy;
        sm[5] += area * div * div;
        sm[6] += area;
        if (ph < 0.0 || ph > 1.0) sm[7] += 1.0;
        mx[0] = fmax(mx[0], ph);
        mx[1] = fmax(mx[1], -ph);
        mx[2] = fmax(mx[2], sqrt(s2));
        mx[3] = fmax(mx[3], fabs(div));
        mx[4] = fmax(mx[4], fabs(uL));
        mx[5] = fmax(mx[5], fabs(vB));
        double nan_uv = (isnan(uL) ? 1.0 : 0.0) + (isnan(vB) ? 1.0 : 0.0);
        if (i == p.nx - 1) {  // the last x-face column
            mx[4] = fmax(mx[4], fabs(uR));
            nan_uv += isnan(uR) ? 1.0 : 0.0;
        }
        if (j == p.ny - 1) {  // the last y-face row
            mx[5] = fmax(mx[5], fabs(vT));
            nan_uv += isnan(vT) ? 1.0 : 0.0;
        }
        sm[8] += isnan(ph) ? 1.0 : 0.0;
        sm[9] += (isnan(s2) || isnan(div)) ? 1.0 : 0.0;
        sm[10] += nan_uv;
    }
    diag_put<DG1_S, DG1_M>(sm, mx, buf, p.exnb);
}

// one region of tensor_region_diagnostics (:753-822) with
// tensor_cell_diagnostics_2d (BenchmarkUtilities.jl:62-88): the cells with
// lo <(=) phi <(=) hi, or every cell when all != 0.
// sums: 0 cells, 1 area, 2 sum a |det - 1|, 3 sum a stress, 4 non-SPD cells, 5 NaN count
// maxima: 0 det, 1 -det, 2 |det - 1|, 3 lambda_max, 4 -lambda_min, 5 stress
constexpr int DGT_S = 6, DGT_M = 6;
FSI_KERNEL void k_diag_tensor(Params p, const double* phi, const double* sxx,
                                                       const double* syy, const double* sxy, double G, int all,
                                                       double lo, int lo_in, double hi, int hi_in, double* buf) {
    double sm[DGT_S], mx[DGT_M];
#pragma unroll
    for (int q = 0; q < DGT_S; ++q) sm[q] = 0.0;
#pragma unroll
    for (int q = 0; q < DGT_M; ++q) mx[q] = -INFINITY;
    const double area = p.hx * p.hy;
    const TileId T = tile_id();
    const int i = T.x * TILE_X + threadIdx.x;
    for (int r = 0; r < ROWS; ++r) {
        const int j = T.y * TILE_ROWS + threadIdx.y + r * TILE_Y;
        if (i >= p.nx || j >= p.ny) break;
        const size_t k = at(p, i, j);
        const double ph = phi[k];
        const bool sel = all || ((lo_in ? ph >= lo : ph > lo) && (hi_in ? ph <= hi : ph < hi));
        if (!sel) continue;
        const double xx = sxx[k], yy = syy[k], xy = sxy[k];
        const double det = xx * yy - xy * xy;
        const double tr = xx + yy;
        const double gap = hypot(xx - yy, 2.0 * xy);
        const double lmin = 0.5 * (tr - gap), lmax = 0.5 * (tr + gap);
        const double stress = fabs(ph * G) * sqrt((xx - 1.0) * (xx - 1.0) + (yy - 1.0) * (yy - 1.0) + 2.0 * (xy * xy));
        const double derr = fabs(det - 1.0);
        sm[0] += 1.0;
        sm[1] += area;
        sm[2] += area * derr;
        sm[3] += area * stress;
        if (!(isfinite(lmin) && lmin > 0.0)) sm[4] += 1.0;
        sm[5] += (isnan(det) || isnan(lmin) || isnan(lmax) || isnan(stress)) ? 1.0 : 0.0;
        mx[0] = fmax(mx[0], det);
        mx[1] = fmax(mx[1], -det);
        mx[2] = fmax(mx[2], derr);
        mx[3] = fmax(mx[3], lmax);
        mx[4] = fmax(mx[4], -lmin);
        mx[5] = fmax(mx[5], stress);
    }
    diag_put<DGT_S, DGT_M>(sm, mx, buf, p.exnb);
}

// pass 2 of phase_geometry_diagnostics (:632-707), about the centroid
// (xc, yc): modes n0 + 1 .. n0 + DG2_MODES of the interface-weighted Fourier
// series, and with n0 == 0 the perimeter proxy and the second moments.
// sums: 0 perimeter, 1 cxx, 2 cyy, 3 cxy, 4 + 2q Rc[n0 + q + 1], 5 + 2q Rs[...]
constexpr int DG2_MODES = 8, DG2_S = 4 + 2 * DG2_MODES;
FSI_KERNEL void k_diag_geometry(Params p, const double* phi, double lx, double ly, double xc,
                                                         double yc, int n0, int nmax, double* buf) {
    double sm[DG2_S], mx[1] = {-INFINITY};
#pragma unroll
    for (int q = 0; q < DG2_S; ++q) sm[q] = 0.0;
    const double area = p.hx * p.hy;
    const TileId T = tile_id();
    const int i = T.x * TILE_X + threadIdx.x;
    auto f = [&](int a, int b) { return phi[at(p, a, b)]; };
    for (int r = 0; r < ROWS; ++r) {
        const int j = T.y * TILE_ROWS + threadIdx.y + r * TILE_Y;
        if (i >= p.nx || j >= p.ny) break;
        double gx, gy;
        grad_center(p, p.bPhi, f, i, j, gx, gy);
        const double iw = area * hypot(gx, gy);
        const double dx = centre_x(lx, p.nx, i) - xc, dy = centre_x(ly, p.ny, j) - yc;
        if (n0 == 0) {
            const double mw = area * phi[at(p, i, j)];
            sm[0] += iw;
            sm[1] += mw * dx * dx;
            sm[2] += mw * dy * dy;
            sm[3] += mw * dx * dy;
        }
        const double theta = atan2(dy, dx);
#pragma unroll
        for (int q = 0; q < DG2_MODES; ++q) {
            const int n = n0 + q + 1;
            if (n > nmax) break;
            double sn, cn;
            sincos((double)n * theta, &sn, &cn);
            sm[4 + 2 * q] += iw * cn / M_PI;
            sm[5 + 2 * q] += iw * sn / M_PI;
        }
    }
    diag_put<DG2_S, 1>(sm, mx, buf, p.exnb);
}

// one block: out[q] = total of row q (sums in a fixed order, then maxima)
template <int NS, int NM>
FSI_KERNEL void k_diag_final(const double* buf, int nb, double* out) {
    double sm[NS], mx[NM];
#pragma unroll
    for (int q = 0; q < NS; ++q) sm[q] = 0.0;
#pragma unroll
    for (int q = 0; q < NM; ++q) mx[q] = -INFINITY;
    for (int b = threadIdx.x; b < nb; b += BLOCK) {
#pragma unroll
        for (int q = 0; q < NS; ++q) sm[q] += buf[(size_t)q * nb + b];
#pragma unroll
        for (int q = 0; q < NM; ++q) mx[q] = fmax(mx[q], buf[(size_t)(NS + q) * nb + b]);
    }
    __shared__ double ls[NS * 16];
    __shared__ double lm[NM][BLOCK / 64];
    block_sum<NS>(sm, ls);
#pragma unroll
    for (int q = 0; q < NM; ++q) mx[q] = wave_max(mx[q]);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < NM; ++q) lm[q][threadIdx.x >> 6] = mx[q];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NS; ++q) out[q] = sm[q];
#pragma unroll
        for (int q = 0; q < NM; ++q) {
            double v = lm[q][0];
            for (int w = 1; w < BLOCK / 64; ++w) v = fmax(v, lm[q][w]);
            out[NS + q] = v;
        }
    }
}

void launch_diag_state(hipStream_t s, const Params& p, const double* U, const double* V, const double* phi,
                       double lx, double ly, double density, double* buf, double* out) {
    fsi_launch<&k_diag_state, BLOCK, 0>(grid2d(p), dim3(TILE_X, TILE_Y), 0, s, p, U, V, phi, lx, ly, density, buf);
    fsi_launch<&k_diag_final<DG1_S, DG1_M>, BLOCK, 0>(1, BLOCK, 0, s, buf, p.exnb, out);
}
void launch_diag_tensor(hipStream_t s, const Params& p, const double* phi, const double* sxx, const double* syy,
                        const double* sxy, double G, const DiagRegion& rg, double* buf, double* out) {
    fsi_launch<&k_diag_tensor, BLOCK, 0>(grid2d(p), dim3(TILE_X, TILE_Y), 0, s, p, phi, sxx, syy, sxy, G, rg.all, rg.lo, rg.lo_in,
                                                             rg.hi, rg.hi_in, buf);
    fsi_launch<&k_diag_final<DGT_S, DGT_M>, BLOCK, 0>(1, BLOCK, 0, s, buf, p.exnb, out);
}
void launch_diag_geometry(hipStream_t s, const Params& p, const double* phi, double lx, double ly, double xc,
                          double yc, int n0, int nmax, double* buf, double* out) {
    fsi_launch<&k_diag_geometry, BLOCK, 0>(grid2d(p), dim3(TILE_X, TILE_Y), 0, s, p, phi, lx, ly, xc, yc, n0, nmax, buf);
    fsi_launch<&k_diag_final<DG2_S, 1>, BLOCK, 0>(1, BLOCK, 0, s, buf, p.exnb, out);
}
int diag_modes_per_launch() { return DG2_MODES; }

// phase_centroid (run_sugiyama_lid_driven_cavity.jl:496-509) from the sums
// of k_diag_state: mass = sum phi dA, x = sum x phi dA / mass, y likewise.
FSI_KERNEL void k_centroid_result(const double* __restrict__ v, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    const double mass = v[0];
    const bool ok = mass > 2.220446049250313e-16;
    out[0] = ok ? v[1] / mass : NAN;
    out[1] = ok ? v[2] / mass : NAN;
    out[2] = mass;
}
void launch_centroid_result(hipStream_t s, const double* reduced, double* result) {
    fsi_launch<&k_centroid_result, 1024, 0>(1, 64, 0, s, reduced, result);
}

// The Status a step starts from: every counter, flag and sum zero, the
// extrema at their identities.  A kernel rather than a copy from the host,
// so the step's first launch does not wait on a host-to-device transfer.
FSI_KERNEL void k_status_reset(Status* st) {
    constexpr int W = (int)(sizeof(Status) / sizeof(int));
    int* w = reinterpret_cast<int*>(st);
    for (int k = threadIdx.x; k < W; k += blockDim.x) w[k] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        st->phi_min = INFINITY;
        st->phi_max = -INFINITY;
        st->min_eig = INFINITY;
    }
}
void launch_status_reset(hipStream_t s, Status* st) { fsi_launch<&k_status_reset, 256, 0>(1, 256, 0, s, st); }

}  // namespace fsi
