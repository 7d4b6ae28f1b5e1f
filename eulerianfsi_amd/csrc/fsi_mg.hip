// fsi_mg.hip -- fused multigrid level kernels of the projection V-cycle.
//
// Per fine level the V-cycle needs, from a zero initial guess,
//   down:  x1 = w D^-1 b;  x2 = x1 + w D^-1 (b - A x1);  b_c = R (b - A x2)
//   up:    x3 = x2 + P e_c;  x4 = x3 + w D^-1 (b - A x3);  x5 = x4 + w D^-1 (b - A x4)
// (A = minus the unit Laplacian, R = P^T/4).  Each is one kernel on an LDS
// tile with the halo the chained stencils need (3 cells down, 2 up), so a
// level costs 2 launches and reads b once, writes x2 once, and writes x5 once.
//
// Tiles are 64 x 16 fine cells for 256 threads; halo cells outside a
// non-periodic wall are never read (the wall enters the stencil through its
// coefficient), halo cells across a periodic side hold the wrapped values.
#include "fsi_kernels.hpp"

namespace fsi {

constexpr int FT_X = 64;  // fine tile
constexpr int FT_Y = 16;
constexpr int MG_THREADS = 256;

// stencil of -A at a cell from its four neighbours; `has` flags which
// neighbours exist (interior or periodic), walls add their coefficient.
struct Nb {
    int wx_lo, wx_hi, wy_lo, wy_hi;  // 0: neighbour exists, 1: Neumann wall, 2: Dirichlet wall
};

__device__ __forceinline__ Nb neighbours(const MgGeom& g, int gi, int gj) {
    Nb n;
    n.wx_lo = (gi > 0 || g.px) ? 0 : (g.dl ? 2 : 1);
    n.wx_hi = (gi < g.nx - 1 || g.px) ? 0 : (g.dr ? 2 : 1);
    n.wy_lo = (gj > 0 || g.py) ? 0 : (g.db ? 2 : 1);
    n.wy_hi = (gj < g.ny - 1 || g.py) ? 0 : (g.dt ? 2 : 1);
    return n;
}

__device__ __forceinline__ double nb_diag(const MgGeom& g, const Nb& n) {
    double cx = (n.wx_lo == 0 ? 1.0 : (n.wx_lo == 2 ? 2.0 : 0.0)) + (n.wx_hi == 0 ? 1.0 : (n.wx_hi == 2 ? 2.0 : 0.0));
    double cy = (n.wy_lo == 0 ? 1.0 : (n.wy_lo == 2 ? 2.0 : 0.0)) + (n.wy_hi == 0 ? 1.0 : (n.wy_hi == 2 ? 2.0 : 0.0));
    return g.ihx2 * cx + g.ihy2 * cy;
}

// -A x at the LDS point (li, lj) of an array with row length W.
__device__ __forceinline__ double nb_apply(const MgGeom& g, const Nb& n, const double* a, int W, int li, int lj,
                                          double d) {
    const double c = a[lj * W + li];
    double sx = 0.0, sy = 0.0;
    if (n.wx_lo == 0) sx += a[lj * W + li - 1];
    if (n.wx_hi == 0) sx += a[lj * W + li + 1];
    if (n.wy_lo == 0) sy += a[(lj - 1) * W + li];
    if (n.wy_hi == 0) sy += a[(lj + 1) * W + li];
    return d * c - g.ihx2 * sx - g.ihy2 * sy;
}

// global index of tile-local coordinate, wrapped across periodic sides;
// returns false when the point lies outside a non-periodic domain.
__device__ __forceinline__ bool gcoord(int n, int per, int k, int& out) {
    if (k >= 0 && k < n) {
        out = k;
        return true;
    }
    if (!per) return false;
    out = ((k % n) + n) % n;
    return true;
}

__device__ __forceinline__ void prol_1d_mg(int f, int nf, int per, int dlo, int dhi, int& I0, double& w0, int& I1,
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

// ---------------------------------------------------------------------------
// down: b (fine) -> x2 (fine, stored) and b_c (coarse)
// ---------------------------------------------------------------------------
constexpr int DH = 3;                 // halo of b
constexpr int DW = FT_X + 2 * DH;     // 70
constexpr int DHh = FT_Y + 2 * DH;    // 22

__global__ void __launch_bounds__(MG_THREADS) k_mg_down(MgGeom gf, MgGeom gc, const double* __restrict__ b,
                                                        double* __restrict__ x2out, double* __restrict__ bc,
                                                        double omega, const int* done, Status* st,
                                                        const double* __restrict__ rr, int nb, int m, int slot) {
    if (*done) return;
    if (st) {  // finest level: convergence test on |r|^2 of the previous PCG kernel
        double tot[1];
        get_totals<1>(rr, nb, tot);
        if (tot[0] <= st->pcg.tol2 || m >= st->pcg.maxiter) {
            if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
                st->pcg.done = 1;
                st->pcg.iter = m;
                st->pcg.rr = tot[0];
                if (slot >= 0) st->proj_iters[slot] = m;
            }
            return;
        }
    }
    __shared__ double sb[DHh * DW];                       // b on halo 3
    __shared__ double sx2[(DHh - 2) * (DW - 2)];          // x2 on halo 2 (68 x 20)
    __shared__ double sr[(DHh - 4) * (DW - 4)];           // r on halo 1 (66 x 18)
    const int tid = threadIdx.x;
    const int fi0 = blockIdx.x * FT_X, fj0 = blockIdx.y * FT_Y;
    // load b with halo 3 (points outside a wall stay unused; store 0)
    for (int k = tid; k < DHh * DW; k += MG_THREADS) {
        int li = k % DW, lj = k / DW;
        int gi, gj;
        double v = 0.0;
        if (gcoord(gf.nx, gf.px, fi0 + li - DH, gi) && gcoord(gf.ny, gf.py, fj0 + lj - DH, gj))
            v = b[(size_t)gj * gf.P + gi];
        sb[k] = v;
    }
    __syncthreads();
    // x2 on halo 2: x2 = x1 + w (b - A x1)/d with x1 = w b / d (pointwise)
    constexpr int W2 = DW - 2, H2 = DHh - 2;
    for (int k = tid; k < W2 * H2; k += MG_THREADS) {
        int li = k % W2, lj = k / W2;  // sb coordinates li+1, lj+1
        int gi, gj;
        double v = 0.0;
        if (gcoord(gf.nx, gf.px, fi0 + li - 2, gi) && gcoord(gf.ny, gf.py, fj0 + lj - 2, gj)) {
            Nb n = neighbours(gf, gi, gj);
            const int si = li + 1, sj = lj + 1;
            double d = nb_diag(gf, n);
            // x1 at the centre and the existing neighbours
            auto x1 = [&](int ii, int jj) -> double {
                int ggi, ggj;
                gcoord(gf.nx, gf.px, fi0 + ii - DH, ggi);
                gcoord(gf.ny, gf.py, fj0 + jj - DH, ggj);
                return omega * sb[jj * DW + ii] / nb_diag(gf, neighbours(gf, ggi, ggj));
            };
            double c = x1(si, sj), sx = 0.0, sy = 0.0;
            if (n.wx_lo == 0) sx += x1(si - 1, sj);
            if (n.wx_hi == 0) sx += x1(si + 1, sj);
            if (n.wy_lo == 0) sy += x1(si, sj - 1);
            if (n.wy_hi == 0) sy += x1(si, sj + 1);
            double ax = d * c - gf.ihx2 * sx - gf.ihy2 * sy;
            v = c + omega * (sb[sj * DW + si] - ax) / d;
        }
        sx2[k] = v;
    }
    __syncthreads();
    // r on halo 1, and store x2 on the interior
    constexpr int W1 = DW - 4, H1 = DHh - 4;
    for (int k = tid; k < W1 * H1; k += MG_THREADS) {
        int li = k % W1, lj = k / W1;  // sx2 coordinates li+1, lj+1
        int gi, gj;
        double v = 0.0;
        if (gcoord(gf.nx, gf.px, fi0 + li - 1, gi) && gcoord(gf.ny, gf.py, fj0 + lj - 1, gj)) {
            Nb n = neighbours(gf, gi, gj);
            double d = nb_diag(gf, n);
            double ax = nb_apply(gf, n, sx2, W2, li + 1, lj + 1, d);
            v = sb[(lj + 2) * DW + li + 2] - ax;
            const bool interior = li >= 1 && li <= FT_X && lj >= 1 && lj <= FT_Y;
            if (interior && fi0 + li - 1 < gf.nx && fj0 + lj - 1 < gf.ny)
                x2out[(size_t)(fj0 + lj - 1) * gf.P + fi0 + li - 1] = sx2[(lj + 1) * W2 + li + 1];
        }
        sr[k] = v;
    }
    __syncthreads();
    // restriction to the coarse tile (FT_X/2 x FT_Y/2), b_c = sum 0.25 P(f,I) r_f
    const int ci0 = fi0 / 2, cj0 = fj0 / 2;
    for (int k = tid; k < (FT_X / 2) * (FT_Y / 2); k += MG_THREADS) {
        int I = ci0 + k % (FT_X / 2), J = cj0 + k / (FT_X / 2);
        if (I >= gc.nx || J >= gc.ny) continue;
        double s = 0.0;
        for (int dy = -1; dy <= 2; ++dy) {
            int fy = 2 * J + dy, gfy;
            if (!gcoord(gf.ny, gf.py, fy, gfy)) continue;
            int I0, I1;
            double w0, w1;
            prol_1d_mg(gfy, gf.ny, gf.py, gf.db, gf.dt, I0, w0, I1, w1);
            double wy = (I0 == J ? w0 : 0.0) + (I1 == J ? w1 : 0.0);
            if (wy == 0.0) continue;
            for (int dx = -1; dx <= 2; ++dx) {
                int fx = 2 * I + dx, gfx;
                if (!gcoord(gf.nx, gf.px, fx, gfx)) continue;
                prol_1d_mg(gfx, gf.nx, gf.px, gf.dl, gf.dr, I0, w0, I1, w1);
                double wx = (I0 == I ? w0 : 0.0) + (I1 == I ? w1 : 0.0);
                if (wx == 0.0) continue;
                // local position of fine (fx, fy) in sr (halo 1)
                int li = fx - fi0 + 1, lj = fy - fj0 + 1;
                s += 0.25 * wx * wy * sr[lj * W1 + li];
            }
        }
        bc[(size_t)J * gc.P + I] = s;
    }
}

// ---------------------------------------------------------------------------
// up: x2 (fine), e_c (coarse) -> x5 (fine); on the finest level also the
// partial of r.z (b = r there) with the last block forming rz and beta.
// ---------------------------------------------------------------------------
constexpr int UH = 2;
constexpr int UW = FT_X + 2 * UH;   // 68
constexpr int UHh = FT_Y + 2 * UH;  // 20
constexpr int CW = FT_X / 2 + 4;    // coarse tile with halo 2: 36
constexpr int CH = FT_Y / 2 + 4;    // 12

__global__ void __launch_bounds__(MG_THREADS) k_mg_up(MgGeom gf, MgGeom gc, const double* __restrict__ x2,
                                                      const double* __restrict__ b, const double* __restrict__ ec,
                                                      double* __restrict__ xout, double omega, const int* done,
                                                      double* rz, int nb) {
    if (*done) return;
    __shared__ double sx3[UHh * UW];              // x3 on halo 2
    __shared__ double sx4[(UHh - 2) * (UW - 2)];  // x4 on halo 1
    __shared__ double sbb[(UHh - 2) * (UW - 2)];  // b on halo 1
    __shared__ double sec[CH * CW];               // e_c around the tile
    const int tid = threadIdx.x;
    const int fi0 = blockIdx.x * FT_X, fj0 = blockIdx.y * FT_Y;
    const int ci0 = fi0 / 2 - 2, cj0 = fj0 / 2 - 2;
    for (int k = tid; k < CH * CW; k += MG_THREADS) {
        int li = k % CW, lj = k / CW, gi, gj;
        double v = 0.0;
        if (gcoord(gc.nx, gc.px, ci0 + li, gi) && gcoord(gc.ny, gc.py, cj0 + lj, gj)) v = ec[(size_t)gj * gc.P + gi];
        sec[k] = v;
    }
    constexpr int W1 = UW - 2, H1 = UHh - 2;
    for (int k = tid; k < W1 * H1; k += MG_THREADS) {
        int li = k % W1, lj = k / W1, gi, gj;
        double v = 0.0;
        if (gcoord(gf.nx, gf.px, fi0 + li - 1, gi) && gcoord(gf.ny, gf.py, fj0 + lj - 1, gj))
            v = b[(size_t)gj * gf.P + gi];
        sbb[k] = v;
    }
    __syncthreads();
    // x3 = x2 + P e_c on halo 2
    for (int k = tid; k < UHh * UW; k += MG_THREADS) {
        int li = k % UW, lj = k / UW, gi, gj;
        double v = 0.0;
        if (gcoord(gf.nx, gf.px, fi0 + li - UH, gi) && gcoord(gf.ny, gf.py, fj0 + lj - UH, gj)) {
            int X0, X1, Y0, Y1;
            double wx0, wx1, wy0, wy1;
            prol_1d_mg(gi, gf.nx, gf.px, gf.dl, gf.dr, X0, wx0, X1, wx1);
            prol_1d_mg(gj, gf.ny, gf.py, gf.db, gf.dt, Y0, wy0, Y1, wy1);
            // parent / neighbour of the unwrapped fine coordinate, relative
            // to the coarse tile loaded at unwrapped coordinates ci0, cj0
            const int fx = fi0 + li - UH, fy = fj0 + lj - UH;
            const int pxc = (fx >= 0) ? fx / 2 : -((-fx + 1) / 2);
            const int pyc = (fy >= 0) ? fy / 2 : -((-fy + 1) / 2);
            const int ox = (fx & 1) ? pxc + 1 : pxc - 1;
            const int oy = (fy & 1) ? pyc + 1 : pyc - 1;
            const int a0 = pxc - ci0, a1 = (wx1 != 0.0) ? ox - ci0 : a0;
            const int b0 = pyc - cj0, b1 = (wy1 != 0.0) ? oy - cj0 : b0;
            (void)X0; (void)X1; (void)Y0; (void)Y1;
            double e = wy0 * (wx0 * sec[b0 * CW + a0] + wx1 * sec[b0 * CW + a1]) +
                       wy1 * (wx0 * sec[b1 * CW + a0] + wx1 * sec[b1 * CW + a1]);
            v = x2[(size_t)gj * gf.P + gi] + e;
        }
        sx3[k] = v;
    }
    __syncthreads();
    // x4 on halo 1
    for (int k = tid; k < W1 * H1; k += MG_THREADS) {
        int li = k % W1, lj = k / W1, gi, gj;
        double v = 0.0;
        if (gcoord(gf.nx, gf.px, fi0 + li - 1, gi) && gcoord(gf.ny, gf.py, fj0 + lj - 1, gj)) {
            Nb n = neighbours(gf, gi, gj);
            double d = nb_diag(gf, n);
            double ax = nb_apply(gf, n, sx3, UW, li + 1, lj + 1, d);
            v = sx3[(lj + 1) * UW + li + 1] + omega * (sbb[k] - ax) / d;
        }
        sx4[k] = v;
    }
    __syncthreads();
    // x5 on the interior, plus r.z on the finest level
    double acc = 0.0;
    for (int k = tid; k < FT_X * FT_Y; k += MG_THREADS) {
        int li = k % FT_X, lj = k / FT_X;
        int gi = fi0 + li, gj = fj0 + lj;
        if (gi >= gf.nx || gj >= gf.ny) continue;
        Nb n = neighbours(gf, gi, gj);
        double d = nb_diag(gf, n);
        double ax = nb_apply(gf, n, sx4, W1, li + 1, lj + 1, d);
        double bv = sbb[(lj + 1) * W1 + li + 1];
        double v = sx4[(lj + 1) * W1 + li + 1] + omega * (bv - ax) / d;
        xout[(size_t)gj * gf.P + gi] = v;
        acc += bv * v;
    }
    if (!rz) return;
    double v[1] = {acc};
    put_partials<1>(v, rz, nb);
}

void launch_mg_down(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const double* b, double* x2, double* bc,
                    double omega, const int* done, Status* st, const double* rr, int nb, int m, int slot) {
    dim3 g((gf.nx + FT_X - 1) / FT_X, (gf.ny + FT_Y - 1) / FT_Y);
    k_mg_down<<<g, MG_THREADS, 0, s>>>(gf, gc, b, x2, bc, omega, done, st, rr, nb, m, slot);
}

void launch_mg_up(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const double* x2, const double* b,
                  const double* ec, double* xout, double omega, const int* done, double* rz, int nb) {
    dim3 g((gf.nx + FT_X - 1) / FT_X, (gf.ny + FT_Y - 1) / FT_Y);
    k_mg_up<<<g, MG_THREADS, 0, s>>>(gf, gc, x2, b, ec, xout, omega, done, rz, nb);
}

}  // namespace fsi
