// fsi_mg.hip -- multigrid V-cycle of the projection preconditioner.
//
// Per fine level the V-cycle needs, from a zero initial guess,
//   down:  x1 = w1 D^-1 b;  x2 = x1 + w2 D^-1 (b - A x1);  b_c = R (b - A x2)
//   up:    x3 = x2 + P e_c;  x4 = x3 + w2 D^-1 (b - A x3);  x5 = x4 + w1 D^-1 (b - A x4)
// (A = minus the unit Laplacian, R = P^T/4, P bilinear).  The two smoothing
// weights are the reciprocal roots of the degree-2 Chebyshev polynomial on
// [1/3, 2] of the spectrum of D^-1 A (which lies in [0, 2]), so the smoother
// is a fixed polynomial in D^-1 A and the V-cycle stays symmetric.  Each is one kernel
// on an LDS tile with the halo the chained stencils need (3 cells down, 2
// up), so a level costs 2 launches, reads b once and writes x2, b_c and x5
// once; 256 threads work a tile of 64 x 16 cells on the finest level and of
// 16 x 16 or 16 x 8 on coarse ones.  The levels of at most coarse_max_cells() cells run in one workgroup
// (k_mg_coarse), whose bottom level is solved with a precomputed dense
// (pseudo-)inverse.
//
// The kernels are templated on the stored and computed floating type: FP64
// throughout, or (mixed precision) FP32 levels, tiles and arithmetic with the
// finest level reading the FP64 PCG residual and writing the FP64 z.
//
// Tiles hold 0 at points outside a non-periodic wall, so -A x at any cell is
// d x - (x_W + x_E)/hx^2 - (x_S + x_N)/hy^2 with the wall's contribution in
// the diagonal d alone; across a periodic side the tile holds the wrapped
// values.  1/d is the precomputed interior constant except on wall rows and
// columns.
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "fsi_kernels.hpp"

namespace fsi {

constexpr int MG_THREADS = 256;

// Loop over the N points of a tile phase with a compile-time trip count, so
// the unrolled iterations issue their global loads back to back.
#define TILE_LOOP(N)                                                               \
    _Pragma("unroll") for (int it_ = 0; it_ < ((N) + MG_THREADS - 1) / MG_THREADS; \
                           ++it_) if (const int k = tid + it_ * MG_THREADS; k < (N))

// diagonal of -A at a cell and its inverse
template <typename T>
struct DiagT {
    T d, id;
};
template <typename T>
__device__ __forceinline__ DiagT<T> mg_diag(const MgGeom& g, int gi, int gj) {
    const bool bx = !g.px && (gi == 0 || gi == g.nx - 1);
    const bool by = !g.py && (gj == 0 || gj == g.ny - 1);
    if (!(bx || by)) return {(T)g.d0, (T)g.id0};
    double cx = 2.0, cy = 2.0;
    if (!g.px) {
        if (gi == 0) cx += g.dl ? 1.0 : -1.0;
        if (gi == g.nx - 1) cx += g.dr ? 1.0 : -1.0;
    }
    if (!g.py) {
        if (gj == 0) cy += g.db ? 1.0 : -1.0;
        if (gj == g.ny - 1) cy += g.dt ? 1.0 : -1.0;
    }
    const double d = g.ihx2 * cx + g.ihy2 * cy;
    return {(T)d, (T)(1.0 / d)};
}

// tile coordinate -> global index, wrapped across a periodic side; false
// outside a non-periodic domain
__device__ __forceinline__ bool wrap_idx(int n, int per, int k, int& out) {
    if (k >= 0 && k < n) {
        out = k;
        return true;
    }
    if (!per) return false;
    out = k < 0 ? k + n : k - n;
    if (out < 0 || out >= n) out = ((k % n) + n) % n;
    return true;
}

// Where a tile sits.  IN = true: every point of the tile and its halo is a
// cell strictly inside the domain (no wall row or column, no periodic wrap),
// so indices are direct, the diagonal is the interior constant and the
// transfer weights are the interior ones; the kernels pick this variant per
// block (about 85 % of the finest-level tiles at N = 1024).
template <bool IN>
struct Where {
    __device__ static __forceinline__ bool locate(const MgGeom& g, int ux, int uy, int& gi, int& gj) {
        if (IN) {
            gi = ux;
            gj = uy;
            return true;
        }
        const bool okx = wrap_idx(g.nx, g.px, ux, gi);
        const bool oky = wrap_idx(g.ny, g.py, uy, gj);
        return okx && oky;
    }
    template <typename T>
    __device__ static __forceinline__ DiagT<T> diag(const MgGeom& g, int gi, int gj) {
        if (IN) return {(T)g.d0, (T)g.id0};
        return mg_diag<T>(g, gi, gj);
    }
};
// is the fine region [x0, x1] x [y0, y1] strictly inside (see Where<true>)?
__device__ __forceinline__ bool inside(const MgGeom& g, int x0, int x1, int y0, int y1) {
    const int lo_x = g.px ? 0 : 1, hi_x = g.px ? g.nx - 1 : g.nx - 2;
    const int lo_y = g.py ? 0 : 1, hi_y = g.py ? g.ny - 1 : g.ny - 2;
    return x0 >= lo_x && x1 <= hi_x && y0 >= lo_y && y1 <= hi_y;
}

// Global loads of the N points of a tile (row length W, origin (x0, y0) in
// unwrapped coordinates) into registers: indices first, then every load back
// to back, so a thread has all of its loads in flight at once.  Points
// outside a non-periodic wall read as 0; val[it] belongs to tile point
// tid + it * MG_THREADS (clamped to the last point).
template <int N, int W, typename TS>
struct TileGather {
    static constexpr int NIT = (N + MG_THREADS - 1) / MG_THREADS;
    TS val[NIT];
    int gi[NIT], gj[NIT];
    bool in[NIT];
    template <bool IN>
    __device__ __forceinline__ void load(const TS* __restrict__ a, const MgGeom& g, int x0, int y0, int tid) {
        int idx[NIT];
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int k = min(tid + it * MG_THREADS, N - 1);
            int i = 0, j = 0;
            const bool ok = Where<IN>::locate(g, x0 + k % W, y0 + k / W, i, j);
            in[it] = ok;
            gi[it] = i;
            gj[it] = j;
            idx[it] = ok ? j * g.P + i : 0;
        }
#pragma unroll
        for (int it = 0; it < NIT; ++it) val[it] = a[idx[it]];
        if (!IN) {
#pragma unroll
            for (int it = 0; it < NIT; ++it) val[it] = in[it] ? val[it] : TS(0);
        }
    }
};

// -A x at tile position c of an array with row length W (zero outside walls)
template <typename T>
__device__ __forceinline__ T tile_negA(const MgGeom& g, const T* a, int c, int W, T d) {
    return d * a[c] - (T)g.ihx2 * (a[c - 1] + a[c + 1]) - (T)g.ihy2 * (a[c - W] + a[c + W]);
}

// 1-D restriction weights of fine 2I and 2I+1 onto coarse I (R = P^T/4 up to
// the factor 1/4): 3/4 inside; at a non-periodic wall the fine cell's ghost
// is the mirrored parent (Neumann, weight 1) or its negative (Dirichlet, 1/2).
template <bool IN, typename T>
__device__ __forceinline__ void rweights(int I, int nf, int per, int dlo, int dhi, T& w0, T& w1) {
    w0 = T(0.75);
    w1 = T(0.75);
    if (!IN && !per) {
        if (2 * I == 0) w0 = dlo ? T(0.5) : T(1);
        if (2 * I + 1 == nf - 1) w1 = dhi ? T(0.5) : T(1);
    }
}

// ---------------------------------------------------------------------------
// down: b (fine) -> x2 (fine, stored) and b_c (coarse)
// ---------------------------------------------------------------------------
template <int FX, int FY>
struct DownTile {
    static constexpr int D3W = FX + 6, D3H = FY + 6;  // halo 3
    static constexpr int D2W = FX + 4, D2H = FY + 4;  // halo 2
    static constexpr int D1W = FX + 2, D1H = FY + 2;  // halo 1
    static constexpr int CTX = FX / 2, CTY = FY / 2;  // coarse tile
    static_assert(D1H * CTX <= D2H * D2W, "row sums fit the x2 buffer");
};

template <int FX, int FY, bool IN, typename TB, typename T>
__device__ __forceinline__ void down_tile(const MgGeom& gf, const MgGeom& gc, const TB* __restrict__ b,
                                          T* __restrict__ x2out, T* __restrict__ bc, int fi0, int fj0, int tid, T* sb,
                                          T* sx, T* s2) {
    using DT = DownTile<FX, FY>;
    constexpr int D3W = DT::D3W, D3H = DT::D3H, D2W = DT::D2W, D2H = DT::D2H, D1W = DT::D1W, D1H = DT::D1H;
    constexpr int CTX = DT::CTX, CTY = DT::CTY;
    const T om1 = (T)gf.om1, om2 = (T)gf.om2;
    {  // b and x1 = w1 D^-1 b on halo 3
        TileGather<D3H * D3W, D3W, TB> tb;
        tb.template load<IN>(b, gf, fi0 - 3, fj0 - 3, tid);
#pragma unroll
        for (int it = 0; it < tb.NIT; ++it) {
            const int k = tid + it * MG_THREADS;
            if (k >= D3H * D3W) break;
            const T bv = (T)tb.val[it];
            sb[k] = bv;
            sx[k] = tb.in[it] ? om1 * Where<IN>::template diag<T>(gf, tb.gi[it], tb.gj[it]).id * bv : T(0);
        }
    }
    __syncthreads();
    // x2 = x1 + w2 D^-1 (b - A x1) on halo 2
    TILE_LOOP(D2H * D2W) {
        const int li = k % D2W, lj = k / D2W;
        int gi, gj;
        T v = T(0);
        if (Where<IN>::locate(gf, fi0 + li - 2, fj0 + lj - 2, gi, gj)) {
            const DiagT<T> D = Where<IN>::template diag<T>(gf, gi, gj);
            const int c = (lj + 1) * D3W + li + 1;
            v = sx[c] + om2 * D.id * (sb[c] - tile_negA<T>(gf, sx, c, D3W, D.d));
        }
        s2[k] = v;
    }
    __syncthreads();
    // r = b - A x2 on halo 1 (into sx), x2 stored on the tile
    TILE_LOOP(D1H * D1W) {
        const int li = k % D1W, lj = k / D1W;
        const int ui = fi0 + li - 1, uj = fj0 + lj - 1;
        int gi, gj;
        T v = T(0);
        if (Where<IN>::locate(gf, ui, uj, gi, gj)) {
            const DiagT<T> D = Where<IN>::template diag<T>(gf, gi, gj);
            const int c = (lj + 1) * D2W + li + 1;
            v = sb[(lj + 2) * D3W + li + 2] - tile_negA<T>(gf, s2, c, D2W, D.d);
            if (li >= 1 && li <= FX && lj >= 1 && lj <= FY && (IN || (ui < gf.nx && uj < gf.ny)))
                x2out[(size_t)uj * gf.P + ui] = s2[c];
        }
        sx[k] = v;
    }
    __syncthreads();
    // restriction, x then y: row sums of r over fine 2I-1 .. 2I+2
    T* sh = s2;
    const T q = T(0.25);
    TILE_LOOP(D1H * CTX) {
        const int ci = k % CTX, lj = k / CTX;
        T w0, w1;
        rweights<IN, T>(fi0 / 2 + ci, gf.nx, gf.px, gf.dl, gf.dr, w0, w1);
        const T* row = sx + lj * D1W + 2 * ci;  // fine 2I-1 at halo-1 index 2 ci
        sh[lj * CTX + ci] = q * row[0] + w0 * row[1] + w1 * row[2] + q * row[3];
    }
    __syncthreads();
    TILE_LOOP(CTY * CTX) {
        const int ci = k % CTX, cj = k / CTX;
        const int I = fi0 / 2 + ci, J = fj0 / 2 + cj;
        if (!IN && (I >= gc.nx || J >= gc.ny)) continue;
        T w0, w1;
        rweights<IN, T>(J, gf.ny, gf.py, gf.db, gf.dt, w0, w1);
        const T* col = sh + 2 * cj * CTX + ci;
        bc[(size_t)J * gc.P + I] = q * (q * col[0] + w0 * col[CTX] + w1 * col[2 * CTX] + q * col[3 * CTX]);
    }
}

// Tiles are FX x FY fine cells (64 x 16 on the finest level, whose partials
// must line up with the PCG kernels' grid; smaller on coarser levels so that
// they still spread over many CUs).
template <int FX, int FY, typename TB, typename T>
FSI_KERNEL void k_mg_down(MgGeom gf, MgGeom gc, const TB* __restrict__ b,
                                                        T* __restrict__ x2out, T* __restrict__ bc, const int* done,
                                                        Status* st, const double* __restrict__ rr, int nb, int m,
                                                        int slot) {
    using DT = DownTile<FX, FY>;
    if (*done) return;
    if (st) {  // finest level: convergence test on |r|^2 of the previous PCG kernel
        double tot[1];
        get_totals<1>(rr, nb, tot);
        if (tot[0] <= st->pcg.tol2 || m >= st->pcg.maxiter || !(tot[0] <= 1.0e300)) {
            if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
                st->pcg.done = 1;
                st->pcg.iter = m;
                st->pcg.rr = tot[0];
                if (slot >= 0) st->proj_iters[slot] = m;
                if (slot >= 0 && tot[0] <= st->pcg.tol2) st->proj_ok[slot] = 1;
            }
            return;
        }
    }
    __shared__ T sb[DT::D3H * DT::D3W];  // b on halo 3
    __shared__ T sx[DT::D3H * DT::D3W];  // x1 on halo 3, then r on halo 1
    __shared__ T s2[DT::D2H * DT::D2W];  // x2 on halo 2, then row sums of the restriction
    const int fi0 = tile_id().x * FX, fj0 = tile_id().y * FY;
    if (inside(gf, fi0 - 3, fi0 + FX + 2, fj0 - 3, fj0 + FY + 2))
        down_tile<FX, FY, true, TB, T>(gf, gc, b, x2out, bc, fi0, fj0, threadIdx.x, sb, sx, s2);
    else
        down_tile<FX, FY, false, TB, T>(gf, gc, b, x2out, bc, fi0, fj0, threadIdx.x, sb, sx, s2);
}

// ---------------------------------------------------------------------------
// up: x2 (fine), e_c (coarse) -> x5 (fine); on the finest level also the
// block partials of r.z (b = r there).
// ---------------------------------------------------------------------------

// bilinear weights of the parent (w0) and the nearer coarse neighbour (1/4)
// of fine f; the tile stores 0 beyond a non-periodic wall, where the ghost
// is the mirrored parent (Neumann, w0 = 1) or its negative (Dirichlet, 1/2).
template <bool IN, typename T>
__device__ __forceinline__ void pweights(int f, int nf, int per, int dlo, int dhi, int& parent, int& other, T& w0) {
    parent = f >> 1;  // floor for negative tile coordinates as well
    other = (f & 1) ? parent + 1 : parent - 1;
    w0 = T(0.75);
    if (!IN && !per) {
        if (f == 0) w0 = dlo ? T(0.5) : T(1);
        if (f == nf - 1) w0 = dhi ? T(0.5) : T(1);
    }
}

template <int FX, int FY>
struct UpTile {
    static constexpr int D2W = FX + 4, D2H = FY + 4;         // halo 2
    static constexpr int D1W = FX + 2, D1H = FY + 2;         // halo 1
    static constexpr int CW = FX / 2 + 4, CH = FY / 2 + 4;   // coarse values around the tile
};

// sbc (TB, the tile's own cells) keeps b at its stored precision for r.z
// when the tile works in a lower one.
template <int FX, int FY, bool IN, typename TB, typename T, typename TO>
__device__ __forceinline__ double up_tile(const MgGeom& gf, const MgGeom& gc, const T* __restrict__ x2,
                                          const TB* __restrict__ b, const T* __restrict__ ec, TO* __restrict__ xout,
                                          int fi0, int fj0, int tid, T* sx3, T* sx4, T* sbb, T* sec, TB* sbc) {
    using UT = UpTile<FX, FY>;
    constexpr int D2W = UT::D2W, D2H = UT::D2H, D1W = UT::D1W, D1H = UT::D1H, CW = UT::CW, CH = UT::CH;
    constexpr bool kSplit = !std::is_same<TB, T>::value;
    const T om1 = (T)gf.om1, om2 = (T)gf.om2, q = T(0.25);
    const int ci0 = fi0 / 2 - 2, cj0 = fj0 / 2 - 2;
    {
        TileGather<CH * CW, CW, T> tc;
        TileGather<D1H * D1W, D1W, TB> tbb;
        TileGather<D2H * D2W, D2W, T> tx2;
        tc.template load<IN>(ec, gc, ci0, cj0, tid);
        tbb.template load<IN>(b, gf, fi0 - 1, fj0 - 1, tid);
        tx2.template load<IN>(x2, gf, fi0 - 2, fj0 - 2, tid);
#pragma unroll
        for (int it = 0; it < tc.NIT; ++it) {
            const int k = tid + it * MG_THREADS;
            if (k < CH * CW) sec[k] = tc.val[it];
        }
#pragma unroll
        for (int it = 0; it < tbb.NIT; ++it) {
            const int k = tid + it * MG_THREADS;
            if (k < D1H * D1W) {
                sbb[k] = (T)tbb.val[it];
                if (kSplit) {
                    const int li = k % D1W, lj = k / D1W;
                    if (li >= 1 && li <= FX && lj >= 1 && lj <= FY) sbc[(lj - 1) * FX + li - 1] = tbb.val[it];
                }
            }
        }
        __syncthreads();
        // x3 = x2 + P e_c on halo 2
#pragma unroll
        for (int it = 0; it < tx2.NIT; ++it) {
            const int k = tid + it * MG_THREADS;
            if (k >= D2H * D2W) break;
            const int li = k % D2W, lj = k / D2W;
            const int fx = fi0 + li - 2, fy = fj0 + lj - 2;
            T v = T(0);
            if (tx2.in[it]) {
                int px0, px1, py0, py1;
                T wx0, wy0;
                pweights<IN, T>(fx, gf.nx, gf.px, gf.dl, gf.dr, px0, px1, wx0);
                pweights<IN, T>(fy, gf.ny, gf.py, gf.db, gf.dt, py0, py1, wy0);
                const T* r0 = sec + (py0 - cj0) * CW;
                const T* r1 = sec + (py1 - cj0) * CW;
                const int a0 = px0 - ci0, a1 = px1 - ci0;
                const T e = wy0 * (wx0 * r0[a0] + q * r0[a1]) + q * (wx0 * r1[a0] + q * r1[a1]);
                v = tx2.val[it] + e;
            }
            sx3[k] = v;
        }
    }
    __syncthreads();
    // x4 on halo 1
    TILE_LOOP(D1H * D1W) {
        const int li = k % D1W, lj = k / D1W;
        int gi, gj;
        T v = T(0);
        if (Where<IN>::locate(gf, fi0 + li - 1, fj0 + lj - 1, gi, gj)) {
            const DiagT<T> D = Where<IN>::template diag<T>(gf, gi, gj);
            const int c = (lj + 1) * D2W + li + 1;
            v = sx3[c] + om2 * D.id * (sbb[k] - tile_negA<T>(gf, sx3, c, D2W, D.d));
        }
        sx4[k] = v;
    }
    __syncthreads();
    // x5 on the tile, and this thread's part of r.z
    double acc = 0.0;
    TILE_LOOP(FX * FY) {
        const int li = k % FX, lj = k / FX;
        const int gi = fi0 + li, gj = fj0 + lj;
        if (!IN && (gi >= gf.nx || gj >= gf.ny)) continue;
        const DiagT<T> D = Where<IN>::template diag<T>(gf, gi, gj);
        const int c = (lj + 1) * D1W + li + 1;
        const T bv = sbb[c];
        const T v = sx4[c] + om1 * D.id * (bv - tile_negA<T>(gf, sx4, c, D1W, D.d));
        xout[(size_t)gj * gf.P + gi] = (TO)v;
        const double bd = kSplit ? (double)sbc[lj * FX + li] : (double)bv;
        acc += bd * (double)(TO)v;
    }
    return acc;
}

template <int FX, int FY, typename TB, typename T, typename TO>
FSI_KERNEL void k_mg_up(MgGeom gf, MgGeom gc, const T* __restrict__ x2,
                                                      const TB* __restrict__ b, const T* __restrict__ ec,
                                                      TO* __restrict__ xout, const int* done, double* rz, int nb) {
    using UT = UpTile<FX, FY>;
    constexpr bool kSplit = !std::is_same<TB, T>::value;
    if (*done) return;
    __shared__ T sx3[UT::D2H * UT::D2W];         // x3 on halo 2
    __shared__ T sx4[UT::D1H * UT::D1W];         // x4 on halo 1
    __shared__ T sbb[UT::D1H * UT::D1W];         // b on halo 1
    __shared__ T sec[UT::CH * UT::CW];           // e_c around the tile
    __shared__ TB sbc[kSplit ? FX * FY : 1];     // b of the tile's cells at its stored precision
    const int fi0 = tile_id().x * FX, fj0 = tile_id().y * FY;
    const int ci0 = fi0 / 2 - 2, cj0 = fj0 / 2 - 2;
    const bool in = inside(gf, fi0 - 2, fi0 + FX + 1, fj0 - 2, fj0 + FY + 1) && ci0 >= 0 && cj0 >= 0 &&
                    ci0 + UT::CW <= gc.nx && cj0 + UT::CH <= gc.ny;
    const double acc =
        in ? up_tile<FX, FY, true, TB, T, TO>(gf, gc, x2, b, ec, xout, fi0, fj0, threadIdx.x, sx3, sx4, sbb, sec, sbc)
           : up_tile<FX, FY, false, TB, T, TO>(gf, gc, x2, b, ec, xout, fi0, fj0, threadIdx.x, sx3, sx4, sbb, sec, sbc);
    if (!rz) return;
    double v[1] = {acc};
    put_partials<1>(v, rz, nb);
}

// ---------------------------------------------------------------------------
// Coarse levels in one workgroup: a top level of at most coarse_max_cells() cells,
// halving while both sizes stay even, down to a bottom of at most
// CO_DENSE_MAX cells solved with its dense (pseudo-)inverse (otherwise:
// CO_SWEEPS pairs of damped-Jacobi sweeps).  Level l keeps x, t and b (n_l
// doubles each) consecutively in LDS.
// ---------------------------------------------------------------------------
constexpr int CO_THREADS = 1024;
constexpr int CO_SWEEPS = 64;

__device__ __forceinline__ MgGeom co_geom(const MgGeom& g0, int l) {
    MgGeom g = g0;
    g.nx = g0.nx >> l;
    g.ny = g0.ny >> l;
    g.P = g.nx;
    const double s = 1.0 / (double)(1 << (2 * l));
    g.ihx2 *= s;
    g.ihy2 *= s;
    g.d0 *= s;
    g.id0 /= s;
    return g;
}

// -A a at (i, j) of a compact level array, with its diagonal
template <typename T>
__device__ __forceinline__ T co_negA(const MgGeom& g, const T* a, int i, int j, T& d) {
    const int k = j * g.nx + i;
    T sx = T(0), sy = T(0), cx = T(0), cy = T(0);
    if (i > 0) { sx += a[k - 1]; cx += T(1); }
    else if (g.px) { sx += a[k + g.nx - 1]; cx += T(1); }
    else if (g.dl) cx += T(2);
    if (i < g.nx - 1) { sx += a[k + 1]; cx += T(1); }
    else if (g.px) { sx += a[k - g.nx + 1]; cx += T(1); }
    else if (g.dr) cx += T(2);
    if (j > 0) { sy += a[k - g.nx]; cy += T(1); }
    else if (g.py) { sy += a[k + (g.ny - 1) * g.nx]; cy += T(1); }
    else if (g.db) cy += T(2);
    if (j < g.ny - 1) { sy += a[k + g.nx]; cy += T(1); }
    else if (g.py) { sy += a[k - (g.ny - 1) * g.nx]; cy += T(1); }
    else if (g.dt) cy += T(2);
    const T ihx2 = (T)g.ihx2, ihy2 = (T)g.ihy2;
    d = ihx2 * cx + ihy2 * cy;
    return d * a[k] - ihx2 * sx - ihy2 * sy;
}
template <typename T>
__device__ __forceinline__ T co_id(const MgGeom& g, T d) {
    return d == (T)g.d0 ? (T)g.id0 : T(1) / d;
}
constexpr double CO_OMEGA = 0.8;  // damped-Jacobi weight of the fallback bottom sweeps

// restriction weight of fine f = 2I + dx onto coarse I, with f wrapped
__device__ __forceinline__ double co_rw(int I, int dx, int nf, int per, int dlo, int dhi, int& f) {
    f = 2 * I + dx;
    if (f < 0 || f >= nf) {
        if (!per) return 0.0;
        f = f < 0 ? f + nf : f - nf;
        return 0.25;
    }
    if (dx == -1 || dx == 2) return 0.25;
    if (!per) {
        if (f == 0) return dlo ? 0.5 : 1.0;
        if (f == nf - 1) return dhi ? 0.5 : 1.0;
    }
    return 0.75;
}

// prolongation of fine f: parent I0 (w0), nearer neighbour I1 (w1), wrapped
template <typename T>
__device__ __forceinline__ void co_pw(int f, int nf, int per, int dlo, int dhi, int& I0, T& w0, int& I1, T& w1) {
    const int nc = nf >> 1;
    I0 = f >> 1;
    I1 = (f & 1) ? I0 + 1 : I0 - 1;
    w0 = T(0.75);
    w1 = T(0.25);
    if (I1 < 0 || I1 >= nc) {
        if (per) I1 = I1 < 0 ? I1 + nc : I1 - nc;
        else {
            w0 = ((I1 < 0 ? dlo : dhi) ? T(0.5) : T(1));
            I1 = I0;
            w1 = T(0);
        }
    }
}

#define CO_FOR(g) \
    for (int j = ty; j < (g).ny; j += 32) \
        for (int i = tx; i < (g).nx; i += 32)

template <typename T>
FSI_KERNEL void k_mg_coarse(MgGeom g0, int levels, const T* __restrict__ dense,
                                                          int dense_n, const T* __restrict__ bglob,
                                                          T* __restrict__ tglob, const int* done) {
    extern __shared__ __align__(16) unsigned char co_smem[];
    T* s = reinterpret_cast<T*>(co_smem);
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    // this thread's entries of the dense bottom inverse (row tid/16, columns
    // tid%16 + 16q), loaded now so they are in registers at the bottom
    const int drow = threadIdx.x >> 4, dlane = threadIdx.x & 15;
    T dm[CO_DENSE_MAX / 16];
#pragma unroll
    for (int q = 0; q < CO_DENSE_MAX / 16; ++q) {
        const int c = dlane + 16 * q;
        dm[q] = (drow < dense_n && c < dense_n) ? dense[drow * dense_n + c] : T(0);
    }
    {  // b of the top level (loads issued with the done-flag read) and x = w1 D^-1 b
        const int finished = *done;
        const MgGeom g = co_geom(g0, 0);
        const int n = g.nx * g.ny;
        T* x = s;
        T* bb = s + 2 * n;
        T bv0 = T(0);
        if (ty < g.ny && tx < g.nx) bv0 = bglob[(size_t)ty * g0.P + tx];
        if (finished) return;
        const T om1 = (T)g.om1;
        CO_FOR(g) {
            const T bv = (i == tx && j == ty) ? bv0 : bglob[(size_t)j * g0.P + i];
            const DiagT<T> D = mg_diag<T>(g, i, j);
            bb[j * g.nx + i] = bv;
            x[j * g.nx + i] = om1 * D.id * bv;
        }
    }
    __syncthreads();
    int off = 0;
    for (int l = 0; l < levels - 1; ++l) {
        const MgGeom g = co_geom(g0, l), gc = co_geom(g0, l + 1);
        const int n = g.nx * g.ny, nc = gc.nx * gc.ny;
        const T om2 = (T)g.om2;
        T* x = s + off;
        T* t = x + n;
        T* bb = x + 2 * n;
        CO_FOR(g) {  // t = x + w2 D^-1 (b - A x)
            T d;
            const T ax = co_negA<T>(g, x, i, j, d);
            const int k = j * g.nx + i;
            t[k] = x[k] + om2 * co_id<T>(g, d) * (bb[k] - ax);
        }
        __syncthreads();
        CO_FOR(g) {  // r = b - A t, into x
            T d;
            const T at = co_negA<T>(g, t, i, j, d);
            x[j * g.nx + i] = bb[j * g.nx + i] - at;
        }
        __syncthreads();
        T* xc = x + 3 * n;
        T* bc = xc + 2 * nc;
        CO_FOR(gc) {  // b_c = R r, x_c = w1 D^-1 b_c
            T acc = T(0);
            for (int dy = -1; dy <= 2; ++dy) {
                int fy;
                const T wy = (T)co_rw(j, dy, g.ny, g.py, g.db, g.dt, fy);
                if (wy == T(0)) continue;
                T row = T(0);
                for (int dx = -1; dx <= 2; ++dx) {
                    int fx;
                    const T wx = (T)co_rw(i, dx, g.nx, g.px, g.dl, g.dr, fx);
                    if (wx != T(0)) row += wx * x[fy * g.nx + fx];
                }
                acc += wy * row;
            }
            const T bv = T(0.25) * acc;
            const DiagT<T> D = mg_diag<T>(gc, i, j);
            bc[j * gc.nx + i] = bv;
            xc[j * gc.nx + i] = (T)gc.om1 * D.id * bv;
        }
        __syncthreads();
        off += 3 * n;
    }
    {  // bottom
        const MgGeom g = co_geom(g0, levels - 1);
        const int n = g.nx * g.ny;
        T* x = s + off;
        T* t = x + n;
        T* bb = x + 2 * n;
        if (dense_n) {  // t = M b, 16 lanes per row
            T acc = T(0);
#pragma unroll
            for (int q = 0; q < CO_DENSE_MAX / 16; ++q) {
                const int c = dlane + 16 * q;
                if (c < n) acc += dm[q] * bb[c];
            }
            for (int msk = 8; msk >= 1; msk >>= 1) acc += __shfl_xor(acc, msk, 16);
            if (drow < n && dlane == 0) t[drow] = acc;
            __syncthreads();
        } else {
            for (int it = 0; it <= 2 * CO_SWEEPS; ++it) {
                T* in = (it & 1) ? t : x;
                T* out = (it & 1) ? x : t;
                CO_FOR(g) {
                    T d;
                    const T ax = co_negA<T>(g, in, i, j, d);
                    const int k = j * g.nx + i;
                    out[k] = in[k] + (T)CO_OMEGA * co_id<T>(g, d) * (bb[k] - ax);
                }
                __syncthreads();
            }
        }
    }
    // up: t_l += P t_{l+1}; x = t + w2 D^-1 (b - A t); t = x + w1 D^-1 (b - A x)
    for (int l = levels - 2; l >= 0; --l) {
        const MgGeom g = co_geom(g0, l), gc = co_geom(g0, l + 1);
        const int n = g.nx * g.ny;
        const T om1 = (T)g.om1, om2 = (T)g.om2;
        off -= 3 * n;
        T* x = s + off;
        T* t = x + n;
        T* bb = x + 2 * n;
        const T* tc = x + 3 * n + gc.nx * gc.ny;
        CO_FOR(g) {
            int X0, X1, Y0, Y1;
            T wx0, wx1, wy0, wy1;
            co_pw<T>(i, g.nx, g.px, g.dl, g.dr, X0, wx0, X1, wx1);
            co_pw<T>(j, g.ny, g.py, g.db, g.dt, Y0, wy0, Y1, wy1);
            t[j * g.nx + i] += wy0 * (wx0 * tc[Y0 * gc.nx + X0] + wx1 * tc[Y0 * gc.nx + X1]) +
                               wy1 * (wx0 * tc[Y1 * gc.nx + X0] + wx1 * tc[Y1 * gc.nx + X1]);
        }
        __syncthreads();
        CO_FOR(g) {
            T d;
            const T at = co_negA<T>(g, t, i, j, d);
            const int k = j * g.nx + i;
            x[k] = t[k] + om2 * co_id<T>(g, d) * (bb[k] - at);
        }
        __syncthreads();
        CO_FOR(g) {
            T d;
            const T ax = co_negA<T>(g, x, i, j, d);
            const int k = j * g.nx + i;
            const T v = x[k] + om1 * co_id<T>(g, d) * (bb[k] - ax);
            if (l == 0) tglob[(size_t)j * g0.P + i] = v;
            else t[k] = v;
        }
        if (l > 0) __syncthreads();
    }
    if (levels == 1) {  // the top level is the bottom
        const int n = g0.nx * g0.ny;
        const T* t = s + n;
        CO_FOR(g0) tglob[(size_t)j * g0.P + i] = t[j * g0.nx + i];
    }
}

// ---- host side -------------------------------------------------------------

// -A of a level as a dense row-major matrix (the operator of co_negA)
static std::vector<double> dense_operator(const MgGeom& g) {
    const int nx = g.nx, ny = g.ny, n = nx * ny;
    std::vector<double> A((size_t)n * n, 0.0);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int k = j * nx + i;
            double* row = &A[(size_t)k * n];
            double cx = 0.0, cy = 0.0;
            auto nb = [&](int ii, int jj, double h) { row[jj * nx + ii] -= h; };
            if (i > 0) { nb(i - 1, j, g.ihx2); cx += 1.0; }
            else if (g.px) { nb(nx - 1, j, g.ihx2); cx += 1.0; }
            else if (g.dl) cx += 2.0;
            if (i < nx - 1) { nb(i + 1, j, g.ihx2); cx += 1.0; }
            else if (g.px) { nb(0, j, g.ihx2); cx += 1.0; }
            else if (g.dr) cx += 2.0;
            if (j > 0) { nb(i, j - 1, g.ihy2); cy += 1.0; }
            else if (g.py) { nb(i, ny - 1, g.ihy2); cy += 1.0; }
            else if (g.db) cy += 2.0;
            if (j < ny - 1) { nb(i, j + 1, g.ihy2); cy += 1.0; }
            else if (g.py) { nb(i, 0, g.ihy2); cy += 1.0; }
            else if (g.dt) cy += 2.0;
            row[k] += g.ihx2 * cx + g.ihy2 * cy;
        }
    return A;
}

// Inverse of -A at the bottom level.  With only Neumann/periodic sides -A is
// singular with the constants as null space e, and its pseudo-inverse is
// (A + a e e^T)^-1 - e e^T / (a n^2).  Gauss-Jordan with partial pivoting.
static bool dense_inverse(const MgGeom& g, std::vector<double>& M) {
    const int n = g.nx * g.ny;
    std::vector<double> A = dense_operator(g);
    const bool singular = (g.px || (!g.dl && !g.dr)) && (g.py || (!g.db && !g.dt));
    const double a = g.d0 / n;
    if (singular)
        for (double& v : A) v += a;
    M.assign((size_t)n * n, 0.0);
    for (int k = 0; k < n; ++k) M[(size_t)k * n + k] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)piv * n + c])) piv = r;
        if (A[(size_t)piv * n + c] == 0.0) return false;
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(A[(size_t)piv * n + k], A[(size_t)c * n + k]);
                std::swap(M[(size_t)piv * n + k], M[(size_t)c * n + k]);
            }
        const double inv = 1.0 / A[(size_t)c * n + c];
        for (int k = 0; k < n; ++k) {
            A[(size_t)c * n + k] *= inv;
            M[(size_t)c * n + k] *= inv;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = A[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
                M[(size_t)r * n + k] -= f * M[(size_t)c * n + k];
            }
        }
    }
    if (singular) {
        const double e = 1.0 / (a * (double)n * (double)n);
        for (double& v : M) v -= e;
    }
    return true;
}

bool coarse_plan(const MgGeom& top, int max_cells, CoarsePlan& cp, std::vector<double>& dense) {
    cp = CoarsePlan{};
    dense.clear();
    if ((long)top.nx * top.ny > max_cells) return false;
    int levels = 1, nx = top.nx, ny = top.ny;
    size_t cells = (size_t)nx * ny;
    while ((long)nx * ny > CO_DENSE_MAX && !(nx & 1) && !(ny & 1) && nx >= 4 && ny >= 4) {
        nx >>= 1;
        ny >>= 1;
        cells += (size_t)nx * ny;
        ++levels;
    }
    cp.levels = levels;
    cp.lds_cells = 3 * cells;
    if ((long)nx * ny <= CO_DENSE_MAX) {
        MgGeom gb = top;
        gb.nx = nx;
        gb.ny = ny;
        const double sc = 1.0 / (double)(1 << (2 * (levels - 1)));
        gb.ihx2 *= sc;
        gb.ihy2 *= sc;
        gb.d0 *= sc;
        gb.id0 /= sc;
        if (dense_inverse(gb, dense)) cp.dense_n = nx * ny;
        else dense.clear();
    }
    return true;
}

int coarse_max_cells() {
    // 1024 cells: measured faster on the 1024^2 cavity than a 4096-cell top
    // level (one CU then works 4 cells per thread in every phase, slower
    // than two extra level launches).  FSI_MG_COARSE_CELLS=4096 selects the
    // larger top level, which needs 3 x 8 x (4096 + 1024 + 256 + 64) bytes of
    // LDS in one workgroup.
    const char* e = std::getenv("FSI_MG_COARSE_CELLS");
    if (!e || std::atoi(e) != 4096) return 1024;
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 1024;
    const size_t need = 3 * sizeof(double) * (4096 + 1024 + 256 + 64);
    if ((size_t)lds < need) return 1024;
    if (fsi_allow_dynamic_lds<&k_mg_coarse<double>, CO_THREADS>((int)need) != hipSuccess ||
        fsi_allow_dynamic_lds<&k_mg_coarse<float>, CO_THREADS>((int)need) != hipSuccess) {
        (void)hipGetLastError();
        return 1024;
    }
    return 4096;
}

void launch_mg_coarse(hipStream_t s, const MgGeom& top, const CoarsePlan& cp, const double* b, double* t,
                      const int* done) {
    fsi_launch<&k_mg_coarse<double>, CO_THREADS, 0>(1, CO_THREADS, cp.lds_cells * sizeof(double), s, top, cp.levels, cp.dense, cp.dense_n,
                                                                             b, t, done);
}
void launch_mg_coarse(hipStream_t s, const MgGeom& top, const CoarsePlan& cp, const float* b, float* t,
                      const int* done) {
    fsi_launch<&k_mg_coarse<float>, CO_THREADS, 0>(1, CO_THREADS, cp.lds_cells * sizeof(float), s, top, cp.levels, cp.dense_f, cp.dense_n,
                                                                           b, t, done);
}

// tile of a level: 64 x 16 on the finest (rz/rr partials follow the PCG
// grid), else the largest of 64 x 16, 16 x 16, 16 x 8 that still gives 1024
// blocks (four per CU), or 16 x 8
static int level_tile(const MgGeom& g, bool finest) {
    if (finest) return 0;
    auto blocks = [&](int fx, int fy) { return (long)((g.nx + fx - 1) / fx) * ((g.ny + fy - 1) / fy); };
    if (blocks(64, 16) >= 1024) return 0;
    if (blocks(16, 16) >= 1024) return 1;
    return 2;
}

template <int FX, int FY, typename TB, typename T>
static void down_t(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const TB* b, T* x2, T* bc, const int* done,
                   Status* st, const double* rr, int nb, int m, int slot) {
    dim3 g((gf.nx + FX - 1) / FX, (gf.ny + FY - 1) / FY);
    fsi_launch<&k_mg_down<FX, FY, TB, T>, MG_THREADS, 0>(g, MG_THREADS, 0, s, gf, gc, b, x2, bc, done, st, rr, nb, m, slot);
}
template <int FX, int FY, typename TB, typename T, typename TO>
static void up_t(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const T* x2, const TB* b, const T* ec, TO* xout,
                 const int* done, double* rz, int nb) {
    dim3 g((gf.nx + FX - 1) / FX, (gf.ny + FY - 1) / FY);
    fsi_launch<&k_mg_up<FX, FY, TB, T, TO>, MG_THREADS, 0>(g, MG_THREADS, 0, s, gf, gc, x2, b, ec, xout, done, rz, nb);
}

template <typename TB, typename T>
static void down_any(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const TB* b, T* x2, T* bc, const int* done,
                     Status* st, const double* rr, int nb, int m, int slot) {
    switch (level_tile(gf, st != nullptr)) {
        case 0: down_t<64, 16>(s, gf, gc, b, x2, bc, done, st, rr, nb, m, slot); break;
        case 1: down_t<16, 16>(s, gf, gc, b, x2, bc, done, st, rr, nb, m, slot); break;
        default: down_t<16, 8>(s, gf, gc, b, x2, bc, done, st, rr, nb, m, slot); break;
    }
}
template <typename TB, typename T, typename TO>
static void up_any(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const T* x2, const TB* b, const T* ec,
                   TO* xout, const int* done, double* rz, int nb) {
    switch (level_tile(gf, rz != nullptr)) {
        case 0: up_t<64, 16>(s, gf, gc, x2, b, ec, xout, done, rz, nb); break;
        case 1: up_t<16, 16>(s, gf, gc, x2, b, ec, xout, done, rz, nb); break;
        default: up_t<16, 8>(s, gf, gc, x2, b, ec, xout, done, rz, nb); break;
    }
}

void launch_mg_down(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const double* b, double* x2, double* bc,
                    const int* done, Status* st, const double* rr, int nb, int m, int slot) {
    down_any(s, gf, gc, b, x2, bc, done, st, rr, nb, m, slot);
}
void launch_mg_down(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const double* b, float* x2, float* bc,
                    const int* done, Status* st, const double* rr, int nb, int m, int slot) {
    down_any(s, gf, gc, b, x2, bc, done, st, rr, nb, m, slot);
}
void launch_mg_down(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const float* b, float* x2, float* bc,
                    const int* done, Status* st, const double* rr, int nb, int m, int slot) {
    down_any(s, gf, gc, b, x2, bc, done, st, rr, nb, m, slot);
}
void launch_mg_up(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const double* x2, const double* b,
                  const double* ec, double* xout, const int* done, double* rz, int nb) {
    up_any(s, gf, gc, x2, b, ec, xout, done, rz, nb);
}
void launch_mg_up(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const float* x2, const double* b,
                  const float* ec, double* xout, const int* done, double* rz, int nb) {
    up_any(s, gf, gc, x2, b, ec, xout, done, rz, nb);
}
void launch_mg_up(hipStream_t s, const MgGeom& gf, const MgGeom& gc, const float* x2, const float* b,
                  const float* ec, float* xout, const int* done, double* rz, int nb) {
    up_any(s, gf, gc, x2, b, ec, xout, done, rz, nb);
}

}  // namespace fsi
