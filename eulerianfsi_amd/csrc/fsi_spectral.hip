// fsi_spectral.hip -- direct solve of the projection Poisson system
// (build_pressure_laplacian, src/Operators.jl:361-390, solved in project!,
// src/Projection.jl:170-307) on uniform grids by fast real transforms.
//
// The cell-centred 5-point operator -A (wall ghosts folded into the
// diagonal: 1/h^2 at a Neumann wall, 3/h^2 at a Dirichlet wall) is the
// Kronecker sum of two 1-D operators.  Per axis, with n cells of width h:
//   SPEC_COS  Neumann/Neumann:      basis cos(pi k (i+1/2)/n),
//             eigenvalue 4 sin^2(pi k / 2n) / h^2
//   SPEC_ALT  Dirichlet/Dirichlet:  (-1)^i times the same basis (the sign
//             flip maps the Dirichlet operator onto 4/h^2 - Neumann),
//             eigenvalue 4 cos^2(pi k / 2n) / h^2
//   SPEC_RDFT periodic:             real DFT in half-complex order (c <= n/2:
//             Re X_c, c > n/2: Im X_{n-c}), eigenvalue 4 sin^2(pi f / n) / h^2
//             with f = min(c, n - c)
// The solve is three passes over the field, each reading and writing it once:
//   rows:    forward transform along x                    (k_spec_rows_fwd)
//   columns: forward along y, divide by the eigenvalue,
//            inverse along y                              (k_spec_cols)
//   rows:    inverse along x                              (k_spec_rows_inv)
// The (0, 0) mode is zeroed where both axes have the constant null vector,
// which is the zero-mean gauge of the singular Neumann/periodic system.
//
// Every transform of two real lines is one complex FFT of length n = 2^m or
// n = 3 * 2^m in LDS: the lines are packed as z = x1 + i x2, loaded in
// bit-reversed order (for 3 * 2^m: input i = r + 2^m q at 3 brev(r) + q, so
// that each triple holds one length-3 sub-transform) and transformed in place
// by decimation-in-time stages: radix 3 on the triples when n = 3 * 2^m,
// then radix 4 (one radix-2 stage first when m is odd) on spans b 2^s with
// b = 1 or 3; then separated with X1 = (Z_k + conj Z_{n-k})/2,
// X2 = (Z_k - conj Z_{n-k})/(2i).  The cosine transform of a real line uses
// Makhoul's reordering v = (x0, x2, x4, ..., x5, x3, x1): C_k = Re(e^{-i pi
// k/2n} V_k); its inverse rebuilds V_k = e^{i pi k/2n} (C_k - i C_{n-k}) and
// applies the inverse FFT (the forward one on conjugated data).
#include <cmath>
#include <vector>

#include "fsi_kernels.hpp"

namespace fsi {


__device__ __forceinline__ int brev(int j, int m) { return (int)(__brev((unsigned)j) >> (32 - m)); }
// LDS position of input index j of a transform of length 2^m (r3 = 0) or 3 * 2^m
__device__ __forceinline__ int fft_pos(int j, int m, int r3) {
    if (!r3) return brev(j, m);
    return 3 * brev(j & ((1 << m) - 1), m) + (j >> m);
}

// LDS slot of index j with one pad double per 32: the power-of-two strides
// of the bit-reversed scatter, of the first butterfly spans and of the
// twiddle reads then fall on distinct banks
__host__ __device__ constexpr int pd(int j) { return j + (j >> 5); }
// LDS doubles that indices [0, v) occupy (odd for v >= 32, which also
// offsets the banks of an array placed right after)
__host__ __device__ constexpr int pdn(int v) { return v + (v >> 5) + 1; }

// position of x_i in Makhoul's reordered sequence v
__device__ __forceinline__ int mk_pos(int i, int n) { return (i & 1) ? n - 1 - (i >> 1) : (i >> 1); }

// Per-axis tables staged in LDS at kernel start: FFT twiddles (n/2 complex),
// the cosine kinds' post-twiddles (n complex) and, for the column pass, the
// eigenvalues of the axis (n).
struct SpecLds {
    double *wr, *wi, *pr, *pi, *mu;
};
template <int NT, int KIND>
__device__ __forceinline__ SpecLds load_tables(const SpecAxis& a, double* base, bool with_mu, int tid) {
    const int n = a.n;
    SpecLds t;
    t.wr = base;
    t.wi = t.wr + pdn(n >> 1);
    t.pr = t.wi + pdn(n >> 1);
    t.pi = t.pr + (KIND == SPEC_RDFT ? 0 : n);
    t.mu = t.pi + (KIND == SPEC_RDFT ? 0 : n);
    for (int k = tid; k < (n >> 1); k += NT) {
        const double2 w = a.wfft[k];
        t.wr[pd(k)] = w.x;
        t.wi[pd(k)] = w.y;
    }
    if (KIND != SPEC_RDFT)
        for (int k = tid; k < n; k += NT) {
            const double2 w = a.wpost[k];
            t.pr[k] = w.x;
            t.pi[k] = w.y;
        }
    if (with_mu)
        for (int k = tid; k < n; k += NT) t.mu[k] = a.mu[k];
    return t;
}
// LDS values of those tables
__host__ __device__ constexpr int spec_table_values(int n, int kind, bool with_mu) {
    return 2 * pdn(n >> 1) + (kind == SPEC_RDFT ? 0 : 2 * n) + (with_mu ? n : 0);
}

// The radix-3 stage and the radix-2/4 stages of spans 3 * 2^s of F complex
// FFTs of length n = 3 * 2^m (the layout and twiddles of lds_fft).
template <int NT>
__device__ void lds_fft3(double* re, double* im, const double* wr, const double* wi, int m, int F, int S, int tid) {
    const int n = 3 << m, n3 = 1 << m;
    constexpr double c3 = -0.5, s3 = -0.86602540378443864676;  // e^{-2 pi i/3}
    // X0 = A + (B + C), X1,2 = A + c3 (B + C) +- i s3 (B - C)
    for (int q = tid; q < F * n3; q += NT) {
        const int f = q >> m, l = 3 * (q & (n3 - 1)), b = f * S;
        const int j0 = b + pd(l), j1 = b + pd(l + 1), j2 = b + pd(l + 2);
        const double Ar = re[j0], Ai = im[j0], Br = re[j1], Bi = im[j1], Cr = re[j2], Ci = im[j2];
        const double tr = Br + Cr, ti = Bi + Ci, ur = Br - Cr, ui = Bi - Ci;
        const double mr = Ar + c3 * tr, mi = Ai + c3 * ti;
        const double pr = -s3 * ui, pi = s3 * ur;
        re[j0] = Ar + tr;
        im[j0] = Ai + ti;
        re[j1] = mr + pr;
        im[j1] = mi + pi;
        re[j2] = mr - pr;
        im[j2] = mi - pi;
    }
    __syncthreads();
    int s = 0;
    if (m & 1) {  // radix 2, span 3: pairs (l, l+3) in blocks of 6, twiddle e^{-2 pi i t/6}
        const int nh = n >> 1;
        for (int q = tid; q < F * nh; q += NT) {
            const int f = q / nh, r = q - f * nh;
            const int blk = r / 3, t = r - 3 * blk;
            const int l = 6 * blk + t, b = f * S;
            const int iw = pd(t << (m - 1));
            const int j0 = b + pd(l), j1 = b + pd(l + 3);
            const double w_r = wr[iw], w_i = wi[iw];
            const double Br = re[j1], Bi = im[j1];
            const double bR = w_r * Br - w_i * Bi, bI = w_r * Bi + w_i * Br;
            const double Ar = re[j0], Ai = im[j0];
            re[j0] = Ar + bR;
            im[j0] = Ai + bI;
            re[j1] = Ar - bR;
            im[j1] = Ai - bI;
        }
        s = 1;
        __syncthreads();
    }
    const int nq = n >> 2;
    for (; s < m; s += 2) {  // radix 4 on spans h = 3 * 2^s and 2h, as in lds_fft
        const int h = 3 << s;
        for (int q = tid; q < F * nq; q += NT) {
            const int f = q / nq, r = q - f * nq;
            const int blk = (r >> s) / 3, t = r - blk * h;
            const int l = blk * 4 * h + t, b = f * S;
            const int i1 = pd(t << (m - s - 1)), i2 = pd(t << (m - s - 2));
            const int ja = b + pd(l), jb = b + pd(l + h), jc = b + pd(l + 2 * h), jd = b + pd(l + 3 * h);
            const double w1r = wr[i1], w1i = wi[i1], w2r = wr[i2], w2i = wi[i2];
            const double Ar = re[ja], Ai = im[ja], Br = re[jb], Bi = im[jb];
            const double Cr = re[jc], Ci = im[jc], Dr = re[jd], Di = im[jd];
            const double bR = w1r * Br - w1i * Bi, bI = w1r * Bi + w1i * Br;
            const double dR = w1r * Dr - w1i * Di, dI = w1r * Di + w1i * Dr;
            const double A1r = Ar + bR, A1i = Ai + bI, B1r = Ar - bR, B1i = Ai - bI;
            const double C1r = Cr + dR, C1i = Ci + dI, D1r = Cr - dR, D1i = Ci - dI;
            const double cR = w2r * C1r - w2i * C1i, cI = w2r * C1i + w2i * C1r;
            const double gR = w2r * D1r - w2i * D1i, gI = w2r * D1i + w2i * D1r;  // w2 D1
            const double eR = gI, eI = -gR;                                         // -i w2 D1
            re[ja] = A1r + cR;
            im[ja] = A1i + cI;
            re[jc] = A1r - cR;
            im[jc] = A1i - cI;
            re[jb] = B1r + eR;
            im[jb] = B1i + eI;
            re[jd] = B1r - eR;
            im[jd] = B1i - eI;
        }
        __syncthreads();
    }
}

// F complex FFTs of length n = 2^m (r3 = 0) or 3 * 2^m (r3 = 1) in place on
// re/im (transform f at f * S, S >= pdn(n), input in the order of fft_pos,
// output in natural order); twiddles wr + i wi = e^{-2 pi i t/n}, t < n/2.
// Every line is addressed through pd().  Ends with a barrier.
template <int NT>
__device__ void lds_fft(double* re, double* im, const double* wr, const double* wi, int m, int r3, int F, int S,
                        int tid) {
    if (r3) {
        lds_fft3<NT>(re, im, wr, wi, m, F, S, tid);
        return;
    }
    const int n = 1 << m;
    int s = 0;
    if (m & 1) {  // radix 2, span 1: butterflies (2q, 2q+1), twiddle 1
        for (int q = tid; q < F * (n >> 1); q += NT) {
            const int l = 2 * (q & ((n >> 1) - 1)), b = (q >> (m - 1)) * S;
            const int j0 = b + pd(l), j1 = b + pd(l + 1);
            const double ar = re[j0], ai = im[j0], br = re[j1], bi = im[j1];
            re[j0] = ar + br;
            im[j0] = ai + bi;
            re[j1] = ar - br;
            im[j1] = ai - bi;
        }
        s = 1;
        __syncthreads();
    }
    // radix 4 = the radix-2 stages of spans h and 2h on j, j+h, j+2h, j+3h:
    // span h pairs (j, j+h), (j+2h, j+3h) with w1 = e^{-2 pi i t/2h}; span 2h
    // pairs (j, j+2h) with w2 = e^{-2 pi i t/4h} and (j+h, j+3h) with -i w2
    for (; s < m; s += 2) {
        const int h = 1 << s;
        for (int q = tid; q < F * (n >> 2); q += NT) {
            const int f = q >> (m - 2), r = q & ((n >> 2) - 1);
            const int t = r & (h - 1);
            const int l = ((r >> s) << (s + 2)) + t, b = f * S;
            const int i1 = pd(t << (m - s - 1)), i2 = pd(t << (m - s - 2));
            const int ja = b + pd(l), jb = b + pd(l + h), jc = b + pd(l + 2 * h), jd = b + pd(l + 3 * h);
            const double w1r = wr[i1], w1i = wi[i1], w2r = wr[i2], w2i = wi[i2];
            const double Ar = re[ja], Ai = im[ja], Br = re[jb], Bi = im[jb];
            const double Cr = re[jc], Ci = im[jc], Dr = re[jd], Di = im[jd];
            const double bR = w1r * Br - w1i * Bi, bI = w1r * Bi + w1i * Br;
            const double dR = w1r * Dr - w1i * Di, dI = w1r * Di + w1i * Dr;
            const double A1r = Ar + bR, A1i = Ai + bI, B1r = Ar - bR, B1i = Ai - bI;
            const double C1r = Cr + dR, C1i = Ci + dI, D1r = Cr - dR, D1i = Ci - dI;
            const double cR = w2r * C1r - w2i * C1i, cI = w2r * C1i + w2i * C1r;
            const double gR = w2r * D1r - w2i * D1i, gI = w2r * D1i + w2i * D1r;  // w2 D1
            const double eR = gI, eI = -gR;                                         // -i w2 D1
            re[ja] = A1r + cR;
            im[ja] = A1i + cI;
            re[jc] = A1r - cR;
            im[jc] = A1i - cI;
            re[jb] = B1r + eR;
            im[jb] = B1i + eI;
            re[jd] = B1r - eR;
            im[jd] = B1i - eI;
        }
        __syncthreads();
    }
}

// the two real coefficient lines of output index k from the packed spectrum
// (re, im: one transform, addressed through pd())
template <int KIND>
__device__ __forceinline__ void split_coeffs(int n, const SpecLds& t, const double* re, const double* im, int k,
                                             double& o1, double& o2) {
    const int kr = k ? n - k : 0;
    const double zr = re[pd(k)], zi = im[pd(k)], cr = re[pd(kr)], ci = im[pd(kr)];
    const double x1r = 0.5 * (zr + cr), x1i = 0.5 * (zi - ci);
    const double x2r = 0.5 * (zi + ci), x2i = -0.5 * (zr - cr);
    if (KIND == SPEC_RDFT) {
        const bool lo = k <= (n >> 1);
        o1 = lo ? x1r : -x1i;
        o2 = lo ? x2r : -x2i;
    } else {
        const double tr = t.pr[k], ti = t.pi[k];
        o1 = tr * x1r - ti * x1i;
        o2 = tr * x2r - ti * x2i;
    }
}

// the spectrum V_k (or X_k) of one real line from its coefficient line c
// (addressed through pd())
template <int KIND>
__device__ __forceinline__ void line_spectrum(int n, const SpecLds& t, const double* c, int k, double& vr,
                                              double& vi) {
    const int h = n >> 1;
    if (KIND == SPEC_RDFT) {
        vr = k <= h ? c[pd(k)] : c[pd(n - k)];
        vi = (k == 0 || k == h) ? 0.0 : (k < h ? c[pd(n - k)] : -c[pd(k)]);
    } else {
        const double ck = c[pd(k)], cr = k ? c[pd(n - k)] : 0.0;
        const double tr = t.pr[k], ti = t.pi[k];  // V = conj(t) (ck - i cr)
        vr = tr * ck - ti * cr;
        vi = -tr * cr - ti * ck;
    }
}

// The right-hand side of the projection on the points i = tid + it NT of
// rows j0 and j0 + 1 (for it in [C0, C0 + 4)), as k_proj_rhs forms it:
// rhs = (div(v*) - G) - b_bc with G = 0, after _inject_compact_predictor!
// (:543-579) has given the far faces the boundary data.  Every load of the
// chunk is issued before the first store; acc gathers the compatibility sums
// (rhs dA, |rhs| dA, rhs^2, div^2 dA) in the order of the thread's points.
template <int IT, int NT, int C0>
__device__ __forceinline__ void proj_rhs_chunk(const Params& p, const ProjBufs& pb, int tid, int j0, bool two,
                                               double (&a)[IT], double (&b)[IT], double (&acc)[4]) {
    constexpr int CH = IT - C0 < 4 ? IT - C0 : 4;
    const int n = p.nx;
    if (C0 * NT >= n) {
#pragma unroll
        for (int c = 0; c < CH; ++c) a[C0 + c] = b[C0 + c] = 0.0;
        return;
    }
    double uL[2][CH], uR[2][CH], vB[2][CH], vT[2][CH], lb[2][CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const int i = tid + (C0 + c) * NT;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const bool in = i < n && (h == 0 || two);
            const size_t k = in ? (size_t)(j0 + h) * p.P + i : 0;
            uL[h][c] = pb.Uin[k];
            uR[h][c] = pb.Uin[k + 1];
            vB[h][c] = pb.Vin[k];
            vT[h][c] = pb.Vin[k + p.P];
            lb[h][c] = pb.lap_bc[k];
        }
    }
    const double area = p.hx * p.hy;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const int i = tid + (C0 + c) * NT;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            double& out = h == 0 ? a[C0 + c] : b[C0 + c];
            out = 0.0;
            if (!(i < n && (h == 0 || two))) continue;
            const int j = j0 + h;
            const size_t k = (size_t)j * p.P + i;
            double ur = uR[h][c], vt = vT[h][c];
            if (pb.inject && i == n - 1) {
                ur = p.px ? pb.Uin[(size_t)j * p.P] : p.bU.r.value;
                pb.Uin[k + 1] = ur;
            }
            if (pb.inject && j == p.ny - 1) {
                vt = p.py ? pb.Vin[i] : p.bV.t.value;
                pb.Vin[k + p.P] = vt;
            }
            const double div = (ur - uL[h][c]) * p.ihx + (vt - vB[h][c]) * p.ihy;
            const double rhs = (div - (pb.G ? pb.G[k] : 0.0)) - lb[h][c];
            pb.rhs[k] = rhs;
            acc[0] += rhs * area;
            acc[1] += fabs(rhs) * area;
            acc[2] += rhs * rhs;
            acc[3] += div * div * area;
            out = rhs;
        }
    }
}

// ---- rows: forward transform along x of two rows per workgroup ----------
// RHS: the rows are the projection's right-hand side, formed here from the
// velocity (proj_rhs_at) and stored, with one set of partials per workgroup.
template <int KIND, int NT, bool RHS>
FSI_KERNEL void k_spec_rows_fwd(SpecAxis ax, int nrows, int P, const double* __restrict__ in,
                                                      double* __restrict__ out, double scale, Params p, ProjBufs pb) {
    constexpr int IT = 2048 / NT;
    extern __shared__ double sp_smem[];
    const int n = ax.n, m = ax.m, tid = threadIdx.x;
    double* re = sp_smem;
    double* im = re + pdn(n);
    const int j0 = 2 * blockIdx.x;
    const bool two = j0 + 1 < nrows;
    const double* r0 = in + (size_t)j0 * P;
    double a[IT], b[IT];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (RHS) {
        static_assert(IT <= 8, "two chunks of four points");
        proj_rhs_chunk<IT, NT, 0>(p, pb, tid, j0, two, a, b, acc);
        if (IT > 4) proj_rhs_chunk<IT, NT, (IT > 4 ? 4 : 0)>(p, pb, tid, j0, two, a, b, acc);
    } else {
#pragma unroll
        for (int it = 0; it < IT; ++it) {  // every load in flight before the first use
            const int i = tid + it * NT;
            a[it] = i < n ? r0[i] : 0.0;
            b[it] = (i < n && two) ? r0[i + P] : 0.0;
        }
    }
    const SpecLds t = load_tables<NT, KIND>(ax, im + pdn(n), false, tid);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int i = tid + it * NT;
        if (i >= n) break;
        const int jv = KIND == SPEC_RDFT ? i : mk_pos(i, n);
        const double s = (KIND == SPEC_ALT && (i & 1)) ? -scale : scale;
        const int p = pd(fft_pos(jv, m, ax.r3));
        re[p] = s * a[it];
        im[p] = s * b[it];
    }
    __syncthreads();
    lds_fft<NT>(re, im, t.wr, t.wi, m, ax.r3, 1, pdn(n), tid);
    double* o0 = out + (size_t)j0 * P;
    for (int k = tid; k < n; k += NT) {
        double o1, o2;
        split_coeffs<KIND>(n, t, re, im, k, o1, o2);
        o0[k] = o1;
        if (two) o0[k + P] = o2;
    }
    if (RHS) put_partials<4>(acc, pb.rhs_part, gridDim.x);
}

// ---- rows: inverse transform along x, in place ---------------------------
template <int KIND, int NT>
// mpart: one partial per workgroup of the area-weighted sum of the two rows
// written (the mean of the solution, :255-259), or null.
FSI_KERNEL void k_spec_rows_inv(SpecAxis ax, int nrows, int P, double* x, double* mpart,
                                                      double area) {
    constexpr int IT = 2048 / NT;
    extern __shared__ double sp_smem[];
    const int n = ax.n, m = ax.m, tid = threadIdx.x;
    double* re = sp_smem;  // the two coefficient lines, then the packed spectrum
    double* im = re + pdn(n);
    const int j0 = 2 * blockIdx.x;
    const bool two = j0 + 1 < nrows;
    double* r0 = x + (size_t)j0 * P;
    double a[IT], b[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int i = tid + it * NT;
        a[it] = i < n ? r0[i] : 0.0;
        b[it] = (i < n && two) ? r0[i + P] : 0.0;
    }
    const SpecLds t = load_tables<NT, KIND>(ax, im + pdn(n), false, tid);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int i = tid + it * NT;
        if (i >= n) break;
        re[pd(i)] = a[it];
        im[pd(i)] = b[it];
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {  // conj(V1 + i V2), V from the coefficient lines
        const int k = tid + it * NT;
        if (k >= n) break;
        double v1r, v1i, v2r, v2i;
        line_spectrum<KIND>(n, t, re, k, v1r, v1i);
        line_spectrum<KIND>(n, t, im, k, v2r, v2i);
        a[it] = v1r - v2i;
        b[it] = -(v1i + v2r);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int k = tid + it * NT;
        if (k >= n) break;
        const int p = pd(fft_pos(k, m, ax.r3));
        re[p] = a[it];
        im[p] = b[it];
    }
    __syncthreads();
    lds_fft<NT>(re, im, t.wr, t.wi, m, ax.r3, 1, pdn(n), tid);
    const double inv_n = 1.0 / (double)n;
    double acc[1] = {0.0};
    for (int i = tid; i < n; i += NT) {  // z = conj(FFT(conj Z)) / n
        const int jv = pd(KIND == SPEC_RDFT ? i : mk_pos(i, n));
        const double s = (KIND == SPEC_ALT && (i & 1)) ? -inv_n : inv_n;
        const double x0 = s * re[jv], x1 = -s * im[jv];
        r0[i] = x0;
        acc[0] += x0 * area;
        if (two) {
            r0[i + P] = x1;
            acc[0] += x1 * area;
        }
    }
    if (mpart) put_partials<1>(acc, mpart, gridDim.x);
}

// ---- columns: forward along y, divide, inverse along y; 2C columns -------
template <int KIND, int C, int NT>
FSI_KERNEL void k_spec_cols(SpecAxis ay, const double* __restrict__ mux, int P, double* x) {
    constexpr int IT = 2048 / NT;  // C n <= 2048 line-pair values per block
    extern __shared__ double sp_smem[];
    const int n = ay.n, m = ay.m, tid = threadIdx.x;
    // C transforms of n, transform f at f S: the odd stride puts the lines of
    // one row (read and written together) on different banks
    const int S = pdn(n);
    double* re = sp_smem;
    double* im = re + C * S;
    // XCD-aware: the 2C-column strips of one XCD are contiguous, so the
    // 128-byte lines that neighbouring strips share are fetched by one L2
    const int x0 = tile_id().x * 2 * C;
    constexpr int W = 2 * C;  // columns per row of the block
    const int total = n * W;
    double v[2 * IT];
#pragma unroll
    for (int it = 0; it < 2 * IT; ++it) {  // element e: row e / W, column e % W
        const int e = tid + it * NT;
        v[it] = e < total ? x[(size_t)(e / W) * P + x0 + (e % W)] : 0.0;
    }
    double lx[W];  // eigenvalues of the block's columns
#pragma unroll
    for (int c = 0; c < W; ++c) lx[c] = mux[x0 + c];
    const SpecLds t = load_tables<NT, KIND>(ay, im + C * S, true, tid);
#pragma unroll
    for (int it = 0; it < 2 * IT; ++it) {
        const int e = tid + it * NT;
        if (e >= total) break;
        const int y = e / W, c = e % W;
        const int jv = KIND == SPEC_RDFT ? y : mk_pos(y, n);
        const double s = (KIND == SPEC_ALT && (y & 1)) ? -1.0 : 1.0;
        const int p = (c >> 1) * S + pd(fft_pos(jv, m, ay.r3));
        if (c & 1) im[p] = s * v[it];
        else re[p] = s * v[it];
    }
    __syncthreads();
    lds_fft<NT>(re, im, t.wr, t.wi, m, ay.r3, C, S, tid);
    // coefficients of the column pair f at row mode l, divided by the eigenvalue
    double o1[IT], o2[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int q = tid + it * NT;
        if (q >= C * n) break;
        const int f = ay.r3 ? (q >= n) : q >> m, l = q - f * n;  // C <= 2 transforms
        split_coeffs<KIND>(n, t, re + f * S, im + f * S, l, o1[it], o2[it]);
        const double my = t.mu[l];
        const double l1 = lx[2 * f] + my, l2 = lx[2 * f + 1] + my;
        o1[it] = l1 != 0.0 ? o1[it] / l1 : 0.0;
        o2[it] = l2 != 0.0 ? o2[it] / l2 : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int q = tid + it * NT;
        if (q >= C * n) break;
        const int f = ay.r3 ? (q >= n) : q >> m;
        const int p = f * S + pd(q - f * n);
        re[p] = o1[it];  // coefficient line of column 2f in re[f S ...], 2f+1 in im[f S ...]
        im[p] = o2[it];
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int q = tid + it * NT;
        if (q >= C * n) break;
        const int f = ay.r3 ? (q >= n) : q >> m, l = q - f * n;
        double v1r, v1i, v2r, v2i;
        line_spectrum<KIND>(n, t, re + f * S, l, v1r, v1i);
        line_spectrum<KIND>(n, t, im + f * S, l, v2r, v2i);
        o1[it] = v1r - v2i;
        o2[it] = -(v1i + v2r);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int q = tid + it * NT;
        if (q >= C * n) break;
        const int f = ay.r3 ? (q >= n) : q >> m;
        const int p = f * S + pd(fft_pos(q - f * n, m, ay.r3));
        re[p] = o1[it];
        im[p] = o2[it];
    }
    __syncthreads();
    lds_fft<NT>(re, im, t.wr, t.wi, m, ay.r3, C, S, tid);
    const double inv_n = 1.0 / (double)n;
#pragma unroll
    for (int it = 0; it < 2 * IT; ++it) {
        const int e = tid + it * NT;
        if (e >= total) break;
        const int y = e / W, c = e % W;
        const int jv = KIND == SPEC_RDFT ? y : mk_pos(y, n);
        const double s = (KIND == SPEC_ALT && (y & 1)) ? -inv_n : inv_n;
        const int p = (c >> 1) * S + pd(jv);
        x[(size_t)y * P + x0 + c] = (c & 1) ? -s * im[p] : s * re[p];
    }
}

// ---- host side -------------------------------------------------------------

int spec_kind(int n, int periodic, int dlo, int dhi) {
    const int odd = (n % 3 == 0) ? n / 3 : n;  // 2^m or 3 * 2^m
    if (n < SPEC_MIN_N || n > SPEC_MAX_N || (odd & (odd - 1))) return SPEC_NONE;
    if (periodic) return SPEC_RDFT;
    if (!dlo && !dhi) return SPEC_COS;
    if (dlo && dhi) return SPEC_ALT;
    return SPEC_NONE;
}

std::vector<double> spec_tables(int n, int kind, double h2inv) {
    std::vector<double> t((size_t)(n + 2 * n + n), 0.0);  // wfft (n/2 complex), wpost (n complex), mu (n)
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < n / 2; ++k) {
        const long double a = -2.0L * pi * k / n;
        t[2 * k] = (double)std::cos(a);
        t[2 * k + 1] = (double)std::sin(a);
    }
    double* wpost = t.data() + n;
    for (int k = 0; k < n; ++k) {
        const long double a = -pi * k / (2.0L * n);
        wpost[2 * k] = (double)std::cos(a);
        wpost[2 * k + 1] = (double)std::sin(a);
    }
    double* mu = t.data() + 3 * n;
    for (int k = 0; k < n; ++k) {
        long double s;
        if (kind == SPEC_COS) s = std::sin(pi * k / (2.0L * n));
        else if (kind == SPEC_ALT) s = std::cos(pi * k / (2.0L * n));
        else s = std::sin(pi * (k <= n / 2 ? k : n - k) / n);
        mu[k] = (double)(4.0L * s * s * h2inv);
    }
    return t;
}

constexpr int SP_ROW_THREADS = 256, SP_COL_THREADS = 512;

static size_t rows_lds(int n, int kind) {
    return (2 * (size_t)pdn(n) + spec_table_values(n, kind, false)) * sizeof(double);
}
// two column pairs per block where n allows (rows of 32 bytes), else one
static int cols_pairs(int n, int nx) { return (n <= SPEC_MAX_N / 2 && nx % 4 == 0) ? 2 : 1; }
static size_t cols_lds(int n, int kind, int C) {
    return (2 * (size_t)C * pdn(n) + spec_table_values(n, kind, true)) * sizeof(double);
}

template <int KX>
static void rows_fwd(hipStream_t s, const SpecPlan& sp, int P, const double* in, double* out) {
    const int n = sp.x.n, rows = sp.y.n;
    fsi_launch<&k_spec_rows_fwd<KX, SP_ROW_THREADS, false>, SP_ROW_THREADS, 0>((rows + 1) / 2, SP_ROW_THREADS, rows_lds(n, KX), s, 
        sp.x, rows, P, in, out, -1.0, Params{}, ProjBufs{});
}
template <int KX>
static void rows_fwd_rhs(hipStream_t s, const SpecPlan& sp, const Params& p, const ProjBufs& b) {
    const int n = sp.x.n, rows = sp.y.n;
    fsi_launch<&k_spec_rows_fwd<KX, SP_ROW_THREADS, true>, SP_ROW_THREADS, 0>((rows + 1) / 2, SP_ROW_THREADS, rows_lds(n, KX), s, 
        sp.x, rows, p.P, nullptr, b.x, -1.0, p, b);
}
template <int KX>
static void rows_inv(hipStream_t s, const SpecPlan& sp, int P, double* x, double* mpart = nullptr,
                     double area = 0.0) {
    const int n = sp.x.n, rows = sp.y.n;
    fsi_launch<&k_spec_rows_inv<KX, SP_ROW_THREADS>, SP_ROW_THREADS, 0>((rows + 1) / 2, SP_ROW_THREADS, rows_lds(n, KX), s, sp.x, rows, P, x,
                                                                                              mpart, area);
}
template <int KY>
static void cols(hipStream_t s, const SpecPlan& sp, int P, double* x) {
    const int n = sp.y.n, nx = sp.x.n;
    if (cols_pairs(n, nx) == 2)
        fsi_launch<&k_spec_cols<KY, 2, SP_COL_THREADS>, SP_COL_THREADS, 0>(nx / 4, SP_COL_THREADS, cols_lds(n, KY, 2), s, sp.y, sp.x.mu, P, x);
    else
        fsi_launch<&k_spec_cols<KY, 1, SP_COL_THREADS>, SP_COL_THREADS, 0>(nx / 2, SP_COL_THREADS, cols_lds(n, KY, 1), s, sp.y, sp.x.mu, P, x);
}

// dynamic LDS above the 64 KiB default needs the kernel attribute (up to the
// device's per-block maximum)
static bool allow_lds(hipError_t (*allow)(int), size_t bytes) {
    if (bytes <= 65536) return true;
    int dev = 0, mx = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&mx, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        bytes > (size_t)mx ||
        allow((int)bytes) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return true;
}
template <int KX>
static bool prepare_rows(int n) {
    return allow_lds(&fsi_allow_dynamic_lds<&k_spec_rows_fwd<KX, SP_ROW_THREADS, false>, SP_ROW_THREADS>,
                     rows_lds(n, KX)) &&
           allow_lds(&fsi_allow_dynamic_lds<&k_spec_rows_fwd<KX, SP_ROW_THREADS, true>, SP_ROW_THREADS>,
                     rows_lds(n, KX)) &&
           allow_lds(&fsi_allow_dynamic_lds<&k_spec_rows_inv<KX, SP_ROW_THREADS>, SP_ROW_THREADS>, rows_lds(n, KX));
}
template <int KY>
static bool prepare_cols(int n, int nx) {
    const int C = cols_pairs(n, nx);
    hipError_t (*k)(int) = C == 2 ? &fsi_allow_dynamic_lds<&k_spec_cols<KY, 2, SP_COL_THREADS>, SP_COL_THREADS>
                                  : &fsi_allow_dynamic_lds<&k_spec_cols<KY, 1, SP_COL_THREADS>, SP_COL_THREADS>;
    return allow_lds(k, cols_lds(n, KY, C));
}

bool spec_prepare(const SpecPlan& sp) {
    bool ok = sp.x.kind == SPEC_COS ? prepare_rows<SPEC_COS>(sp.x.n)
              : sp.x.kind == SPEC_ALT ? prepare_rows<SPEC_ALT>(sp.x.n)
                                      : prepare_rows<SPEC_RDFT>(sp.x.n);
    ok = ok && (sp.y.kind == SPEC_COS ? prepare_cols<SPEC_COS>(sp.y.n, sp.x.n)
                : sp.y.kind == SPEC_ALT ? prepare_cols<SPEC_ALT>(sp.y.n, sp.x.n)
                                        : prepare_cols<SPEC_RDFT>(sp.y.n, sp.x.n));
    return ok;
}

void launch_spectral_solve(hipStream_t s, const SpecPlan& sp, int P, const double* rhs, double* x) {
    switch (sp.x.kind) {
    case SPEC_COS: rows_fwd<SPEC_COS>(s, sp, P, rhs, x); break;
    case SPEC_ALT: rows_fwd<SPEC_ALT>(s, sp, P, rhs, x); break;
    default: rows_fwd<SPEC_RDFT>(s, sp, P, rhs, x); break;
    }
    switch (sp.y.kind) {
    case SPEC_COS: cols<SPEC_COS>(s, sp, P, x); break;
    case SPEC_ALT: cols<SPEC_ALT>(s, sp, P, x); break;
    default: cols<SPEC_RDFT>(s, sp, P, x); break;
    }
    switch (sp.x.kind) {
    case SPEC_COS: rows_inv<SPEC_COS>(s, sp, P, x); break;
    case SPEC_ALT: rows_inv<SPEC_ALT>(s, sp, P, x); break;
    default: rows_inv<SPEC_RDFT>(s, sp, P, x); break;
    }
}

void launch_spectral_proj(hipStream_t s, const SpecPlan& sp, const Params& p, const ProjBufs& b) {
    switch (sp.x.kind) {
    case SPEC_COS: rows_fwd_rhs<SPEC_COS>(s, sp, p, b); break;
    case SPEC_ALT: rows_fwd_rhs<SPEC_ALT>(s, sp, p, b); break;
    default: rows_fwd_rhs<SPEC_RDFT>(s, sp, p, b); break;
    }
    switch (sp.y.kind) {
    case SPEC_COS: cols<SPEC_COS>(s, sp, p.P, b.x); break;
    case SPEC_ALT: cols<SPEC_ALT>(s, sp, p.P, b.x); break;
    default: cols<SPEC_RDFT>(s, sp, p.P, b.x); break;
    }
    const double area = p.hx * p.hy;
    switch (sp.x.kind) {
    case SPEC_COS: rows_inv<SPEC_COS>(s, sp, p.P, b.x, b.q_part, area); break;
    case SPEC_ALT: rows_inv<SPEC_ALT>(s, sp, p.P, b.x, b.q_part, area); break;
    default: rows_inv<SPEC_RDFT>(s, sp, p.P, b.x, b.q_part, area); break;
    }
}

}  // namespace fsi
