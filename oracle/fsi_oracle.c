/*
 * fsi_oracle.c -- FP64 CPU restatement of the EulerianFSI.jl IMEX step.
 *
 * TEST INFRASTRUCTURE ONLY (see fsi_oracle.h).  Each function restates one
 * reference function and cites it (paths relative to the reference repo).
 * Loops, boundary branches and evaluation order follow the reference; the
 * sparse matrices the reference assembles are applied here matrix-free as
 * the same composition of factors (e.g. Dstress * (w .* (E*x))), so results
 * agree with the reference up to floating-point rounding.
 *
 * Linear solves: the reference uses GMRES / UMFPACK for the velocity
 * predictor and GMRES + Ruge-Stueben AMG for the projection.  Both systems
 * have unique solutions (after the gauge/mean fix of the projection), which
 * the oracle computes with CG (velocity) and a geometric-multigrid
 * preconditioned CG (projection) to the tolerances in fsi_numerics.
 *
 * Gauge of the Neumann/periodic projection: the reference pins q[1] = 0 in
 * row/column 1 (src/Projection.jl:106-113, :226-230) and then removes the
 * area-weighted mean (:255-259).  The oracle solves the singular system with
 * the right-hand side's mean removed and then removes the mean of q.  The two
 * agree exactly when the compatibility defect (:206-221) is zero and differ
 * by a term proportional to that defect otherwise.
 */
#include "fsi_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <malloc.h>

/* The OpenMP build of this file (liboracle_fsi_omp.so, the all-cores CPU
 * baseline of the benchmark; never the parity checker) runs the loops marked
 * OMP_FOR in parallel: each of them writes disjoint entries.  Its dot
 * products sum per thread and its multigrid smoother visits the cells in
 * red-black order, so it agrees with the serial build to the solver
 * tolerances, not bitwise.  The serial build ignores the markers. */
#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
/* the same on a multigrid level, serial below 128 x 128 cells */
#define OMP_FOR_LEVEL _Pragma("omp parallel for schedule(static) if (nx * ny >= 16384)")
#else
#define OMP_FOR
#define OMP_FOR_LEVEL
#endif

/* The tangent build (fsi_tangent.cpp) compiles this file as C++ with a
 * dual-number type in place of double: these hooks give the shear modulus
 * its derivatives there and take the value where a double field of an
 * fsi_api.h struct is written.  Here they are the identity. */
#ifndef FSIO_SHEAR_MODULUS
#define FSIO_SHEAR_MODULUS(c) ((c)->m.shear_modulus)
#endif
#ifndef FSIO_VALUE
#define FSIO_VALUE(x) (x)
#endif

#define UX(a, i, j) (a)[(size_t)(i) + (size_t)(j) * (size_t)(c->nx + 1)]
#define VY(a, i, j) (a)[(size_t)(i) + (size_t)(j) * (size_t)(c->nx)]
#define CC(a, i, j) (a)[(size_t)(i) + (size_t)(j) * (size_t)(c->nx)]

typedef struct {
    double *dx_f, *dx_c, *dy_f, *dy_c;
    double *wL_c2fx, *wR_c2fx, *wB_c2fy, *wT_c2fy;
    double *wL_f2cx, *wR_f2cx, *wB_f2cy, *wT_f2cy;
    double *dx_west_corner, *dx_east_corner, *dy_south_corner, *dy_north_corner;
    double *x_p, *y_p, *x_u, *y_v;
} metrics_t;

/* one level of the projection preconditioner hierarchy */
typedef struct {
    int nx, ny;
    double hx, hy;
    double *x, *b, *r;
} mg_level;

struct fsio_ctx {
    int nx, ny, N;
    double lx, ly;
    int px, py;
    fsi_bcs bcs;
    fsi_models m;
    fsi_numerics num;
    metrics_t mt;
    char err[512];
    /* projection */
    int has_nullspace;
    double* lap_bc; /* b_bc = D * gradient_bc (nx*ny) */
    int nlev;
    mg_level lev[32];
    /* Kolmogorov body force per row (TransportConstitutive.jl:385-397) */
    double* force_row;
    /* growth tensor per unit G (GrowthModels.jl:268-318): K_g = G (kxx, kxy; kxy, kyy) */
    double kxx, kxy, kyy;
};

/* ------------------------------------------------------------------------- */
static void apply_extension_reset(fsio_ctx* c, double* sxx, double* syy, double* sxy, const double* phi);

static int set_err(fsio_ctx* c, int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof(c->err), fmt, ap);
    va_end(ap);
    return code;
}

const char* fsio_last_error(const fsio_ctx* c) { return c ? c->err : "null context"; }

static double* dalloc(size_t n) {
#ifdef _OPENMP
    /* zeroed by the threads (calloc would clear, and fault in, the pages on
       one core) */
    double* p = (double*)malloc((n ? n : 1) * sizeof(double));
    if (p) {
        OMP_FOR
        for (long long k = 0; k < (long long)(n ? n : 1); ++k) p[k] = 0.0;
    }
#else
    double* p = (double*)calloc(n ? n : 1, sizeof(double));
#endif
    if (!p) {
        fprintf(stderr, "fsi_oracle: out of memory\n");
        abort();
    }
    return p;
}

static inline double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }
static inline double sq(double x) { return x * x; }

/* ---------------------------------------------------------------------------
 * Metrics: generate_grid(:uniform) + _axis_metrics
 * src/MeshGeneration.jl:179-233 and :237-335.
 * ------------------------------------------------------------------------- */
static void axis_metrics(int N, double L, int periodic, double* centers, double* faces,
                         double* h_f, double* h_c, double* wm_c2f, double* wp_c2f,
                         double* wm_f2c, double* wp_f2c, double* dm_corner,
                         double* dp_corner) {
    /* generate_grid: x_u = L*((i-1)/N) (i = 1..N+1), x_p = L*((i-0.5)/N)
       (MeshGeneration.jl:203-213), endpoints forced exactly (:206-210). */
    for (int k = 0; k <= N; ++k) faces[k] = L * ((double)k / (double)N);
    faces[0] = 0.0;
    faces[N] = L;
    for (int k = 0; k < N; ++k) centers[k] = L * (((double)k + 0.5) / (double)N);

    for (int k = 0; k < N; ++k) h_c[k] = faces[k + 1] - faces[k]; /* :250 */
    double wrapped = centers[0] + L - centers[N - 1];                /* :254 */
    h_f[0] = periodic ? wrapped : 2.0 * (centers[0] - faces[0]);
    for (int k = 1; k < N; ++k) h_f[k] = centers[k] - centers[k - 1];
    h_f[N] = periodic ? wrapped : 2.0 * (faces[N] - centers[N - 1]);

    if (periodic) { /* :266-272 */
        double dm = faces[0] + L - centers[N - 1], dp = centers[0] - faces[0], tot = dm + dp;
        wm_c2f[0] = dp / tot;
        wp_c2f[0] = dm / tot;
    } else {
        wm_c2f[0] = 0.0;
        wp_c2f[0] = 1.0;
    }
    for (int k = 1; k < N; ++k) {
        double dm = faces[k] - centers[k - 1], dp = centers[k] - faces[k], tot = dm + dp;
        wm_c2f[k] = dp / tot;
        wp_c2f[k] = dm / tot;
    }
    if (periodic) {
        double dm = faces[N] - centers[N - 1], dp = centers[0] + L - faces[N], tot = dm + dp;
        wm_c2f[N] = dp / tot;
        wp_c2f[N] = dm / tot;
    } else {
        wm_c2f[N] = 1.0;
        wp_c2f[N] = 0.0;
    }
    for (int k = 0; k < N; ++k) { /* :296-303 */
        double dm = centers[k] - faces[k], dp = faces[k + 1] - centers[k], tot = dm + dp;
        wm_f2c[k] = dp / tot;
        wp_f2c[k] = dm / tot;
    }
    dp_corner[0] = centers[0] - faces[0]; /* :309-318 */
    dm_corner[0] = periodic ? faces[0] + L - centers[N - 1] : dp_corner[0];
    for (int k = 1; k < N; ++k) {
        dm_corner[k] = faces[k] - centers[k - 1];
        dp_corner[k] = centers[k] - faces[k];
    }
    dm_corner[N] = faces[N] - centers[N - 1];
    dp_corner[N] = periodic ? centers[0] + L - faces[N] : dm_corner[N];
}

static void build_metrics(fsio_ctx* c) {
    metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny;
    m->dx_f = dalloc(nx + 1); m->dx_c = dalloc(nx);
    m->dy_f = dalloc(ny + 1); m->dy_c = dalloc(ny);
    m->wL_c2fx = dalloc(nx + 1); m->wR_c2fx = dalloc(nx + 1);
    m->wB_c2fy = dalloc(ny + 1); m->wT_c2fy = dalloc(ny + 1);
    m->wL_f2cx = dalloc(nx); m->wR_f2cx = dalloc(nx);
    m->wB_f2cy = dalloc(ny); m->wT_f2cy = dalloc(ny);
    m->dx_west_corner = dalloc(nx + 1); m->dx_east_corner = dalloc(nx + 1);
    m->dy_south_corner = dalloc(ny + 1); m->dy_north_corner = dalloc(ny + 1);
    m->x_p = dalloc(nx); m->x_u = dalloc(nx + 1); m->y_p = dalloc(ny); m->y_v = dalloc(ny + 1);
    axis_metrics(nx, c->lx, c->px, m->x_p, m->x_u, m->dx_f, m->dx_c, m->wL_c2fx, m->wR_c2fx,
                 m->wL_f2cx, m->wR_f2cx, m->dx_west_corner, m->dx_east_corner);
    axis_metrics(ny, c->ly, c->py, m->y_p, m->y_v, m->dy_f, m->dy_c, m->wB_c2fy, m->wT_c2fy,
                 m->wB_f2cy, m->wT_f2cy, m->dy_south_corner, m->dy_north_corner);
}

const double* fsio_metric(const fsio_ctx* c, int which, int* len) {
    const metrics_t* m = &c->mt;
    const double* tab[20] = {m->dx_f, m->dx_c, m->dy_f, m->dy_c, m->wL_c2fx, m->wR_c2fx,
                             m->wB_c2fy, m->wT_c2fy, m->wL_f2cx, m->wR_f2cx, m->wB_f2cy,
                             m->wT_f2cy, m->dx_west_corner, m->dx_east_corner,
                             m->dy_south_corner, m->dy_north_corner, m->x_p, m->y_p, m->x_u,
                             m->y_v};
    int lens[20] = {c->nx + 1, c->nx, c->ny + 1, c->ny, c->nx + 1, c->nx + 1, c->ny + 1,
                    c->ny + 1, c->nx, c->nx, c->ny, c->ny, c->nx + 1, c->nx + 1, c->ny + 1,
                    c->ny + 1, c->nx, c->ny, c->nx + 1, c->ny + 1};
    if (which < 0 || which >= 20) return NULL;
    if (len) *len = lens[which];
    return tab[which];
}

/* ---------------------------------------------------------------------------
 * Generic MAC helpers (src/Operators.jl)
 * ------------------------------------------------------------------------- */

/* interpolate_center_to_x_faces, Operators.jl:110-134 */
static void c2fx(fsio_ctx* c, const double* f, double* face) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny;
    OMP_FOR
    for (int j = 0; j < ny; ++j) {
        if (c->px) {
            double v = m->wL_c2fx[0] * CC(f, nx - 1, j) + m->wR_c2fx[0] * CC(f, 0, j);
            UX(face, 0, j) = v;
            UX(face, nx, j) = v;
        } else {
            UX(face, 0, j) = CC(f, 0, j);
            UX(face, nx, j) = CC(f, nx - 1, j);
        }
        for (int i = 1; i < nx; ++i)
            UX(face, i, j) = m->wL_c2fx[i] * CC(f, i - 1, j) + m->wR_c2fx[i] * CC(f, i, j);
    }
}

/* interpolate_center_to_y_faces, Operators.jl:137-159 */
static void c2fy(fsio_ctx* c, const double* f, double* face) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny;
    OMP_FOR
    for (int i = 0; i < nx; ++i) {
        if (c->py) {
            double v = m->wB_c2fy[0] * CC(f, i, ny - 1) + m->wT_c2fy[0] * CC(f, i, 0);
            VY(face, i, 0) = v;
            VY(face, i, ny) = v;
        } else {
            VY(face, i, 0) = CC(f, i, 0);
            VY(face, i, ny) = CC(f, i, ny - 1);
        }
        for (int j = 1; j < ny; ++j)
            VY(face, i, j) = m->wB_c2fy[j] * CC(f, i, j - 1) + m->wT_c2fy[j] * CC(f, i, j);
    }
}

/* divergence, Operators.jl:407-425 */
static void divergence(fsio_ctx* c, const double* U, const double* V, double* out) {
    const metrics_t* m = &c->mt;
    OMP_FOR
    for (int j = 0; j < c->ny; ++j)
        for (int i = 0; i < c->nx; ++i)
            CC(out, i, j) = (UX(U, i + 1, j) - UX(U, i, j)) / m->dx_c[i] +
                            (VY(V, i, j + 1) - VY(V, i, j)) / m->dy_c[j];
}
void fsio_divergence(fsio_ctx* c, const double* U, const double* V, double* out) {
    divergence(c, U, V, out);
}

/* scalar_gradient = gradient matrices + boundary vectors,
 * build_scalar_differential_operators Operators.jl:189-356 and :395-404.
 * Supports Periodic / Dirichlet / Neumann sides. */
static void scalar_gradient(fsio_ctx* c, const fsi_bc_setup* bc, const double* q, double* gx,
                            double* gy) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny;
    OMP_FOR
    for (int j = 0; j < ny; ++j) {
        /* left face (:216-238) */
        if (c->px) {
            double inv_h = 1.0 / m->dx_f[0];
            UX(gx, 0, j) = inv_h * CC(q, 0, j) - inv_h * CC(q, nx - 1, j);
        } else if (bc->left.kind == FSI_BC_DIRICHLET) {
            double dist = m->dx_f[0] / 2.0;
            UX(gx, 0, j) = (1.0 / dist) * CC(q, 0, j) + (-bc->left.value / dist);
        } else {
            UX(gx, 0, j) = -bc->left.value;
        }
        for (int i = 1; i < nx; ++i) { /* :240-246 */
            double inv_h = 1.0 / m->dx_f[i];
            UX(gx, i, j) = inv_h * CC(q, i, j) - inv_h * CC(q, i - 1, j);
        }
        /* right face (:248-272): periodic duplicates face 1 */
        if (c->px) {
            UX(gx, nx, j) = UX(gx, 0, j);
        } else if (bc->right.kind == FSI_BC_DIRICHLET) {
            double dist = m->dx_f[nx] / 2.0;
            UX(gx, nx, j) = (-1.0 / dist) * CC(q, nx - 1, j) + bc->right.value / dist;
        } else {
            UX(gx, nx, j) = bc->right.value;
        }
    }
    OMP_FOR
    for (int i = 0; i < nx; ++i) { /* :275-334 */
        if (c->py) {
            double inv_h = 1.0 / m->dy_f[0];
            VY(gy, i, 0) = inv_h * CC(q, i, 0) - inv_h * CC(q, i, ny - 1);
        } else if (bc->bottom.kind == FSI_BC_DIRICHLET) {
            double dist = m->dy_f[0] / 2.0;
            VY(gy, i, 0) = (1.0 / dist) * CC(q, i, 0) + (-bc->bottom.value / dist);
        } else {
            VY(gy, i, 0) = -bc->bottom.value;
        }
        for (int j = 1; j < ny; ++j) {
            double inv_h = 1.0 / m->dy_f[j];
            VY(gy, i, j) = inv_h * CC(q, i, j) - inv_h * CC(q, i, j - 1);
        }
        if (c->py) {
            VY(gy, i, ny) = VY(gy, i, 0);
        } else if (bc->top.kind == FSI_BC_DIRICHLET) {
            double dist = m->dy_f[ny] / 2.0;
            VY(gy, i, ny) = (-1.0 / dist) * CC(q, i, ny - 1) + bc->top.value / dist;
        } else {
            VY(gy, i, ny) = bc->top.value;
        }
    }
}
void fsio_scalar_gradient(fsio_ctx* c, int which, const double* q, double* gx, double* gy) {
    scalar_gradient(c, which == 0 ? &c->bcs.correction : &c->bcs.phi, q, gx, gy);
}

/* velocity_gradient, Operators.jl:431-586 (Nx, Ny >= 2; scalar wall data). */
static int velocity_gradient(fsio_ctx* c, const double* U, const double* V, double* du_dx,
                             double* du_dy, double* dv_dx, double* dv_dy) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny;
    const fsi_bc_setup* bu = &c->bcs.U;
    const fsi_bc_setup* bv = &c->bcs.V;
    double* Uc = dalloc((size_t)nx * ny);
    double* Vc = dalloc((size_t)nx * ny);
    OMP_FOR
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            CC(du_dx, i, j) = (UX(U, i + 1, j) - UX(U, i, j)) / m->dx_c[i];
            CC(dv_dy, i, j) = (VY(V, i, j + 1) - VY(V, i, j)) / m->dy_c[j];
            CC(Uc, i, j) = m->wL_f2cx[i] * UX(U, i, j) + m->wR_f2cx[i] * UX(U, i + 1, j);
            CC(Vc, i, j) = m->wB_f2cy[j] * VY(V, i, j) + m->wT_f2cy[j] * VY(V, i, j + 1);
        }
    int rc = 0;
    for (int i = 0; i < nx && !rc; ++i) { /* du/dy, :472-522 */
        if (c->py) {
            for (int j = 0; j < ny; ++j) {
                int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
                CC(du_dy, i, j) = (CC(Uc, i, jp) - CC(Uc, i, jm)) / (m->dy_f[j] + m->dy_f[j + 1]);
            }
        } else {
            if (bu->bottom.kind == FSI_BC_DIRICHLET)
                CC(du_dy, i, 0) = (CC(Uc, i, 0) - bu->bottom.value) / m->dy_north_corner[0];
            else if (bu->bottom.kind == FSI_BC_NEUMANN)
                CC(du_dy, i, 0) = -bu->bottom.value;
            else
                rc = set_err(c, FSI_E_ARGUMENT, "Unsupported U-bottom boundary for du/dy.");
            for (int j = 1; j < ny - 1; ++j)
                CC(du_dy, i, j) = (CC(Uc, i, j + 1) - CC(Uc, i, j - 1)) / (m->dy_f[j] + m->dy_f[j + 1]);
            if (bu->top.kind == FSI_BC_DIRICHLET)
                CC(du_dy, i, ny - 1) = (bu->top.value - CC(Uc, i, ny - 1)) / m->dy_south_corner[ny];
            else if (bu->top.kind == FSI_BC_NEUMANN)
                CC(du_dy, i, ny - 1) = bu->top.value;
            else
                rc = set_err(c, FSI_E_ARGUMENT, "Unsupported U-top boundary for du/dy.");
        }
    }
    for (int j = 0; j < ny && !rc; ++j) { /* dv/dx, :525-575 */
        if (c->px) {
            for (int i = 0; i < nx; ++i) {
                int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
                CC(dv_dx, i, j) = (CC(Vc, ip, j) - CC(Vc, im, j)) / (m->dx_f[i] + m->dx_f[i + 1]);
            }
        } else {
            if (bv->left.kind == FSI_BC_DIRICHLET)
                CC(dv_dx, 0, j) = (CC(Vc, 0, j) - bv->left.value) / m->dx_east_corner[0];
            else if (bv->left.kind == FSI_BC_NEUMANN)
                CC(dv_dx, 0, j) = -bv->left.value;
            else
                rc = set_err(c, FSI_E_ARGUMENT, "Unsupported V-left boundary for dv/dx.");
            for (int i = 1; i < nx - 1; ++i)
                CC(dv_dx, i, j) = (CC(Vc, i + 1, j) - CC(Vc, i - 1, j)) / (m->dx_f[i] + m->dx_f[i + 1]);
            if (bv->right.kind == FSI_BC_DIRICHLET)
                CC(dv_dx, nx - 1, j) = (bv->right.value - CC(Vc, nx - 1, j)) / m->dx_west_corner[nx];
            else if (bv->right.kind == FSI_BC_NEUMANN)
                CC(dv_dx, nx - 1, j) = bv->right.value;
            else
                rc = set_err(c, FSI_E_ARGUMENT, "Unsupported V-right boundary for dv/dx.");
        }
    }
    free(Uc);
    free(Vc);
    return rc;
}
int fsio_velocity_gradient(fsio_ctx* c, const double* U, const double* V, double* dudx,
                           double* dudy, double* dvdx, double* dvdy) {
    return velocity_gradient(c, U, V, dudx, dudy, dvdx, dvdy);
}

/* ---------------------------------------------------------------------------
 * symmetric_stress_divergence, Operators.jl:620-735, with the linear
 * extrapolation helpers :738-842.
 * ------------------------------------------------------------------------- */
static double lex_left(fsio_ctx* c, const double* f, int j) { /* :738-744 */
    double d = c->mt.dx_east_corner[0];
    return CC(f, 0, j) - (CC(f, 1, j) - CC(f, 0, j)) * d / c->mt.dx_f[1];
}
static double lex_right(fsio_ctx* c, const double* f, int j) { /* :746-752 */
    int nx = c->nx;
    double d = c->mt.dx_west_corner[nx];
    return CC(f, nx - 1, j) + (CC(f, nx - 1, j) - CC(f, nx - 2, j)) * d / c->mt.dx_f[nx - 1];
}
static double lex_bottom(fsio_ctx* c, const double* f, int i) { /* :754-760 */
    double d = c->mt.dy_north_corner[0];
    return CC(f, i, 0) - (CC(f, i, 1) - CC(f, i, 0)) * d / c->mt.dy_f[1];
}
static double lex_top(fsio_ctx* c, const double* f, int i) { /* :762-768 */
    int ny = c->ny;
    double d = c->mt.dy_south_corner[ny];
    return CC(f, i, ny - 1) + (CC(f, i, ny - 1) - CC(f, i, ny - 2)) * d / c->mt.dy_f[ny - 1];
}

static void stress_divergence(fsio_ctx* c, const double* Txx, const double* Tyy,
                              const double* Txy, double* Fx, double* Fy) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny;
    double* Txy_x = dalloc((size_t)(nx + 1) * ny);
    double* Txy_y = dalloc((size_t)nx * (ny + 1));
    /* _center_to_x_faces_linear_boundary :770-789 */
    OMP_FOR
    for (int j = 0; j < ny; ++j) {
        if (c->px) {
            double v = m->wL_c2fx[0] * CC(Txy, nx - 1, j) + m->wR_c2fx[0] * CC(Txy, 0, j);
            UX(Txy_x, 0, j) = v;
            UX(Txy_x, nx, j) = v;
        } else {
            UX(Txy_x, 0, j) = lex_left(c, Txy, j);
            UX(Txy_x, nx, j) = lex_right(c, Txy, j);
        }
        for (int i = 1; i < nx; ++i)
            UX(Txy_x, i, j) = m->wL_c2fx[i] * CC(Txy, i - 1, j) + m->wR_c2fx[i] * CC(Txy, i, j);
    }
    /* _center_to_y_faces_linear_boundary :791-810 */
    OMP_FOR
    for (int i = 0; i < nx; ++i) {
        if (c->py) {
            double v = m->wB_c2fy[0] * CC(Txy, i, ny - 1) + m->wT_c2fy[0] * CC(Txy, i, 0);
            VY(Txy_y, i, 0) = v;
            VY(Txy_y, i, ny) = v;
        } else {
            VY(Txy_y, i, 0) = lex_bottom(c, Txy, i);
            VY(Txy_y, i, ny) = lex_top(c, Txy, i);
        }
        for (int j = 1; j < ny; ++j)
            VY(Txy_y, i, j) = m->wB_c2fy[j] * CC(Txy, i, j - 1) + m->wT_c2fy[j] * CC(Txy, i, j);
    }
    /* d(Txx)/dx on x-faces :642-656 */
    OMP_FOR
    for (int j = 0; j < ny; ++j) {
        if (c->px) {
            double w = (CC(Txx, 0, j) - CC(Txx, nx - 1, j)) / m->dx_f[0];
            UX(Fx, 0, j) = w;
            UX(Fx, nx, j) = w;
        } else {
            double lw = lex_left(c, Txx, j), rw = lex_right(c, Txx, j);
            UX(Fx, 0, j) = (CC(Txx, 0, j) - lw) / m->dx_east_corner[0];
            UX(Fx, nx, j) = (rw - CC(Txx, nx - 1, j)) / m->dx_west_corner[nx];
        }
        for (int i = 1; i < nx; ++i) UX(Fx, i, j) = (CC(Txx, i, j) - CC(Txx, i - 1, j)) / m->dx_f[i];
    }
    /* + d(Txy)/dy at U locations :659-683 (extrapolate along y of the x-face field) */
    OMP_FOR
    for (int i = 0; i <= nx; ++i) {
        if (c->py) {
            for (int j = 0; j < ny; ++j) {
                int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
                UX(Fx, i, j) += (UX(Txy_x, i, jp) - UX(Txy_x, i, jm)) / (m->dy_f[j] + m->dy_f[j + 1]);
            }
        } else {
            double d0 = m->dy_north_corner[0];
            double bw = UX(Txy_x, i, 0) - (UX(Txy_x, i, 1) - UX(Txy_x, i, 0)) * d0 / m->dy_f[1];
            double d1 = m->dy_south_corner[ny];
            double tw = UX(Txy_x, i, ny - 1) +
                        (UX(Txy_x, i, ny - 1) - UX(Txy_x, i, ny - 2)) * d1 / m->dy_f[ny - 1];
            UX(Fx, i, 0) += (UX(Txy_x, i, 0) - bw) / m->dy_north_corner[0];
            for (int j = 1; j < ny - 1; ++j)
                UX(Fx, i, j) += (UX(Txy_x, i, j + 1) - UX(Txy_x, i, j - 1)) / (m->dy_f[j] + m->dy_f[j + 1]);
            UX(Fx, i, ny - 1) += (tw - UX(Txy_x, i, ny - 1)) / m->dy_south_corner[ny];
        }
    }
    /* d(Tyy)/dy on y-faces :686-700 */
    OMP_FOR
    for (int i = 0; i < nx; ++i) {
        if (c->py) {
            double w = (CC(Tyy, i, 0) - CC(Tyy, i, ny - 1)) / m->dy_f[0];
            VY(Fy, i, 0) = w;
            VY(Fy, i, ny) = w;
        } else {
            double bw = lex_bottom(c, Tyy, i), tw = lex_top(c, Tyy, i);
            VY(Fy, i, 0) = (CC(Tyy, i, 0) - bw) / m->dy_north_corner[0];
            VY(Fy, i, ny) = (tw - CC(Tyy, i, ny - 1)) / m->dy_south_corner[ny];
        }
        for (int j = 1; j < ny; ++j) VY(Fy, i, j) = (CC(Tyy, i, j) - CC(Tyy, i, j - 1)) / m->dy_f[j];
    }
    /* + d(Txy)/dx at V locations :702-724 */
    OMP_FOR
    for (int j = 0; j <= ny; ++j) {
        if (c->px) {
            for (int i = 0; i < nx; ++i) {
                int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
                VY(Fy, i, j) += (VY(Txy_y, ip, j) - VY(Txy_y, im, j)) / (m->dx_f[i] + m->dx_f[i + 1]);
            }
        } else {
            double d0 = m->dx_east_corner[0];
            double lw = VY(Txy_y, 0, j) - (VY(Txy_y, 1, j) - VY(Txy_y, 0, j)) * d0 / m->dx_f[1];
            double d1 = m->dx_west_corner[nx];
            double rw = VY(Txy_y, nx - 1, j) +
                        (VY(Txy_y, nx - 1, j) - VY(Txy_y, nx - 2, j)) * d1 / m->dx_f[nx - 1];
            VY(Fy, 0, j) += (VY(Txy_y, 0, j) - lw) / m->dx_east_corner[0];
            for (int i = 1; i < nx - 1; ++i)
                VY(Fy, i, j) += (VY(Txy_y, i + 1, j) - VY(Txy_y, i - 1, j)) / (m->dx_f[i] + m->dx_f[i + 1]);
            VY(Fy, nx - 1, j) += (rw - VY(Txy_y, nx - 1, j)) / m->dx_west_corner[nx];
        }
    }
    if (c->px) /* :726-731 */
        for (int j = 0; j < ny; ++j) UX(Fx, nx, j) = UX(Fx, 0, j);
    if (c->py)
        for (int i = 0; i < nx; ++i) VY(Fy, i, ny) = VY(Fy, i, 0);
    free(Txy_x);
    free(Txy_y);
}
void fsio_stress_divergence(fsio_ctx* c, const double* Txx, const double* Tyy,
                            const double* Txy, double* Fx, double* Fy) {
    stress_divergence(c, Txx, Tyy, Txy, Fx, Fy);
}

/* ---------------------------------------------------------------------------
 * material_convection, src/MomentumTransport.jl:252-338
 * ------------------------------------------------------------------------- */
/* _tangential_u_boundary_value :68-103 */
static double tangential_u(fsio_ctx* c, const double* U, int ii, int top) {
    const fsi_bc_side* b = top ? &c->bcs.U.top : &c->bcs.U.bottom;
    if (b->kind == FSI_BC_DIRICHLET) return b->value;
    double adjacent = top ? UX(U, ii, c->ny - 1) : UX(U, ii, 0);
    double distance = top ? c->mt.dy_f[c->ny] / 2.0 : c->mt.dy_f[0] / 2.0;
    return adjacent + b->value * distance;
}
/* _tangential_v_boundary_value :105-139 */
static double tangential_v(fsio_ctx* c, const double* V, int jj, int right) {
    const fsi_bc_side* b = right ? &c->bcs.V.right : &c->bcs.V.left;
    if (b->kind == FSI_BC_DIRICHLET) return b->value;
    double adjacent = right ? VY(V, c->nx - 1, jj) : VY(V, 0, jj);
    double distance = right ? c->mt.dx_f[c->nx] / 2.0 : c->mt.dx_f[0] / 2.0;
    return adjacent + b->value * distance;
}
/* _u_at_corner :161-189 */
static double u_at_corner(fsio_ctx* c, const double* U, int i, int j) {
    const metrics_t* m = &c->mt;
    int ny = c->ny;
    int ii = (c->px && i == c->nx) ? 0 : i;
    if (c->py) {
        if (j == 0 || j == ny) return m->wB_c2fy[0] * UX(U, ii, ny - 1) + m->wT_c2fy[0] * UX(U, ii, 0);
        return m->wB_c2fy[j] * UX(U, ii, j - 1) + m->wT_c2fy[j] * UX(U, ii, j);
    }
    if (j == 0) return tangential_u(c, U, ii, 0);
    if (j == ny) return tangential_u(c, U, ii, 1);
    return m->wB_c2fy[j] * UX(U, ii, j - 1) + m->wT_c2fy[j] * UX(U, ii, j);
}
/* _v_at_corner :191-219 */
static double v_at_corner(fsio_ctx* c, const double* V, int i, int j) {
    const metrics_t* m = &c->mt;
    int nx = c->nx;
    int jj = (c->py && j == c->ny) ? 0 : j;
    if (c->px) {
        if (i == 0 || i == nx) return m->wL_c2fx[0] * VY(V, nx - 1, jj) + m->wR_c2fx[0] * VY(V, 0, jj);
        return m->wL_c2fx[i] * VY(V, i - 1, jj) + m->wR_c2fx[i] * VY(V, i, jj);
    }
    if (i == 0) return tangential_v(c, V, jj, 0);
    if (i == nx) return tangential_v(c, V, jj, 1);
    return m->wL_c2fx[i] * VY(V, i - 1, jj) + m->wR_c2fx[i] * VY(V, i, jj);
}

static void material_convection(fsio_ctx* c, const double* U, const double* V, const double* S,
                                 double* Cu, double* Cv) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny;
    size_t nk = (size_t)(nx + 1) * (ny + 1);
    double* u2 = dalloc((size_t)nx * ny);
    double* v2 = dalloc((size_t)nx * ny);
    double* uv = dalloc(nk);
    /* _normal_flux_squares :141-159 */
    OMP_FOR
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double ur = (c->px && i == nx - 1) ? UX(U, 0, j) : UX(U, i + 1, j);
            double uc = m->wL_f2cx[i] * UX(U, i, j) + m->wR_f2cx[i] * ur;
            CC(u2, i, j) = uc * uc;
            double vt = (c->py && j == ny - 1) ? VY(V, i, 0) : VY(V, i, j + 1);
            double vc = m->wB_f2cy[j] * VY(V, i, j) + m->wT_f2cy[j] * vt;
            CC(v2, i, j) = vc * vc;
        }
    /* _cross_flux_corners :221-247 */
#define KK(a, i, j) (a)[(size_t)(i) + (size_t)(j) * (size_t)(nx + 1)]
    OMP_FOR
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) KK(uv, i, j) = u_at_corner(c, U, i, j) * v_at_corner(c, V, i, j);
    if (c->px)
        for (int j = 0; j <= ny; ++j) KK(uv, nx, j) = KK(uv, 0, j);
    if (c->py)
        for (int i = 0; i <= nx; ++i) KK(uv, i, ny) = KK(uv, i, 0);
    /* conservative_convection :268-295 */
    memset(Cu, 0, sizeof(double) * (size_t)(nx + 1) * ny);
    memset(Cv, 0, sizeof(double) * (size_t)nx * (ny + 1));
    OMP_FOR
    for (int j = 0; j < ny; ++j) {
        if (c->px) {
            UX(Cu, 0, j) = (CC(u2, 0, j) - CC(u2, nx - 1, j)) / m->dx_f[0] +
                           (KK(uv, 0, j + 1) - KK(uv, 0, j)) / m->dy_c[j];
            UX(Cu, nx, j) = UX(Cu, 0, j);
        }
        for (int i = 1; i < nx; ++i)
            UX(Cu, i, j) = (CC(u2, i, j) - CC(u2, i - 1, j)) / m->dx_f[i] +
                           (KK(uv, i, j + 1) - KK(uv, i, j)) / m->dy_c[j];
    }
    OMP_FOR
    for (int i = 0; i < nx; ++i) {
        if (c->py) {
            VY(Cv, i, 0) = (KK(uv, i + 1, 0) - KK(uv, i, 0)) / m->dx_c[i] +
                           (CC(v2, i, 0) - CC(v2, i, ny - 1)) / m->dy_f[0];
            VY(Cv, i, ny) = VY(Cv, i, 0);
        }
        for (int j = 1; j < ny; ++j)
            VY(Cv, i, j) = (KK(uv, i + 1, j) - KK(uv, i, j)) / m->dx_c[i] +
                           (CC(v2, i, j) - CC(v2, i, j - 1)) / m->dy_f[j];
    }
#undef KK
    /* material form :303-338: C -= S_face * v, then normal rows */
    double* Sx = dalloc((size_t)(nx + 1) * ny);
    double* Sy = dalloc((size_t)nx * (ny + 1));
    c2fx(c, S, Sx);
    c2fy(c, S, Sy);
    OMP_FOR
    for (size_t k = 0; k < (size_t)(nx + 1) * ny; ++k) Cu[k] -= Sx[k] * U[k];
    OMP_FOR
    for (size_t k = 0; k < (size_t)nx * (ny + 1); ++k) Cv[k] -= Sy[k] * V[k];
    for (int j = 0; j < ny; ++j) {
        if (c->px) UX(Cu, nx, j) = UX(Cu, 0, j);
        else { UX(Cu, 0, j) = 0.0; UX(Cu, nx, j) = 0.0; }
    }
    for (int i = 0; i < nx; ++i) {
        if (c->py) VY(Cv, i, ny) = VY(Cv, i, 0);
        else { VY(Cv, i, 0) = 0.0; VY(Cv, i, ny) = 0.0; }
    }
    free(Sx); free(Sy); free(u2); free(v2); free(uv);
}
void fsio_material_convection(fsio_ctx* c, const double* U, const double* V, const double* S,
                              double* Cu, double* Cv) {
    material_convection(c, U, V, S, Cu, Cv);
}

/* ---------------------------------------------------------------------------
 * Phase-field right-hand sides (src/PhaseFieldMethods.jl,
 * src/InterfaceRegularization.jl, src/InterfaceTransport.jl)
 * ------------------------------------------------------------------------- */

/* _x_boundary_state / _y_boundary_state, InterfaceTransport.jl:112-159 (and
 * the same rule for sigma, StructureTensor.jl:195-265). side: 0 left/bottom,
 * 1 right/top. */
static int upwind_boundary_state(fsio_ctx* c, const fsi_bc_side* b, double velocity, int side,
                                 double interior, double* out) {
    int inflow = side == 0 ? velocity > 0.0 : velocity < 0.0;
    if (velocity == 0.0 || !inflow) { *out = interior; return 0; }
    if (b->kind == FSI_BC_DIRICHLET) { *out = b->value; return 0; }
    return set_err(c, FSI_E_ARGUMENT,
                   "A Neumann condition does not define the incoming advected state; "
                   "prescribe Dirichlet data for inflow.");
}

/* first-order upwind conservative advection -div(v f), InterfaceTransport.jl:161-258
 * (identical to StructureTensor._conservative_upwind_advection :267-343). */
static int upwind_advection(fsio_ctx* c, const double* f, const double* U, const double* V,
                            const fsi_bc_setup* bc, double* rhs) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny;
    double* fxx = dalloc((size_t)(nx + 1) * ny);
    double* fyy = dalloc((size_t)nx * (ny + 1));
    int rc = 0;
    for (int j = 0; j < ny && !rc; ++j) {
        double u = UX(U, 0, j);
        if (c->px) {
            UX(fxx, 0, j) = u * (u >= 0.0 ? CC(f, nx - 1, j) : CC(f, 0, j));
        } else {
            double bs = 0.0;
            rc = upwind_boundary_state(c, &bc->left, u, 0, CC(f, 0, j), &bs);
            UX(fxx, 0, j) = u * (u >= 0.0 ? bs : CC(f, 0, j));
        }
        if (c->px) {
            UX(fxx, nx, j) = UX(fxx, 0, j);
        } else if (!rc) {
            u = UX(U, nx, j);
            double bs = 0.0;
            rc = upwind_boundary_state(c, &bc->right, u, 1, CC(f, nx - 1, j), &bs);
            UX(fxx, nx, j) = u * (u >= 0.0 ? CC(f, nx - 1, j) : bs);
        }
    }
    for (int i = 0; i < nx && !rc; ++i) {
        double v = VY(V, i, 0);
        if (c->py) {
            VY(fyy, i, 0) = v * (v >= 0.0 ? CC(f, i, ny - 1) : CC(f, i, 0));
        } else {
            double bs = 0.0;
            rc = upwind_boundary_state(c, &bc->bottom, v, 0, CC(f, i, 0), &bs);
            VY(fyy, i, 0) = v * (v >= 0.0 ? bs : CC(f, i, 0));
        }
        if (c->py) {
            VY(fyy, i, ny) = VY(fyy, i, 0);
        } else if (!rc) {
            v = VY(V, i, ny);
            double bs = 0.0;
            rc = upwind_boundary_state(c, &bc->top, v, 1, CC(f, i, ny - 1), &bs);
            VY(fyy, i, ny) = v * (v >= 0.0 ? CC(f, i, ny - 1) : bs);
        }
    }
    if (!rc) {
        /* the interior faces (the boundary faces above, in row order, so the
           first inflow error is the one reported) */
        OMP_FOR
        for (int j = 0; j < ny; ++j)
            for (int i = 1; i < nx; ++i) {
                double u = UX(U, i, j);
                UX(fxx, i, j) = u * (u >= 0.0 ? CC(f, i - 1, j) : CC(f, i, j));
            }
        OMP_FOR
        for (int i = 0; i < nx; ++i)
            for (int j = 1; j < ny; ++j) {
                double v = VY(V, i, j);
                VY(fyy, i, j) = v * (v >= 0.0 ? CC(f, i, j - 1) : CC(f, i, j));
            }
        OMP_FOR
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                CC(rhs, i, j) = -((UX(fxx, i + 1, j) - UX(fxx, i, j)) / m->dx_c[i] +
                                  (VY(fyy, i, j + 1) - VY(fyy, i, j)) / m->dy_c[j]);
    }
    free(fxx);
    free(fyy);
    return rc;
}

/* _center_gradient PhaseFieldMethods.jl:219-231 / _cell_centered_phase_gradient
 * InterfaceRegularization.jl:99-121: face gradients averaged to centres. */
static void center_gradient(fsio_ctx* c, const double* f, const fsi_bc_setup* bc, double* gx,
                            double* gy, double* gxf, double* gyf) {
    const metrics_t* m = &c->mt;
    scalar_gradient(c, bc, f, gxf, gyf);
    OMP_FOR
    for (int j = 0; j < c->ny; ++j)
        for (int i = 0; i < c->nx; ++i) {
            CC(gx, i, j) = m->wL_f2cx[i] * UX(gxf, i, j) + m->wR_f2cx[i] * UX(gxf, i + 1, j);
            CC(gy, i, j) = m->wB_f2cy[j] * VY(gyf, i, j) + m->wT_f2cy[j] * VY(gyf, i, j + 1);
        }
}

/* _phase_boundary_value PhaseFieldMethods.jl:174-196 */
static int phase_boundary_value(fsio_ctx* c, const fsi_bc_side* b, double interior, double* out) {
    if (b->kind == FSI_BC_DIRICHLET) { *out = b->value; return 0; }
    if (b->kind == FSI_BC_NEUMANN) {
        if (fabs(b->value) > 100 * 2.220446049250313e-16)
            return set_err(c, FSI_E_ARGUMENT,
                           "Only homogeneous Neumann phase boundaries are supported by the ACDI/UWCDI operators.");
        *out = interior;
        return 0;
    }
    return set_err(c, FSI_E_ARGUMENT, "Unsupported phase boundary condition.");
}

static void enforce_face_flux_walls(fsio_ctx* c, double* jx, double* jy) {
    /* _enforce_interface_flux_boundaries! InterfaceRegularization.jl:147-166 and
       the identical wall rule of _acdi_regularization_fluxes :637-648 */
    int nx = c->nx, ny = c->ny;
    for (int j = 0; j < ny; ++j) {
        if (c->px) UX(jx, nx, j) = UX(jx, 0, j);
        else { UX(jx, 0, j) = 0.0; UX(jx, nx, j) = 0.0; }
    }
    for (int i = 0; i < nx; ++i) {
        if (c->py) VY(jy, i, ny) = VY(jy, i, 0);
        else { VY(jy, i, 0) = 0.0; VY(jy, i, ny) = 0.0; }
    }
}

/* CDI: LegacySplitPhaseFieldMethod path PhaseFieldMethods.jl:408-435, 457-475
 * -> interface_flux InterfaceRegularization.jl:187-228
 * -> interface_explicit_rhs InterfaceTransport.jl:260-276.
 * UPWIND (NoInterfaceFlux) is the same path with j = 0. */
static int phase_rhs_cdi(fsio_ctx* c, const double* U, const double* V, const double* phi,
                         const double* S, double* out, int with_flux) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny;
    size_t nc = (size_t)nx * ny, nfx = (size_t)(nx + 1) * ny, nfy = (size_t)nx * (ny + 1);
    double* jx = dalloc(nfx);
    double* jy = dalloc(nfy);
    if (with_flux) {
        double *gx = dalloc(nc), *gy = dalloc(nc), *gxf = dalloc(nfx), *gyf = dalloc(nfy);
        double *shx = dalloc(nc), *shy = dalloc(nc), *shxf = dalloc(nfx), *shyf = dalloc(nfy);
        center_gradient(c, phi, &c->bcs.phi, gx, gy, gxf, gyf);
        double d2 = sq(c->m.cdi_delta_normal);
        for (size_t k = 0; k < nc; ++k) { /* _sharpening_vector :123-145 */
            double mag = sqrt(gx[k] * gx[k] + gy[k] * gy[k] + d2);
            double amp = phi[k] * (1.0 - phi[k]);
            shx[k] = amp * gx[k] / mag;
            shy[k] = amp * gy[k] / mag;
        }
        c2fx(c, shx, shxf);
        c2fy(c, shy, shyf);
        double ell = 0.5 * c->m.cdi_epsilon, G = c->m.cdi_inverse_peclet;
        for (size_t k = 0; k < nfx; ++k) jx[k] = G * (shxf[k] - ell * gxf[k]);
        for (size_t k = 0; k < nfy; ++k) jy[k] = G * (shyf[k] - ell * gyf[k]);
        enforce_face_flux_walls(c, jx, jy);
        free(gx); free(gy); free(gxf); free(gyf); free(shx); free(shy); free(shxf); free(shyf);
    }
    int rc = upwind_advection(c, phi, U, V, &c->bcs.phi, out);
    if (!rc)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double divj = (UX(jx, i + 1, j) - UX(jx, i, j)) / m->dx_c[i] +
                              (VY(jy, i, j + 1) - VY(jy, i, j)) / m->dy_c[j];
                CC(out, i, j) = CC(out, i, j) - divj + CC(phi, i, j) * CC(S, i, j);
            }
    free(jx);
    free(jy);
    return rc;
}

/* _psi_value PhaseFieldMethods.jl:479-483 */
static double psi_value(fsio_ctx* c, double p0) {
    double p = clamp01(p0), e = c->m.acdi_psi_floor;
    return c->m.acdi_epsilon * log((p + e) / (1.0 - p + e));
}

/* ACDI phase_field_rhs PhaseFieldMethods.jl:652-671 with
 * _acdi_convective_fluxes :531-574, _acdi_effective_gamma :576-596,
 * _acdi_regularization_fluxes :598-650, _psi_boundary_setup :493-529. */
static int phase_rhs_acdi(fsio_ctx* c, const double* U, const double* V, const double* phi,
                          const double* S, double* out) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny, rc = 0;
    size_t nc = (size_t)nx * ny, nfx = (size_t)(nx + 1) * ny, nfy = (size_t)nx * (ny + 1);
    double *fx = dalloc(nfx), *fy = dalloc(nfy);
    /* convective fluxes */
    for (int j = 0; j < ny && !rc; ++j) {
        if (c->px) {
            double pf = m->wL_c2fx[0] * CC(phi, nx - 1, j) + m->wR_c2fx[0] * CC(phi, 0, j);
            UX(fx, 0, j) = UX(U, 0, j) * pf;
            UX(fx, nx, j) = UX(fx, 0, j);
        } else {
            double lb = 0.0, rb = 0.0;
            rc = phase_boundary_value(c, &c->bcs.phi.left, CC(phi, 0, j), &lb);
            if (!rc) rc = phase_boundary_value(c, &c->bcs.phi.right, CC(phi, nx - 1, j), &rb);
            if (rc) break;
            UX(fx, 0, j) = UX(U, 0, j) * lb;
            UX(fx, nx, j) = UX(U, nx, j) * rb;
        }
        for (int i = 1; i < nx; ++i) {
            double pf = m->wL_c2fx[i] * CC(phi, i - 1, j) + m->wR_c2fx[i] * CC(phi, i, j);
            UX(fx, i, j) = UX(U, i, j) * pf;
        }
    }
    for (int i = 0; i < nx && !rc; ++i) {
        if (c->py) {
            double pf = m->wB_c2fy[0] * CC(phi, i, ny - 1) + m->wT_c2fy[0] * CC(phi, i, 0);
            VY(fy, i, 0) = VY(V, i, 0) * pf;
            VY(fy, i, ny) = VY(fy, i, 0);
        } else {
            double bb = 0.0, tb = 0.0;
            rc = phase_boundary_value(c, &c->bcs.phi.bottom, CC(phi, i, 0), &bb);
            if (!rc) rc = phase_boundary_value(c, &c->bcs.phi.top, CC(phi, i, ny - 1), &tb);
            if (rc) break;
            VY(fy, i, 0) = VY(V, i, 0) * bb;
            VY(fy, i, ny) = VY(V, i, ny) * tb;
        }
        for (int j = 1; j < ny; ++j) {
            double pf = m->wB_c2fy[j] * CC(phi, i, j - 1) + m->wT_c2fy[j] * CC(phi, i, j);
            VY(fy, i, j) = VY(V, i, j) * pf;
        }
    }
    if (rc) { free(fx); free(fy); return rc; }

    /* regularisation fluxes */
    fsi_bc_setup bpsi = c->bcs.phi;
    fsi_bc_side* sides[4] = {&bpsi.left, &bpsi.right, &bpsi.bottom, &bpsi.top};
    for (int s = 0; s < 4; ++s) {
        if (sides[s]->kind == FSI_BC_NEUMANN) {
            if (fabs(sides[s]->value) > 100 * 2.220446049250313e-16) {
                free(fx); free(fy);
                return set_err(c, FSI_E_ARGUMENT, "ACDI currently supports only homogeneous Neumann phase gradients.");
            }
            sides[s]->value = 0.0;
        } else if (sides[s]->kind == FSI_BC_DIRICHLET) {
            sides[s]->value = FSIO_VALUE(psi_value(c, sides[s]->value));
        }
    }
    double *psi = dalloc(nc), *gx = dalloc(nc), *gy = dalloc(nc), *gxf = dalloc(nfx), *gyf = dalloc(nfy);
    double *pgx = dalloc(nc), *pgy = dalloc(nc), *pgxf = dalloc(nfx), *pgyf = dalloc(nfy);
    double *nxc = dalloc(nc), *nyc = dalloc(nc), *nxf = dalloc(nfx), *nyf = dalloc(nfy);
    double *psxf = dalloc(nfx), *psyf = dalloc(nfy), *Ax = dalloc(nfx), *Ay = dalloc(nfy);
    OMP_FOR
    for (size_t k = 0; k < nc; ++k) psi[k] = psi_value(c, phi[k]);
    center_gradient(c, phi, &c->bcs.phi, gx, gy, gxf, gyf);
    center_gradient(c, psi, &bpsi, pgx, pgy, pgxf, pgyf);
    double d2 = sq(c->m.acdi_delta_normal);
    OMP_FOR
    for (size_t k = 0; k < nc; ++k) { /* _center_normals :233-244 */
        double mag = sqrt(pgx[k] * pgx[k] + pgy[k] * pgy[k] + d2);
        nxc[k] = pgx[k] / mag;
        nyc[k] = pgy[k] / mag;
    }
    c2fx(c, nxc, nxf);
    c2fy(c, nyc, nyf);
    c2fx(c, psi, psxf);
    c2fy(c, psi, psyf);
    double gam = c->m.acdi_gamma;
    if (c->m.acdi_dynamic_gamma) { /* :576-596 */
        double umax = 0.0;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double uc = m->wL_f2cx[i] * UX(U, i, j) + m->wR_f2cx[i] * UX(U, i + 1, j);
                double vc = m->wB_f2cy[j] * VY(V, i, j) + m->wT_f2cy[j] * VY(V, i, j + 1);
                double h = hypot(uc, vc);
                if (h > umax) umax = h;
            }
        gam = gam * umax;
    }
    double eps = c->m.acdi_epsilon;
    OMP_FOR
    for (size_t k = 0; k < nfx; ++k) {
        double t = tanh(psxf[k] / (2.0 * eps));
        double amp = 0.25 * (1.0 - t * t);
        Ax[k] = gam * (eps * gxf[k] - amp * nxf[k]);
    }
    OMP_FOR
    for (size_t k = 0; k < nfy; ++k) {
        double t = tanh(psyf[k] / (2.0 * eps));
        double amp = 0.25 * (1.0 - t * t);
        Ay[k] = gam * (eps * gyf[k] - amp * nyf[k]);
    }
    enforce_face_flux_walls(c, Ax, Ay);
    OMP_FOR
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double dF = (UX(fx, i + 1, j) - UX(fx, i, j)) / m->dx_c[i] +
                        (VY(fy, i, j + 1) - VY(fy, i, j)) / m->dy_c[j];
            double dA = (UX(Ax, i + 1, j) - UX(Ax, i, j)) / m->dx_c[i] +
                        (VY(Ay, i, j + 1) - VY(Ay, i, j)) / m->dy_c[j];
            CC(out, i, j) = -dF + dA + CC(phi, i, j) * CC(S, i, j);
        }
    free(fx); free(fy); free(psi); free(gx); free(gy); free(gxf); free(gyf);
    free(pgx); free(pgy); free(pgxf); free(pgyf); free(nxc); free(nyc); free(nxf); free(nyf);
    free(psxf); free(psyf); free(Ax); free(Ay);
    return 0;
}

/* UWCDI phase_field_rhs PhaseFieldMethods.jl:852-867 with _uwcdi_fluxes
 * :675-766 and _normal_components_on_faces :246-278 (:standard). */
/* The 9-point filter of _smoothed9_field (PhaseFieldMethods.jl:364-383) and
 * of _smooth_phase_for_capillarity (Capillarity.jl:81-111): index clamped at a
 * wall, wrapped across a periodic side. */
static void smooth9(fsio_ctx* c, const double* f, double* out) {
    int nx = c->nx, ny = c->ny;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int im = i - 1, ip = i + 1, jm = j - 1, jp = j + 1;
            if (c->px) { im = (im + nx) % nx; ip = ip % nx; }
            else { im = im < 0 ? 0 : im; ip = ip > nx - 1 ? nx - 1 : ip; }
            if (c->py) { jm = (jm + ny) % ny; jp = jp % ny; }
            else { jm = jm < 0 ? 0 : jm; jp = jp > ny - 1 ? ny - 1 : jp; }
            CC(out, i, j) = 0.25 * CC(f, i, j) +
                            0.125 * (CC(f, ip, j) + CC(f, im, j) + CC(f, i, jp) + CC(f, i, jm)) +
                            0.0625 * (CC(f, ip, jp) + CC(f, im, jp) + CC(f, ip, jm) + CC(f, im, jm));
        }
}

/* Cell normals of the UWCDI model: _center_normals of phi (:standard), of
 * the 9-point-filtered phi (:smoothed9, :341-404, homogeneous Neumann or
 * periodic phase sides only), or Youngs' 3x3 stencil on every cell off the
 * non-periodic walls (:youngs, :280-339). */
static int uwcdi_center_normals(fsio_ctx* c, const double* phi, double* nxc, double* nyc) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny, method = c->m.uwcdi_normal_method;
    size_t nc = (size_t)nx * ny, nfx = (size_t)(nx + 1) * ny, nfy = (size_t)nx * (ny + 1);
    double dn = c->m.uwcdi_delta_normal, d2 = dn * dn;
    const double* field = phi;
    double* smooth = NULL;
    if (method == FSI_UWCDI_NORMAL_SMOOTHED9) {
        const fsi_bc_side* sides[4] = {&c->bcs.phi.left, &c->bcs.phi.right, &c->bcs.phi.bottom, &c->bcs.phi.top};
        const char* names[4] = {"left", "right", "bottom", "top"};
        for (int k = 0; k < 4; ++k) {
            if (sides[k]->kind == FSI_BC_PERIODIC) continue;
            if (sides[k]->kind == FSI_BC_NEUMANN && fabs(sides[k]->value) <= 100 * 2.220446049250313e-16) continue;
            return set_err(c, FSI_E_ARGUMENT,
                           "smoothed9 normal reconstruction requires homogeneous Neumann or periodic phase "
                           "boundaries; %s does not.", names[k]);
        }
        smooth = dalloc(nc);
        smooth9(c, phi, smooth);
        field = smooth;
    }
    double *gx = dalloc(nc), *gy = dalloc(nc), *gxf = dalloc(nfx), *gyf = dalloc(nfy);
    center_gradient(c, field, &c->bcs.phi, gx, gy, gxf, gyf);
    for (size_t k = 0; k < nc; ++k) {
        double mag = sqrt(gx[k] * gx[k] + gy[k] * gy[k] + d2);
        nxc[k] = gx[k] / mag;
        nyc[k] = gy[k] / mag;
    }
    if (method == FSI_UWCDI_NORMAL_YOUNGS) {
        double hx = m->dx_c[0], hy = m->dy_c[0], ex = 0.0, ey = 0.0;
        for (int i = 0; i < nx; ++i) ex = fmax(ex, fabs(m->dx_c[i] - hx));
        for (int j = 0; j < ny; ++j) ey = fmax(ey, fabs(m->dy_c[j] - hy));
        if (!(ex <= 100 * 2.220446049250313e-16 * fmax(1.0, fabs(hx))) ||
            !(ey <= 100 * 2.220446049250313e-16 * fmax(1.0, fabs(hy)))) {
            free(gx); free(gy); free(gxf); free(gyf);
            return set_err(c, FSI_E_ARGUMENT, "Youngs normal reconstruction currently requires a uniform %s grid.",
                           ex <= 100 * 2.220446049250313e-16 * fmax(1.0, fabs(hx)) ? "y" : "x");
        }
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if ((!c->px && (i == 0 || i == nx - 1)) || (!c->py && (j == 0 || j == ny - 1))) continue;
                int im = c->px ? (i - 1 + nx) % nx : i - 1, ip = c->px ? (i + 1) % nx : i + 1;
                int jm = c->py ? (j - 1 + ny) % ny : j - 1, jp = c->py ? (j + 1) % ny : j + 1;
                double ggx = (CC(phi, ip, jm) + 2.0 * CC(phi, ip, j) + CC(phi, ip, jp) - CC(phi, im, jm) -
                              2.0 * CC(phi, im, j) - CC(phi, im, jp)) / (8.0 * hx);
                double ggy = (CC(phi, im, jp) + 2.0 * CC(phi, i, jp) + CC(phi, ip, jp) - CC(phi, im, jm) -
                              2.0 * CC(phi, i, jm) - CC(phi, ip, jm)) / (8.0 * hy);
                double mag = sqrt(ggx * ggx + ggy * ggy + d2);
                CC(nxc, i, j) = ggx / mag;
                CC(nyc, i, j) = ggy / mag;
            }
    }
    free(gx); free(gy); free(gxf); free(gyf);
    if (smooth) free(smooth);
    return 0;
}

int fsio_uwcdi_center_normals(fsio_ctx* c, const double* phi, double* nxc, double* nyc) {
    return uwcdi_center_normals(c, phi, nxc, nyc);
}
void fsio_smooth9(fsio_ctx* c, const double* f, double* out) { smooth9(c, f, out); }

/* dir: 0 both directions (phase_field_rhs, with the phi S source), 1 the
 * x-fluxes alone, 2 the y-fluxes alone (uwcdi_directional_rhs :768-791) */
static int phase_rhs_uwcdi_dir(fsio_ctx* c, const double* U, const double* V, const double* phi,
                               const double* S, double* out, int dir) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny, rc = 0;
    size_t nc = (size_t)nx * ny, nfx = (size_t)(nx + 1) * ny, nfy = (size_t)nx * (ny + 1);
    double *nxc = dalloc(nc), *nyc = dalloc(nc), *nxf = dalloc(nfx), *nyf = dalloc(nfy);
    double *nyx = dalloc(nfx), *nxy = dalloc(nfy), *fx = dalloc(nfx), *fy = dalloc(nfy);
    double dn = c->m.uwcdi_delta_normal, d2 = dn * dn;
    rc = uwcdi_center_normals(c, phi, nxc, nyc);
    if (rc) {
        free(nxc); free(nyc); free(nxf); free(nyf); free(nyx); free(nxy); free(fx); free(fy);
        return rc;
    }
    c2fx(c, nxc, nxf);
    c2fy(c, nyc, nyf);
    c2fx(c, nyc, nyx);
    c2fy(c, nxc, nxy);
    for (size_t k = 0; k < nfx; ++k) nxf[k] /= sqrt(nxf[k] * nxf[k] + nyx[k] * nyx[k] + d2);
    for (size_t k = 0; k < nfy; ++k) nyf[k] /= sqrt(nxy[k] * nxy[k] + nyf[k] * nyf[k] + d2);
    if (c->px)
        for (int j = 0; j < ny; ++j) UX(nxf, nx, j) = UX(nxf, 0, j);
    if (c->py)
        for (int i = 0; i < nx; ++i) VY(nyf, i, ny) = VY(nyf, i, 0);
    double alpha = c->m.uwcdi_alpha;
    for (int j = 0; j < ny && !rc; ++j)
        for (int i = 0; i <= nx; ++i) {
            double u = UX(U, i, j), pu, bnd = 0.0;
            if (!c->px && i == 0) {
                rc = phase_boundary_value(c, &c->bcs.phi.left, CC(phi, 0, j), &bnd);
                if (rc) break;
                pu = u >= 0.0 ? bnd : CC(phi, 0, j);
            } else if (!c->px && i == nx) {
                rc = phase_boundary_value(c, &c->bcs.phi.right, CC(phi, nx - 1, j), &bnd);
                if (rc) break;
                pu = u >= 0.0 ? CC(phi, nx - 1, j) : bnd;
            } else {
                int iL = (i == 0 || i == nx) ? nx - 1 : i - 1, iR = (i == 0 || i == nx) ? 0 : i;
                pu = u >= 0.0 ? CC(phi, iL, j) : CC(phi, iR, j);
            }
            double r = m->dy_c[j] / m->dx_f[i];
            double ax = alpha * (r > 1.0 ? r : 1.0);
            double beta = fabs(u) / ax;
            UX(fx, i, j) = pu * (u + beta * UX(nxf, i, j) * (1.0 - pu));
        }
    for (int j = 0; j <= ny && !rc; ++j)
        for (int i = 0; i < nx; ++i) {
            double v = VY(V, i, j), pu, bnd = 0.0;
            if (!c->py && j == 0) {
                rc = phase_boundary_value(c, &c->bcs.phi.bottom, CC(phi, i, 0), &bnd);
                if (rc) break;
                pu = v >= 0.0 ? bnd : CC(phi, i, 0);
            } else if (!c->py && j == ny) {
                rc = phase_boundary_value(c, &c->bcs.phi.top, CC(phi, i, ny - 1), &bnd);
                if (rc) break;
                pu = v >= 0.0 ? CC(phi, i, ny - 1) : bnd;
            } else {
                int jB = (j == 0 || j == ny) ? ny - 1 : j - 1, jT = (j == 0 || j == ny) ? 0 : j;
                pu = v >= 0.0 ? CC(phi, i, jB) : CC(phi, i, jT);
            }
            double r = m->dx_c[i] / m->dy_f[j];
            double ay = alpha * (r > 1.0 ? r : 1.0);
            double beta = fabs(v) / ay;
            VY(fy, i, j) = pu * (v + beta * VY(nyf, i, j) * (1.0 - pu));
        }
    if (!rc) {
        if (c->px)
            for (int j = 0; j < ny; ++j) UX(fx, nx, j) = UX(fx, 0, j);
        if (c->py)
            for (int i = 0; i < nx; ++i) VY(fy, i, ny) = VY(fy, i, 0);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double dFx = (UX(fx, i + 1, j) - UX(fx, i, j)) / m->dx_c[i];
                double dFy = (VY(fy, i, j + 1) - VY(fy, i, j)) / m->dy_c[j];
                if (dir == 1) CC(out, i, j) = -(dFx + 0.0);
                else if (dir == 2) CC(out, i, j) = -(0.0 + dFy);
                else CC(out, i, j) = -(dFx + dFy) + CC(phi, i, j) * CC(S, i, j);
            }
    }
    free(nxc); free(nyc); free(nxf); free(nyf);
    free(nyx); free(nxy); free(fx); free(fy);
    return rc;
}

static int phase_rhs_uwcdi(fsio_ctx* c, const double* U, const double* V, const double* phi,
                           const double* S, double* out) {
    return phase_rhs_uwcdi_dir(c, U, V, phi, S, out, 0);
}

/* uwcdi_strang_step PhaseFieldMethods.jl:793-850: Euler sweeps x(dt/2),
 * y(dt), x(dt/2) of the directional right-hand sides (first_direction 1 = :y
 * swaps the roles); phi is replaced by the result. */
int fsio_uwcdi_strang_step(fsio_ctx* c, const double* U, const double* V, double* phi, fsio_scalar dt,
                           int first_direction) {
    if (c->m.phase_method != FSI_PHASE_UWCDI)
        return set_err(c, FSI_E_ARGUMENT, "uwcdi_strang_step needs the UWCDI phase field.");
    if (!(isfinite(dt) && dt >= 0.0)) return set_err(c, FSI_E_ARGUMENT, "dt must be finite and non-negative.");
    if (first_direction != 0 && first_direction != 1)
        return set_err(c, FSI_E_ARGUMENT, "first_direction must be :x or :y.");
    int N = c->N, rc = 0;
    const int d1 = first_direction == 0 ? 1 : 2, d2 = first_direction == 0 ? 2 : 1;
    const int dirs[3] = {d1, d2, d1};
    const double h[3] = {0.5 * dt, dt, 0.5 * dt};
    double *rhs = dalloc(N), *q = dalloc(N);
    memcpy(q, phi, sizeof(double) * N);
    for (int k = 0; k < 3 && !rc; ++k) {
        if (h[k] == 0.0) continue;  /* _uwcdi_directional_euler_sweep returns phi at dt = 0 */
        rc = phase_rhs_uwcdi_dir(c, U, V, q, NULL, rhs, dirs[k]);
        for (int n = 0; n < N && !rc; ++n) q[n] = q[n] + h[k] * rhs[n];
    }
    if (!rc) memcpy(phi, q, sizeof(double) * N);
    free(rhs);
    free(q);
    return rc;
}

/* capillary_force(::DiffuseAreaCapillarity) Capillarity.jl:134-187:
 * f = coefficient kappa grad(phi) on the faces, kappa = -div(n) of the
 * regularised normal of the (optionally filtered) phase field, the normal
 * and the curvature carried to the faces by interpolate_center_to_*_faces. */
static void capillary_force(fsio_ctx* c, const double* phi, double* fxo, double* fyo) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny;
    size_t nc = (size_t)nx * ny, nfx = (size_t)(nx + 1) * ny, nfy = (size_t)nx * (ny + 1);
    double *geo = dalloc(nc), *tmp = dalloc(nc);
    memcpy(geo, phi, sizeof(double) * nc);
    for (int k = 0; k < c->m.capillary_smoothing_passes; ++k) {
        smooth9(c, geo, tmp);
        double* t = geo; geo = tmp; tmp = t;
    }
    double *ggx = dalloc(nfx), *ggy = dalloc(nfy), *gx = dalloc(nfx), *gy = dalloc(nfy);
    scalar_gradient(c, &c->bcs.phi, geo, ggx, ggy);
    scalar_gradient(c, &c->bcs.phi, phi, gx, gy);
    double *nxc = dalloc(nc), *nyc = dalloc(nc), *nxf = dalloc(nfx), *nyf = dalloc(nfy), *kap = dalloc(nc);
    double d2 = c->m.capillary_delta_normal * c->m.capillary_delta_normal;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double gxc = m->wL_f2cx[i] * UX(ggx, i, j) + m->wR_f2cx[i] * UX(ggx, i + 1, j);
            double gyc = m->wB_f2cy[j] * VY(ggy, i, j) + m->wT_f2cy[j] * VY(ggy, i, j + 1);
            double den = sqrt(gxc * gxc + gyc * gyc + d2);
            CC(nxc, i, j) = gxc / den;
            CC(nyc, i, j) = gyc / den;
        }
    c2fx(c, nxc, nxf);
    c2fy(c, nyc, nyf);
    divergence(c, nxf, nyf, kap);
    for (size_t k = 0; k < nc; ++k) kap[k] = -kap[k];
    double *kx = dalloc(nfx), *ky = dalloc(nfy);
    c2fx(c, kap, kx);
    c2fy(c, kap, ky);
    double coef = c->m.capillary_coefficient;
    for (size_t k = 0; k < nfx; ++k) fxo[k] = coef * kx[k] * gx[k];
    for (size_t k = 0; k < nfy; ++k) fyo[k] = coef * ky[k] * gy[k];
    free(geo); free(tmp); free(ggx); free(ggy); free(gx); free(gy);
    free(nxc); free(nyc); free(nxf); free(nyf); free(kap); free(kx); free(ky);
}

void fsio_capillary_force(fsio_ctx* c, const double* phi, double* fx, double* fy) {
    if (c->m.capillarity_kind == FSI_CAPILLARITY_DIFFUSE_AREA) {
        capillary_force(c, phi, fx, fy);
        return;
    }
    memset(fx, 0, sizeof(double) * (size_t)(c->nx + 1) * c->ny);
    memset(fy, 0, sizeof(double) * (size_t)c->nx * (c->ny + 1));
}

static int phase_rhs(fsio_ctx* c, const double* U, const double* V, const double* phi,
                     const double* S, double* out) {
    switch (c->m.phase_method) {
    case FSI_PHASE_UPWIND: return phase_rhs_cdi(c, U, V, phi, S, out, 0);
    case FSI_PHASE_CDI: return phase_rhs_cdi(c, U, V, phi, S, out, 1);
    case FSI_PHASE_ACDI: return phase_rhs_acdi(c, U, V, phi, S, out);
    case FSI_PHASE_UWCDI: return phase_rhs_uwcdi(c, U, V, phi, S, out);
    }
    return set_err(c, FSI_E_UNSUPPORTED, "unknown phase method");
}
int fsio_phase_rhs(fsio_ctx* c, const double* U, const double* V, const double* phi,
                   const double* S, double* out) {
    return phase_rhs(c, U, V, phi, S, out);
}

/* ---------------------------------------------------------------------------
 * Structure tensor (src/StructureTensor.jl, src/StructuralConstitutive.jl)
 * ------------------------------------------------------------------------- */

/* _sigma_log_product_2d StructureTensor.jl:482-515; 1 where the tensor is
 * not SPD (the DomainError of :489-493), else 0 */
static int sigma_log_product_eval(double a, double d, double b, double* ox, double* oy, double* oxy) {
    double mm = 0.5 * (a + d), delta = 0.5 * (a - d);
    double r = hypot(delta, b);
    double lhi = mm + r, llo = mm - r;
    if (!(isfinite(llo) && isfinite(lhi) && llo > 0.0)) return 1;
    double ghi = lhi * log(lhi), glo = llo * log(llo);
    double scale = fabs(mm) > 1.0 ? fabs(mm) : 1.0, alpha, beta;
    if (r <= 1.0e-12 * scale) {
        alpha = log(mm) + 1.0;
        beta = -mm;
    } else {
        double den = lhi - llo;
        alpha = (ghi - glo) / den;
        beta = (lhi * glo - llo * ghi) / den;
    }
    *ox = alpha * a + beta;
    *oy = alpha * d + beta;
    *oxy = alpha * b;
    return 0;
}
static int sigma_log_product(fsio_ctx* c, double a, double d, double b, double* ox, double* oy,
                             double* oxy) {
    if (sigma_log_product_eval(a, d, b, ox, oy, oxy))
        return set_err(c, FSI_E_DOMAIN,
                       "GeometricStructureExtension requires an SPD tensor; encountered a "
                       "non-positive or non-finite eigenvalue.");
    return 0;
}

/* structure_extension_rhs(::GeometricStructureExtension) :522-551 */
static int extension_rhs(fsio_ctx* c, const double* sxx, const double* syy, const double* sxy,
                         const double* phi, size_t n, double* ex, double* ey, double* exy) {
    /* the first failing cell in index order reports the DomainError */
    long long bad = (long long)n;
    OMP_FOR
    for (long long k = 0; k < (long long)n; ++k) {
        ex[k] = ey[k] = exy[k] = 0.0;
        if (c->num.extension_kind != FSI_EXT_GEOMETRIC) continue;
        double kappa = c->num.ext_max_rate * pow(1.0 - clamp01(phi[k]), c->num.ext_exponent);
        if (kappa == 0.0) continue;
        double gx = 0.0, gy = 0.0, gxy = 0.0;
        if (sigma_log_product_eval(sxx[k], syy[k], sxy[k], &gx, &gy, &gxy)) {
#ifdef _OPENMP
#pragma omp critical(fsio_extension)
#endif
            if (k < bad) bad = k;
            continue;
        }
        ex[k] = -kappa * gx;
        ey[k] = -kappa * gy;
        exy[k] = -kappa * gxy;
    }
    if (bad < (long long)n) {
        double gx, gy, gxy;
        return sigma_log_product(c, sxx[bad], syy[bad], sxy[bad], &gx, &gy, &gxy);
    }
    return 0;
}
int fsio_structure_extension_rhs(fsio_ctx* c, const double* sxx, const double* syy,
                                 const double* sxy, const double* phi, int n, double* oxx,
                                 double* oyy, double* oxy) {
    return extension_rhs(c, sxx, syy, sxy, phi, (size_t)n, oxx, oyy, oxy);
}

/* structure_tensor_rhs StructureTensor.jl:377-463 (K_g = 0 under NoGrowth). */
static int structure_rhs(fsio_ctx* c, const double* sxx, const double* syy, const double* sxy,
                         const double* U, const double* V, const double* phi, const double* G,
                         double* oxx, double* oyy, double* oxy) {
    size_t nc = (size_t)c->N;
    double *axx = dalloc(nc), *ayy = dalloc(nc), *axy = dalloc(nc), *dv = dalloc(nc);
    double *ux = dalloc(nc), *uy = dalloc(nc), *vx = dalloc(nc), *vy = dalloc(nc);
    double *ex = dalloc(nc), *ey = dalloc(nc), *exy = dalloc(nc);
    int rc = upwind_advection(c, sxx, U, V, &c->bcs.sigma_xx, axx);
    if (!rc) rc = upwind_advection(c, syy, U, V, &c->bcs.sigma_yy, ayy);
    if (!rc) rc = upwind_advection(c, sxy, U, V, &c->bcs.sigma_xy, axy);
    if (!rc) rc = velocity_gradient(c, U, V, ux, uy, vx, vy);
    if (!rc) rc = extension_rhs(c, sxx, syy, sxy, phi, nc, ex, ey, exy);
    if (!rc) {
        divergence(c, U, V, dv);
        double rate = c->m.relaxation_kind == FSI_RELAX_MAXWELL ? c->m.relaxation_rate : 0.0;
        OMP_FOR
        for (size_t k = 0; k < nc; ++k) {
            /* A = grad v - K_g (StructureTensor.jl:431-434); K_g = 0 without growth */
            double Kxx = G ? c->kxx * G[k] : 0.0, Kxy = G ? c->kxy * G[k] : 0.0, Kyy = G ? c->kyy * G[k] : 0.0;
            double A11 = ux[k] - Kxx, A12 = uy[k] - Kxy, A21 = vx[k] - Kxy, A22 = vy[k] - Kyy;
            double adxx = axx[k] + sxx[k] * dv[k];
            double adyy = ayy[k] + syy[k] * dv[k];
            double adxy = axy[k] + sxy[k] * dv[k];
            double dfxx = 2.0 * (A11 * sxx[k] + A12 * sxy[k]);
            double dfyy = 2.0 * (A21 * sxy[k] + A22 * syy[k]);
            double dfxy = A21 * sxx[k] + A12 * syy[k] + (A11 + A22) * sxy[k];
            /* relaxation_rhs StructuralConstitutive.jl:61-85 */
            double rxx = 0.0, ryy = 0.0, rxy = 0.0;
            if (c->m.relaxation_kind == FSI_RELAX_MAXWELL) {
                rxx = -rate * (sxx[k] - 1.0);
                ryy = -rate * (syy[k] - 1.0);
                rxy = -rate * sxy[k];
            }
            oxx[k] = adxx + dfxx + rxx + ex[k];
            oyy[k] = adyy + dfyy + ryy + ey[k];
            oxy[k] = adxy + dfxy + rxy + exy[k];
        }
    }
    free(axx); free(ayy); free(axy); free(dv); free(ux); free(uy); free(vx); free(vy);
    free(ex); free(ey); free(exy);
    return rc;
}
int fsio_structure_rhs(fsio_ctx* c, const double* sxx, const double* syy, const double* sxy,
                       const double* U, const double* V, const double* phi, double* oxx,
                       double* oyy, double* oxy) {
    return structure_rhs(c, sxx, syy, sxy, U, V, phi, NULL, oxx, oyy, oxy);
}

/* apply_structure_extension! StructureTensor.jl:553-613 (safety reset). */
static void apply_extension_reset(fsio_ctx* c, double* sxx, double* syy, double* sxy,
                                  const double* phi) {
    double cutoff;
    if (c->num.extension_kind == FSI_EXT_IDENTITY) cutoff = c->num.ext_cutoff;
    else if (c->num.extension_kind == FSI_EXT_GEOMETRIC) {
        cutoff = c->num.ext_cutoff;
        if (cutoff <= 0.0) return;
    } else return;
    for (int k = 0; k < c->N; ++k)
        if (phi[k] < cutoff) { sxx[k] = 1.0; syy[k] = 1.0; sxy[k] = 0.0; }
}

/* ---------------------------------------------------------------------------
 * Growth and resource closures.
 * ------------------------------------------------------------------------- */
/* the resource is transported only when a growth, diffusion or consumption
   closure is active (TimeIntegrationIMEX.jl:1004-1009) */
static int growth_active(const fsio_ctx* c) {
    return c->m.growth_source_kind != FSI_GROWTH_NONE || c->m.diffusivity_kind != FSI_DIFFUSION_NONE ||
           c->m.consumption_kind != FSI_CONSUMPTION_NONE;
}

/* growth_source! GrowthModels.jl:214-256 and growth_diagnostics :336-352
   (the pointwise trace error of K_g, growth_tensor! :268-318); returns
   max |K_xx + K_yy - G| */
static double growth_source(fsio_ctx* c, const double* phi, const double* res, double* G) {
    const double rate = c->m.growth_rate, K = c->m.growth_half_saturation;
    double err = 0.0;
    for (int k = 0; k < c->N; ++k) {
        double g = 0.0;
        if (c->m.growth_source_kind == FSI_GROWTH_CONSTANT) g = rate * phi[k];
        else if (c->m.growth_source_kind == FSI_GROWTH_SATURATING) g = rate * phi[k] * res[k] / (K + res[k]);
        G[k] = g;
        double e = fabs(c->kxx * g + c->kyy * g - g);
        if (e > err || isnan(e)) err = e;
    }
    return err;
}
double fsio_growth_source(fsio_ctx* c, const double* phi, const double* res, double* G) {
    return FSIO_VALUE(growth_source(c, phi, res, G));
}

/* diffusivity_field! TransportConstitutive.jl:241-263 (the harmonic law takes
   phi unclamped) with its validation */
static int diffusivity_field(fsio_ctx* c, const double* phi, double* D) {
    const double Df = c->m.diffusivity_fluid, Ds = c->m.diffusivity_structure;
    for (int k = 0; k < c->N; ++k) {
        if (!isfinite(phi[k])) return set_err(c, FSI_E_ARGUMENT, "phi contains non-finite values.");
        D[k] = c->m.diffusivity_kind == FSI_DIFFUSION_HARMONIC ? 1.0 / ((1.0 - phi[k]) / Df + phi[k] / Ds) : Df;
    }
    for (int k = 0; k < c->N; ++k)
        if (!isfinite(D[k])) return set_err(c, FSI_E_ARGUMENT, "evaluated diffusivity contains non-finite values.");
    for (int k = 0; k < c->N; ++k)
        if (!(D[k] > 0.0))
            return set_err(c, FSI_E_ARGUMENT, "evaluated diffusivity must be strictly positive everywhere.");
    return 0;
}

/* build_diffusion_operator Operators.jl:361-384 applied matrix-free:
   L q + b = D_x (D_xf (G_x q + g_x)) + D_y (D_yf (G_y q + g_y)), the face
   diffusivities interpolated from the cells (:110-159), the gradient with the
   resource BC rows (:189-356; their data only `with_boundary`) */
static void diffusion_apply(fsio_ctx* c, const double* Dc, const double* q, double* out, int with_boundary) {
    int nx = c->nx, ny = c->ny;
    size_t nfx = (size_t)(nx + 1) * ny, nfy = (size_t)nx * (ny + 1);
    double *gx = dalloc(nfx), *gy = dalloc(nfy), *dx = dalloc(nfx), *dy = dalloc(nfy);
    fsi_bc_setup bc = c->bcs.resource;
    if (!with_boundary) { bc.left.value = bc.right.value = bc.bottom.value = bc.top.value = 0.0; }
    scalar_gradient(c, &bc, q, gx, gy);
    c2fx(c, Dc, dx);
    c2fy(c, Dc, dy);
    for (size_t k = 0; k < nfx; ++k) gx[k] *= dx[k];
    for (size_t k = 0; k < nfy; ++k) gy[k] *= dy[k];
    divergence(c, gx, gy, out);
    free(gx); free(gy); free(dx); free(dy);
}
int fsio_diffusion_apply(fsio_ctx* c, const double* phi, const double* q, double* out, int with_boundary) {
    double* D = dalloc(c->N);
    int rc = diffusivity_field(c, phi, D);
    if (!rc) diffusion_apply(c, D, q, out, with_boundary);
    free(D);
    return rc;
}

/* resource_sink! TransportConstitutive.jl:303-335 and resource_explicit_rhs
   ResourceTransport.jl:292-342: -div(upwind c v) - C(phi, c), with the
   advective boundary states of :84-159 (the rule upwind_advection follows) */
static int resource_explicit_rhs(fsio_ctx* c, const double* U, const double* V, const double* phi,
                                 const double* res, double* out) {
    for (int k = 0; k < c->N; ++k) {
        if (!isfinite(phi[k])) return set_err(c, FSI_E_ARGUMENT, "phi contains non-finite values.");
    }
    for (int k = 0; k < c->N; ++k) {
        if (!isfinite(res[k])) return set_err(c, FSI_E_ARGUMENT, "resource contains non-finite values.");
    }
    double* sink = dalloc(c->N);
    if (c->m.consumption_kind == FSI_CONSUMPTION_SATURATING) {
        const double rate = c->m.consumption_rate, K = c->m.consumption_half_saturation;
        for (int k = 0; k < c->N; ++k) sink[k] = rate * phi[k] * res[k] / (K + res[k]);
        for (int k = 0; k < c->N; ++k)
            if (!isfinite(sink[k])) {
                free(sink);
                return set_err(c, FSI_E_ARGUMENT, "Evaluated resource sink contains non-finite values.");
            }
    }
    int rc = upwind_advection(c, res, U, V, &c->bcs.resource, out);
    if (!rc)
        for (int k = 0; k < c->N; ++k) out[k] = out[k] - sink[k];
    free(sink);
    return rc;
}
int fsio_resource_explicit_rhs(fsio_ctx* c, const double* U, const double* V, const double* phi,
                               const double* res, double* out) {
    return resource_explicit_rhs(c, U, V, phi, res, out);
}

static double dot(const double* a, const double* b, size_t n);
/* The resource stage solve, _build_resource_stage! :742-774: (I - alpha L) c =
 * rhs + alpha b, L the diffusion operator of phi_s.  L is symmetric negative
 * semidefinite on uniform grids (the face terms are symmetric, a Dirichlet
 * face adds to the diagonal only), so the system is SPD and solved by CG with
 * the stopping rules of _solve_linear_system! :602-641 (x0 = c_n, method
 * resource_method): :direct to |r| <= 1e-14 max(|b|, eps); :gmres to
 * |r| <= reltol |b - A x0| within maxiter, then accepted below
 * fallback_reltol, else the direct target (direct_fallback) or the
 * ErrorException of :636. */
static int resource_solve(fsio_ctx* c, double alpha, const double* Dc, const double* rhs_in, const double* x0,
                          double* x, int* iters) {
    int N = c->N;
    double *b = dalloc(N), *r = dalloc(N), *p = dalloc(N), *q = dalloc(N), *tmp = dalloc(N), *zero = dalloc(N);
    diffusion_apply(c, Dc, zero, tmp, 1); /* the boundary vector */
    for (int k = 0; k < N; ++k) b[k] = rhs_in[k] + alpha * tmp[k];
    diffusion_apply(c, Dc, x0, tmp, 0);
    double bn2 = 0.0, ref2 = 0.0;
    for (int k = 0; k < N; ++k) {
        double e = b[k] - (x0[k] - alpha * tmp[k]);
        bn2 += b[k] * b[k];
        ref2 += e * e;
    }
    const double bscale = sqrt(bn2) > 2.220446049250313e-16 ? sqrt(bn2) : 2.220446049250313e-16;
    const double fb = c->num.fallback_rtol * bscale;
    double dtol = 1e-14 * bscale;
    if (dtol > fb) dtol = fb;
    const int maxiter = c->num.velocity_maxiter;
    const int cap = maxiter + (10 * maxiter > 1000 ? 10 * maxiter : 1000);
    const int direct = c->num.resource_method == FSI_SOLVER_DIRECT;
    double tol = direct ? dtol : c->num.velocity_rtol * sqrt(ref2);
    int limit = direct ? cap : maxiter, phase = direct;
    memcpy(x, x0, sizeof(double) * N);
    for (int k = 0; k < N; ++k) r[k] = b[k] - (x[k] - alpha * tmp[k]);
    memcpy(p, r, sizeof(double) * N);
    double rr = dot(r, r, N);
    int it = 0, rc = 0;
    for (;;) {
        if (!isfinite(rr)) {
            rc = set_err(c, FSI_E_SOLVER, "resource stage residual is not finite");
            break;
        }
        if (sqrt(rr) <= tol || it >= limit) {
            if (phase) {
                if (sqrt(rr) <= tol) break;
                rc = set_err(c, FSI_E_SOLVER, "resource GMRES residual is too large: %.6e", FSIO_VALUE(sqrt(rr) / bscale));
                break;
            }
            if (sqrt(rr) <= fb) break;
            if (c->num.direct_fallback) {
                phase = 1;
                tol = dtol;
                limit = cap;
                continue;
            }
            rc = set_err(c, FSI_E_SOLVER, "resource GMRES residual is too large: %.6e", FSIO_VALUE(sqrt(rr) / bscale));
            break;
        }
        diffusion_apply(c, Dc, p, tmp, 0);
        for (int k = 0; k < N; ++k) q[k] = p[k] - alpha * tmp[k];
        double pq = dot(p, q, N);
        double a = rr / pq;
        for (int k = 0; k < N; ++k) { x[k] += a * p[k]; r[k] -= a * q[k]; }
        double rr2 = dot(r, r, N);
        double beta = rr2 / rr;
        rr = rr2;
        for (int k = 0; k < N; ++k) p[k] = r[k] + beta * p[k];
        ++it;
    }
    if (iters) *iters = it;
    free(b); free(r); free(p); free(q); free(tmp); free(zero);
    return rc;
}
int fsio_resource_solve(fsio_ctx* c, fsio_scalar alpha, const double* phi, const double* rhs, const double* x0,
                        double* x, int* iters) {
    double* D = dalloc(c->N);
    int rc = diffusivity_field(c, phi, D);
    if (!rc) rc = resource_solve(c, alpha, D, rhs, x0, x, iters);
    free(D);
    return rc;
}

/* ---------------------------------------------------------------------------
 * Viscous operator (build_velocity_gradient_operators Operators.jl:848-924,
 * build_viscous_operator :929-1083) applied matrix-free on compact vectors.
 * The viscosity is constant (ConstantViscosity), so mu_center = mu_corner = mu.
 * ------------------------------------------------------------------------- */
static void viscous_apply(fsio_ctx* c, const double* uv, double* out, int with_boundary) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny, N = c->N;
    const double* u = uv;
    const double* v = uv + N;
    double mu = c->m.viscosity;
    double *exx = dalloc(N), *exy = dalloc(N), *eyy = dalloc(N);
    /* E * uv (+ E_bc) :868-913, 966-979 */
    OMP_FOR
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int row = i + j * nx;
            double inv_dx = 1.0 / m->dx_c[i];
            double a = -inv_dx * u[row];
            if (i < nx - 1) a += inv_dx * u[row + 1];
            else if (c->px) a += inv_dx * u[j * nx];
            else if (with_boundary) a += c->bcs.U.right.value * inv_dx;
            exx[row] = a;
            double inv_dy = 1.0 / m->dy_c[j];
            double b = -inv_dy * v[row];
            if (j < ny - 1) b += inv_dy * v[row + nx];
            else if (c->py) b += inv_dy * v[i];
            else if (with_boundary) b += c->bcs.V.top.value * inv_dy;
            eyy[row] = b;
            double duy = 0.0, dvx = 0.0;
            if (j > 0) { double ih = 1.0 / m->dy_f[j]; duy = -ih * u[row - nx] + ih * u[row]; }
            else if (c->py) { double ih = 1.0 / m->dy_f[0]; duy = -ih * u[i + (ny - 1) * nx] + ih * u[row]; }
            if (i > 0) { double ih = 1.0 / m->dx_f[i]; dvx = -ih * v[row - 1] + ih * v[row]; }
            else if (c->px) { double ih = 1.0 / m->dx_f[0]; dvx = -ih * v[nx - 1 + j * nx] + ih * v[row]; }
            exy[row] = 0.5 * duy + 0.5 * dvx;
        }
    /* stress weights 2*mu (:960-965), then Dstress (:1000-1034) */
    OMP_FOR
    for (int k = 0; k < N; ++k) { exx[k] *= 2.0 * mu; exy[k] *= 2.0 * mu; eyy[k] *= 2.0 * mu; }
    OMP_FOR
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int row = i + j * nx;
            double ih = 1.0 / m->dx_f[i];
            double fu = ih * exx[row];
            if (i > 0) fu += -ih * exx[row - 1];
            else if (c->px) fu += -ih * exx[nx - 1 + j * nx];
            double ihy = 1.0 / m->dy_c[j];
            fu += -ihy * exy[row];
            if (j < ny - 1) fu += ihy * exy[row + nx];
            else if (c->py) fu += ihy * exy[i];
            out[row] = fu;
            double ihx = 1.0 / m->dx_c[i];
            double fv = -ihx * exy[row];
            if (i < nx - 1) fv += ihx * exy[row + 1];
            else if (c->px) fv += ihx * exy[j * nx];
            double ihyy = 1.0 / m->dy_f[j];
            fv += ihyy * eyy[row];
            if (j > 0) fv += -ihyy * eyy[row - nx];
            else if (c->py) fv += -ihyy * eyy[i + (ny - 1) * nx];
            out[N + row] = fv;
        }
    /* tangential Dirichlet wall lifting :1036-1080 */
    if (!c->py) {
        double db = m->dy_f[0] / 2.0, dt = m->dy_f[ny] / 2.0;
        for (int i = 0; i < nx; ++i) {
            if (c->bcs.U.bottom.kind == FSI_BC_DIRICHLET) {
                int row = i;
                double w = mu / (m->dy_c[0] * db);
                out[row] += -w * u[row] + (with_boundary ? w * c->bcs.U.bottom.value : 0.0);
            }
            if (c->bcs.U.top.kind == FSI_BC_DIRICHLET) {
                int row = i + (ny - 1) * nx;
                double w = mu / (m->dy_c[ny - 1] * dt);
                out[row] += -w * u[row] + (with_boundary ? w * c->bcs.U.top.value : 0.0);
            }
        }
    }
    if (!c->px) {
        double dl = m->dx_f[0] / 2.0, dr = m->dx_f[nx] / 2.0;
        for (int j = 0; j < ny; ++j) {
            if (c->bcs.V.left.kind == FSI_BC_DIRICHLET) {
                int row = j * nx;
                double w = mu / (m->dx_c[0] * dl);
                out[N + row] += -w * v[row] + (with_boundary ? w * c->bcs.V.left.value : 0.0);
            }
            if (c->bcs.V.right.kind == FSI_BC_DIRICHLET) {
                int row = nx - 1 + j * nx;
                double w = mu / (m->dx_c[nx - 1] * dr);
                out[N + row] += -w * v[row] + (with_boundary ? w * c->bcs.V.right.value : 0.0);
            }
        }
    }
    free(exx); free(exy); free(eyy);
}
void fsio_viscous_apply(fsio_ctx* c, const double* uv, double* out, int with_boundary) {
    viscous_apply(c, uv, out, with_boundary);
}

/* velocity_dirichlet_data Operators.jl:1085-1110: U-left column, V-bottom row. */
static void dirichlet_mask(fsio_ctx* c, unsigned char* mask, double* values) {
    int nx = c->nx, ny = c->ny, N = c->N;
    memset(mask, 0, 2 * (size_t)N);
    if (c->bcs.U.left.kind == FSI_BC_DIRICHLET)
        for (int j = 0; j < ny; ++j) { mask[j * nx] = 1; values[j * nx] = c->bcs.U.left.value; }
    if (c->bcs.V.bottom.kind == FSI_BC_DIRICHLET)
        for (int i = 0; i < nx; ++i) { mask[N + i] = 1; values[N + i] = c->bcs.V.bottom.value; }
}

static double dot(const double* a, const double* b, size_t n) {
    double s = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : s) schedule(static)
#endif
    for (size_t k = 0; k < n; ++k) s += a[k] * b[k];
    return s;
}

/* (I - alpha M) x = rhs + alpha b with Dirichlet rows: _build_velocity_stage!
 * :794-812, _shifted_identity :643-647, _impose_row_dirichlet :649-672.  The
 * identity rows fix x_D = g; the remaining rows form an SPD system in x_F
 * (M is symmetric negative semidefinite on uniform grids) solved by CG.
 *
 * Stopping rules of _solve_linear_system! :602-641 (x0 = uv_n, :807):
 *   :direct  the exact solve A \ b, here CG to |r| <= 1e-14 max(|b|, eps);
 *   :gmres   IterativeSolvers' reltol rule, |r| <= reltol |b - A x0| (its
 *            initial residual), within maxiter iterations; where it stops
 *            (the rule met or the limit reached) the iterate is accepted
 *            when |r| / max(|b|, eps) <= fallback_reltol (:629-631), else the
 *            direct solve (direct_fallback, :632-634: the direct target within
 *            maxiter + max(10 maxiter, 1000) iterations in all) or the
 *            ErrorException of :636.
 * b and A are the whole system of _impose_row_dirichlet: the Dirichlet rows
 * enter |b| with their value and |b - A x0| with g - x0.  A non-finite
 * residual is an error at once. */
int fsio_velocity_solve(fsio_ctx* c, fsio_scalar alpha, const double* rhs_in, const double* x0,
                        double* x, int* iters) {
    int N2 = 2 * c->N;
    unsigned char* mask = (unsigned char*)calloc(N2, 1);
    double *g = dalloc(N2), *b = dalloc(N2), *r = dalloc(N2), *p = dalloc(N2), *q = dalloc(N2);
    double* tmp = dalloc(N2);
    dirichlet_mask(c, mask, g);
    /* rhs + alpha * boundary */
    double* zero = dalloc(N2);
    viscous_apply(c, zero, tmp, 1); /* boundary vector b = M*0 + b */
    OMP_FOR
    for (int k = 0; k < N2; ++k) b[k] = rhs_in[k] + alpha * tmp[k];
    /* |b| and |b - A x0| of the system with identity Dirichlet rows */
    double bn2 = 0.0, ref2 = 0.0;
    viscous_apply(c, x0, tmp, 0);
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : bn2, ref2) schedule(static)
#endif
    for (int k = 0; k < N2; ++k) {
        double bk = mask[k] ? g[k] : b[k];
        double ek = mask[k] ? g[k] - x0[k] : b[k] - (x0[k] - alpha * tmp[k]);
        bn2 += bk * bk;
        ref2 += ek * ek;
    }
    const double bscale = sqrt(bn2) > 2.220446049250313e-16 ? sqrt(bn2) : 2.220446049250313e-16;
    const double fb = c->num.fallback_rtol * bscale;
    double dtol = 1e-14 * bscale;
    if (dtol > fb) dtol = fb;
    const int maxiter = c->num.velocity_maxiter;
    const int cap = maxiter + (10 * maxiter > 1000 ? 10 * maxiter : 1000);
    const int direct = c->num.velocity_method == FSI_SOLVER_DIRECT;
    double tol = direct ? dtol : c->num.velocity_rtol * sqrt(ref2);
    int limit = direct ? cap : maxiter, phase = direct;
    OMP_FOR
    for (int k = 0; k < N2; ++k) x[k] = mask[k] ? g[k] : x0[k];
    viscous_apply(c, x, tmp, 0);
    OMP_FOR
    for (int k = 0; k < N2; ++k) r[k] = mask[k] ? 0.0 : b[k] - (x[k] - alpha * tmp[k]);
    memcpy(p, r, sizeof(double) * N2);
    double rr = dot(r, r, N2);
    int it = 0, rc = 0;
    for (;;) {
        if (!isfinite(rr)) {
            rc = set_err(c, FSI_E_SOLVER, "velocity predictor residual is not finite");
            break;
        }
        if (sqrt(rr) <= tol || it >= limit) {
            if (phase) {
                if (sqrt(rr) <= tol) break;
                rc = set_err(c, FSI_E_SOLVER, "velocity GMRES residual is too large: %.6e", FSIO_VALUE(sqrt(rr) / bscale));
                break;
            }
            /* GMRES has stopped (the rule met or the limit reached): the check
               of :629-636 on |b - A x| / max(|b|, eps) */
            if (sqrt(rr) <= fb) break;                     /* accepted (:629-631) */
            if (c->num.direct_fallback) {                 /* the direct solve (:632-634) */
                phase = 1;
                tol = dtol;
                limit = cap;
                continue;
            }
            rc = set_err(c, FSI_E_SOLVER, "velocity GMRES residual is too large: %.6e", FSIO_VALUE(sqrt(rr) / bscale));
            break;
        }
        viscous_apply(c, p, tmp, 0);
        OMP_FOR
        for (int k = 0; k < N2; ++k) q[k] = mask[k] ? 0.0 : p[k] - alpha * tmp[k];
        double pq = dot(p, q, N2);
        double a = rr / pq;
        OMP_FOR
        for (int k = 0; k < N2; ++k) { x[k] += a * p[k]; r[k] -= a * q[k]; }
        double rr2 = dot(r, r, N2);
        double beta = rr2 / rr;
        rr = rr2;
        OMP_FOR
        for (int k = 0; k < N2; ++k) p[k] = r[k] + beta * p[k];
        ++it;
    }
    if (iters) *iters = it;
    free(mask); free(g); free(b); free(r); free(p); free(q); free(tmp); free(zero);
    return rc;
}

/* ---------------------------------------------------------------------------
 * Projection (src/Projection.jl) with a geometric-multigrid-preconditioned CG.
 * ------------------------------------------------------------------------- */

/* fine-level A q = D*(G*q) (build_diffusion_operator Operators.jl:361-384 with
 * unit coefficient, build_pressure_laplacian :387-390); b_bc = D*gbc. */
static void laplacian_apply(fsio_ctx* c, const double* q, double* out, int with_boundary) {
    int nx = c->nx, ny = c->ny;
    double* gx = dalloc((size_t)(nx + 1) * ny);
    double* gy = dalloc((size_t)nx * (ny + 1));
    fsi_bc_setup bc = c->bcs.correction;
    if (!with_boundary) { bc.left.value = bc.right.value = bc.bottom.value = bc.top.value = 0.0; }
    scalar_gradient(c, &bc, q, gx, gy);
    divergence(c, gx, gy, out);
    free(gx);
    free(gy);
}
void fsio_laplacian_apply(fsio_ctx* c, const double* q, double* out, int with_boundary) {
    laplacian_apply(c, q, out, with_boundary);
}

/* Coarse-level operator: rediscretised unit Laplacian with the same side kinds
 * (homogeneous data), applied as -A (SPD/SPSD), uniform spacing. */
static double neg_lap_point(fsio_ctx* c, const mg_level* L, const double* x, int i, int j,
                            double* diag) {
    int nx = L->nx, ny = L->ny;
    double ihx2 = 1.0 / (L->hx * L->hx), ihy2 = 1.0 / (L->hy * L->hy);
    double xc = x[i + j * nx], s = 0.0, d = 0.0;
    const fsi_bc_setup* b = &c->bcs.correction;
    /* x-direction */
    if (i > 0) { s -= ihx2 * x[i - 1 + j * nx]; d += ihx2; }
    else if (c->px) { s -= ihx2 * x[nx - 1 + j * nx]; d += ihx2; }
    else if (b->left.kind == FSI_BC_DIRICHLET) { d += 2.0 * ihx2; }
    if (i < nx - 1) { s -= ihx2 * x[i + 1 + j * nx]; d += ihx2; }
    else if (c->px) { s -= ihx2 * x[j * nx]; d += ihx2; }
    else if (b->right.kind == FSI_BC_DIRICHLET) { d += 2.0 * ihx2; }
    if (j > 0) { s -= ihy2 * x[i + (j - 1) * nx]; d += ihy2; }
    else if (c->py) { s -= ihy2 * x[i + (ny - 1) * nx]; d += ihy2; }
    else if (b->bottom.kind == FSI_BC_DIRICHLET) { d += 2.0 * ihy2; }
    if (j < ny - 1) { s -= ihy2 * x[i + (j + 1) * nx]; d += ihy2; }
    else if (c->py) { s -= ihy2 * x[i]; d += ihy2; }
    else if (b->top.kind == FSI_BC_DIRICHLET) { d += 2.0 * ihy2; }
    if (diag) *diag = d;
    return s + d * xc;
}

/* symmetric Gauss-Seidel sweep (forward then backward) on -A x = b */
static void sgs(fsio_ctx* c, mg_level* L, int forward) {
    int nx = L->nx, ny = L->ny;
#ifdef _OPENMP
    /* red-black order (forward: red then black; backward: black then red),
       where the colouring is consistent across periodic sides */
    if (!(c->px && (nx & 1)) && !(c->py && (ny & 1))) {
        for (int h = 0; h < 2; ++h) {
            const int colour = forward ? h : 1 - h;
#pragma omp parallel for schedule(static) if (nx * ny >= 16384)
            for (int j = 0; j < ny; ++j)
                for (int i = (j + colour) & 1; i < nx; i += 2) {
                    double d, ax = neg_lap_point(c, L, L->x, i, j, &d);
                    L->x[i + j * nx] += (L->b[i + j * nx] - ax) / d;
                }
        }
        return;
    }
#endif
    if (forward) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double d, ax = neg_lap_point(c, L, L->x, i, j, &d);
                L->x[i + j * nx] += (L->b[i + j * nx] - ax) / d;
            }
    } else {
        for (int j = ny - 1; j >= 0; --j)
            for (int i = nx - 1; i >= 0; --i) {
                double d, ax = neg_lap_point(c, L, L->x, i, j, &d);
                L->x[i + j * nx] += (L->b[i + j * nx] - ax) / d;
            }
    }
}

/* 1-D bilinear prolongation weights of coarse cell I onto fine cell f, with the
 * boundary ghost rule: Neumann mirror, Dirichlet odd mirror, periodic wrap. */
static void transfer_1d(int nf, int periodic, int lo_kind, int hi_kind, int f, int* I, double* w,
                        int* n) {
    int nc = nf / 2, parent = f / 2, other = (f % 2 == 0) ? parent - 1 : parent + 1;
    *n = 0;
    I[*n] = parent; w[*n] = 0.75; (*n)++;
    if (other < 0 || other >= nc) {
        if (periodic) { I[*n] = (other + nc) % nc; w[*n] = 0.25; (*n)++; }
        else {
            int kind = other < 0 ? lo_kind : hi_kind;
            if (kind == FSI_BC_DIRICHLET) w[0] -= 0.25; /* ghost = -value */
            else w[0] += 0.25;                           /* ghost = value */
        }
    } else { I[*n] = other; w[*n] = 0.25; (*n)++; }
}

static void mg_vcycle(fsio_ctx* c, int l) {
    mg_level* L = &c->lev[l];
    int nx = L->nx, ny = L->ny;
    size_t n = (size_t)nx * ny;
    if (l == c->nlev - 1) {
        memset(L->x, 0, sizeof(double) * n);
        for (int s = 0; s < 200; ++s) { sgs(c, L, 1); sgs(c, L, 0); }
        return;
    }
    memset(L->x, 0, sizeof(double) * n);
    sgs(c, L, 1);
    sgs(c, L, 1);
    OMP_FOR_LEVEL
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) L->r[i + j * nx] = L->b[i + j * nx] - neg_lap_point(c, L, L->x, i, j, NULL);
    mg_level* C = &c->lev[l + 1];
    const fsi_bc_setup* b = &c->bcs.correction;
#ifdef _OPENMP
    if (nx * ny >= 16384) {
        /* R = P^T / 4 gathered per coarse cell from the fine cells whose
           prolongation stencil holds it (fine 2I - 1 .. 2I + 2, wrapped
           across periodic sides) */
        const int cnx = C->nx, cny = C->ny;
#pragma omp parallel for schedule(static)
        for (int J = 0; J < cny; ++J)
            for (int I = 0; I < cnx; ++I) {
                double acc = 0.0;
                for (int dy = -1; dy <= 2; ++dy) {
                    int fy = 2 * J + dy;
                    if (fy < 0 || fy >= ny) {
                        if (!c->py) continue;
                        fy = (fy + ny) % ny;
                    }
                    int Jy[2], ny_; double wy[2];
                    transfer_1d(ny, c->py, b->bottom.kind, b->top.kind, fy, Jy, wy, &ny_);
                    double wj = 0.0;
                    for (int a = 0; a < ny_; ++a) if (Jy[a] == J) wj += wy[a];
                    if (wj == 0.0) continue;
                    for (int dx = -1; dx <= 2; ++dx) {
                        int fx = 2 * I + dx;
                        if (fx < 0 || fx >= nx) {
                            if (!c->px) continue;
                            fx = (fx + nx) % nx;
                        }
                        int Ix[2], nx_; double wx[2];
                        transfer_1d(nx, c->px, b->left.kind, b->right.kind, fx, Ix, wx, &nx_);
                        double wi = 0.0;
                        for (int e = 0; e < nx_; ++e) if (Ix[e] == I) wi += wx[e];
                        if (wi != 0.0) acc += 0.25 * wi * wj * L->r[fx + fy * nx];
                    }
                }
                C->b[I + J * cnx] = acc;
            }
    } else
#endif
    {
    memset(C->b, 0, sizeof(double) * (size_t)C->nx * C->ny);
    for (int j = 0; j < ny; ++j) {
        int Jy[2], ny_; double wy[2];
        transfer_1d(ny, c->py, b->bottom.kind, b->top.kind, j, Jy, wy, &ny_);
        for (int i = 0; i < nx; ++i) {
            int Ix[2], nx_; double wx[2];
            transfer_1d(nx, c->px, b->left.kind, b->right.kind, i, Ix, wx, &nx_);
            double rv = L->r[i + j * nx];
            for (int a = 0; a < ny_; ++a)
                for (int e = 0; e < nx_; ++e) C->b[Ix[e] + Jy[a] * C->nx] += 0.25 * wx[e] * wy[a] * rv;
        }
    }
    }
    mg_vcycle(c, l + 1);
    OMP_FOR_LEVEL
    for (int j = 0; j < ny; ++j) {
        int Jy[2], ny_; double wy[2];
        transfer_1d(ny, c->py, b->bottom.kind, b->top.kind, j, Jy, wy, &ny_);
        for (int i = 0; i < nx; ++i) {
            int Ix[2], nx_; double wx[2];
            transfer_1d(nx, c->px, b->left.kind, b->right.kind, i, Ix, wx, &nx_);
            double s = 0.0;
            for (int a = 0; a < ny_; ++a)
                for (int e = 0; e < nx_; ++e) s += wx[e] * wy[a] * C->x[Ix[e] + Jy[a] * C->nx];
            L->x[i + j * nx] += s;
        }
    }
    sgs(c, L, 0);
    sgs(c, L, 0);
}

static void build_mg(fsio_ctx* c) {
    int nx = c->nx, ny = c->ny;
    double hx = c->lx / nx, hy = c->ly / ny;
    c->nlev = 0;
    for (;;) {
        mg_level* L = &c->lev[c->nlev++];
        L->nx = nx; L->ny = ny; L->hx = hx; L->hy = hy;
        L->x = dalloc((size_t)nx * ny); L->b = dalloc((size_t)nx * ny); L->r = dalloc((size_t)nx * ny);
        if (nx % 2 || ny % 2 || nx < 8 || ny < 8 || c->nlev >= 30) break;
        nx /= 2; ny /= 2; hx *= 2.0; hy *= 2.0;
    }
}

/* PCG on -A q = -rhs (SPD, or SPSD with the constant nullspace when no side is
 * Dirichlet), preconditioned by one symmetric V-cycle. */
static int projection_solve(fsio_ctx* c, const double* rhs, double* q, int* iters) {
    int N = c->N;
    double *b = dalloc(N), *r = dalloc(N), *z = dalloc(N), *p = dalloc(N), *Ap = dalloc(N);
    for (int k = 0; k < N; ++k) b[k] = -rhs[k];
    if (c->has_nullspace) {
        double mean = 0.0;
        for (int k = 0; k < N; ++k) mean += b[k];
        mean /= N;
        for (int k = 0; k < N; ++k) b[k] -= mean;
    }
    laplacian_apply(c, q, Ap, 0);
    for (int k = 0; k < N; ++k) r[k] = b[k] + Ap[k];
    double bn = sqrt(dot(b, b, N));
    double tol = c->num.projection_rtol * (bn > 0 ? bn : 1.0);
    mg_level* L0 = &c->lev[0];
    int it = 0;
    double rz = 0.0;
    double rn = sqrt(dot(r, r, N));
    while (rn > tol && it < c->num.projection_maxiter) {
        memcpy(L0->b, r, sizeof(double) * N);
        mg_vcycle(c, 0);
        memcpy(z, L0->x, sizeof(double) * N);
        double rz2 = dot(r, z, N);
        if (it == 0) memcpy(p, z, sizeof(double) * N);
        else {
            double beta = rz2 / rz;
            OMP_FOR
            for (int k = 0; k < N; ++k) p[k] = z[k] + beta * p[k];
        }
        rz = rz2;
        laplacian_apply(c, p, Ap, 0);
        OMP_FOR
        for (int k = 0; k < N; ++k) Ap[k] = -Ap[k];
        double a = rz / dot(p, Ap, N);
        OMP_FOR
        for (int k = 0; k < N; ++k) { q[k] += a * p[k]; r[k] -= a * Ap[k]; }
        rn = sqrt(dot(r, r, N));
        ++it;
    }
    if (iters) *iters = it;
    free(b); free(r); free(z); free(p); free(Ap);
    return 0;
}

/* project! src/Projection.jl:170-307 with target divergence G. */
static int project(fsio_ctx* c, double* U, double* V, const double* G, double* correction,
                   fsi_projection_result* res) {
    const metrics_t* m = &c->mt;
    int nx = c->nx, ny = c->ny, N = c->N;
    for (size_t k = 0; k < (size_t)(nx + 1) * ny; ++k)
        if (!isfinite(U[k])) return set_err(c, FSI_E_ARGUMENT, "U contains non-finite values.");
    for (size_t k = 0; k < (size_t)nx * (ny + 1); ++k)
        if (!isfinite(V[k])) return set_err(c, FSI_E_ARGUMENT, "V contains non-finite values.");
    double *pdiv = dalloc(N), *rhs = dalloc(N), *Aq = dalloc(N);
    divergence(c, U, V, pdiv);
    OMP_FOR
    for (int k = 0; k < N; ++k) rhs[k] = (pdiv[k] - G[k]) - c->lap_bc[k]; /* :201-204 */
    double defect = 0.0;
    if (c->has_nullspace) { /* :206-221 */
        double scale = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : defect, scale) schedule(static)
#endif
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double area = m->dx_c[i] * m->dy_c[j];
                defect += CC(rhs, i, j) * area;
                scale += fabs(CC(rhs, i, j)) * area;
            }
        double tol = 1e-12 + 1e-10 * (scale > 2.220446049250313e-16 ? scale : 2.220446049250313e-16);
        if (!(fabs(defect) <= tol)) {
            free(pdiv); free(rhs); free(Aq);
            return set_err(c, FSI_E_ARGUMENT,
                           "The prescribed divergence constraint is incompatible with the "
                           "Neumann/periodic correction boundaries (defect %.3e, tolerance %.3e).",
                           FSIO_VALUE(defect), FSIO_VALUE(tol));
        }
    }
    for (int k = 0; k < N; ++k)
        if (!isfinite(correction[k])) {
            free(pdiv); free(rhs); free(Aq);
            return set_err(c, FSI_E_ARGUMENT, "correction_guess contains non-finite values.");
        }
    int iters = 0;
    projection_solve(c, rhs, correction, &iters);
    if (c->has_nullspace) { /* :255-259 */
        double tot = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : tot) schedule(static)
#endif
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) tot += CC(correction, i, j) * (m->dx_c[i] * m->dy_c[j]);
        double mean = tot / (c->lx * c->ly);
        OMP_FOR
        for (int k = 0; k < N; ++k) correction[k] -= mean;
    }
    laplacian_apply(c, correction, Aq, 0); /* :261-264 */
    double rn = 0.0, bn = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : rn, bn) schedule(static)
#endif
    for (int k = 0; k < N; ++k) { rn += sq(Aq[k] - rhs[k]); bn += rhs[k] * rhs[k]; }
    double lin = sqrt(rn) / (sqrt(bn) > 2.220446049250313e-16 ? sqrt(bn) : 2.220446049250313e-16);
    double* gx = dalloc((size_t)(nx + 1) * ny);
    double* gy = dalloc((size_t)nx * (ny + 1));
    scalar_gradient(c, &c->bcs.correction, correction, gx, gy); /* :266-273 */
    OMP_FOR
    for (size_t k = 0; k < (size_t)(nx + 1) * ny; ++k) U[k] -= gx[k];
    OMP_FOR
    for (size_t k = 0; k < (size_t)nx * (ny + 1); ++k) V[k] -= gy[k];
    divergence(c, U, V, Aq);
    double linf = 0.0, l2 = 0.0, pl2 = 0.0, tl2 = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(max : linf) reduction(+ : l2, pl2, tl2) schedule(static)
#endif
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double area = m->dx_c[i] * m->dy_c[j];
            double rr = CC(Aq, i, j) - CC(G, i, j);
            if (fabs(rr) > linf) linf = fabs(rr);
            l2 += rr * rr * area;
            pl2 += sq(CC(pdiv, i, j)) * area;
            tl2 += sq(CC(G, i, j)) * area;
        }
    double LA = c->lx * c->ly;
    l2 = sqrt(l2 / LA);
    pl2 = sqrt(pl2 / LA);
    tl2 = sqrt(tl2 / LA);
    free(pdiv); free(rhs); free(Aq); free(gx); free(gy);
    if (res) {
        res->linear_relative_residual = FSIO_VALUE(lin);
        res->constraint_linf = FSIO_VALUE(linf);
        res->constraint_l2 = FSIO_VALUE(l2);
        res->compatibility_defect = FSIO_VALUE(defect);
        res->iterations = iters;
    }
    if (!(lin <= 1e-8)) /* :279-285 */
        return set_err(c, FSI_E_SOLVER,
                       "Projection Poisson solve did not reach the required true-residual "
                       "accuracy. Relative residual = %.3e, tolerance = 1e-08.", FSIO_VALUE(lin));
    double cs = pl2 > tl2 ? pl2 : tl2;
    if (cs < 1.0) cs = 1.0;
    double ctol = 1e-12 + 1e-8 * cs;
    if (!(l2 <= ctol)) /* :287-297 */
        return set_err(c, FSI_E_SOLVER,
                       "Projection did not satisfy the prescribed divergence constraint. "
                       "Weighted L2 residual = %.3e, tolerance = %.3e.", FSIO_VALUE(l2), FSIO_VALUE(ctol));
    return 0;
}

int fsio_project(fsio_ctx* c, double* U, double* V, double* correction,
                 fsi_projection_result* res) {
    double* G = dalloc(c->N);
    int rc = project(c, U, V, G, correction, res);
    free(G);
    return rc;
}

/* ---------------------------------------------------------------------------
 * Context
 * ------------------------------------------------------------------------- */
static int validate_setup(fsio_ctx* c, const fsi_bc_setup* b, const char* name) {
    int lp = b->left.kind == FSI_BC_PERIODIC, rp = b->right.kind == FSI_BC_PERIODIC;
    int bp = b->bottom.kind == FSI_BC_PERIODIC, tp = b->top.kind == FSI_BC_PERIODIC;
    if (lp != rp || bp != tp) /* BoundaryConditions.jl:74-84 */
        return set_err(c, FSI_E_ARGUMENT, "Periodic boundaries must be specified as an opposing pair (%s).", name);
    if (lp != c->px) /* TimeIntegrationIMEX.jl:447-466 */
        return set_err(c, FSI_E_ARGUMENT, "%s boundary conditions do not match MeshMetrics periodicity in x.", name);
    if (bp != c->py)
        return set_err(c, FSI_E_ARGUMENT, "%s boundary conditions do not match MeshMetrics periodicity in y.", name);
    return 0;
}

int fsio_create(const fsi_grid* g, const fsi_bcs* bcs, const fsi_models* models,
                const fsi_numerics* num, fsio_ctx** out) {
    /* the work arrays of every call are allocated and freed again: keep the
       freed blocks in the heap instead of returning them to the system, so
       the next call does not fault its pages in again */
    mallopt(M_MMAP_THRESHOLD, 1 << 30);
    mallopt(M_TRIM_THRESHOLD, 1 << 30);
    fsio_ctx* c = (fsio_ctx*)calloc(1, sizeof(fsio_ctx));
    *out = c;
    c->nx = g->nx; c->ny = g->ny; c->N = g->nx * g->ny;
    c->lx = g->lx; c->ly = g->ly; c->px = g->periodic_x != 0; c->py = g->periodic_y != 0;
    c->bcs = *bcs; c->m = *models; c->num = *num;
    if (c->nx < 2 || c->ny < 2) return set_err(c, FSI_E_ARGUMENT, "Nx and Ny must be at least 2.");
    build_metrics(c);
    const fsi_bc_setup* all[8] = {&bcs->U, &bcs->V, &bcs->correction, &bcs->phi, &bcs->resource,
                                  &bcs->sigma_xx, &bcs->sigma_yy, &bcs->sigma_xy};
    const char* names[8] = {"U", "V", "correction", "phi", "resource", "sigma_xx", "sigma_yy", "sigma_xy"};
    for (int k = 0; k < 8; ++k) {
        int rc = validate_setup(c, all[k], names[k]);
        if (rc) return rc;
    }
    const fsi_bc_setup* q = &bcs->correction;
    c->has_nullspace = !(q->left.kind == FSI_BC_DIRICHLET || q->right.kind == FSI_BC_DIRICHLET ||
                         q->bottom.kind == FSI_BC_DIRICHLET || q->top.kind == FSI_BC_DIRICHLET);
    c->lap_bc = dalloc(c->N);
    double* zero = dalloc(c->N);
    laplacian_apply(c, zero, c->lap_bc, 1);
    free(zero);
    build_mg(c);
    /* growth tensor per unit G: G/2 I, or G t (x) t with t normalised as the
       DirectionalGrowthTensor constructor does (GrowthModels.jl:150-165) */
    c->kxx = 0.5; c->kxy = 0.0; c->kyy = 0.5;
    if (models->growth_tensor_kind == FSI_GROWTH_TENSOR_DIRECTIONAL) {
        double tx = models->growth_direction_x, ty = models->growth_direction_y, mag = hypot(tx, ty);
        if (!(isfinite(tx) && isfinite(ty))) return set_err(c, FSI_E_ARGUMENT, "Directional-growth components must be finite.");
        if (!(mag > 0.0)) return set_err(c, FSI_E_ARGUMENT, "Directional growth requires a non-zero direction vector.");
        tx = tx / mag;
        ty = ty / mag;
        c->kxx = tx * tx; c->kxy = tx * ty; c->kyy = ty * ty;
    }
    c->force_row = dalloc(c->ny);
    if (models->body_force_kind == FSI_FORCE_KOLMOGOROV) { /* TransportConstitutive.jl:385-397 */
        double y = 0.0;
        for (int j = 0; j < c->ny; ++j) {
            y += 0.5 * c->mt.dy_c[j];
            c->force_row[j] = models->force_amplitude * sin(models->force_wavenumber * y + models->force_phase);
            y += 0.5 * c->mt.dy_c[j];
        }
    }
    return 0;
}

void fsio_destroy(fsio_ctx* c) {
    if (!c) return;
    metrics_t* m = &c->mt;
    double* ptrs[] = {m->dx_f, m->dx_c, m->dy_f, m->dy_c, m->wL_c2fx, m->wR_c2fx, m->wB_c2fy,
                      m->wT_c2fy, m->wL_f2cx, m->wR_f2cx, m->wB_f2cy, m->wT_f2cy,
                      m->dx_west_corner, m->dx_east_corner, m->dy_south_corner,
                      m->dy_north_corner, m->x_p, m->y_p, m->x_u, m->y_v, c->lap_bc, c->force_row};
    for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
    for (int l = 0; l < c->nlev; ++l) { free(c->lev[l].x); free(c->lev[l].b); free(c->lev[l].r); }
    free(c);
}

/* ---------------------------------------------------------------------------
 * Stage tendencies: _evaluate_stage_tendencies! TimeIntegrationIMEX.jl:944-1071
 * ------------------------------------------------------------------------- */
static int tendencies(fsio_ctx* c, const double* U, const double* V, const double* phi,
                      const double* sxx, const double* syy, const double* sxy, const double* G,
                      double t_stage, double* E_u, double* E_v, double* I_u, double* I_v,
                      double* E_phi, double* E_sxx, double* E_syy, double* E_sxy,
                      const double* res, const double* Dc, double* E_c, double* I_c) {
    int nx = c->nx, ny = c->ny, N = c->N;
    size_t nfx = (size_t)(nx + 1) * ny, nfy = (size_t)nx * (ny + 1);
    (void)t_stage; /* the supported body forces are steady */
    double *Cu = dalloc(nfx), *Cv = dalloc(nfy), *Fx = dalloc(nfx), *Fy = dalloc(nfy);
    material_convection(c, U, V, G, Cu, Cv);
    for (size_t k = 0; k < nfx; ++k) E_u[k] = -Cu[k];
    for (size_t k = 0; k < nfy; ++k) E_v[k] = -Cv[k];
    /* body_force! TransportConstitutive.jl:369-397 */
    if (c->m.body_force_kind == FSI_FORCE_CONSTANT) {
        for (size_t k = 0; k < nfx; ++k) E_u[k] += c->m.force_fx;
        for (size_t k = 0; k < nfy; ++k) E_v[k] += c->m.force_fy;
    } else if (c->m.body_force_kind == FSI_FORCE_KOLMOGOROV) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i) UX(E_u, i, j) += c->force_row[j];
        for (size_t k = 0; k < nfy; ++k) E_v[k] += 0.0;
    }
    /* capillary_force (Capillarity.jl:134-187), :975-977 */
    if (c->m.capillarity_kind == FSI_CAPILLARITY_DIFFUSE_AREA) {
        capillary_force(c, phi, Fx, Fy);
        for (size_t k = 0; k < nfx; ++k) E_u[k] += Fx[k];
        for (size_t k = 0; k < nfy; ++k) E_v[k] += Fy[k];
    }
    /* _accumulate_elastic_force! :903-924 with elastic_stress
       StructuralConstitutive.jl:100-113 */
    if (c->m.elastic_kind == FSI_ELASTIC_LINEAR) {
        double *Txx = dalloc(N), *Tyy = dalloc(N), *Txy = dalloc(N), Gm = FSIO_SHEAR_MODULUS(c);
        for (int k = 0; k < N; ++k) {
            Txx[k] = phi[k] * (Gm * (sxx[k] - 1.0));
            Tyy[k] = phi[k] * (Gm * (syy[k] - 1.0));
            Txy[k] = phi[k] * (Gm * sxy[k]);
        }
        stress_divergence(c, Txx, Tyy, Txy, Fx, Fy);
        for (size_t k = 0; k < nfx; ++k) E_u[k] += Fx[k];
        for (size_t k = 0; k < nfy; ++k) E_v[k] += Fy[k];
        free(Txx); free(Tyy); free(Txy);
    }
    /* _accumulate_growth_viscous_force! :926-942: the divergence of the
       stress -2 mu K_g (exactly zero under NoGrowth, so skipped there).
       */
    if (c->m.viscosity_kind == FSI_VISC_CONSTANT && growth_active(c)) {
        double *Txx = dalloc(N), *Tyy = dalloc(N), *Txy = dalloc(N), mu = c->m.viscosity;
        for (int k = 0; k < N; ++k) {
            double Kxx = c->kxx * G[k], Kxy = c->kxy * G[k], Kyy = c->kyy * G[k];
            Txx[k] = -2.0 * mu * Kxx;
            Tyy[k] = -2.0 * mu * Kyy;
            Txy[k] = -mu * (Kxy + Kxy);
        }
        stress_divergence(c, Txx, Tyy, Txy, Fx, Fy);
        for (size_t k = 0; k < nfx; ++k) E_u[k] += Fx[k];
        for (size_t k = 0; k < nfy; ++k) E_v[k] += Fy[k];
        free(Txx); free(Tyy); free(Txy);
    }
    int rc = phase_rhs(c, U, V, phi, G, E_phi);
    if (!rc) rc = structure_rhs(c, sxx, syy, sxy, U, V, phi, G, E_sxx, E_syy, E_sxy);
    /* resource transport and consumption :1002-1026; the implicit history
       L c + b of the diffusion operator :1064-1069 */
    if (!rc && E_c) {
        if (growth_active(c)) rc = resource_explicit_rhs(c, U, V, phi, res, E_c);
        else memset(E_c, 0, sizeof(double) * N);
    }
    if (!rc && I_c) {
        if (c->m.diffusivity_kind != FSI_DIFFUSION_NONE) diffusion_apply(c, Dc, res, I_c, 1);
        else memset(I_c, 0, sizeof(double) * N);
    }
    if (!rc && (I_u || I_v)) {
        double *uv = dalloc(2 * (size_t)N), *iu = dalloc(2 * (size_t)N);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) { uv[i + j * nx] = UX(U, i, j); uv[N + i + j * nx] = VY(V, i, j); }
        if (c->m.viscosity_kind == FSI_VISC_CONSTANT) viscous_apply(c, uv, iu, 1);
        if (I_u) {
            memset(I_u, 0, sizeof(double) * nfx);
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) UX(I_u, i, j) = iu[i + j * nx];
        }
        if (I_v) {
            memset(I_v, 0, sizeof(double) * nfy);
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) VY(I_v, i, j) = iu[N + i + j * nx];
        }
        free(uv); free(iu);
    }
    free(Cu); free(Cv); free(Fx); free(Fy);
    return rc;
}

/* Tendencies of the stage-1 state: sigma is taken after the safety reset of
 * _copy_time_n_state! (:686-702), as the first stage of integrate_step! sees it. */
int fsio_eval_tendencies(fsio_ctx* c, const double* U, const double* V, const double* phi,
                         const double* sxx, const double* syy, const double* sxy, fsio_scalar t,
                         double* E_u, double* E_v, double* I_u, double* I_v, double* E_phi,
                         double* E_sxx, double* E_syy, double* E_sxy) {
    double* G = dalloc(c->N);
    double *a = dalloc(c->N), *d = dalloc(c->N), *b = dalloc(c->N);
    memcpy(a, sxx, sizeof(double) * c->N);
    memcpy(d, syy, sizeof(double) * c->N);
    memcpy(b, sxy, sizeof(double) * c->N);
    apply_extension_reset(c, a, d, b, phi);
    int rc = tendencies(c, U, V, phi, a, d, b, G, t, E_u, E_v, I_u, I_v, E_phi, E_sxx, E_syy, E_sxy, NULL, NULL,
                        NULL, NULL);
    free(G);
    free(a);
    free(d);
    free(b);
    return rc;
}

/* The momentum tendencies of a stage with target divergence / growth source G
 * and resource res (sigma taken as given): the terms S v and the growth
 * viscous stress of _evaluate_stage_tendencies! :944-985 included. */
int fsio_eval_tendencies_growth(fsio_ctx* c, const double* U, const double* V, const double* phi,
                                const double* sxx, const double* syy, const double* sxy, const double* res,
                                const double* G, double* E_u, double* E_v) {
    double *Ep = dalloc(c->N), *Exx = dalloc(c->N), *Eyy = dalloc(c->N), *Exy = dalloc(c->N), *Ec = dalloc(c->N);
    int rc = tendencies(c, U, V, phi, sxx, syy, sxy, G, 0.0, E_u, E_v, NULL, NULL, Ep, Exx, Eyy, Exy, res, NULL,
                        Ec, NULL);
    free(Ep); free(Exx); free(Eyy); free(Exy); free(Ec);
    return rc;
}

/* ---------------------------------------------------------------------------
 * integrate_step! TimeIntegrationIMEX.jl:1152-1256
 * ------------------------------------------------------------------------- */
static const double AE[4][4] = {{0, 0, 0, 0},
                                {8.0 / 15.0, 0, 0, 0},
                                {1.0 / 4.0, 5.0 / 12.0, 0, 0},
                                {1.0 / 4.0, 0, 3.0 / 4.0, 0}};
static const double AI[4][4] = {{0, 0, 0, 0},
                                {4.0 / 15.0, 4.0 / 15.0, 0, 0},
                                {4.0 / 15.0, 1.0 / 3.0, 1.0 / 15.0, 0},
                                {4.0 / 15.0, 1.0 / 3.0, 7.0 / 30.0, 1.0 / 6.0}};
static const double CABS[4] = {0.0, 8.0 / 15.0, 2.0 / 3.0, 1.0};

static int validate_state(fsio_ctx* c, const double* U, const double* V, const double* phi,
                          const double* res, const double* sxx, const double* syy,
                          const double* sxy) {
    /* _validate_state :495-523 */
    const double* cells[5] = {phi, res, sxx, syy, sxy};
    const char* names[5] = {"state.phi", "state.resource", "state.sigma_xx", "state.sigma_yy", "state.sigma_xy"};
    for (int f = 0; f < 5; ++f)
        for (int k = 0; k < c->N; ++k)
            if (!isfinite(cells[f][k])) return set_err(c, FSI_E_ARGUMENT, "%s contains non-finite values.", names[f]);
    for (size_t k = 0; k < (size_t)(c->nx + 1) * c->ny; ++k)
        if (!isfinite(U[k])) return set_err(c, FSI_E_ARGUMENT, "state.U contains non-finite values.");
    for (size_t k = 0; k < (size_t)c->nx * (c->ny + 1); ++k)
        if (!isfinite(V[k])) return set_err(c, FSI_E_ARGUMENT, "state.V contains non-finite values.");
    return 0;
}

int fsio_integrate_step(fsio_ctx* c, double* U, double* V, double* phi, double* resource,
                        double* sxx, double* syy, double* sxy, double* correction, fsio_scalar t,
                        fsio_scalar dt, fsi_step_report* rep, fsi_step_stats* stats) {
    int nx = c->nx, ny = c->ny, N = c->N, rc;
    size_t nfx = (size_t)(nx + 1) * ny, nfy = (size_t)nx * (ny + 1);
    if ((rc = validate_state(c, U, V, phi, resource, sxx, syy, sxy))) return rc;
    if (!isfinite(t)) return set_err(c, FSI_E_ARGUMENT, "t must be finite.");
    if (!isfinite(dt)) return set_err(c, FSI_E_ARGUMENT, "dt must be finite.");
    if (!(dt > 0)) return set_err(c, FSI_E_ARGUMENT, "dt must be strictly positive.");
    if (stats) memset(stats, 0, sizeof(*stats));

    /* workspace (IMEXWorkspace :333-441) */
    double *uv_n = dalloc(2 * (size_t)N), *phi_n = dalloc(N), *sxx_n = dalloc(N), *syy_n = dalloc(N), *sxy_n = dalloc(N);
    double *Us = dalloc(nfx), *Vs = dalloc(nfy), *phis = dalloc(N), *sxxs = dalloc(N), *syys = dalloc(N), *sxys = dalloc(N);
    double *rhs = dalloc(2 * (size_t)N), *pred = dalloc(2 * (size_t)N), *G = dalloc(N), *dv = dalloc(N);
    double *Eu[3], *Ev[3], *Iu[3], *Iv[3], *Ep[3], *Exx[3], *Eyy[3], *Exy[3], *Ec[3], *Ic[3];
    for (int s = 0; s < 3; ++s) {
        Eu[s] = dalloc(nfx); Ev[s] = dalloc(nfy); Iu[s] = dalloc(nfx); Iv[s] = dalloc(nfy);
        Ep[s] = dalloc(N); Exx[s] = dalloc(N); Eyy[s] = dalloc(N); Exy[s] = dalloc(N);
        Ec[s] = dalloc(N); Ic[s] = dalloc(N);
    }
    /* the resource of the stage (c_n, c_stage, rhs_c) and the diffusivity of phi_s */
    double *c_n = dalloc(N), *cs = dalloc(N), *rhs_c = dalloc(N), *Dc = dalloc(N);
    const int growth = growth_active(c), diffusion = c->m.diffusivity_kind != FSI_DIFFUSION_NONE;
    double max_trace = 0.0;
    memcpy(c_n, resource, sizeof(double) * N);
    /* _copy_time_n_state! :686-702 */
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) { uv_n[i + j * nx] = UX(U, i, j); uv_n[N + i + j * nx] = VY(V, i, j); }
    memcpy(phi_n, phi, sizeof(double) * N);
    memcpy(sxx_n, sxx, sizeof(double) * N);
    memcpy(syy_n, syy, sizeof(double) * N);
    memcpy(sxy_n, sxy, sizeof(double) * N);
    apply_extension_reset(c, sxx_n, syy_n, sxy_n, phi_n);

    double max_linf = 0.0;
    for (int s = 0; s < 4 && !rc; ++s) {
        if (s == 0) { /* _set_stage_one! :704-714 */
            memcpy(Us, U, sizeof(double) * nfx);
            memcpy(Vs, V, sizeof(double) * nfy);
            memcpy(phis, phi_n, sizeof(double) * N);
            memcpy(sxxs, sxx_n, sizeof(double) * N);
            memcpy(syys, syy_n, sizeof(double) * N);
            memcpy(sxys, sxy_n, sizeof(double) * N);
            memcpy(cs, c_n, sizeof(double) * N);
        } else { /* _build_explicit_stage! :716-740 */
            memcpy(phis, phi_n, sizeof(double) * N);
            memcpy(sxxs, sxx_n, sizeof(double) * N);
            memcpy(syys, syy_n, sizeof(double) * N);
            memcpy(sxys, sxy_n, sizeof(double) * N);
            for (int j = 0; j < s; ++j) {
                double a = AE[s][j];
                if (a == 0.0) continue;
                OMP_FOR
                for (int k = 0; k < N; ++k) {
                    phis[k] += dt * a * Ep[j][k];
                    sxxs[k] += dt * a * Exx[j][k];
                    syys[k] += dt * a * Eyy[j][k];
                    sxys[k] += dt * a * Exy[j][k];
                }
            }
            apply_extension_reset(c, sxxs, syys, sxys, phis);
        }
        /* _build_diffusion_operator! :867-880 (phi validated even without
           diffusion), then the resource stage :742-774 for s > 1 */
        if (growth || diffusion) {
            if (diffusion) {
                if ((rc = diffusivity_field(c, phis, Dc))) break;
            } else {
                for (int k = 0; k < N; ++k)
                    if (!isfinite(phis[k])) { rc = set_err(c, FSI_E_ARGUMENT, "phi contains non-finite values."); break; }
                if (rc) break;
            }
        }
        if (s > 0 && growth) {
            memcpy(rhs_c, c_n, sizeof(double) * N);
            for (int q = 0; q < s; ++q) {
                double ae = AE[s][q], ai = AI[s][q];
                if (ae == 0.0 && ai == 0.0) continue;
                for (int k = 0; k < N; ++k) rhs_c[k] += dt * (ae * Ec[q][k] + ai * Ic[q][k]);
            }
            if (diffusion) {
                int it = 0;
                if ((rc = resource_solve(c, FSIO_VALUE(dt * AI[s][s]), Dc, rhs_c, c_n, cs, &it))) break;
            } else {
                memcpy(cs, rhs_c, sizeof(double) * N);
            }
        }
        /* _evaluate_growth! :842-865: G_s, K_g,s and the trace error */
        if (growth) {
            double te = growth_source(c, phis, cs, G);
            if (te > max_trace || isnan(te)) max_trace = te;
        }
        if (s == 0) { /* stage-one constraint :1194-1201, :1111-1127 */
            divergence(c, Us, Vs, dv);
            double linf = 0.0, gmax = 0.0;
            for (int k = 0; k < N; ++k) {
                double r = fabs(dv[k] - G[k]);
                if (r > linf) linf = r;
                if (fabs(G[k]) > gmax) gmax = fabs(G[k]);
            }
            if (c->num.check_stage1) {
                double tol = c->num.stage1_atol + c->num.stage1_rtol * (gmax > 1.0 ? gmax : 1.0);
                if (!(linf <= tol)) {
                    rc = set_err(c, FSI_E_ARGUMENT,
                                 "The velocity entering stage 1 does not satisfy the prescribed "
                                 "divergence constraint. Maximum |div(v)-G| = %.3e, tolerance = %.3e.",
                                 FSIO_VALUE(linf), FSIO_VALUE(tol));
                    break;
                }
            }
            if (linf > max_linf) max_linf = linf;
        } else { /* _build_velocity_stage! :776-836 */
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) { rhs[i + j * nx] = uv_n[i + j * nx]; rhs[N + i + j * nx] = uv_n[N + i + j * nx]; }
            for (int q = 0; q < s; ++q) {
                double ae = AE[s][q], ai = AI[s][q];
                if (ae == 0.0 && ai == 0.0) continue;
                OMP_FOR
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        int k = i + j * nx;
                        rhs[k] += dt * (ae * UX(Eu[q], i, j) + ai * UX(Iu[q], i, j));
                        rhs[N + k] += dt * (ae * VY(Ev[q], i, j) + ai * VY(Iv[q], i, j));
                    }
            }
            if (c->m.viscosity_kind == FSI_VISC_NONE) {
                unsigned char* mask = (unsigned char*)calloc(2 * (size_t)N, 1);
                double* gv = dalloc(2 * (size_t)N);
                dirichlet_mask(c, mask, gv);
                for (int k = 0; k < 2 * N; ++k) pred[k] = mask[k] ? gv[k] : rhs[k];
                free(mask); free(gv);
            } else {
                int it = 0;
                rc = fsio_velocity_solve(c, FSIO_VALUE(dt * AI[s][s]), rhs, uv_n, pred, &it);
                if (stats) stats->velocity_iterations[s - 1] = it;
                if (rc) break;
            }
            /* _inject_compact_predictor! :543-579 */
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) { UX(Us, i, j) = pred[i + j * nx]; VY(Vs, i, j) = pred[N + i + j * nx]; }
            if (c->px) for (int j = 0; j < ny; ++j) UX(Us, nx, j) = UX(Us, 0, j);
            else if (c->bcs.U.right.kind == FSI_BC_DIRICHLET) for (int j = 0; j < ny; ++j) UX(Us, nx, j) = c->bcs.U.right.value;
            else { rc = set_err(c, FSI_E_ARGUMENT, "The compact predictor currently requires a Dirichlet or Periodic U-right boundary."); break; }
            if (c->py) for (int i = 0; i < nx; ++i) VY(Vs, i, ny) = VY(Vs, i, 0);
            else if (c->bcs.V.top.kind == FSI_BC_DIRICHLET) for (int i = 0; i < nx; ++i) VY(Vs, i, ny) = c->bcs.V.top.value;
            else { rc = set_err(c, FSI_E_ARGUMENT, "The compact predictor currently requires a Dirichlet or Periodic V-top boundary."); break; }
            fsi_projection_result pr;
            rc = project(c, Us, Vs, G, correction, &pr);
            if (rc) break;
            if (stats) stats->projection_iterations[s - 1] = pr.iterations;
            if (pr.constraint_linf > max_linf) max_linf = pr.constraint_linf;
        }
        if (s <= 2) {
            rc = tendencies(c, Us, Vs, phis, sxxs, syys, sxys, G, t + CABS[s] * dt, Eu[s], Ev[s],
                            Iu[s], Iv[s], Ep[s], Exx[s], Eyy[s], Exy[s], cs, Dc, Ec[s], Ic[s]);
        }
    }
    if (!rc) {
        memcpy(U, Us, sizeof(double) * nfx);
        memcpy(V, Vs, sizeof(double) * nfy);
        memcpy(phi, phis, sizeof(double) * N);
        memcpy(sxx, sxxs, sizeof(double) * N);
        memcpy(syy, syys, sizeof(double) * N);
        memcpy(sxy, sxys, sizeof(double) * N);
        memcpy(resource, cs, sizeof(double) * N);
        if (rep) { /* structure_tensor_diagnostics StructureTensor.jl:627-662 */
            double mineig = INFINITY, maxdet = 0.0, pmin = INFINITY, pmax = -INFINITY, rmin = INFINITY;
            long long nonspd = 0;
            for (int k = 0; k < N; ++k) {
                double a = sxx[k], d = syy[k], b = sxy[k];
                double det = a * d - b * b;
                double lmin = 0.5 * (a + d - hypot(a - d, 2.0 * b));
                if (fabs(det - 1.0) > maxdet) maxdet = fabs(det - 1.0);
                if (lmin < mineig) mineig = lmin;
                if (!(isfinite(lmin) && lmin > 0.0)) ++nonspd;
                if (phi[k] < pmin) pmin = phi[k];
                if (phi[k] > pmax) pmax = phi[k];
                if (resource[k] < rmin) rmin = resource[k];
            }
            rep->maximum_projection_constraint_linf = FSIO_VALUE(max_linf);
            rep->maximum_growth_trace_error = FSIO_VALUE(max_trace);
            rep->minimum_phi = FSIO_VALUE(pmin);
            rep->maximum_phi = FSIO_VALUE(pmax);
            rep->minimum_resource = FSIO_VALUE(rmin);
            rep->minimum_sigma_eigenvalue = FSIO_VALUE(mineig);
            rep->maximum_abs_sigma_determinant_error = FSIO_VALUE(maxdet);
            rep->non_spd_cells = nonspd;
        }
    }
    free(uv_n); free(phi_n); free(sxx_n); free(syy_n); free(sxy_n);
    free(Us); free(Vs); free(phis); free(sxxs); free(syys); free(sxys);
    free(rhs); free(pred); free(G); free(dv);
    free(c_n); free(cs); free(rhs_c); free(Dc);
    for (int s = 0; s < 3; ++s) {
        free(Eu[s]); free(Ev[s]); free(Iu[s]); free(Iv[s]); free(Ep[s]); free(Exx[s]); free(Eyy[s]); free(Exy[s]);
        free(Ec[s]); free(Ic[s]);
    }
    return rc;
}

/* make_velocity_compatible! TimeIntegrationIMEX.jl:1088-1109 */
int fsio_make_velocity_compatible(fsio_ctx* c, double* U, double* V, const double* phi,
                                  const double* resource, double* correction,
                                  fsi_projection_result* res) {
    double* G = dalloc(c->N);
    growth_source(c, phi, resource, G); /* growth_source! :1094 */
    int rc = project(c, U, V, G, correction, res);
    free(G);
    return rc;
}
