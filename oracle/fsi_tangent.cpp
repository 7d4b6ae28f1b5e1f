/*
 * fsi_tangent.cpp -- the oracle's FP64 tangent (forward-mode derivatives of
 * the CPU restatement), built as liboracle_fsi_tan.so.
 *
 * TEST INFRASTRUCTURE ONLY.  It is the checker of the device adjoint: the
 * same source as liboracle_fsi.so (fsi_oracle.c) compiled as C++ with every
 * double of the oracle replaced by fsio_real (fsio_dual.hpp), a value with
 * FSIO_NDIR directional derivatives.  The boundary structs of fsi_api.h keep
 * their doubles (they are included first); a state array element is
 * {value, d_1 .. d_FSIO_NDIR}.  The shear modulus of the context can carry
 * derivatives too (fsio_tangent_set_shear_modulus).  Comparisons see values
 * only, so every branch (upwind side, clamp, safety reset, solver stopping
 * test) is taken as in the FP64 oracle: the result is the derivative of the
 * discrete step along the branches the primal took, which is what a
 * hand-written discrete adjoint with frozen switches transposes.
 */
#include <malloc.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>

#include "../include/fsi_api.h"
#include "fsio_dual.hpp"

static fsio_real g_shear_modulus_dual;
static int g_shear_modulus_set = 0;

/* the oracle's declarations are re-made below with the dual type */
#define FSI_ORACLE_H
typedef struct fsio_ctx fsio_ctx;
typedef double fsio_scalar;
#define FSIO_SHEAR_MODULUS(c) (g_shear_modulus_set ? g_shear_modulus_dual : fsio_real((c)->m.shear_modulus))
static inline double fsio_value_of(const fsio_real& x) { return x.v; }
#define FSIO_VALUE(x) fsio_value_of(x)

#define double fsio_real
extern "C" {
#include "fsi_oracle.c"
}
#undef double

/* d(shear modulus) along each direction; value must equal the context's */
extern "C" void fsio_tangent_set_shear_modulus(double value, const double* derivs) {
    g_shear_modulus_dual = fsio_real(value);
    for (int k = 0; k < FSIO_NDIR; ++k) g_shear_modulus_dual.d[k] = derivs ? derivs[k] : 0.0;
    g_shear_modulus_set = derivs != nullptr;
}

extern "C" int fsio_tangent_ndir(void) { return FSIO_NDIR; }
