/*
 * fsio_dual.hpp -- forward-mode dual numbers for the oracle's tangent build.
 *
 * TEST INFRASTRUCTURE ONLY (see fsi_tangent.cpp).  A value carries
 * FSIO_NDIR directional derivatives; every operation propagates them by the
 * chain rule, and every comparison looks at the value alone, so branches,
 * upwind choices and clamps are taken as in the FP64 oracle and the
 * derivative is the one of the branch taken (a frozen switch).
 */
#ifndef FSIO_DUAL_HPP
#define FSIO_DUAL_HPP

#include <cmath>

#ifndef FSIO_NDIR
#define FSIO_NDIR 4
#endif

struct fsio_real {
    double v;
    double d[FSIO_NDIR];
    fsio_real() = default;  // trivial: calloc'ed and memset arrays are zero
    fsio_real(double x) : v(x), d{} {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    explicit operator int() const { return (int)v; }
    explicit operator long() const { return (long)v; }
    explicit operator long long() const { return (long long)v; }
    explicit operator bool() const { return v != 0.0; }
    fsio_real& operator+=(const fsio_real& b) {
        v += b.v;
        for (int k = 0; k < FSIO_NDIR; ++k) d[k] += b.d[k];
        return *this;
    }
    fsio_real& operator-=(const fsio_real& b) {
        v -= b.v;
        for (int k = 0; k < FSIO_NDIR; ++k) d[k] -= b.d[k];
        return *this;
    }
    fsio_real& operator*=(const fsio_real& b) {
        for (int k = 0; k < FSIO_NDIR; ++k) d[k] = d[k] * b.v + v * b.d[k];
        v *= b.v;
        return *this;
    }
    fsio_real& operator/=(const fsio_real& b) {
        const double q = v / b.v;
        for (int k = 0; k < FSIO_NDIR; ++k) d[k] = (d[k] - q * b.d[k]) / b.v;
        v = q;
        return *this;
    }
};

/* value and derivative scale: f(a) with f'(a) = s */
static inline fsio_real fsio_chain(const fsio_real& a, double f, double s) {
    fsio_real r(f);
    for (int k = 0; k < FSIO_NDIR; ++k) r.d[k] = s * a.d[k];
    return r;
}

static inline fsio_real operator+(fsio_real a, const fsio_real& b) { return a += b; }
static inline fsio_real operator-(fsio_real a, const fsio_real& b) { return a -= b; }
static inline fsio_real operator*(fsio_real a, const fsio_real& b) { return a *= b; }
static inline fsio_real operator/(fsio_real a, const fsio_real& b) { return a /= b; }
static inline fsio_real operator+(fsio_real a, double b) { a.v += b; return a; }
static inline fsio_real operator+(double b, fsio_real a) { a.v += b; return a; }
static inline fsio_real operator-(fsio_real a, double b) { a.v -= b; return a; }
static inline fsio_real operator-(double b, const fsio_real& a) { return fsio_chain(a, b - a.v, -1.0); }
static inline fsio_real operator*(const fsio_real& a, double b) { return fsio_chain(a, a.v * b, b); }
static inline fsio_real operator*(double b, const fsio_real& a) { return fsio_chain(a, b * a.v, b); }
static inline fsio_real operator/(const fsio_real& a, double b) {
    fsio_real r(a.v / b);
    for (int k = 0; k < FSIO_NDIR; ++k) r.d[k] = a.d[k] / b;
    return r;
}
static inline fsio_real operator/(double b, const fsio_real& a) {
    const double q = b / a.v;
    return fsio_chain(a, q, -q / a.v);
}
static inline fsio_real operator-(const fsio_real& a) { return fsio_chain(a, -a.v, -1.0); }
static inline fsio_real operator+(const fsio_real& a) { return a; }

#define FSIO_CMP(op)                                                                          \
    static inline bool operator op(const fsio_real& a, const fsio_real& b) { return a.v op b.v; } \
    static inline bool operator op(const fsio_real& a, double b) { return a.v op b; }           \
    static inline bool operator op(double a, const fsio_real& b) { return a op b.v; }
FSIO_CMP(<)
FSIO_CMP(>)
FSIO_CMP(<=)
FSIO_CMP(>=)
FSIO_CMP(==)
FSIO_CMP(!=)
#undef FSIO_CMP
static inline bool operator!(const fsio_real& a) { return !(a.v != 0.0); }

static inline fsio_real sqrt(const fsio_real& a) {
    const double r = std::sqrt(a.v);
    return fsio_chain(a, r, 0.5 / r);
}
static inline fsio_real fabs(const fsio_real& a) { return a.v < 0.0 ? -a : a; }
static inline fsio_real log(const fsio_real& a) { return fsio_chain(a, std::log(a.v), 1.0 / a.v); }
static inline fsio_real exp(const fsio_real& a) {
    const double e = std::exp(a.v);
    return fsio_chain(a, e, e);
}
static inline fsio_real tanh(const fsio_real& a) {
    const double t = std::tanh(a.v);
    return fsio_chain(a, t, 1.0 - t * t);
}
static inline fsio_real sin(const fsio_real& a) { return fsio_chain(a, std::sin(a.v), std::cos(a.v)); }
static inline fsio_real pow(const fsio_real& a, double e) {
    const double p = std::pow(a.v, e);
    return fsio_chain(a, p, e == 0.0 ? 0.0 : e * std::pow(a.v, e - 1.0));
}
static inline fsio_real hypot(const fsio_real& a, const fsio_real& b) {
    const double h = std::hypot(a.v, b.v);
    fsio_real r(h);
    if (h > 0.0)
        for (int k = 0; k < FSIO_NDIR; ++k) r.d[k] = (a.v * a.d[k] + b.v * b.d[k]) / h;
    return r;
}
static inline fsio_real hypot(const fsio_real& a, double b) { return hypot(a, fsio_real(b)); }
static inline fsio_real hypot(double a, const fsio_real& b) { return hypot(fsio_real(a), b); }
/* fmax/fmin: the larger (smaller) operand with its derivative; NaN skipped as C does */
static inline fsio_real fmax(const fsio_real& a, const fsio_real& b) {
    if (std::isnan(a.v)) return b;
    if (std::isnan(b.v)) return a;
    return a.v >= b.v ? a : b;
}
static inline fsio_real fmin(const fsio_real& a, const fsio_real& b) {
    if (std::isnan(a.v)) return b;
    if (std::isnan(b.v)) return a;
    return a.v <= b.v ? a : b;
}
static inline fsio_real fmax(const fsio_real& a, double b) { return fmax(a, fsio_real(b)); }
static inline fsio_real fmax(double a, const fsio_real& b) { return fmax(fsio_real(a), b); }
static inline fsio_real fmin(const fsio_real& a, double b) { return fmin(a, fsio_real(b)); }
static inline fsio_real fmin(double a, const fsio_real& b) { return fmin(fsio_real(a), b); }
static inline bool isfinite(const fsio_real& a) { return std::isfinite(a.v); }
static inline bool isnan(const fsio_real& a) { return std::isnan(a.v); }

#endif
