/*
 * dualmap.cpp -- first-order truncated power series build of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (loaded by tests/test_dual_maps.py).  The
 * reference obtains its one-turn and element maps by passing tps objects
 * through the same templated maps (DAPass, Cal_OneTurnMap
 * simtrack.cpp:15122-15179, Cal_ElementMap cpp:15205-15253); to first order
 * a tps is a dual number: a value and its six derivatives with respect to
 * the initial coordinates.  st_oracle.c is compiled here as C++ with every
 * `double` replaced by Dual, so each map is differentiated exactly (to
 * rounding), without the finite-difference step the GPU's
 * st_one_turn_matrix and st_element_maps use.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <cmath>
#include <cstddef>

#include "simtrack_amd.h"

#define ST_FN static inline
#define ST_CONST_TABLE static const
#define ST_FMA(a, b, c) fma((a), (b), (c))
#define ST_RINT(x) nearbyint(x)
#define ST_SQRT(x) sqrt(x)
#include "st_libm.h"
#include "st_rng.h"

struct Dual {
    double v;
    double d[6];
    Dual() : v(0.), d{0., 0., 0., 0., 0., 0.} {}
    Dual(double x) : v(x), d{0., 0., 0., 0., 0., 0.} {}
    operator double() const { return v; }
};

/* r = f(a) with f'(a) = fa */
static inline Dual chain1(double v, const Dual &a, double fa)
{
    Dual r(v);
    for (int k = 0; k < 6; k++) r.d[k] = fa * a.d[k];
    return r;
}
/* r = f(a, b) with partial derivatives fa, fb */
static inline Dual chain2(double v, const Dual &a, double fa, const Dual &b, double fb)
{
    Dual r(v);
    for (int k = 0; k < 6; k++) r.d[k] = fa * a.d[k] + fb * b.d[k];
    return r;
}

static inline Dual operator+(const Dual &a, const Dual &b) { return chain2(a.v + b.v, a, 1., b, 1.); }
static inline Dual operator-(const Dual &a, const Dual &b) { return chain2(a.v - b.v, a, 1., b, -1.); }
static inline Dual operator*(const Dual &a, const Dual &b) { return chain2(a.v * b.v, a, b.v, b, a.v); }
static inline Dual operator/(const Dual &a, const Dual &b)
{
    double q = a.v / b.v;
    return chain2(q, a, 1. / b.v, b, -q / b.v);
}
static inline Dual operator-(const Dual &a) { return chain1(-a.v, a, -1.); }

#define DUAL_MIXED(op)                                                                                  \
    static inline Dual operator op(const Dual &a, double b) { return a op Dual(b); }                     \
    static inline Dual operator op(double a, const Dual &b) { return Dual(a) op b; }                     \
    static inline Dual operator op(const Dual &a, int b) { return a op Dual((double)b); }                \
    static inline Dual operator op(int a, const Dual &b) { return Dual((double)a) op b; }
DUAL_MIXED(+)
DUAL_MIXED(-)
DUAL_MIXED(*)
DUAL_MIXED(/)
#undef DUAL_MIXED

#define DUAL_CMP(op)                                                                                    \
    static inline bool operator op(const Dual &a, const Dual &b) { return a.v op b.v; }                  \
    static inline bool operator op(const Dual &a, double b) { return a.v op b; }                         \
    static inline bool operator op(double a, const Dual &b) { return a op b.v; }                         \
    static inline bool operator op(const Dual &a, int b) { return a.v op b; }                            \
    static inline bool operator op(int a, const Dual &b) { return a op b.v; }
DUAL_CMP(<)
DUAL_CMP(>)
DUAL_CMP(<=)
DUAL_CMP(>=)
DUAL_CMP(==)
DUAL_CMP(!=)
#undef DUAL_CMP

static inline Dual &operator+=(Dual &a, const Dual &b) { return a = a + b; }
static inline Dual &operator-=(Dual &a, const Dual &b) { return a = a - b; }
static inline Dual &operator*=(Dual &a, const Dual &b) { return a = a * b; }
static inline Dual &operator/=(Dual &a, const Dual &b) { return a = a / b; }

static inline Dual sqrt(const Dual &a)
{
    double s = std::sqrt(a.v);
    return chain1(s, a, 0.5 / s);
}
static inline Dual fabs(const Dual &a) { return chain1(std::fabs(a.v), a, a.v < 0. ? -1. : 1.); }
static inline bool isnan(const Dual &a) { return std::isnan(a.v); }
static inline Dual st_sin(const Dual &a)
{
    double s, c;
    st_sincos(a.v, &s, &c);
    return chain1(s, a, c);
}
static inline Dual st_cos(const Dual &a)
{
    double s, c;
    st_sincos(a.v, &s, &c);
    return chain1(c, a, -s);
}
static inline Dual st_tan(const Dual &a)
{
    double t = st_tan(a.v);
    return chain1(t, a, 1. + t * t);
}
static inline Dual st_asin(const Dual &a) { return chain1(st_asin(a.v), a, 1. / std::sqrt(1. - a.v * a.v)); }
static inline Dual st_exp(const Dual &a)
{
    double e = st_exp(a.v);
    return chain1(e, a, e);
}
static inline Dual st_powi(const Dual &a, int n) { return chain1(st_powi(a.v, n), a, n * st_powi(a.v, n - 1)); }
static inline Dual pow(const Dual &a, const Dual &b)
{
    double p = std::pow(a.v, b.v);
    return chain2(p, a, b.v * std::pow(a.v, b.v - 1.), b, a.v > 0. ? p * std::log(a.v) : 0.);
}

/* st_element with Dual members, field for field (constants: zero gradient) */
struct dual_element {
    int32_t type, nint, norder, nslice, treatment, reserved[3];
    Dual l, angle, e1, e2, k0l, knl[11], knsl[11], dx, dy, ds, dtheta, dphi, dpsi, apx, apy, vrf, frf, phase0,
        hkick, vkick, theta, np, bbscale, sigmal, emitx, betax, alfax, emity, betay, alfay, cc[8], rot_n[3],
        rot_angle, m66[36], xco_in[6], xco_out[6], ks, psi, phi, alpha, sepx, sepy, sigmax, sigmay, ne, betae,
        rout, rin, nud, phid, turns, turne, klmax, tturns, phi0, b0, kx, kz, phiz0;
};
static const int DUAL_NDOUBLE = (int)((offsetof(st_element, phiz0) - offsetof(st_element, l)) / sizeof(double)) + 1;
static_assert((offsetof(dual_element, phiz0) - offsetof(dual_element, l)) / sizeof(Dual) + 1 ==
                  (offsetof(st_element, phiz0) - offsetof(st_element, l)) / sizeof(double) + 1,
              "dual_element must mirror st_element");

static void to_dual(const st_element *e, dual_element *o)
{
    memcpy(&o->type, &e->type, 8 * sizeof(int32_t));
    const double *src = &e->l;
    Dual *dst = &o->l;
    for (int i = 0; i < DUAL_NDOUBLE; i++) dst[i] = Dual(src[i]);
}

#define ORC_VALUE(f) ((f).v)
#define ORC_SCALAR_BUILD 1   /* st_oracle.c: leave out the exports that take plain double pointers */
#define st_element dual_element
#define double Dual
#include "st_oracle.c"
#undef double
#undef st_element

/* x0 plus the identity as derivatives */
static void seed(Dual x[6], const double x0[6])
{
    for (int i = 0; i < 6; i++) {
        x[i] = Dual(x0[i]);
        x[i].d[i] = 1.;
    }
}

/*
 * The one-turn matrix of Cal_OneTurnMap (its linear part) about x_co, with
 * the RF cavities' voltages off as the reference's map is taken: m[6 i + k]
 * = d x_i(after one turn) / d x_k.  Returns the maps' rc.
 */
extern "C" __attribute__((visibility("default"))) int dual_one_turn_matrix(const st_element *line, int n,
                                                                              const st_globals *gp,
                                                                              const double x_co[6], double m[36])
{
    orc_set_globals(gp);
    dual_element e;
    Dual x[6];
    seed(x, x_co);
    orc_rng rng = {gp->rng_seed, 0u, 0u};
    for (int j = 0; j < n; j++) {
        to_dual(&line[j], &e);
        if (e.type == ST_RFCAV) e.vrf = Dual(0.);
        int rc = elem_pass(&e, x, &rng);
        if (rc) return rc;
    }
    for (int i = 0; i < 6; i++)
        for (int k = 0; k < 6; k++) m[6 * i + k] = x[i].d[k];
    return 0;
}

/*
 * Every element's matrix about the orbit X (n x 6, the orbit after each
 * element, as st_element_maps returns it): element k is expanded about the
 * orbit after element k - 1 (the last for k = 0) with z = 0 and x[5] = x5,
 * RF voltages off, as Cal_ElementMap (cpp:15205-15253).  T: n x 36.
 */
extern "C" __attribute__((visibility("default"))) int dual_element_maps(const st_element *line, int n,
                                                                           const st_globals *gp, const double *X,
                                                                           double x5, double *T)
{
    orc_set_globals(gp);
    dual_element e;
    for (int k = 0; k < n; k++) {
        const double *xo = X + 6 * (k > 0 ? k - 1 : n - 1);
        double x0[6];
        for (int j = 0; j < 6; j++) x0[j] = xo[j];
        x0[4] = 0.;
        x0[5] = x5;
        Dual x[6];
        seed(x, x0);
        to_dual(&line[k], &e);
        if (e.type == ST_RFCAV) e.vrf = Dual(0.);
        orc_rng rng = {gp->rng_seed, 0u, 0u};
        int rc = elem_pass(&e, x, &rng);
        if (rc) return rc;
        for (int i = 0; i < 6; i++)
            for (int a = 0; a < 6; a++) T[36 * (size_t)k + 6 * i + a] = x[i].d[a];
    }
    return 0;
}
