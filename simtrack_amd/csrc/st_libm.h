/*
 * st_libm.h -- project-owned FP64 elementary functions for the tracking path.
 *
 * The reference calls the host C library for sin, cos, asin, exp and pow
 * inside its element maps (simtrack.cpp:3854 asin in sbend_exact_pass,
 * :6009 sin in RFCAV_Pass, :6021-6022 CRABRF_Pass, :6040-6041 CCMULT_Pass,
 * :6273 pow and :6313-6314 exp/cos/sin in cerrf, :6345 and :6647 exp in
 * BB4D/BB6D, :4612 cos/sin in spin()).  Device and host libraries differ by
 * an ulp now and then, and one ulp is enough to change the loss turn of a
 * chaotic particle.  This header therefore implements those functions once,
 * from published methods, in a subset of C that both gcc (the CPU oracle)
 * and hipcc (the gfx950 kernels) compile to the same IEEE operation
 * sequence: only +, -, *, /, sqrt, explicit fma and integer arithmetic,
 * with floating-point contraction disabled on both sides.
 *
 * Methods:
 *   sin/cos  argument reduction x - k*pi/2 by Cody-Waite with a four-part
 *            pi/2 and FMA for |x| < 2^20, Payne-Hanek with an exact
 *            integer product against the bits of 2/pi beyond; kernels are
 *            the fdlibm-style compensated forms with Taylor coefficients
 *            (truncation < 2^-62 relative on |r| <= pi/4).
 *   asin     Maclaurin series of asin for |x| <= 1/2 and the half-angle
 *            identity asin(x) = pi/2 - 2 asin(sqrt((1-x)/2)) above, with a
 *            fma-exact square-root residual.
 *   exp      correctly rounded: Tang's table-driven reduction with Ziv's
 *            two-phase rounding test (details at st_exp).
 *   cosh, sinh  from exp: (e^x +- e^-x)/2, sinh below |x| = 1/2 by its
 *            Maclaurin series (no cancellation); within 2 ulp.
 *   powi     h^n for integer n by binary powering in double-double
 *            arithmetic (exact products via fma) and one final rounding.
 * Accuracy (tested against mpmath at 50 digits): < 1 ulp (cosh, sinh and
 * tan < 2 ulp); exp is correctly rounded.
 *
 * The includer defines, before including:
 *   ST_FN            function qualifiers ("static inline" for C,
 *                    "__device__ __forceinline__" for HIP)
 *   ST_CONST_TABLE   qualifiers for the constant tables
 *   ST_FMA(a,b,c)    exactly rounded fused multiply-add
 *   ST_RINT(x)       round to nearest integer, ties to even
 *   ST_SQRT(x)       correctly rounded square root
 */
#ifndef ST_LIBM_H
#define ST_LIBM_H

#include <stdint.h>
#include "st_libm_tables.h"

#ifndef ST_FMA_K
#define ST_FMA_K(a, b, c) ST_FMA(a, b, c)
#endif
#if !defined(ST_FN) || !defined(ST_CONST_TABLE) || !defined(ST_FMA) || \
    !defined(ST_RINT) || !defined(ST_SQRT)
#error "define ST_FN, ST_CONST_TABLE, ST_FMA, ST_RINT, ST_SQRT before including st_libm.h"
#endif
/* functions on rare paths (huge trigonometric arguments); a GPU build makes
   them out-of-line so the hot code stays small */
#ifndef ST_FN_COLD
#define ST_FN_COLD ST_FN
#endif

ST_CONST_TABLE uint32_t st_two_over_pi_words[ST_TWO_OVER_PI_NWORDS] = ST_TWO_OVER_PI_WORDS;

/* ---------------------------------------------------------------- bits */

ST_FN uint64_t st_d2u(double d) { uint64_t u; __builtin_memcpy(&u, &d, 8); return u; }
ST_FN double st_u2d(uint64_t u) { double d; __builtin_memcpy(&d, &u, 8); return d; }
ST_FN double st_fabs(double x) { return st_u2d(st_d2u(x) & 0x7fffffffffffffffull); }
ST_FN int st_isnan(double x) { return x != x; }

/* 2^k for -1022 <= k <= 1023, exact */
ST_FN double st_pow2i(int k) { return st_u2d((uint64_t)(k + 1023) << 52); }

/* ------------------------------------------------------ double-double */

typedef struct { double hi, lo; } st_dd;

ST_FN st_dd st_dd_make(double hi, double lo)
{
    st_dd r;
    r.hi = hi;
    r.lo = lo;
    return r;
}

ST_FN st_dd st_two_sum(double a, double b)
{
    st_dd r;
    double s = a + b;
    double bb = s - a;
    r.hi = s;
    r.lo = (a - (s - bb)) + (b - bb);
    return r;
}

ST_FN st_dd st_fast_two_sum(double a, double b) /* |a| >= |b| */
{
    st_dd r;
    double s = a + b;
    r.hi = s;
    r.lo = b - (s - a);
    return r;
}

ST_FN st_dd st_two_prod(double a, double b)
{
    st_dd r;
    double p = a * b;
    r.hi = p;
    r.lo = ST_FMA(a, b, -p);
    return r;
}

ST_FN st_dd st_dd_add(st_dd a, st_dd b)
{
    st_dd s = st_two_sum(a.hi, b.hi);
    double e = s.lo + (a.lo + b.lo);
    return st_fast_two_sum(s.hi, e);
}

ST_FN st_dd st_dd_mul(st_dd a, st_dd b)
{
    st_dd p = st_two_prod(a.hi, b.hi);
    double e = p.lo + (a.hi * b.lo + a.lo * b.hi);
    return st_fast_two_sum(p.hi, e);
}

/* ---------------------------------------------------- sin/cos kernels */

/* sin(x+y), |x| <= pi/4, |y| <= ulp(x)/2 */
ST_FN double st_ksin(double x, double y, int has_y)
{
    double z = x * x;
    double v = z * x;
    double r = ST_FMA_K(z, ST_SIN_8, ST_SIN_7);
    r = ST_FMA_K(z, r, ST_SIN_6);
    r = ST_FMA_K(z, r, ST_SIN_5);
    r = ST_FMA_K(z, r, ST_SIN_4);
    r = ST_FMA_K(z, r, ST_SIN_3);
    r = ST_FMA_K(z, r, ST_SIN_2);
    if (!has_y)
        return x + v * ST_FMA_K(z, r, ST_SIN_1);
    return x - ((z * (0.5 * y - v * r) - y) - v * ST_SIN_1);
}

/* cos(x+y), |x| <= pi/4 */
ST_FN double st_kcos(double x, double y)
{
    double z = x * x;
    double r = ST_FMA_K(z, ST_COS_8, ST_COS_7);
    r = ST_FMA_K(z, r, ST_COS_6);
    r = ST_FMA_K(z, r, ST_COS_5);
    r = ST_FMA_K(z, r, ST_COS_4);
    r = ST_FMA_K(z, r, ST_COS_3);
    r = ST_FMA_K(z, r, ST_COS_2);
    r = ST_FMA_K(z, r, ST_COS_1);
    r = z * r;
    double hz = 0.5 * z;
    double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * r - x * y));
}

/* ------------------------------------------------ argument reduction */

/* Payne-Hanek: |x| >= 2^20, finite.  Returns the quadrant q (mod 4) and
   r = x - q*pi/2 as a double-double.  With x = m*2^e (m a 53-bit integer),
   the bits of 2/pi of weight above 2^(2-e) only add multiples of 4 to
   x*2/pi, so a 192-bit window of 2/pi starting at fraction bit e-1 gives
   x*2/pi mod 4 to 2^-137 absolute. */
ST_FN_COLD int st_rem_pio2_large(double ax, st_dd *r)
{
    uint64_t bits = st_d2u(ax);
    int e = (int)((bits >> 52) & 0x7ff) - 1075;       /* ax = m * 2^e, e >= -32 */
    uint64_t m = (bits & 0x000fffffffffffffull) | 0x0010000000000000ull;
    /* fraction bit i (weight 2^-i) sits at table bit offset i-1+64 counted
       from the msb of word 0 (two zero pad words cover i <= 0) */
    int off = e + 62;                                 /* offset of bit e-1 */
    int j0 = off >> 5, s = off & 31;
    uint64_t wv[3];
    for (int k = 0; k < 3; ++k) {
        uint64_t hi = ((uint64_t)st_two_over_pi_words[j0 + 2 * k] << 32) |
                      (uint64_t)st_two_over_pi_words[j0 + 2 * k + 1];
        uint64_t nx = (uint64_t)st_two_over_pi_words[j0 + 2 * k + 2];
        wv[k] = s ? ((hi << s) | (nx >> (32 - s))) : hi;
    }
    /* W = wv[0]*2^128 + wv[1]*2^64 + wv[2];  P = m*W mod 2^192 */
    unsigned __int128 t2 = (unsigned __int128)m * wv[2];
    unsigned __int128 t1 = (unsigned __int128)m * wv[1];
    unsigned __int128 t0 = (unsigned __int128)m * wv[0];
    uint64_t p0 = (uint64_t)t2;
    unsigned __int128 c1 = (t2 >> 64) + (unsigned __int128)(uint64_t)t1;
    uint64_t p1 = (uint64_t)c1;
    uint64_t p2 = (uint64_t)(c1 >> 64) + (uint64_t)(t1 >> 64) + (uint64_t)t0;
    /* x*2/pi mod 4 = P * 2^-190: the quadrant is bits 63..62 of p2 */
    int q = (int)(p2 >> 62);
    uint64_t F = (p2 << 2) | (p1 >> 62);              /* fraction bits 1..64   */
    uint64_t G = (p1 << 2) | (p0 >> 62);              /* fraction bits 65..128 */
    double sign = 1.0;
    if (F >> 63) {                                    /* fraction >= 1/2: f - 1 */
        q = (q + 1) & 3;
        G = ~G + 1u;
        F = ~F + (G == 0u ? 1u : 0u);
        sign = -1.0;
    }
    /* |f| = F*2^-64 + G*2^-128 split into three exact doubles */
    double fa = (double)(F >> 11) * 0x1p-53;
    double fb = (double)(((F & 0x7ffu) << 42) | (G >> 22)) * 0x1p-106;
    double fc = (double)(G & 0x3fffffu) * 0x1p-128;
    st_dd f = st_two_sum(fa, fb);
    f = st_fast_two_sum(f.hi, f.lo + fc);
    st_dd pio2;
    pio2.hi = ST_PIO2_1;
    pio2.lo = ST_PIO2_2;
    st_dd rr = st_dd_mul(f, pio2);
    r->hi = sign * rr.hi;
    r->lo = sign * rr.lo;
    return q;
}

/* r = x - q*pi/2 with |r| <~ pi/4; returns q mod 4.  x finite. */
ST_FN int st_rem_pio2(double x, st_dd *r)
{
    double ax = st_fabs(x);
    if (ax < 0x1p20) {
        double k = ST_RINT(x * ST_TWO_OVER_PI);
        double r1 = ST_FMA(-k, ST_PIO2_1, x);            /* exact */
        st_dd t = st_two_prod(k, ST_PIO2_2);
        st_dd s = st_two_sum(r1, -t.hi);
        s.lo = s.lo - t.lo;
        st_dd u = st_two_prod(k, ST_PIO2_3);
        s.lo = s.lo - u.hi;
        s.lo = s.lo - (u.lo + k * ST_PIO2_4);
        *r = st_fast_two_sum(s.hi, s.lo);
        int64_t ki = (int64_t)k;
        return (int)(ki & 3);
    }
    int q = st_rem_pio2_large(ax, r);
    if (x < 0.0) {
        r->hi = -r->hi;
        r->lo = -r->lo;
        q = (4 - q) & 3;
    }
    return q;
}

/* ------------------------------------------------------------ sin/cos */

ST_FN void st_sincos(double x, double *s, double *c)
{
    double ax = st_fabs(x);
    if (!(ax <= 0x1.fffffffffffffp+1023)) {          /* inf or nan */
        *s = x - x;
        *c = x - x;
        return;
    }
    if (ax <= ST_PIO4) {
        if (ax < 0x1p-27) {
            *s = x;
            *c = 1.0;
            return;
        }
        *s = st_ksin(x, 0.0, 0);
        *c = st_kcos(x, 0.0);
        return;
    }
    st_dd r;
    int q = st_rem_pio2(x, &r);
    double sr = st_ksin(r.hi, r.lo, 1);
    double cr = st_kcos(r.hi, r.lo);
    switch (q) {
    case 0: *s = sr;  *c = cr;  break;
    case 1: *s = cr;  *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

/* The |x| <= pi/4 path of st_sincos without its tiny-argument branch: for
   0 < |x| < 2^-27 the kernels already return sin = x and cos = 1 exactly
   (the cubic term is below a quarter ulp of x, 1 - x^2/2 rounds to 1); at
   x = -0 the sine kernel gives +0, so zero is passed through by a select.
   For every |x| <= pi/4, including subnormals, this equals st_sincos. */
ST_FN void st_sincos_small(double x, double *s, double *c)
{
    double sr = st_ksin(x, 0.0, 0);
    *s = x == 0.0 ? x : sr;
    *c = st_kcos(x, 0.0);
}

ST_FN double st_sin(double x)
{
    double s, c;
    st_sincos(x, &s, &c);
    return s;
}

ST_FN double st_cos(double x)
{
    double s, c;
    st_sincos(x, &s, &c);
    return c;
}

/* --------------------------------------------------------------- asin */

/* Horner evaluation c0 + t*(c1 + t*(...)) of the first 10 and all 27
   series coefficients, written out so the coefficients are immediate
   operands rather than table loads. */
#define ST_ASIN_STEP(p, t, i) p = ST_FMA_K(p, t, ST_ASIN_C##i)
ST_FN double st_asin_poly10(double t)
{
    double p = ST_ASIN_C9;
    ST_ASIN_STEP(p, t, 8); ST_ASIN_STEP(p, t, 7); ST_ASIN_STEP(p, t, 6); ST_ASIN_STEP(p, t, 5);
    ST_ASIN_STEP(p, t, 4); ST_ASIN_STEP(p, t, 3); ST_ASIN_STEP(p, t, 2); ST_ASIN_STEP(p, t, 1);
    ST_ASIN_STEP(p, t, 0);
    return p;
}

ST_FN double st_asin_poly27(double t)
{
    double p = ST_ASIN_C26;
    ST_ASIN_STEP(p, t, 25); ST_ASIN_STEP(p, t, 24); ST_ASIN_STEP(p, t, 23); ST_ASIN_STEP(p, t, 22);
    ST_ASIN_STEP(p, t, 21); ST_ASIN_STEP(p, t, 20); ST_ASIN_STEP(p, t, 19); ST_ASIN_STEP(p, t, 18);
    ST_ASIN_STEP(p, t, 17); ST_ASIN_STEP(p, t, 16); ST_ASIN_STEP(p, t, 15); ST_ASIN_STEP(p, t, 14);
    ST_ASIN_STEP(p, t, 13); ST_ASIN_STEP(p, t, 12); ST_ASIN_STEP(p, t, 11); ST_ASIN_STEP(p, t, 10);
    ST_ASIN_STEP(p, t, 9); ST_ASIN_STEP(p, t, 8); ST_ASIN_STEP(p, t, 7); ST_ASIN_STEP(p, t, 6);
    ST_ASIN_STEP(p, t, 5); ST_ASIN_STEP(p, t, 4); ST_ASIN_STEP(p, t, 3); ST_ASIN_STEP(p, t, 2);
    ST_ASIN_STEP(p, t, 1); ST_ASIN_STEP(p, t, 0);
    return p;
}

/* The |x| < 0.125 path of st_asin without its tiny-argument branch: for
   |x| < 2^-26, x*t*p is below 0.17 * 2^-52 |x|, less than half an ulp of x
   even at a binade boundary, so fma(x*t, p, x) = x; for every |x| < 0.125,
   including +-0 and subnormals, this equals st_asin. */
ST_FN double st_asin_small(double x)
{
    double t = x * x;
    return ST_FMA(x * t, st_asin_poly10(t), x);
}

ST_FN double st_asin(double x)
{
    double ax = st_fabs(x);
    if (!(ax <= 1.0)) {
        if (st_isnan(x))
            return x + x;
        return (x - x) / (x - x);                    /* |x| > 1: nan */
    }
    if (ax < 0x1p-26)
        return x;
    if (ax <= 0.5) {
        double t = x * x;
        double p = ax < 0.125 ? st_asin_poly10(t) : st_asin_poly27(t);
        return ST_FMA(x * t, p, x);
    }
    /* asin(ax) = pi/2 - 2*asin(s), s = sqrt((1-ax)/2) <= 1/2 */
    double t = (1.0 - ax) * 0.5;                     /* exact */
    double s = ST_SQRT(t);
    double slo = ST_FMA(-s, s, t) / (s + s);          /* sqrt(t) ~ s + slo */
    double p = st_asin_poly27(t);
    st_dd r = st_two_sum(ST_PIO2_1, -(s + s));
    double corr = ST_PIO2_2 - 2.0 * (slo + (s * t) * p);
    double res = r.hi + (r.lo + corr);
    return x < 0.0 ? -res : res;
}

/* ---------------------------------------------------------------- exp */

/*
 * exp, correctly rounded to nearest.  Method: Tang's table-driven reduction
 * (ACM TOMS 15 (1989) 144-157) x = (128 m + j) ln2/128 + r, |r| <= ln2/256,
 * exp(x) = 2^m 2^(j/128) exp(r), with Ziv's two-phase rounding test
 * (ACM TOMS 17 (1991) 410-423):
 *   fast phase      r as a double-double, exp(r)-1 to degree 6, 2^(j/128)
 *                   as a double-double; the result yh + yl carries a
 *                   relative error below ST_EXP_EPS1 = 2^-64 (the terms are
 *                   bounded in the comments below).  If the exact value
 *                   cannot lie on the other side of a rounding boundary of
 *                   yh, yh is the correctly rounded result.
 *   accurate phase  (one call in ~2^10) exp(r)-1 to degree 11 by Horner in
 *                   double-double, 2^(j/128) as a triple; relative error
 *                   below ST_EXP_EPS2 = 2^-100, the same test.  An argument
 *                   whose exp lies within 2^-100 of a midpoint between two
 *                   doubles (a measure-2^-47 set) keeps the double-double's
 *                   rounding.
 * Subnormal results are rounded once, on the 2^-1074 grid.
 */
#define ST_EXP_EPS1 0x1p-64
#define ST_EXP_EPS2 0x1p-100
ST_CONST_TABLE double st_exp_tab[384] = ST_EXP_TAB;

/* Round 2^m (yh + yl) to the nearest double, given that the exact value is
   within eps*|yh| of yh + yl (|yl| <= ulp(yh)/2, yh in [0.99, 2)).  Returns 1
   and sets *res when the rounding is decided, 0 when the exact value may lie
   on either side of a rounding boundary. */
ST_FN int st_exp_round(double yh, double yl, int m, double eps, double *res)
{
    double err = eps * yh;
    if (m > -1022 || (m == -1022 && yh >= 1.0)) {
        uint64_t b = st_d2u(yh);
        double u = st_u2d(b & 0x7ff0000000000000ull) * 0x1p-52;       /* ulp(yh) */
        /* half the gap to the neighbour on yl's side; below a power of two
           the gap is half as wide */
        double half = ((b & 0x000fffffffffffffull) == 0 && yl < 0.0) ? 0.25 * u : 0.5 * u;
        if (st_fabs(yl) + err >= half)
            return 0;
        if (m > 1023)
            *res = (yh * 2.0) * st_pow2i(m - 1);                      /* may overflow to inf */
        else
            *res = yh * st_pow2i(m);
        return 1;
    }
    /* subnormal (or zero) result: n = rint(2^(m+1074) (yh + yl)), value n 2^-1074 */
    int s = m + 1074;
    if (s < -60) {                                   /* far below 2^-1075 */
        *res = 0.0;
        return 1;
    }
    double sc = s >= 0 ? st_pow2i(s) : 0x1p-60 * st_pow2i(s + 60);
    double V = yh * sc, Vl = yl * sc;                /* exact scalings */
    double n = ST_RINT(V);
    double d = (V - n) + Vl;
    if (d > 0.5) {
        n += 1.0;
        d -= 1.0;
    } else if (d < -0.5) {
        n -= 1.0;
        d += 1.0;
    }
    if (0.5 - st_fabs(d) <= err * sc)
        return 0;
    *res = n * st_u2d(1);                            /* n * 2^-1074, exact */
    return 1;
}

/* The accurate phase: exp(rh + rl) 2^(j/128) as a double-double with
   relative error below ST_EXP_EPS2, rounded by st_exp_round. */
ST_FN_COLD double st_exp_accurate(double rh, double rl, int j, int m)
{
    /* degrees 7..11 in double: |r^7/7!| < 2^-71.9, so their rounding
       errors are below 2^-124 */
    double hp = ST_FMA(rh, ST_EXP_11, ST_EXP_10);
    hp = ST_FMA(rh, hp, ST_EXP_9);
    hp = ST_FMA(rh, hp, ST_EXP_8);
    hp = ST_FMA(rh, hp, ST_EXP_7);
    st_dd r;
    r.hi = rh;
    r.lo = rl;
    /* A = 1/6! + r*(1/7! + ...), then A = 1/n! + r*A for n = 5..2, 1 + r*A,
       e = r*A: Horner in double-double */
    st_dd A = st_dd_add(st_dd_make(ST_EXPC6_HI, ST_EXPC6_LO), st_dd_mul(r, st_dd_make(hp, 0.0)));
    A = st_dd_add(st_dd_make(ST_EXPC5_HI, ST_EXPC5_LO), st_dd_mul(r, A));
    A = st_dd_add(st_dd_make(ST_EXPC4_HI, ST_EXPC4_LO), st_dd_mul(r, A));
    A = st_dd_add(st_dd_make(ST_EXPC3_HI, ST_EXPC3_LO), st_dd_mul(r, A));
    A = st_dd_add(st_dd_make(ST_EXPC2_HI, ST_EXPC2_LO), st_dd_mul(r, A));
    A = st_dd_add(st_dd_make(1.0, 0.0), st_dd_mul(r, A));
    st_dd e = st_dd_mul(r, A);                       /* exp(r) - 1 */
    st_dd T;
    T.hi = st_exp_tab[3 * j];
    T.lo = st_exp_tab[3 * j + 1];
    double tl = st_exp_tab[3 * j + 2];
    /* T (1 + e) + tl (1 + e) = T + (T e + tl) */
    st_dd p = st_dd_mul(T, e);
    p = st_dd_add(p, st_dd_make(tl, 0.0));
    st_dd y = st_dd_add(T, p);
    double res;
    if (st_exp_round(y.hi, y.lo, m, ST_EXP_EPS2, &res))
        return res;
    /* within 2^-100 of a midpoint: the double-double's own rounding */
    if (m > 1023)
        return ((y.hi + y.lo) * 2.0) * st_pow2i(m - 1);
    if (m >= -1021)
        return (y.hi + y.lo) * st_pow2i(m);
    return ((y.hi + y.lo) * st_pow2i(m + 60)) * 0x1p-60;
}

/* phase = 0: both phases (st_exp); 1: the accurate phase only; 2: the fast
   phase only, nan where it leaves the rounding undecided (phases 1 and 2
   let tests exercise each phase on its own) */
ST_FN double st_exp_phase(double x, int phase)
{
    if (st_isnan(x))
        return x + x;
    if (x > 709.79)
        return 0x1p1023 * 2.0;                       /* +inf: exp(709.79) > 2^1024 */
    if (x < -745.2)
        return 0.0;                                  /* exp(-745.2) < 2^-1075 */
    if (st_fabs(x) < 0x1p-54)
        return 1.0 + x;                              /* exp(x) within a quarter ulp of 1 */
    double kd = ST_RINT(x * ST_EXP_INVL);            /* |kd| <= 137612 < 2^18 */
    int k = (int)kd;
    int j = k & 127;
    int m = (k - j) / 128;
    /* r = x - kd ln2/128 as rh + rl.  x - kd*L1 is exact: kd*L1 has at most
       53 significant bits, and both terms are multiples of 2^-61 (or kd = 0)
       with a difference below 2^-8. */
    double r1 = x - kd * ST_EXP_L1;
    st_dd p2 = st_two_prod(kd, ST_EXP_L2);          /* exact */
    st_dd s = st_two_sum(r1, -p2.hi);
    double lo = (s.lo - p2.lo) - kd * ST_EXP_L3;     /* error < 2^-112 */
    st_dd r = st_fast_two_sum(s.hi, lo);
    double rh = r.hi, rl = r.lo;                     /* |rh| < 2^-8.52 */
    if (phase == 1)
        return st_exp_accurate(rh, rl, j, m);
    /* q = rh^2/2 + ... + rh^6/720: |q| < 2^-18, error < 2^-69.4; the
       omitted r^7/7! is below 2^-71.9 */
    double q = ST_FMA(rh, ST_EXP_6, ST_EXP_5);
    q = ST_FMA(rh, q, ST_EXP_4);
    q = ST_FMA(rh, q, ST_EXP_3);
    q = ST_FMA(rh, q, ST_EXP_2);
    q = (rh * rh) * q;
    /* exp(r) - 1 = rh + sl, sl = rl + rh*rl + q (rl^2/2 < 2^-124 omitted) */
    double sl = rl + (rh * rl + q);
    double th = st_exp_tab[3 * j], tm = st_exp_tab[3 * j + 1];
    /* (th + tm)(1 + rh + sl) = th + th*rh + [th*sl + tm + tm*(rh + sl)] */
    st_dd ph = st_two_prod(th, rh);                  /* exact */
    st_dd hs = st_fast_two_sum(th, ph.hi);           /* exact, th >= 1 > |ph| */
    double small = hs.lo + (ph.lo + (th * sl + (tm + tm * rh)));   /* |small| < 2^-16.9, error < 2^-66.3 */
    st_dd y = st_fast_two_sum(hs.hi, small);
    double res;
    if (st_exp_round(y.hi, y.lo, m, ST_EXP_EPS1, &res))
        return res;
    if (phase == 2)
        return x - x + (x - x) / (x - x);            /* nan: undecided */
    return st_exp_accurate(rh, rl, j, m);
}

ST_FN double st_exp(double x) { return st_exp_phase(x, 0); }

/* ---------------------------------------------------------- cosh, sinh */

/* cosh(x) = (e^x + e^-x)/2 from the correctly rounded exp (within 2 ulp);
   sinh the same for |x| >= 1/2, and below its Maclaurin series to x^17
   (truncation < 2^-60 relative), where the difference of exponentials
   would cancel.  The wiggler's field expansion calls them. */
#define ST_SINH_C(k) (1.0 / (k))
ST_FN double st_cosh(double x)
{
    double e = st_exp(__builtin_fabs(x));
    return 0.5 * (e + 1.0 / e);
}
ST_FN double st_sinh(double x)
{
    double a = __builtin_fabs(x);
    double r;
    if (a < 0.5) {
        double t = x * x;
        double p = ST_SINH_C(355687428096000.);        /* 1/17! */
        p = ST_FMA(p, t, ST_SINH_C(1307674368000.));   /* 1/15! */
        p = ST_FMA(p, t, ST_SINH_C(6227020800.));      /* 1/13! */
        p = ST_FMA(p, t, ST_SINH_C(39916800.));        /* 1/11! */
        p = ST_FMA(p, t, ST_SINH_C(362880.));          /* 1/9!  */
        p = ST_FMA(p, t, ST_SINH_C(5040.));            /* 1/7!  */
        p = ST_FMA(p, t, ST_SINH_C(120.));             /* 1/5!  */
        p = ST_FMA(p, t, ST_SINH_C(6.));               /* 1/3!  */
        return ST_FMA(x * t, p, x);
    }
    double e = st_exp(a);
    r = 0.5 * (e - 1.0 / e);
    return x < 0 ? -r : r;
}
/* Both from one exponential: the same bits as st_cosh(x) and st_sinh(x). */
ST_FN void st_coshsinh(double x, double *ch, double *sh)
{
    double a = __builtin_fabs(x);
    double e = st_exp(a), ie = 1.0 / e;
    *ch = 0.5 * (e + ie);
    if (a < 0.5) {
        *sh = st_sinh(x);
    } else {
        double r = 0.5 * (e - ie);
        *sh = x < 0 ? -r : r;
    }
}

/* --------------------------------------------------------------- powi */

/* h^n for integer n (the reference calls pow(h, 1-nc) with integer 1-nc). */
ST_FN double st_powi(double h, int n)
{
    if (n == 0)
        return 1.0;
    unsigned int e = (unsigned int)(n < 0 ? -n : n);
    st_dd base;
    base.hi = h;
    base.lo = 0.0;
    st_dd acc;
    acc.hi = 1.0;
    acc.lo = 0.0;
    while (e) {
        if (e & 1u)
            acc = st_dd_mul(acc, base);
        e >>= 1;
        if (e)
            base = st_dd_mul(base, base);
    }
    double ah = st_fabs(acc.hi);
    if (!(ah <= 0x1.fffffffffffffp+1023) || ah < 0x1p-969) {
        /* overflow, underflow or nan: plain products decide */
        double v = 1.0, b = h;
        e = (unsigned int)(n < 0 ? -n : n);
        while (e) {
            if (e & 1u)
                v *= b;
            e >>= 1;
            if (e)
                b *= b;
        }
        return n < 0 ? 1.0 / v : v;
    }
    if (n > 0)
        return acc.hi + acc.lo;
    /* 1 / (acc.hi + acc.lo) */
    double q = 1.0 / acc.hi;
    double r = ST_FMA(-q, acc.hi, 1.0) - q * acc.lo;
    return ST_FMA(r, q, q);
}

/* ---------------------------------------------------------------- log */

ST_FN double st_log(double x)
{
    if (!(x > 0.0)) {
        if (x == 0.0)
            return -(0x1p1023 * 2.0);                 /* -inf */
        return (x - x) / (x - x);                     /* x < 0 or nan: nan */
    }
    if (x > 0x1.fffffffffffffp+1023)
        return x;                                     /* +inf */
    uint64_t u = st_d2u(x);
    int k = 0;
    if (u < 0x0010000000000000ull) {                  /* subnormal */
        x *= 0x1p54;
        u = st_d2u(x);
        k = -54;
    }
    k += (int)(u >> 52) - 1023;
    double m = st_u2d((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 2.0 * ST_SQRT_HALF) {                     /* m in [sqrt(2)/2, sqrt(2)) */
        m *= 0.5;
        k += 1;
    }
    double f = m - 1.0;                               /* exact */
    double s = f / (2.0 + f);
    double z = s * s;
    double R = ST_FMA(z, ST_LOG_14, ST_LOG_13);
    R = ST_FMA(z, R, ST_LOG_12);
    R = ST_FMA(z, R, ST_LOG_11);
    R = ST_FMA(z, R, ST_LOG_10);
    R = ST_FMA(z, R, ST_LOG_9);
    R = ST_FMA(z, R, ST_LOG_8);
    R = ST_FMA(z, R, ST_LOG_7);
    R = ST_FMA(z, R, ST_LOG_6);
    R = ST_FMA(z, R, ST_LOG_5);
    R = ST_FMA(z, R, ST_LOG_4);
    R = ST_FMA(z, R, ST_LOG_3);
    R = ST_FMA(z, R, ST_LOG_2);
    R = ST_FMA(z, R, ST_LOG_1);
    R = z * R;
    double hfsq = 0.5 * f * f;
    /* log(1+f) = 2 atanh(s) = f - (hfsq - s*(hfsq + R)) */
    if (k == 0)
        return f - (hfsq - s * (hfsq + R));
    double dk = (double)k;
    double t = hfsq - (s * (hfsq + R) + dk * ST_LN2_LO);
    st_dd p = st_two_prod(dk, ST_LN2_HI);             /* exact k*ln2_hi */
    st_dd a = st_two_sum(p.hi, f);                    /* exact */
    return a.hi + ((a.lo + p.lo) - t);
}

/* --------------------------------------------------------------- atan */

/*
 * atan by the argument reduction of fdlibm's s_atan.c (a published
 * method): |x| in [0, 7/16) directly; [7/16, 11/16) about atan(1/2) with
 * t = (2x - 1)/(2 + x); [11/16, 19/16) about pi/4 with t = (x - 1)/(x + 1);
 * [19/16, 39/16) about atan(3/2) with t = (x - 1.5)/(1 + 1.5x); beyond, about
 * pi/2 with t = -1/x.  atan(t) for |t| < 7/16 is this project's odd Taylor
 * series t + t z P(z), z = t^2, P = sum_{k=1..22} (-1)^k z^(k-1) / (2k+1)
 * (truncation below 4e-18 relative), the anchors as hi + lo pairs.
 */
ST_FN double st_atan(double x)
{
    const double a = st_fabs(x);
    if (!(a <= 0x1.fffffffffffffp+1023)) {
        if (a != a) return x + x;                         /* nan */
        return x > 0 ? 0x1.921fb54442d18p+0 : -0x1.921fb54442d18p+0;
    }
    if (a < 0x1p-27) return x;
    if (a >= 0x1p66) return x > 0 ? 0x1.921fb54442d18p+0 : -0x1.921fb54442d18p+0;
    double hi = 0., lo = 0., t;
    int anchored = 1;
    if (a < 0.4375) {
        t = a;
        anchored = 0;
    } else if (a < 0.6875) {
        t = (2.0 * a - 1.0) / (2.0 + a);
        hi = 0x1.dac670561bb4fp-2;
        lo = 0x1.a2b7f222f65e2p-56;
    } else if (a < 1.1875) {
        t = (a - 1.0) / (a + 1.0);
        hi = 0x1.921fb54442d18p-1;
        lo = 0x1.1a62633145c07p-55;
    } else if (a < 2.4375) {
        t = (a - 1.5) / (1.0 + 1.5 * a);
        hi = 0x1.f730bd281f69bp-1;
        lo = 0x1.007887af0cbbdp-56;
    } else {
        t = -1.0 / a;
        hi = 0x1.921fb54442d18p+0;
        lo = 0x1.1a62633145c07p-54;
    }
    const double z = t * t;
    double P = 0x1.6c16c16c16c17p-6;
    P = ST_FMA(z, P, -0x1.7d05f417d05f4p-6);
    P = ST_FMA(z, P, 0x1.8f9c18f9c18fap-6);
    P = ST_FMA(z, P, -0x1.a41a41a41a41ap-6);
    P = ST_FMA(z, P, 0x1.bacf914c1bad0p-6);
    P = ST_FMA(z, P, -0x1.d41d41d41d41dp-6);
    P = ST_FMA(z, P, 0x1.f07c1f07c1f08p-6);
    P = ST_FMA(z, P, -0x1.0842108421084p-5);
    P = ST_FMA(z, P, 0x1.1a7b9611a7b96p-5);
    P = ST_FMA(z, P, -0x1.2f684bda12f68p-5);
    P = ST_FMA(z, P, 0x1.47ae147ae147bp-5);
    P = ST_FMA(z, P, -0x1.642c8590b2164p-5);
    P = ST_FMA(z, P, 0x1.8618618618618p-5);
    P = ST_FMA(z, P, -0x1.af286bca1af28p-5);
    P = ST_FMA(z, P, 0x1.e1e1e1e1e1e1ep-5);
    P = ST_FMA(z, P, -0x1.1111111111111p-4);
    P = ST_FMA(z, P, 0x1.3b13b13b13b14p-4);
    P = ST_FMA(z, P, -0x1.745d1745d1746p-4);
    P = ST_FMA(z, P, 0x1.c71c71c71c71cp-4);
    P = ST_FMA(z, P, -0x1.2492492492492p-3);
    P = ST_FMA(z, P, 0x1.999999999999ap-3);
    P = ST_FMA(z, P, -0x1.5555555555555p-2);
    const double r = t * z * P;
    const double res = anchored ? hi + (t + (r + lo)) : t + r;
    return x < 0 ? -res : res;
}

/* ---------------------------------------------------------------- tan */

/* tan = sin/cos of one reduction (element edge angles, host-side constants) */
ST_FN double st_tan(double x)
{
    double s, c;
    st_sincos(x, &s, &c);
    if (st_fabs(x) < 0x1p-27)
        return x;
    return s / c;
}

/* ------------------------------------------------ Gaussian quantile */

/*
 * Inverse of the standard normal CDF.  The reference obtains beam-beam
 * slice quantiles from GSL's gsl_cdf_ugaussian_Pinv (simtrack.cpp:6614,
 * 6958; GSL not vendored, version unpinned).  GSL implements Wichura's
 * algorithm AS 241 (PPND16, Applied Statistics 37 (1988) 477-484):
 * rational approximations in q = p - 1/2 for |q| <= 0.425, and in
 * r = sqrt(-log(min(p,1-p))) for r <= 5 and beyond, with numerator and
 * denominator each evaluated by Horner's rule (no fma) and one division.
 */
ST_FN double st_ratpoly8(const double *a, const double *b, double x)
{
    double u = a[7], v = b[7];
    for (int i = 7; i > 0; --i)
        u = x * u + a[i - 1];
    for (int j = 7; j > 0; --j)
        v = x * v + b[j - 1];
    return u / v;
}

ST_FN double st_ugaussian_pinv(double P)
{
    const double a[8] = {3.387132872796366608, 133.14166789178437745,
                         1971.5909503065514427, 13731.693765509461125,
                         45921.953931549871457, 67265.770927008700853,
                         33430.575583588128105, 2509.0809287301226727};
    const double b[8] = {1.0, 42.313330701600911252, 687.1870074920579083,
                         5394.1960214247511077, 21213.794301586595867,
                         39307.89580009271061, 28729.085735721942674,
                         5226.495278852545925};
    const double c[8] = {1.42343711074968357734, 4.6303378461565452959,
                         5.7694972214606914055, 3.64784832476320460504,
                         1.27045825245236838258, 0.24178072517745061177,
                         0.0227238449892691845833, 7.7454501427834140764e-4};
    const double d[8] = {1.0, 2.05319162663775882187, 1.6763848301838038494,
                         0.68976733498510000455, 0.14810397642748007459,
                         0.0151986665636164571966, 5.475938084995344946e-4,
                         1.05075007164441684324e-9};
    const double e[8] = {6.6579046435011037772, 5.4637849111641143699,
                         1.7848265399172913358, 0.29656057182850489123,
                         0.026532189526576123093, 0.0012426609473880784386,
                         2.71155556874348757815e-5, 2.01033439929228813265e-7};
    const double f[8] = {1.0, 0.59983220655588793769, 0.13692988092273580531,
                         0.0148753612908506148525, 7.868691311456132591e-4,
                         1.8463183175100546818e-5, 1.4215117583164458887e-7,
                         2.04426310338993978564e-15};
    double dP = P - 0.5;
    if (P == 1.0)
        return 0x1p1023 * 2.0;
    if (P == 0.0)
        return -(0x1p1023 * 2.0);
    if (st_fabs(dP) <= 0.425) {
        double r = 0.180625 - dP * dP;
        return dP * st_ratpoly8(a, b, r);
    }
    double pp = (P < 0.5) ? P : 1.0 - P;
    double r = ST_SQRT(-st_log(pp));
    double x = (r <= 5.0) ? st_ratpoly8(c, d, r - 1.6) : st_ratpoly8(e, f, r - 5.0);
    return (P < 0.5) ? -x : x;
}

#endif /* ST_LIBM_H */
