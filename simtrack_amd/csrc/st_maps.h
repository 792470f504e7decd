/*
 * st_maps.h -- element maps of the tracking path, gfx950 device code.
 *
 * One thread owns one particle; its state (x, px, y, py, z, pt[, Sx, Sy,
 * Ss]) lives in registers for the whole launch.  Element constants come from
 * the packed table (st_table_format.h) through scalar loads: every lane of a
 * wavefront works on the same element at the same time.
 *
 * Each map performs the reference's arithmetic operation for operation
 * (same expression trees and evaluation order; the library is compiled with
 * -ffp-contract=off), so results equal the reference's CPU maps bit for bit
 * given the same elementary functions (st_libm.h).  What changes is where
 * constants come from: values the reference recomputes per call from
 * element parameters are read precomputed from the table, and the
 * pt-derived quantities delta1, gamma1, beta1 (recomputed by every DRIFT_Pass,
 * simtrack.cpp:3755-3757) are cached per particle and refreshed only after a
 * map that changes pt.
 */
#ifndef ST_MAPS_H
#define ST_MAPS_H

#include "st_table_format.h"
#include "st_fastdiv.h"

#define ST_DEV __device__ __forceinline__

typedef double st_real;

/* Element payload pointer.  The packed table is read-only for the whole
   launch, so it is addressed in the constant address space: reads at
   wave-uniform offsets compile to scalar loads into SGPRs (one per wave, no
   VGPRs, scalar-cache latency) instead of per-lane vector loads. */
#if defined(__HIP_DEVICE_COMPILE__)
#define ST_KCONST const __attribute__((address_space(4)))
#else
#define ST_KCONST const /* host pass: kernel bodies are parsed only */
#endif
typedef ST_KCONST double *st_pp;
typedef ST_KCONST st_op *st_opp;
/* the reference-particle globals, read through a constant-address reference
   (scalar loads where a map uses them, instead of SGPRs held for the whole
   element loop) */
typedef ST_KCONST st_kglobals &st_gref;
typedef ST_KCONST st_kglobals *st_gp;

/* The element's payload base pointer, made opaque to the optimiser.  Every
   P[k] then becomes a scalar load at an immediate offset from one SGPR pair;
   without it the compiler hoists par + 8k for every k used by any element
   type out of the element loop, and those loop-invariant addresses spill. */
ST_DEV st_pp st_launder(st_pp p)
{
    __asm__ volatile("" : "+s"(p));
    return p;
}

/* the same for the globals: laundered once per element, their fields are
   loaded where the element's map reads them and not kept live across the
   element loop (where they spilled SGPRs into VGPR lanes) */
ST_DEV st_gp st_launder_g(st_gp p)
{
    __asm__ volatile("" : "+s"(p));
    return p;
}

/* particle state: x[0..5] orbit, x[6..8] spin */
struct St {
    double x[9];
};

/* per-particle cache of quantities that depend only on pt */
struct Ptc {
    double d1;     /* delta1 = sqrt(1 + 2 pt/beta + pt^2) - 1   (cpp:3755)   */
    double opd;    /* 1 + delta1                                             */
    double opd2;   /* (1 + delta1)*(1 + delta1)                              */
    double g1;     /* gamma1 = gamma + sqrt(gamma^2 - 1) pt      (cpp:3756)  */
    double b1;     /* beta1 = sqrt(1 - 1/gamma1/gamma1)          (cpp:3757)  */
    StRcp ropd;    /* 1 + delta1 and beta1 prepared as divisors              */
    StRcp rb1;
};

ST_DEV void ptc_update(Ptc &c, double pt, st_gref g)
{
    c.d1 = st_fsqrt(1.0 + 2 * pt / g.beta + pt * pt) - 1.0;
    c.g1 = g.gamma + g.sg * pt;
    c.b1 = st_fsqrt(1.0 - 1.0 / c.g1 / c.g1);
    c.opd = 1.0 + c.d1;
    c.opd2 = c.opd * c.opd;
    c.ropd = st_rcp_prep(c.opd);
    c.rb1 = st_rcp_prep(c.b1);
}

/* ------------------------------------------------ arithmetic policies */

/*
 * A map step is written once, templated on how it divides, takes square
 * roots, asin and sin/cos.  ArFast evaluates the common case without a
 * single branch (shared reciprocals, guard-free sqrt, the small-argument
 * libm paths) and records whether every precondition of those shortcuts
 * held; ArExact uses the operators and the full functions.  Where every
 * precondition holds the two give identical bits, so a caller runs ArFast,
 * and replays the step with ArExact from the saved input state in the
 * (rare) lanes where bad() is true.  A step is thereby one basic block that
 * the scheduler can interleave freely -- the point on a GPU running one
 * wavefront per SIMD, whose speed is set by instruction-level parallelism.
 *
 * The preconditions are tracked cheaply, tested once per step.  The
 * compiler's division sequence (st_fastdiv.h) reduces to y = refined 1/b,
 * q0 = a*y, r = fma(-b, q0, a), q = fma(r, y, q0) unless an exponent is
 * extreme: b, 1/b or a/b subnormal, exp(a) - exp(b) >= 768, |a| < 2^-969,
 * or a zero, infinite or NaN operand or result; sqrt reduces to its
 * refinement for 2^-767 <= x < inf.  One window W = [2^-262, 2^263) serves
 * every check: with |q0| (within 2 ulp of |a/b|) and |b| in W, |a| lies in
 * [2^-525, 2^526), so none of those conditions can hold; a divisor prepared
 * before the step (StRcp::ok, 2^-60 < |b| < 2^60) is inside W as well.  The
 * window is far wider than any coordinate, quotient or sqrt argument of a
 * particle that is still in the machine, so it only sends zeros and
 * diverging lanes to the exact replay.
 * Membership is a test on the biased exponent E, 761 <= E < 1286: the high
 * word shifted left by one (sign dropped) minus 761 << 21 is below
 * (1286 - 761) << 21 as an unsigned number exactly then, and smaller
 * exponents wrap around to large values.  So one 32-bit running maximum
 * covers both ends and every kind of value (one shift-add per value, one
 * three-way max per two values), where FP64 min/max took two instructions
 * per value and a divisor's range two compares.  Zero, subnormal, infinite
 * and NaN values fall outside the window; the sign of a general sqrt
 * argument is kept apart, as the OR of its high words.
 */
#ifdef ST_PERF_MODE
/* Perf mode: a division by a prepared divisor is a * (refined 1/b), within an
   ulp or so of a / b; the square root skips the range bookkeeping.  Only the
   asin/sincos argument ranges are still checked (their small-argument forms
   are inaccurate beyond), and a lane outside them replays the step with the
   full functions. */
struct ArFast {
    bool ok = true;
    ST_DEV bool bad() const { return !ok; }
    ST_DEV StRcp rcp(double b) { return st_rcp_y(b); }
    ST_DEV double div(double a, const StRcp &r) { return a * r.y; }
    ST_DEV double sqrt(double x)
    {
        double y = __builtin_amdgcn_rsq(x);
        double g = x * y;
        double h = y * 0.5;
        double r = __builtin_fma(-h, g, 0.5);
        g = __builtin_fma(g, r, g);
        h = __builtin_fma(h, r, h);
        double d = __builtin_fma(-g, g, x);
        g = __builtin_fma(d, h, g);
        return x > 0.0 && x < __builtin_inf() ? g : __builtin_sqrt(x);
    }
    ST_DEV double sqrt_sq(double x) { return sqrt(x); }
    ST_DEV double asin(double x)
    {
        ok = ok & (__builtin_fabs(x) < 0.125);
        return st_asin_small(x);
    }
    ST_DEV void sincos(double x, double *sn, double *cs)
    {
        ok = ok & (__builtin_fabs(x) <= ST_PIO4);
        st_sincos_small(x, sn, cs);
    }
};
#else
#define ST_WIN_BASE 0x5F200000u  /* 761 << 21: 2^-262 */
#define ST_WIN_SPAN 0x41A00000u  /* (1286 - 761) << 21: 2^263 excluded */
struct ArFast {
    unsigned int w = 0;     /* max of (high word << 1) - ST_WIN_BASE over the checked magnitudes */
    unsigned int sg = 0;    /* OR of the high words of general sqrt arguments: bit 31 = a negative one */
    bool ok = true;         /* every divisor in range, every asin/sincos argument small */
    ST_DEV bool bad() const { return !(ok && w < ST_WIN_SPAN && (int)sg >= 0); }
    ST_DEV void mag(double v)
    {
        /* the high word taken as a vector element: one v_lshl_add_u32 (from
           a 64-bit shift the compiler would shift both words and mask) */
        typedef unsigned int st_u2 __attribute__((ext_vector_type(2)));
        unsigned int t = (__builtin_bit_cast(st_u2, v).y << 1) - ST_WIN_BASE;
        w = t > w ? t : w;
    }
    /* a divisor prepared inside the step: its range joins the window */
    ST_DEV StRcp rcp(double b)
    {
        mag(b);
        return st_rcp_y(b);
    }
    ST_DEV double div(double a, const StRcp &r)
    {
        double q0 = a * r.y;
        double rr = __builtin_fma(-r.b, q0, a);
        double q = __builtin_fma(rr, r.y, q0);
        mag(q0);
        ok = ok & r.ok;
        return q;
    }
    /* sqrt of a sum of squares (never negative or -0; NaN falls outside the window) */
    ST_DEV double sqrt_sq(double x)
    {
        double y = __builtin_amdgcn_rsq(x);
        double g = x * y;
        double h = y * 0.5;
        double r = __builtin_fma(-h, g, 0.5);
        g = __builtin_fma(g, r, g);
        h = __builtin_fma(h, r, h);
        double d = __builtin_fma(-g, g, x);
        g = __builtin_fma(d, h, g);
        d = __builtin_fma(-g, g, x);
        g = __builtin_fma(d, h, g);
        mag(x);
        return g;
    }
    ST_DEV double sqrt(double x)
    {
        typedef unsigned int st_u2 __attribute__((ext_vector_type(2)));
        sg |= __builtin_bit_cast(st_u2, x).y;
        return sqrt_sq(x);
    }
    ST_DEV double asin(double x)
    {
        ok = ok & (__builtin_fabs(x) < 0.125);
        return st_asin_small(x);
    }
    ST_DEV void sincos(double x, double *sn, double *cs)
    {
        ok = ok & (__builtin_fabs(x) <= ST_PIO4);
        st_sincos_small(x, sn, cs);
    }
};

#endif

struct ArExact {
    ST_DEV StRcp rcp(double b) { return st_rcp_y(b); }
    ST_DEV double div(double a, const StRcp &r) { return a / r.b; }
    ST_DEV double sqrt(double x) { return __builtin_sqrt(x); }
    ST_DEV double sqrt_sq(double x) { return __builtin_sqrt(x); }
    ST_DEV double asin(double x) { return st_asin(x); }
    ST_DEV void sincos(double x, double *sn, double *cs) { st_sincos(x, sn, cs); }
};

/* ------------------------------------------------------------ drift */

/* DRIFT_Pass, simtrack.cpp:3743-3766.  Lb = L/GP.beta. */
template <bool HEXP, class A>
ST_DEV void drift_t(A &ar, St &s, double L, double Lb, const Ptc &c)
{
    double *x = s.x;
    if (HEXP) {
        StRcp r5 = ar.rcp(1.0 + x[5]);
        double u = ar.div(L, r5);
        x[0] = x[0] + x[1] * u;
        x[2] = x[2] + x[3] * u;
        x[4] = x[4] - ar.div((x[1] * x[1] + x[3] * x[3]) * u / 2.0, r5);
    } else {
        double pz = ar.sqrt(c.opd2 - x[1] * x[1] - x[3] * x[3]);
        StRcp rz = ar.rcp(pz);
        double betaz = ar.div(pz * c.b1, c.ropd);
        x[0] = x[0] + ar.div(x[1] * L, rz);
        x[2] = x[2] + ar.div(x[3] * L, rz);
        x[4] = x[4] - (L / betaz - Lb);
    }
}

template <bool HEXP>
ST_DEV void drift(St &s, double L, double Lb, const Ptc &c)
{
    double x0 = s.x[0], x2 = s.x[2], x4 = s.x[4];
    ArFast af;
    drift_t<HEXP>(af, s, L, Lb, c);
    if (__builtin_expect(af.bad(), 0)) {
        ST_FD_COLD_PATH();
        s.x[0] = x0;
        s.x[2] = x2;
        s.x[4] = x4;
        ArExact ae;
        drift_t<HEXP>(ae, s, L, Lb, c);
    }
}

/*
 * DRIFT_Pass with L = +0 (markers, zero-length RF halves).  For
 * pz^2 = (1+delta)^2 - px^2 - py^2 > 0 and finite px, py the exact map
 * reduces to x += px*0, y += py*0 (adds a signed zero), z unchanged; any
 * other case takes the full formula so NaN/inf propagate as in the
 * reference.
 */
template <bool HEXP>
ST_DEV void drift0(St &s, const Ptc &c)
{
    double *x = s.x;
    if (HEXP) {
        drift<true>(s, 0.0, 0.0, c);
        return;
    }
    double pz2 = c.opd2 - x[1] * x[1] - x[3] * x[3];
    if (pz2 > 0.0 && pz2 <= 0x1.fffffffffffffp+1023 && c.b1 > 0.0 && c.opd > 0.0) {
        x[0] = x[0] + x[1] * 0.0;
        x[2] = x[2] + x[3] * 0.0;
    } else {
        drift<false>(s, 0.0, 0.0, c);
    }
}

/* ------------------------------------------------------ misalignment */

/* GtoL1 (cpp:3443-3461) and GtoL2 (cpp:3567-3626) */
ST_DEV void mis_in(St &s, int flags, st_pp M)
{
    double *x = s.x;
    if (flags & ST_F_MIS1) {
        if (flags & ST_F_MIS1_OFS) {
            x[0] = x[0] - M[0];
            x[2] = x[2] - M[1];
        }
        if (flags & ST_F_MIS1_ROT) {
            double c = M[2], sn = M[3];
            double t0 = x[0], t1 = x[1], t2 = x[2], t3 = x[3];
            x[0] = c * t0 + sn * t2;
            x[1] = c * t1 + sn * t3;
            x[2] = c * t2 - sn * t0;
            x[3] = c * t3 - sn * t1;
        }
    } else if (flags & ST_F_MIS2) {
        double xt[6];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double a = 0.;
#pragma unroll
            for (int j = 0; j < 6; j++) a = a + M[i * 6 + j] * x[j];
            xt[i] = a;
        }
#pragma unroll
        for (int i = 0; i < 5; i++) x[i] = xt[i] + M[36 + i];
        x[5] = xt[5];
    }
}

/* LtoG1 (cpp:3463-3481) and LtoG2 (cpp:3628-3699) */
ST_DEV void mis_out(St &s, int flags, st_pp M)
{
    double *x = s.x;
    if (flags & ST_F_MIS1) {
        if (flags & ST_F_MIS1_ROT) {
            double c = M[2], sn = M[3];
            double t0 = x[0], t1 = x[1], t2 = x[2], t3 = x[3];
            x[0] = c * t0 - sn * t2;
            x[1] = c * t1 - sn * t3;
            x[2] = sn * t0 + c * t2;
            x[3] = sn * t1 + c * t3;
        }
        if (flags & ST_F_MIS1_OFS) {
            x[0] = x[0] + M[0];
            x[2] = x[2] + M[1];
        }
    } else if (flags & ST_F_MIS2) {
        st_pp R = M + 41;
        double xt[6];
#pragma unroll
        for (int i = 0; i < 5; i++) xt[i] = x[i] + R[36 + i];
        xt[5] = x[5];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double a = 0.;
#pragma unroll
            for (int j = 0; j < 6; j++) a = a + R[i * 6 + j] * xt[j];
            x[i] = a;
        }
    }
}

/* ------------------------------------------------------------ kicks */

/* quad_kick_pass, simtrack.cpp:4046-4058 */
ST_DEV void quad_kick(St &s, double k1l, double k1sl)
{
    double *x = s.x;
    double preal = x[0], pimag = x[2];
    double Bx = k1sl * preal + k1l * pimag;
    double By = k1l * preal - k1sl * pimag;
    x[1] = x[1] - By;
    x[3] = x[3] + Bx;
}

/* sext_kick_pass, simtrack.cpp:4060-4071 */
ST_DEV void sext_kick(St &s, double k2l, double k2sl)
{
    double *x = s.x;
    double preal = x[0] * x[0] - x[2] * x[2];
    double pimag = 2. * x[0] * x[2];
    double Bx = (k2sl * preal + k2l * pimag) / 2.;
    double By = (k2l * preal - k2sl * pimag) / 2.;
    x[1] = x[1] - By;
    x[3] = x[3] + Bx;
}

/* oct_kick_pass, simtrack.cpp:4073-4084 */
ST_DEV void oct_kick(St &s, double k3l, double k3sl)
{
    double *x = s.x;
    double preal = x[0] * x[0] * x[0] - 3 * x[0] * x[2] * x[2];
    double pimag = 3 * x[0] * x[0] * x[2] - x[2] * x[2] * x[2];
    double Bx = (k3sl * preal + k3l * pimag) / 6.;
    double By = (k3l * preal - k3sl * pimag) / 6.;
    x[1] = x[1] - By;
    x[3] = x[3] + Bx;
}

/* n! as double for n = 0..10 (the reference's int fac, converted at the divide) */
__device__ __constant__ const double st_fact[11] = {1., 1., 2., 6., 24., 120., 720., 5040., 40320., 362880.,
                                                    3628800.};
/* 1/n! rounded to nearest.  With the correctly rounded reciprocal y of b,
   q0 = a*y, r = fma(-b, q0, a) (exact) and q = fma(r, y, q0) is the
   correctly rounded a/b whenever nothing underflows or overflows
   (Markstein's residual correction), so ArFast::div with {n!, 1/n!} gives
   the bits of a / n! inside its window on |q0|; outside it the step is
   replayed with the division itself. */
__device__ __constant__ const double st_rfact[11] = {1., 1., 1. / 2., 1. / 6., 1. / 24., 1. / 120., 1. / 720.,
                                                     1. / 5040., 1. / 40320., 1. / 362880., 1. / 3628800.};

/* order I of the field sum: Xn + i Yn = (x + i y)^I, and the order's term
   when its coefficients kni, ksi are not both zero (mask bit I) */
template <int I, class A>
ST_DEV void mult_order(A &ar, const double *x, int mask, double kni, double ksi, double &Xn, double &Yn, double &By,
                       double &Bx)
{
    double Xn0 = Xn, Yn0 = Yn;
    Xn = Xn0 * x[0] - Yn0 * x[2];
    Yn = Xn0 * x[2] + Yn0 * x[0];
    if ((mask >> I) & 1) {
        StRcp rf;
        rf.b = st_fact[I];
        rf.y = st_rfact[I];
        rf.ok = true;
        By = By + ar.div(kni * Xn - ksi * Yn, rf);
        Bx = Bx + ar.div(kni * Yn + ksi * Xn, rf);
    }
}

/*
 * Field sum of mult_kick_pass / bend_mult_kick_pass / cal_Bfield
 * (simtrack.cpp:4086-4111, 4113-4137, 4169-4196): By + i Bx accumulated over
 * orders 1..Norder whose coefficients are not both zero (mask bit i).  The
 * orders (norder <= 10, checked when the table is built) are written out one
 * after the other: n! and 1/n! are constants and the coefficients load at
 * fixed offsets, with no loop counter or pointer arithmetic.
 */
template <class A>
ST_DEV void mult_field_t(A &ar, const St &s, int norder, int mask, st_pp kn, st_pp ks, double &By, double &Bx)
{
    const double *x = s.x;
    double Xn = 1., Yn = 0.;
    By = kn[0];
    Bx = ks[0];
#define ST_ORDER(I)                                                                                                    \
    if (norder < I) return;                                                                                            \
    mult_order<I>(ar, x, mask, kn[I], ks[I], Xn, Yn, By, Bx);
    ST_ORDER(1) ST_ORDER(2) ST_ORDER(3) ST_ORDER(4) ST_ORDER(5)
    ST_ORDER(6) ST_ORDER(7) ST_ORDER(8) ST_ORDER(9) ST_ORDER(10)
#undef ST_ORDER
}

/* the field sum with the factorial divisions as reciprocal steps, replayed
   with the divisions in the lanes outside their window */
ST_DEV void mult_field(const St &s, int norder, int mask, st_pp kn, st_pp ks, double &By, double &Bx)
{
    ArFast af;
    mult_field_t(af, s, norder, mask, kn, ks, By, Bx);
    if (__builtin_expect(af.bad(), 0)) {
        ST_FD_COLD_PATH();
        ArExact ae;
        mult_field_t(ae, s, norder, mask, kn, ks, By, Bx);
    }
}

ST_DEV void mult_kick(St &s, int norder, int mask, st_pp kn, st_pp ks)
{
    double By, Bx;
    mult_field(s, norder, mask, kn, ks, By, Bx);
    s.x[1] = s.x[1] - By;
    s.x[3] = s.x[3] + Bx;
}

/* bend_kick_pass, simtrack.cpp:3790-3794 */
ST_DEV void bend_kick(St &s, double L, double href)
{
    double *x = s.x;
    x[1] = x[1] + (href * x[5] - href * href * x[0]) * L;
    x[4] = x[4] - href * x[0] * L;
}

/* general_bend_kick_pass, simtrack.cpp:3796-3800 */
ST_DEV void general_bend_kick(St &s, double L, double href, double hreal)
{
    double *x = s.x;
    x[1] = x[1] + (-(hreal - href) + href * x[5] - href * hreal * x[0]) * L;
    x[4] = x[4] - href * x[0] * L;
}

/* bend_mult_kick_pass, simtrack.cpp:4113-4137 */
ST_DEV void bend_mult_kick(St &s, double L, double href, int norder, int mask, st_pp kn,
                           st_pp ks)
{
    double By, Bx;
    mult_field(s, norder, mask, kn, ks, By, Bx);
    double *x = s.x;
    x[1] = x[1] - By + (href * x[5] - href * href * x[0]) * L;
    x[3] = x[3] + Bx;
    x[4] = x[4] - href * x[0] * L;
}

/* general_bend_mult_kick_pass, simtrack.cpp:4139-4163 */
ST_DEV void general_bend_mult_kick(St &s, double L, double href, double hreal, int norder, int mask,
                                   st_pp kn, st_pp ks)
{
    double By, Bx;
    mult_field(s, norder, mask, kn, ks, By, Bx);
    double *x = s.x;
    x[1] = x[1] - By + (-(hreal - href) + href * x[5] - href * hreal * x[0]) * L;
    x[3] = x[3] + Bx;
    x[4] = x[4] - href * x[0] * L;
}

/* ------------------------------------------------------------ bends */

/* sbend_exact_pass, simtrack.cpp:3833-3858.  ih = 1.0/href, Lb = L/GP.beta,
   rh = href prepared as a divisor */
template <class A>
ST_DEV void sbend_exact_t(A &ar, St &s, double L, double href, double ih, double cosA, double sinA, double Lb,
                          const Ptc &c, const StRcp &rh)
{
    double *x = s.x;
    double x0 = x[0], x1 = x[1], x3 = x[3];
    double pz0 = ar.sqrt(c.opd2 - x1 * x1 - x3 * x3);
    double nx1 = x1 * cosA + (pz0 - 1. - href * x0) * sinA;
    double pz1 = ar.sqrt(c.opd2 - nx1 * nx1 - x3 * x3);
    double derivative = -1.0 * x1 * sinA * href + (pz0 - 1. - x0 * href) * cosA * href;
    x[1] = nx1;
    x[0] = ar.div(pz1, rh) - ar.div(ar.div(derivative, rh), rh) - ih;
    double temp = ar.sqrt(c.opd2 - x3 * x3);
    StRcp rt = ar.rcp(temp);
    double arcsin1 = ar.asin(ar.div(x1, rt));
    double arcsin2 = ar.asin(ar.div(nx1, rt));
    x[2] = x[2] + x3 * L + ar.div(x3 * (arcsin1 - arcsin2), rh);
    temp = c.opd * L + ar.div(c.opd * (arcsin1 - arcsin2), rh);
    x[4] = x[4] - (ar.div(temp, c.rb1) - Lb);
}

ST_DEV void sbend_exact(St &s, double L, double href, double ih, double cosA, double sinA, double Lb, const Ptc &c,
                        const StRcp &rh)
{
    double x0 = s.x[0], x1 = s.x[1], x2 = s.x[2], x4 = s.x[4];
    ArFast af;
    sbend_exact_t(af, s, L, href, ih, cosA, sinA, Lb, c, rh);
    if (__builtin_expect(af.bad(), 0)) {
        ST_FD_COLD_PATH();
        s.x[0] = x0;
        s.x[1] = x1;
        s.x[2] = x2;
        s.x[4] = x4;
        ArExact ae;
        sbend_exact_t(ae, s, L, href, ih, cosA, sinA, Lb, c, rh);
    }
}

/* general_sbend_exact_pass, simtrack.cpp:3974-4002.
   P: L, rhoc, b0, b0*rhoc, cosA, sinA, Lb */
ST_DEV void general_sbend_exact(St &s, st_pp P, const Ptc &c)
{
    double *x = s.x;
    double L = P[0], rhoc = P[1], b0 = P[2], b0r = P[3], cosA = P[4], sinA = P[5], Lb = P[6];
    double x0 = x[0], x1 = x[1], x3 = x[3];
    double pz0 = st_fsqrt(c.opd2 - x1 * x1 - x3 * x3);
    double nx1 = x1 * cosA + (pz0 - b0r - b0 * x0) * sinA;
    double pz1 = st_fsqrt(c.opd2 - nx1 * nx1 - x3 * x3);
    double derivative = -x1 * sinA / rhoc + (pz0 - b0r - b0 * x0) * cosA / rhoc;
    x[1] = nx1;
    x[0] = pz1 / b0 - derivative * rhoc / b0 - rhoc;
    double temp = st_fsqrt(c.opd2 - x3 * x3);
    double arcsin1 = st_asin(x1 / temp);
    double arcsin2 = st_asin(nx1 / temp);
    x[2] = x[2] + x3 * L / b0 / rhoc + x3 * (arcsin1 - arcsin2) / b0;
    temp = c.opd * L / b0 / rhoc + c.opd * (arcsin1 - arcsin2) / b0;
    x[4] = x[4] - (temp / c.b1 - Lb);
}

/* edge focusing of SBEND_Pass et al. (cpp:4706-4708): ef = tan(E)*h */
ST_DEV void edge(St &s, double ef)
{
    double *x = s.x;
    x[1] = x[1] + ef * x[0];
    x[3] = x[3] - ef * x[2];
}

/* ------------------------------------------------------------- spin */

/*
 * spin(), simtrack.cpp:4574-4630: Thomas-BMT rotation of the spin (x[6..8])
 * for one integration step with integrated field BLbrho.  The reference
 * calls it on the step's saved entry state x0 (e.g. cpp:7571-7575), so the
 * entry x, px, py are passed explicitly; pt is unchanged inside a step, so
 * the entry delta1/gamma1 are those of the cache c.
 */
template <class A>
ST_DEV void spin_rotate_t(A &ar, St &s, double xe, double pxe, double pye, double pze, double angle, double href,
                          double B0, double B1, double B2, const Ptc &c, st_gref g)
{
    double *x = s.x;
    {
        double Gr = g.G * c.g1;
        StRcp rt = ar.rcp(pze);
        double u0 = ar.div(pxe, rt), u1 = ar.div(pye, rt), u2 = 1.0;
        double temp = ar.sqrt_sq(u0 * u0 + u1 * u1 + u2 * u2);
        StRcp rn = ar.rcp(temp);
        u0 = ar.div(u0, rn);
        u1 = ar.div(u1, rn);
        u2 = ar.div(u2, rn);
        temp = (1 + href * xe) * temp;
        double udotB = u0 * B0 + u1 * B1 + u2 * B2;
        double Bp0 = udotB * u0, Bp1 = udotB * u1, Bp2 = udotB * u2;
        double k = g.G * (1 - c.g1);
        double F0 = (1 + Gr) * B0 + k * Bp0;
        double F1 = (1 + Gr) * B1 + k * Bp1;
        double F2 = (1 + Gr) * B2 + k * Bp2;
        F0 = ar.div(F0 * temp, c.ropd);
        F1 = ar.div(F1 * temp, c.ropd);
        F2 = ar.div(F2 * temp, c.ropd);
        F1 = F1 - angle;
        double Fabs = ar.sqrt_sq(F0 * F0 + F1 * F1 + F2 * F2);
        StRcp rf = ar.rcp(Fabs);
        double a = ar.div(F0, rf), b = ar.div(F1, rf), d = ar.div(F2, rf);
        double sn, cs;
        ar.sincos(-Fabs, &sn, &cs);
        double omc = 1 - cs;
        double r0 = a * a * omc + cs, r1 = a * b * omc - d * sn, r2 = a * d * omc + b * sn;
        double r3 = a * b * omc + d * sn, r4 = b * b * omc + cs, r5 = b * d * omc - a * sn;
        double r6 = a * d * omc - b * sn, r7 = b * d * omc + a * sn, r8 = d * d * omc + cs;
        double s0 = x[6], s1 = x[7], s2 = x[8];
        x[6] = ((0. + r0 * s0) + r1 * s1) + r2 * s2;
        x[7] = ((0. + r3 * s0) + r4 * s1) + r5 * s2;
        x[8] = ((0. + r6 * s0) + r7 * s1) + r8 * s2;
    }
}

/* spin() on the entry state (the reference's temp = sqrt((1+delta1)^2 -
   px^2 - py^2) is the entry pz, evaluated only for a non-zero field), fast
   pass with an exact replay of the rotation where needed */
ST_DEV void spin_rotate(St &s, double xe, double pxe, double pye, double angle, double href, double B0, double B1,
                        double B2, const Ptc &c, st_gref g)
{
    if (B0 * B0 + B1 * B1 + B2 * B2 != 0.) {
        double s6 = s.x[6], s7 = s.x[7], s8 = s.x[8];
        ArFast af;
        double pze = af.sqrt(c.opd2 - pxe * pxe - pye * pye);
        spin_rotate_t(af, s, xe, pxe, pye, pze, angle, href, B0, B1, B2, c, g);
        if (__builtin_expect(af.bad(), 0)) {
            ST_FD_COLD_PATH();
            s.x[6] = s6;
            s.x[7] = s7;
            s.x[8] = s8;
            ArExact ae;
            pze = ae.sqrt(c.opd2 - pxe * pxe - pye * pye);
            spin_rotate_t(ae, s, xe, pxe, pye, pze, angle, href, B0, B1, B2, c, g);
        }
    }
}

/* ------------------------------------------------------- radiation */

/* vector_cross_product(BLbrho, u) and the energy loss of radiate/radiate1 */
ST_DEV void rad_common(const St &s, double href, double B0, double B1, double B2, const Ptc &c, double &temp1,
                       double &BLper2)
{
    const double *x = s.x;
    double pz = st_fsqrt(c.opd2 - x[1] * x[1] - x[3] * x[3]);
    double xp = (1 + x[0] * href) * x[1] / pz;
    double yp = (1 + x[0] * href) * x[3] / pz;
    temp1 = st_fsqrt((1 + x[0] * href) * (1 + x[0] * href) + xp * xp + yp * yp);
    double u0 = xp / temp1, u1 = yp / temp1, u2 = (1 + x[0] * href) / temp1;
    double c0 = B1 * u2 - B2 * u1;
    double c1 = -1.0 * B0 * u2 + B2 * u0;
    double c2 = B0 * u1 - B1 * u0;
    BLper2 = c0 * c0 + c1 * c1 + c2 * c2;
}

/* radiate, simtrack.cpp:4198-4234 (classical energy loss) */
ST_DEV void radiate(St &s, double L, double href, double B0, double B1, double B2, Ptc &c, st_gref g)
{
    double *x = s.x;
    double px0 = x[1], py0 = x[3];
    double temp1, BLper2;
    rad_common(s, href, B0, B1, B2, c, temp1, BLper2);
    double temp2 = g.Cr * c.opd * c.opd * g.e3 / 2 / 3.14159265358979323;
    x[5] = x[5] - temp2 * BLper2 * temp1 / L;
    double delta2 = st_fsqrt(1.0 + 2 * x[5] / g.beta + x[5] * x[5]) - 1.0;
    double ratio = (1 + delta2) / c.opd;
    x[1] = px0 * ratio;
    x[3] = py0 * ratio;
    ptc_update(c, x[5], g);
}

/* radiate1, simtrack.cpp:4236-4288 (loss + Gaussian quantum excitation) */
ST_DEV void radiate1(St &s, double L, double href, double B0, double B1, double B2, Ptc &c, st_gref g,
                     uint32_t pid, uint64_t &ctr)
{
    double *x = s.x;
    double px0 = x[1], py0 = x[3];
    double temp1, BLper2;
    rad_common(s, href, B0, B1, B2, c, temp1, BLper2);
    double temp2 = g.Cr * c.opd * c.opd * g.e3 / 2 / 3.14159265358979323;
    double ptloss = temp2 * BLper2 * temp1 / L;
    double dtdl = temp1 / c.b1 / 299792458.;
    double g1 = st_fsqrt(BLper2 / L / L) / c.opd;
    double randgaus = st_rng_normal(g.rng_seed, pid, ctr);
    ctr++;
    double gm = c.g1;
    double ptfluct = st_fsqrt(g.Cf * gm * gm * gm * gm * gm * g1 * g1 * g1 * dtdl * L) * randgaus;
    x[5] = x[5] - ptloss + ptfluct;
    double delta2 = st_fsqrt(1.0 + 2 * x[5] / g.beta + x[5] * x[5]) - 1.0;
    double ratio = (1.0 + delta2) / c.opd;
    x[1] = px0 * ratio;
    x[3] = py0 * ratio;
    ptc_update(c, x[5], g);
}

/*
 * radiate2, simtrack.cpp:4290-4366 (photon emission by Monte Carlo), for a
 * field BLbrho = (0, B1, 0).  As in the reference, yp is formed from x[1]
 * (cpp:4331) and the bending radius from 1 + pt (cpp:4339).  One draw
 * number per call; the photons come from its sub-stream (st_rad2_photons).
 */
ST_DEV void radiate2(St &s, double L, double href, double B1, Ptc &c, st_gref g, uint32_t pid, uint64_t &ctr)
{
    double *x = s.x;
    double px0 = x[1], py0 = x[3];
    const double B0 = 0., B2 = 0.;
    double pz = st_fsqrt(c.opd2 - x[1] * x[1] - x[3] * x[3]);
    double xp = (1 + x[0] * href) * x[1] / pz;
    double yp = (1 + x[0] * href) * x[1] / pz;
    double temp1 = st_fsqrt((1 + x[0] * href) * (1 + x[0] * href) + xp * xp + yp * yp);
    double u0 = xp / temp1, u1 = yp / temp1, u2 = (1 + x[0] * href) / temp1;
    double c0 = B1 * u2 - B2 * u1;
    double c1 = -1.0 * B0 * u2 + B2 * u0;
    double c2 = B0 * u1 - B1 * u0;
    double BLper2 = c0 * c0 + c1 * c1 + c2 * c2;
    double gm = c.g1;
    double rho1 = (1 + x[5]) / st_fsqrt(BLper2) * L;
    double Ec = 1.5 * 6.582119e-16 * 2.99792458e8 * gm * gm * gm / rho1 / 1.0e6;
    double Nr = 5. * (1.0 / 137.0359997) * gm / 2. / 0x1.bb67ae8584caap+0 / rho1 * L * temp1 / c.b1;
    double Etot = st_rad2_photons(g.rng_seed, pid, ctr, Nr);
    ctr++;
    Etot = Etot * Ec;
    x[5] = x[5] - Etot / g.energy;
    double delta2 = st_fsqrt(1.0 + 2 * x[5] / g.beta + x[5] * x[5]) - 1.0;
    double ratio = (1 + delta2) / c.opd;
    x[1] = px0 * ratio;
    x[3] = py0 * ratio;
    ptc_update(c, x[5], g);
}

/* ------------------------------------------------------- beam-beam */

/*
 * cerrf, simtrack.cpp:6258-6323 (Gautschi's continued fraction for the
 * Faddeeva function).  The reference fills rx[n], ry[n] for n = nu..1 and
 * then sums n = nc..1 (nc <= nu); both loops run in descending n and the sum
 * at n uses only rx[n], ry[n], so they are fused into one pass with no
 * arrays.  Every operation is the reference's.
 */
ST_DEV void cerrf(double xx, double yy, double &wx, double &wy)
{
    const double cc = 1.12837916709551, zero = 0, one = 1, two = 2, half = .5, xlim = 5.33, ylim = 4.29,
                 fac1 = 3.2, fac2 = 23, fac3 = 21;
    double x = __builtin_fabs(xx), y = __builtin_fabs(yy);
    if (y < ylim && x < xlim) {
        double q = (one - y / ylim) * st_fsqrt(one - (x / xlim) * (x / xlim));
        double h = one / (fac1 * q);
        int nc = 7 + (int)(fac2 * q);
        double xl = st_powi(h, 1 - nc);
        double xh = y + half / h;
        double yh = x;
        int nu = 10 + (int)(fac3 * q);
        double rx = zero, ry = zero, sx = zero, sy = zero;
        for (int n = nu; n >= 1; n--) {
            double tx = xh + n * rx;
            double ty = yh - n * ry;
            double tn = tx * tx + ty * ty;
            rx = half * tx / tn;
            ry = half * ty / tn;
            if (n <= nc) {
                double saux = sx + xl;
                double nsx = rx * saux - ry * sy;
                sy = rx * sy + ry * saux;
                sx = nsx;
                xl = h * xl;
            }
        }
        wx = cc * sx;
        wy = cc * sy;
    } else {
        double xh = y, yh = x, rx = zero, ry = zero;
        for (int n = 9; n >= 1; n--) {
            double tx = xh + n * rx;
            double ty = yh - n * ry;
            double tn = tx * tx + ty * ty;
            rx = half * tx / tn;
            ry = half * ty / tn;
        }
        wx = cc * rx;
        wy = cc * ry;
    }
    if (yy < zero) {
        double sn, cs;
        double e = st_exp(y * y - x * x);
        st_sincos(two * x * y, &sn, &cs);
        wx = two * e * cs - wx;
        wy = -two * e * sn - wy;
        if (xx > zero) wy = -wy;
    } else {
        if (xx < zero) wy = -wy;
    }
}

/*
 * BB4D (scalar form), simtrack.cpp:6325-6382.  c2Nrg = 2*N*rp/gamma with
 * rp = 1.53469826769e-18 (the reference's BB4D literal).
 */
ST_DEV void bb4d(double x, double y, double c2Nrg, double sigmax, double sigmay, double &Dpx, double &Dpy)
{
    const double PI = 3.14159265358979323;
    int signx = (x >= 0.0) ? 1 : -1;
    int signy = (y >= 0.0) ? 1 : -1;
    if (x == 0. && y == 0.) {
        Dpx = 0.;
        Dpy = 0.;
        return;
    }
    if (__builtin_fabs(sigmax - sigmay) / sigmax < 1.0e-6) {
        double r2 = x * x + y * y;
        double temp1 = c2Nrg / r2;
        double temp2 = 1 - st_exp(-r2 / 2 / sigmax / sigmax);
        Dpx = temp1 * x * temp2;
        Dpy = temp1 * y * temp2;
    } else if (sigmax > sigmay) {
        double temp1 = c2Nrg * st_fsqrt(PI / 2. / (sigmax * sigmax - sigmay * sigmay));
        double temp2 = st_exp(-x * x / 2. / sigmax / sigmax - y * y / 2. / sigmay / sigmay);
        double temp3 = st_fsqrt(2. * (sigmax * sigmax - sigmay * sigmay));
        double w1r, w1i, w2r, w2i;
        double x1 = __builtin_fabs(x) / temp3;
        double y1 = __builtin_fabs(y) / temp3;
        double x2 = __builtin_fabs(x) * sigmay / sigmax / temp3;
        double y2 = __builtin_fabs(y) * sigmax / sigmay / temp3;
        cerrf(x1, y1, w1r, w1i);
        cerrf(x2, y2, w2r, w2i);
        Dpy = temp1 * (w1r - temp2 * w2r) * signy;
        Dpx = temp1 * (w1i - temp2 * w2i) * signx;
    } else if (sigmax < sigmay) {
        double temp1 = c2Nrg * st_fsqrt(PI / 2 / (sigmay * sigmay - sigmax * sigmax));
        double temp2 = st_exp(-x * x / 2 / sigmax / sigmax - y * y / 2 / sigmay / sigmay);
        double temp3 = st_fsqrt(2 * (sigmay * sigmay - sigmax * sigmax));
        double w1r, w1i, w2r, w2i;
        double x1 = __builtin_fabs(x) / temp3;
        double y1 = __builtin_fabs(y) / temp3;
        double x2 = __builtin_fabs(x) * sigmay / sigmax / temp3;
        double y2 = __builtin_fabs(y) * sigmax / sigmay / temp3;
        cerrf(y1, x1, w1r, w1i);
        cerrf(y2, x2, w2r, w2i);
        Dpy = temp1 * (w1i - temp2 * w2i) * signy;
        Dpx = temp1 * (w1r - temp2 * w2r) * signx;
    } else {
        /* sigmax or sigmay NaN: the reference leaves Dpx, Dpy uninitialised;
           NaN makes the particle lost at this element */
        Dpx = sigmax - sigmay;
        Dpy = Dpx;
    }
}

/*
 * One slice of BB6D (cpp:6622-6677) / BB6D_Angle (cpp:6980-7043): the
 * kick given S, the slice beam sizes and dsigma^2/ds.  z (x[4]) unchanged.
 * Q: c2Nrg, temp3, bbscale.
 */
ST_DEV void bb_slice_kick(double *x, double S, double sigmax, double sigmay, double dsx2, double dsy2, double X,
                          double Y, double c2Nrg, double temp3, double bbscale)
{
    double Dpx, Dpy;
    bb4d(X, Y, c2Nrg, sigmax, sigmay, Dpx, Dpy);
    double temp1 = 1.0 / 2. / (sigmax * sigmax - sigmay * sigmay);
    double temp2 = X * Dpx + Y * Dpy;
    double temp4 = st_exp(-X * X / 2. / sigmax / sigmax - Y * Y / 2. / sigmay / sigmay);
    double x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3], x5 = x[5];
    x[0] = x0 - S * Dpx * bbscale;
    x[1] = x1 + Dpx * bbscale;
    x[2] = x2 - S * Dpy * bbscale;
    x[3] = x3 + Dpy * bbscale;
    if (__builtin_fabs(sigmax - sigmay) / sigmax < 1.0e-6) {
        x[5] = x5 + 0.5 * Dpx * bbscale * (x1 + 0.5 * Dpx * bbscale) + 0.5 * Dpy * bbscale * (x3 + 0.5 * Dpy * bbscale) +
               (1.0 / sigmax / sigmax) * dsx2 * (temp3 * bbscale / 4.) * temp4;
    } else if (sigmax > sigmay) {
        double dUdsigmax2 = temp1 * bbscale * (temp2 + temp3 * (sigmay / sigmax * temp4 - 1.));
        double dUdsigmay2 = -temp1 * bbscale * (temp2 + temp3 * (sigmax / sigmay * temp4 - 1.));
        x[5] = x5 + 0.5 * Dpx * bbscale * (x1 + 0.5 * Dpx * bbscale) + 0.5 * Dpy * bbscale * (x3 + 0.5 * Dpy * bbscale) -
               (0.5 * dsx2 * dUdsigmax2 + 0.5 * dsy2 * dUdsigmay2);
    } else {
        double dUdsigmay2 = -temp1 * bbscale * (temp2 + temp3 * (sigmax / sigmay * temp4 - 1.));
        double dUdsigmax2 = temp1 * bbscale * (temp2 + temp3 * (sigmay / sigmax * temp4 - 1.));
        x[5] = x5 + 0.5 * Dpx * bbscale * (x1 + 0.5 * Dpx * bbscale) + 0.5 * Dpy * bbscale * (x3 + 0.5 * Dpy * bbscale) -
               (0.5 * dsx2 * dUdsigmax2 + 0.5 * dsy2 * dUdsigmay2);
    }
}

/*
 * BB6D, simtrack.cpp:6579-6678.  P: c2Nrg, temp3, bbscale, gx*, gy*, betax*,
 * alfx*, betay*, alfy*, emitx, emity, 2*emitx, 2*emity, z*[nslice]
 */
ST_DEV void bb6d(St &s, int nslice, st_pp P)
{
    double *x = s.x;
    double c2Nrg = P[0], temp3 = P[1], bbscale = P[2], gx = P[3], gy = P[4];
    double bx = P[5], ax = P[6], by = P[7], ay = P[8], ex = P[9], ey = P[10], ex2 = P[11], ey2 = P[12];
    for (int i = 0; i < nslice; i++) {
        double S = (x[4] - P[13 + i]) / 2.0;
        double betax = bx + 2 * S * ax + S * S * gx;
        double betay = by + 2 * S * ay + S * S * gy;
        double sigmax = st_fsqrt(ex * betax);
        double sigmay = st_fsqrt(ey * betay);
        double dsx2 = ex2 * (ax + S * gx);
        double dsy2 = ey2 * (ay + S * gy);
        double X = x[0] + x[1] * S;
        double Y = x[2] + x[3] * S;
        bb_slice_kick(x, S, sigmax, sigmay, dsx2, dsy2, X, Y, c2Nrg, temp3, bbscale);
    }
}

/*
 * BB6D_Angle, simtrack.cpp:6917-7079 (boost, head-on slices, inverse boost).
 * P: c2Nrg, temp3, bbscale, sin, cos, tan, cos(theta), gx*, gy*, betax*,
 * alfx*, betay*, alfy*, emitx, emity (all in the boosted frame),
 * Z0*[nslice] (= Z0/cos).
 */
ST_DEV void bb6d_angle(St &s, int nslice, st_pp P)
{
    double *x = s.x;
    double c2Nrg = P[0], temp3 = P[1], bbscale = P[2], sintheta = P[3], costheta = P[4], tantheta = P[5],
           cos_theta = P[6];
    double gx = P[7], gy = P[8], bx = P[9], ax = P[10], by = P[11], ay = P[12], ex = P[13], ey = P[14];
    double x0 = x[0], px0 = x[1], y0 = x[2], py0 = x[3], z0 = x[4], pz0 = x[5];
    double ps0 = st_fsqrt((1 + pz0) * (1 + pz0) - px0 * px0 - py0 * py0);
    double h0 = (1 + pz0) - ps0;
    double px_star = (px0 - tantheta * h0) / cos_theta;
    double py_star = py0 / costheta;
    double pz_star = pz0 - tantheta * px0 + tantheta * tantheta * h0;
    double ps_star = st_fsqrt((1 + pz_star) * (1 + pz_star) - px_star * px_star - py_star * py_star);
    double h_star = (1 + pz_star) - ps_star;
    double x_star = tantheta * z0 + (1 + (px_star / ps_star) * sintheta) * x0;
    double y_star = y0 + sintheta * (py_star / ps_star) * x0;
    double z_star = z0 / costheta + (-h_star / ps_star) * sintheta * x0;
    double w[6] = {x_star, px_star, y_star, py_star, z_star, pz_star};
    for (int i = 0; i < nslice; i++) {
        double Z0s = P[15 + i];
        double S = (w[4] - Z0s) / 2.0;
        double betax = bx - 2 * S * ax + S * S * gx;
        double alfx = (ax - gx * S);
        double betay = by - 2 * S * ay + S * S * gy;
        double alfy = (ay - gy * S);
        double sigmax = st_fsqrt(ex * betax);
        double sigmay = st_fsqrt(ey * betay);
        double dsx2 = -2.0 * alfx * (ex);
        double dsy2 = -2.0 * alfy * (ey);
        double X = w[0] + w[1] * S - Z0s * sintheta;
        double Y = w[2] + w[3] * S - 0;
        bb_slice_kick(w, S, sigmax, sigmay, dsx2, dsy2, X, Y, c2Nrg, temp3, bbscale);
    }
    x_star = w[0];
    px_star = w[1];
    y_star = w[2];
    py_star = w[3];
    z_star = w[4];
    pz_star = w[5];
    ps_star = st_fsqrt((1 + pz_star) * (1 + pz_star) - px_star * px_star - py_star * py_star);
    h_star = 1 + pz_star - ps_star;
    double hx_star = px_star / ps_star;
    double hy_star = py_star / ps_star;
    double hz_star = -h_star / ps_star;
    px0 = (px_star + h_star * sintheta) * costheta;
    py0 = py_star * costheta;
    pz0 = pz_star + px_star * sintheta;
    x0 = (x_star - z_star * sintheta) / (1 + hx_star * sintheta - hz_star * sintheta * sintheta);
    z0 = (z_star - hz_star * sintheta * x0) * costheta;
    y0 = y_star - sintheta * hy_star * x0;
    x[0] = x0;
    x[1] = px0;
    x[2] = y0;
    x[3] = py0;
    x[4] = z0;
    x[5] = pz0;
}

/* LBT_Pass, simtrack.cpp:6069-6093.  P: sin, cos, tan, cos(theta) */
ST_DEV void lbt(St &s, st_pp P)
{
    double *x = s.x;
    double sintheta = P[0], costheta = P[1], tantheta = P[2], cos_theta = P[3];
    double x0 = x[0], px0 = x[1], y0 = x[2], py0 = x[3], z0 = x[4], pz0 = x[5];
    double ps0 = st_fsqrt((1 + pz0) * (1 + pz0) - px0 * px0 - py0 * py0);
    double h0 = (1 + pz0) - ps0;
    double px_star = (px0 - tantheta * h0) / cos_theta;
    double py_star = py0 / costheta;
    double pz_star = pz0 - tantheta * px0 + tantheta * tantheta * h0;
    double ps_star = st_fsqrt((1 + pz_star) * (1 + pz_star) - px_star * px_star - py_star * py_star);
    double h_star = (1 + pz_star) - ps_star;
    x[0] = tantheta * z0 + (1 + (px_star / ps_star) * sintheta) * x0;
    x[1] = px_star;
    x[2] = y0 + sintheta * (py_star / ps_star) * x0;
    x[3] = py_star;
    x[4] = z0 / costheta + (-h_star / ps_star) * sintheta * x0;
    x[5] = pz_star;
}

/* ILBT_Pass, simtrack.cpp:6095-6123.  P: sin, cos */
ST_DEV void ilbt(St &s, st_pp P)
{
    double *x = s.x;
    double sintheta = P[0], costheta = P[1];
    double x_star = x[0], px_star = x[1], y_star = x[2], py_star = x[3], z_star = x[4], pz_star = x[5];
    double ps_star = st_fsqrt((1 + pz_star) * (1 + pz_star) - px_star * px_star - py_star * py_star);
    double h_star = 1 + pz_star - ps_star;
    double hx_star = px_star / ps_star;
    double hy_star = py_star / ps_star;
    double hz_star = -h_star / ps_star;
    double px0 = (px_star + h_star * sintheta) * costheta;
    double py0 = py_star * costheta;
    double pz0 = pz_star + px_star * sintheta;
    double x0 = (x_star - z_star * sintheta) / (1 + hx_star * sintheta - hz_star * sintheta * sintheta);
    double z0 = (z_star - hz_star * sintheta * x0) * costheta;
    double y0 = y_star - sintheta * hy_star * x0;
    x[0] = x0;
    x[1] = px0;
    x[2] = y0;
    x[3] = py0;
    x[4] = z0;
    x[5] = pz0;
}

#endif /* ST_MAPS_H */
