/*
 * st_fastdiv.h -- correctly rounded FP64 division and square root for gfx950
 * with the shared parts hoisted, bit-identical to the compiler's a / b and
 * sqrt(x).
 *
 * The compiler lowers a / b to
 *     bs = div_scale(b); y = rcp(bs); as = div_scale(a)
 *     y = fma(y, fma(-bs, y, 1), y)            (twice)
 *     q = as*y; r = fma(-bs, q, as); q = div_fmas(r, y, q); div_fixup(q, b, a)
 * div_scale leaves its operand unchanged (and div_fmas is a plain fma) unless
 * an exponent is extreme: b or 1/b or a/b subnormal, exp(a) - exp(b) >= 768,
 * or |a| < 2^-969; div_fixup only changes the result for zero, infinite or
 * NaN operands and for overflow/underflow.  So for 2^-60 < |b| < 2^60 and a
 * quotient with 2^-900 < |q| < 2^700 the whole sequence reduces to
 *     y (a function of b only), q0 = a*y, r = fma(-b, q0, a), q = fma(r, y, q0),
 * and y can be computed once for every quotient with the same divisor.  Any
 * lane outside those bounds (including zero, inf and NaN operands) takes the
 * full division instead, so results equal a / b for every input.
 *
 * sqrt(x) is lowered to a range check (x < 2^-767 is scaled by 2^256), rsq,
 * a Goldschmidt/Newton refinement, the inverse scaling and a class check
 * (+-0 and +inf return x).  For 2^-767 <= x < inf the scalings are identities
 * and the class check passes, so the refinement alone gives the same bits.
 */
#ifndef ST_FASTDIV_H
#define ST_FASTDIV_H

#define ST_FD __device__ __forceinline__
/* keeps the full operation on a real (exec-masked, usually skipped) branch:
   without it the compiler evaluates both paths and selects */
#define ST_FD_COLD_PATH() __asm__ volatile("")

/* a divisor prepared for several quotients */
struct StRcp {
    double b;   /* the divisor                                              */
    double y;   /* its reciprocal as refined by the division sequence       */
    bool ok;    /* 2^-60 < |b| < 2^60                                       */
};

/* the refined reciprocal alone, for a caller that checks the range of b itself */
ST_FD StRcp st_rcp_y(double b)
{
    StRcp r;
    r.b = b;
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    r.y = y;
    r.ok = true;
    return r;
}

ST_FD StRcp st_rcp_prep(double b)
{
    StRcp r = st_rcp_y(b);
    double ab = __builtin_fabs(b);
    r.ok = ab > 0x1p-60 && ab < 0x1p60;
    return r;
}

/* a / r.b, bit-identical to the division operator */
ST_FD double st_qdiv(double a, const StRcp &r)
{
    double q0 = a * r.y;
    double rr = __builtin_fma(-r.b, q0, a);
    double q = __builtin_fma(rr, r.y, q0);
    double aq = __builtin_fabs(q);
    if (__builtin_expect(!(r.ok && aq > 0x1p-900 && aq < 0x1p700), 0)) {
        ST_FD_COLD_PATH();
        q = a / r.b;
    }
    return q;
}

/* sqrt(x), bit-identical to __builtin_sqrt */
ST_FD double st_fsqrt(double x)
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
    if (__builtin_expect(!(x >= 0x1p-767 && x < __builtin_inf()), 0)) {
        ST_FD_COLD_PATH();
        g = __builtin_sqrt(x);
    }
    return g;
}

#endif
