/*
 * st_kernels.hip -- gfx950 kernels of the tracking engine.
 *
 * k_track: one thread per particle, state in VGPRs, a wavefront-uniform loop
 * over the packed element table (scalar loads), loss masking per element
 * with the reference's test, whole-wavefront exit once every lane is lost.
 * A launch covers a block of turns; between blocks the host compacts the
 * surviving particle indices (k_compact_*), so later launches only spend
 * wavefronts on particles that are still alive.
 *
 * Compiled with -ffp-contract=off: no FMA contraction of the element maps
 * (explicit fma() only inside st_libm.h, identical on the host).
 */
#include <hip/hip_runtime.h>

#include <stdint.h>

#define ST_FN __device__ __forceinline__
#define ST_FN_COLD __device__ __noinline__
#define ST_CONST_TABLE static __device__ __constant__ const
#define ST_FMA(a, b, c) __builtin_fma((a), (b), (c))
/* a fused multiply-add with a constant addend, written as the three-address
   v_fma_f64: the compiler's two-address v_fmac_f64 would first copy the
   constant (kept in a VGPR across the loop) into the destination */
__device__ __forceinline__ double st_fma_k(double a, double b, double c)
{
    double r;
    __asm__("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
#define ST_FMA_K(a, b, c) st_fma_k((a), (b), (c))
#define ST_RINT(x) __builtin_rint(x)
#include "st_fastdiv.h"
#define ST_SQRT(x) st_fsqrt(x)
#include "st_libm.h"
#include "st_rng.h"
#include "st_maps.h"
#include "st_kernels.h"

/* st_kernels_perf.hip compiles this file a second time as the perf-mode
   kernels, in their own namespace so their symbols stay apart */
#ifdef ST_PERF_MODE
namespace st_perf {
#endif

/* ----------------------------------------------------- element dispatch */

/* quadrupole hard-edge fringe of QUAD_Pass, simtrack.cpp:4895-4911 (entry, sgn=+1)
   and cpp:4942-4957 (exit, sgn=-1) */
ST_DEV void quad_fringe(St &s, double k, bool entry)
{
    double *x = s.x;
    double x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3], x4 = x[4], x5 = x[5];
    if (entry) {
        x[0] = x0 + k * (x0 * x0 * x0 + 3 * x0 * x2 * x2) / 12 / (1 + x5);
        x[1] = x1 - k * ((x0 * x0 + x2 * x2) * x1 - 2 * x0 * x2 * x3) / 4 / (1 + x5);
        x[2] = x2 - k * (3 * x0 * x0 * x2 + x2 * x2 * x2) / 12 / (1 + x5);
        x[3] = x3 + k * ((x0 * x0 + x2 * x2) * x3 - 2 * x0 * x2 * x1) / 4 / (1 + x5);
        x[4] = x4 + k * (x2 * x2 * x2 * x3 - x0 * x0 * x0 * x1 + 3 * x0 * x0 * x2 * x3 - 3 * x2 * x2 * x0 * x1) / 12 /
                        (1 + x5) / (1 + x5);
    } else {
        x[0] = x0 - k * (x0 * x0 * x0 + 3 * x0 * x2 * x2) / 12 / (1 + x5);
        x[1] = x1 + k * ((x0 * x0 + x2 * x2) * x1 - 2 * x0 * x2 * x3) / 4 / (1 + x5);
        x[2] = x2 + k * (3 * x0 * x0 * x2 + x2 * x2 * x2) / 12 / (1 + x5);
        x[3] = x3 - k * ((x0 * x0 + x2 * x2) * x3 - 2 * x0 * x2 * x1) / 4 / (1 + x5);
        x[4] = x4 - k * (x2 * x2 * x2 * x3 - x0 * x0 * x0 * x1 + 3 * x0 * x0 * x2 * x3 - 3 * x2 * x2 * x0 * x1) / 12 /
                        (1 + x5) / (1 + x5);
    }
}

/* the thin kick of a QUAD / SEXT / OCT op */
template <int KK>
ST_DEV void thin_kick(St &s, double kn, double ks)
{
    if (KK == K_QUAD) quad_kick(s, kn, ks);
    if (KK == K_SEXT) sext_kick(s, kn, ks);
    if (KK == K_OCT) oct_kick(s, kn, ks);
}

/* midpoint field of the spin passes of QUAD/SEXT/OCT (cpp:7574-7579, 7631-7636, 7686-7691) */
template <int KK>
ST_DEV void thin_field(double xm, double ym, double kn, double ks, double &B0, double &B1)
{
    if (KK == K_QUAD) {
        double preal = xm, pimag = ym;
        B0 = ks * preal + kn * pimag;
        B1 = kn * preal - ks * pimag;
    }
    if (KK == K_SEXT) {
        double preal = xm * xm - ym * ym;
        double pimag = 2. * xm * ym;
        B0 = (ks * preal + kn * pimag) / 2.;
        B1 = (kn * preal - ks * pimag) / 2.;
    }
    if (KK == K_OCT) {
        double preal = xm * xm * xm - 3 * xm * ym * ym;
        double pimag = 3 * xm * xm * ym - ym * ym * ym;
        B0 = (ks * preal + kn * pimag) / 6;
        B1 = (kn * preal - ks * pimag) / 6;
    }
}

/*
 * QUAD_Pass / SEXT_Pass / OCT_Pass thick body (cpp:4913-4935, 5017-5033,
 * 5090-5106) and their spin passes QUAD_sPass / SEXT_sPass / OCT_sPass
 * (cpp:7561-7589, 7618-7644, 7673-7699).
 */
template <int KK, bool SPIN, bool HEXP>
ST_DEV void thick_multipole(St &s, Ptc &c, int nint, int flags, st_pp P, st_gref g)
{
    double d1 = P[0], d1b = P[1], d2 = P[2], d2b = P[3];
    double kn1 = P[4], ks1 = P[5], kn2 = P[6], ks2 = P[7];
    if (!SPIN && KK == K_QUAD && (flags & ST_F_FRINGE)) quad_fringe(s, P[11], true);
    for (int i = 0; i < nint; i++) {
        double xe = s.x[0], pxe = s.x[1], ye = s.x[2], pye = s.x[3];
        drift<HEXP>(s, d1, d1b, c);
        thin_kick<KK>(s, kn1, ks1);
        drift<HEXP>(s, d2, d2b, c);
        thin_kick<KK>(s, kn2, ks2);
        drift<HEXP>(s, d2, d2b, c);
        thin_kick<KK>(s, kn1, ks1);
        drift<HEXP>(s, d1, d1b, c);
        if (SPIN) {
            double xm = (s.x[0] + xe) / 2., ym = (s.x[2] + ye) / 2.;
            double B0, B1;
            thin_field<KK>(xm, ym, P[8], P[9], B0, B1);
            B0 = B0 / P[10];
            B1 = B1 / P[10];
            spin_rotate(s, xe, pxe, pye, 0, 0, B0, B1, 0., c, g);
        }
    }
    if (!SPIN && KK == K_QUAD && (flags & ST_F_FRINGE)) quad_fringe(s, P[11], false);
}

/* thin (L = 0) QUAD/SEXT/OCT and their spin passes (cpp:7546-7559, 7603-7616, 7658-7671) */
template <int KK, bool SPIN>
ST_DEV void thin_multipole(St &s, Ptc &c, st_pp P, st_gref g)
{
    double xe = s.x[0], pxe = s.x[1], ye = s.x[2], pye = s.x[3];
    thin_kick<KK>(s, P[0], P[1]);
    if (SPIN) {
        double xm = (s.x[0] + xe) / 2., ym = (s.x[2] + ye) / 2.;
        double B0, B1;
        thin_field<KK>(xm, ym, P[0], P[1], B0, B1);
        if (KK == K_OCT) {          /* OCT_sPass divides by Nint in its L == 0 branch too */
            B0 = B0 / P[2];
            B1 = B1 / P[2];
        }
        spin_rotate(s, xe, pxe, pye, 0, 0, B0, B1, 0., c, g);
    }
}

/* MULT_Pass (cpp:5152-5182) and MULT_sPass (cpp:7702-7740), thick */
template <bool SPIN, bool HEXP>
ST_DEV void thick_mult(St &s, Ptc &c, const st_op &op, st_pp P, st_gref g)
{
    int n1 = op.norder + 1;
    st_pp kn1 = P + 6, ks1 = kn1 + n1, kn2 = ks1 + n1, ks2 = kn2 + n1, KNL = ks2 + n1, KNSL = KNL + n1;
    double d1 = P[0], d1b = P[1], d2 = P[2], d2b = P[3], nint_d = P[4];
    int mask0 = (int)P[5];
    for (int i = 0; i < op.nint; i++) {
        double xe = s.x[0], pxe = s.x[1], ye = s.x[2], pye = s.x[3];
        drift<HEXP>(s, d1, d1b, c);
        mult_kick(s, op.norder, op.mask1, kn1, ks1);
        drift<HEXP>(s, d2, d2b, c);
        mult_kick(s, op.norder, op.mask2, kn2, ks2);
        drift<HEXP>(s, d2, d2b, c);
        mult_kick(s, op.norder, op.mask1, kn1, ks1);
        drift<HEXP>(s, d1, d1b, c);
        if (SPIN) {
            St m;
            m.x[0] = (s.x[0] + xe) / 2.;
            m.x[2] = (s.x[2] + ye) / 2.;
            double By, Bx;
            mult_field(m, op.norder, mask0, KNL, KNSL, By, Bx);
            spin_rotate(s, xe, pxe, pye, 0, 0, Bx / nint_d, (0. + By) / nint_d, 0., c, g);
        }
    }
}

/* thin MULT: mult_kick_pass; spin: cal_Bfield + spin on the entry state, then kick (cpp:7708-7713) */
template <bool SPIN>
ST_DEV void thin_mult(St &s, Ptc &c, const st_op &op, st_pp P, st_gref g)
{
    int n1 = op.norder + 1;
    st_pp KNL = P, KNSL = P + n1;
    if (SPIN) {
        double By, Bx;
        mult_field(s, op.norder, op.mask1, KNL, KNSL, By, Bx);
        spin_rotate(s, s.x[0], s.x[1], s.x[3], 0, 0, Bx / 1, (0. + By) / 1, 0., c, g);
    }
    mult_kick(s, op.norder, op.mask1, KNL, KNSL);
}

/*
 * SBENDMULT_Pass / GMULT_Pass / SMULT_Pass (cpp:5348-5470; identical
 * bodies) and their spin passes (cpp:7835-7928), L > 0.
 */
template <bool SPIN, bool HEXP, bool RAD>
ST_DEV void bendmult(St &s, Ptc &c, const st_op &op, st_pp P, st_gref g)
{
    int n1 = op.norder + 1;
    st_pp kn1 = P + ST_BM_ARR, ks1 = kn1 + n1, kn2 = ks1 + n1, ks2 = kn2 + n1, KNL = ks2 + n1, KNSL = KNL + n1;
    double href = P[2], Lint = P[3], nint_d = P[4], Angle = P[5], An = P[6];
    int mask0 = (int)P[7];
    StRcp rh1 = st_rcp_prep(P[9]), rh2 = st_rcp_prep(P[15]), rn = st_rcp_prep(nint_d);
    edge(s, P[0]);
    for (int i = 0; i < op.nint; i++) {
        double xe = s.x[0], pxe = s.x[1], ye = s.x[2], pye = s.x[3];
        if (HEXP) {
            double d1 = P[8], d1b = P[9], d2 = P[10], d2b = P[11], kl1 = P[12], kl2 = P[13];
            drift<HEXP>(s, d1, d1b, c);
            bend_mult_kick(s, kl1, href, op.norder, op.mask1, kn1, ks1);
            drift<HEXP>(s, d2, d2b, c);
            bend_mult_kick(s, kl2, href, op.norder, op.mask2, kn2, ks2);
            drift<HEXP>(s, d2, d2b, c);
            bend_mult_kick(s, kl1, href, op.norder, op.mask1, kn1, ks1);
            drift<HEXP>(s, d1, d1b, c);
        } else {
            sbend_exact(s, P[8], P[9], P[10], P[11], P[12], P[13], c, rh1);
            mult_kick(s, op.norder, op.mask1, kn1, ks1);
            sbend_exact(s, P[14], P[15], P[16], P[17], P[18], P[19], c, rh2);
            mult_kick(s, op.norder, op.mask2, kn2, ks2);
            sbend_exact(s, P[14], P[15], P[16], P[17], P[18], P[19], c, rh2);
            mult_kick(s, op.norder, op.mask1, kn1, ks1);
            sbend_exact(s, P[8], P[9], P[10], P[11], P[12], P[13], c, rh1);
        }
        if (SPIN || RAD) {
            St m;
            m.x[0] = (s.x[0] + xe) / 2.;
            m.x[2] = (s.x[2] + ye) / 2.;
            double By, Bx;
            mult_field(m, op.norder, mask0, KNL, KNSL, By, Bx);
            double B0 = st_qdiv(Bx, rn), B1 = st_qdiv(Angle + By, rn);
            if (SPIN) spin_rotate(s, xe, pxe, pye, An, href, B0, B1, 0., c, g);
            if (RAD) radiate(s, Lint, href, B0, B1, 0., c, g);
        }
    }
    edge(s, P[1]);
}

/* L = 0 branch of the bend-multipole passes (cpp:5350-5352; spin cpp:7846-7854) */
template <bool SPIN>
ST_DEV void bendmult_thin(St &s, Ptc &c, const st_op &op, st_pp P, st_gref g)
{
    int n1 = op.norder + 1;
    st_pp KNL = P + 1, KNSL = KNL + n1;
    double xe = s.x[0], pxe = s.x[1], ye = s.x[2], pye = s.x[3];
    mult_kick(s, op.norder, op.mask1, KNL, KNSL);
    if (SPIN) {
        St m;
        m.x[0] = (s.x[0] + xe) / 2.;
        m.x[2] = (s.x[2] + ye) / 2.;
        double By, Bx;
        mult_field(m, op.norder, op.mask1, KNL, KNSL, By, Bx);
        spin_rotate(s, xe, pxe, pye, P[0], 0, Bx / 1, (P[0] + By) / 1, 0., c, g);
    }
}

/* GSBENDMULT_Pass, cpp:5597-5723 (orbit only; its sPass is the SBENDMULT one) */
template <bool HEXP, bool RAD>
ST_DEV void gsbendmult(St &s, Ptc &c, const st_op &op, st_pp P, st_gref g)
{
    int n1 = op.norder + 1;
    st_pp kn1 = P + ST_GBM_ARR, ks1 = kn1 + n1, kn2 = ks1 + n1, ks2 = kn2 + n1;
    double href = P[2], hreal = P[3], Lint = P[4], BLy = P[6];
    edge(s, P[0]);
    for (int i = 0; i < op.nint; i++) {
        if (HEXP) {
            double d1 = P[8], d1b = P[9], d2 = P[10], d2b = P[11], kl1 = P[12], kl2 = P[13];
            drift<HEXP>(s, d1, d1b, c);
            general_bend_mult_kick(s, kl1, href, hreal, op.norder, op.mask1, kn1, ks1);
            drift<HEXP>(s, d2, d2b, c);
            general_bend_mult_kick(s, kl2, href, hreal, op.norder, op.mask2, kn2, ks2);
            drift<HEXP>(s, d2, d2b, c);
            general_bend_mult_kick(s, kl1, href, hreal, op.norder, op.mask1, kn1, ks1);
            drift<HEXP>(s, d1, d1b, c);
        } else {
            general_sbend_exact(s, P + 8, c);
            mult_kick(s, op.norder, op.mask1, kn1, ks1);
            general_sbend_exact(s, P + 15, c);
            mult_kick(s, op.norder, op.mask2, kn2, ks2);
            general_sbend_exact(s, P + 15, c);
            mult_kick(s, op.norder, op.mask1, kn1, ks1);
            general_sbend_exact(s, P + 8, c);
        }
        if (RAD) radiate(s, Lint, href, 0., BLy, 0., c, g);
    }
    edge(s, P[1]);
}

/*
 * Exact-body slices of SBEND_sPass (cpp:7486-7536) without radiation:
 * Nint times sbend_exact_pass (cpp:3833-3858) followed by spin() on the
 * slice's entry state.  Across slices of one element x[3] and pt do not
 * change, so quantities the reference recomputes per slice from identical
 * operands are carried instead (same values, same bits): sqrt((1+d)^2-py^2)
 * once per element, the entry pz of slice i+1 is the exit pz of slice i
 * (also spin()'s temp), and asin(px_entry/temp) of slice i+1 is the second
 * asin of slice i.
 */
template <bool SPIN, class A>
ST_DEV void sbend_slice_t(A &ar, St &s, double &pz0, double &as1, double x3, const StRcp &rt, const StRcp &rh,
                          double L, double hs, double ih, double cosA, double sinA, double Lb, double An, double href,
                          const Ptc &c, st_gref g)
{
    double *x = s.x;
    const double x0 = x[0], x1 = x[1];
    double nx1 = x1 * cosA + (pz0 - 1. - hs * x0) * sinA;
    double pz1 = ar.sqrt(c.opd2 - nx1 * nx1 - x3 * x3);
    double derivative = -1.0 * x1 * sinA * hs + (pz0 - 1. - x0 * hs) * cosA * hs;
    x[1] = nx1;
    x[0] = ar.div(pz1, rh) - ar.div(ar.div(derivative, rh), rh) - ih;
    double as2 = ar.asin(ar.div(nx1, rt));
    x[2] = x[2] + x3 * L + ar.div(x3 * (as1 - as2), rh);
    double t = c.opd * L + ar.div(c.opd * (as1 - as2), rh);
    x[4] = x[4] - (ar.div(t, c.rb1) - Lb);
    if (SPIN) spin_rotate_t(ar, s, x0, x1, x3, pz0, An, href, 0., An, 0., c, g);
    pz0 = pz1;
    as1 = as2;
}

/* nint slices with or without the spin rotation: the loop is written once
   per case, so each slice body (orbit and spin) is one scheduling region */
template <bool ROT>
ST_DEV void sbend_slice_loop(St &s, const Ptc &c, int nint, double &pz0, double &as1, double x3, const StRcp &rt,
                             const StRcp &rh, double L, double hs, double ih, double cosA, double sinA, double Lb,
                             double An, double href, st_gref g)
{
    for (int i = 0; i < nint; i++) {
        St s0 = s;
        double pz00 = pz0, as10 = as1;
        ArFast af;
        sbend_slice_t<ROT>(af, s, pz0, as1, x3, rt, rh, L, hs, ih, cosA, sinA, Lb, An, href, c, g);
        if (__builtin_expect(af.bad(), 0)) {
            ST_FD_COLD_PATH();
            s = s0;
            pz0 = pz00;
            as1 = as10;
            ArExact ae;
            sbend_slice_t<ROT>(ae, s, pz0, as1, x3, rt, rh, L, hs, ih, cosA, sinA, Lb, An, href, c, g);
        }
    }
}

template <bool SPIN>
ST_DEV void sbend_slices_exact(St &s, Ptc &c, int nint, double L, double hs, double ih, double cosA, double sinA,
                               double Lb, double An, double href, st_gref g)
{
    double *x = s.x;
    StRcp rh = st_rcp_prep(hs);
    const double x3 = x[3];
    const double temp = __builtin_sqrt(c.opd2 - x3 * x3);
    StRcp rt = st_rcp_prep(temp);
    double pz0 = __builtin_sqrt(c.opd2 - x[1] * x[1] - x3 * x3);
    double as1 = st_asin(x[1] / temp);
    /* spin() runs only for a non-zero field; B = (0, An, 0) here */
    const bool rot = SPIN && (0. * 0. + An * An + 0. * 0. != 0.);
    if (rot)
        sbend_slice_loop<true>(s, c, nint, pz0, as1, x3, rt, rh, L, hs, ih, cosA, sinA, Lb, An, href, g);
    else
        sbend_slice_loop<false>(s, c, nint, pz0, as1, x3, rt, rh, L, hs, ih, cosA, sinA, Lb, An, href, g);
}

/*
 * SBEND_Pass with slices (H_expand or radiate, cpp:4781-4813) and
 * SBEND_sPass (cpp:7486-7536).
 */
template <bool SPIN, bool HEXP, bool RAD>
ST_DEV void sbend_sliced(St &s, Ptc &c, int nint, st_pp P, st_gref g,
                         uint32_t pid, uint64_t &rctr)
{
    double href = P[2], Lint = P[3], An = P[4];
    edge(s, P[0]);
    if (!HEXP && !RAD) {
        sbend_slices_exact<SPIN>(s, c, nint, Lint, P[6], P[7], P[8], P[9], P[10], An, href, g);
        edge(s, P[1]);
        return;
    }
    StRcp rh = st_rcp_prep(P[6]);
    for (int i = 0; i < nint; i++) {
        double xe = s.x[0], pxe = s.x[1], pye = s.x[3];
        if (HEXP) {
            double d1 = P[6], d1b = P[7], d2 = P[8], d2b = P[9], kl1 = P[10], kl2 = P[11];
            drift<HEXP>(s, d1, d1b, c);
            bend_kick(s, kl1, href);
            drift<HEXP>(s, d2, d2b, c);
            bend_kick(s, kl2, href);
            drift<HEXP>(s, d2, d2b, c);
            bend_kick(s, kl1, href);
            drift<HEXP>(s, d1, d1b, c);
        } else {
            sbend_exact(s, Lint, P[6], P[7], P[8], P[9], P[10], c, rh);
        }
        if (SPIN) spin_rotate(s, xe, pxe, pye, An, href, 0., An, 0., c, g);
        if (RAD) {
            if (SPIN)
                radiate(s, Lint, href, 0., An, 0., c, g);
            else
                radiate1(s, Lint, href, 0., An, 0., c, g, pid, rctr);
        }
    }
    edge(s, P[1]);
}

/* GSBEND_Pass, cpp:4818-4880.  With RAD the exact branch runs nint slices
   of general_sbend_exact_pass, and both branches call radiate2 after each
   slice with BLbrho = (0, K0L/Nint, 0) */
template <bool HEXP, bool RAD>
ST_DEV void gsbend(St &s, Ptc &c, int kind, int nint, st_pp P, st_gref g, uint32_t pid, uint64_t &rctr)
{
    edge(s, P[0]);
    if (kind == K_GSBEND_EXACT) {
        if (RAD) {
            for (int i = 0; i < nint; i++) {
                general_sbend_exact(s, P + 2, c);
                radiate2(s, P[11], P[9], P[10], c, g, pid, rctr);
            }
        } else {
            general_sbend_exact(s, P + 2, c);
        }
    } else {
        double d1 = P[2], d1b = P[3], d2 = P[4], d2b = P[5], kl1 = P[6], kl2 = P[7], href = P[8], hreal = P[9];
        for (int i = 0; i < nint; i++) {
            drift<HEXP>(s, d1, d1b, c);
            general_bend_kick(s, kl1, href, hreal);
            drift<HEXP>(s, d2, d2b, c);
            general_bend_kick(s, kl2, href, hreal);
            drift<HEXP>(s, d2, d2b, c);
            general_bend_kick(s, kl1, href, hreal);
            drift<HEXP>(s, d1, d1b, c);
            if (RAD) radiate2(s, P[11], href, P[10], c, g, pid, rctr);
        }
    }
    edge(s, P[1]);
}

/* CCMULT_Pass body, cpp:6034-6064 */
ST_DEV void ccmult_kick(St &s, st_pp P, st_gref g)
{
    double *x = s.x;
    double scale = P[4], brho = P[5];
    double B2 = P[6], A2 = P[7], B3 = P[8], A3 = P[9], B4 = P[10], A4 = P[11], B5 = P[12], A5 = P[13];
    double sinangle, cosangle;
    st_sincos(P[2] * x[4] / 2.99792458e8 + P[3], &sinangle, &cosangle);
    double x1 = x[0], y1 = x[2];
    double x2 = x1 * x1, x3 = x2 * x1, x4 = x3 * x1, x5 = x4 * x1;
    double y2 = y1 * y1, y3 = y2 * y1, y4 = y3 * y1, y5 = y4 * y1;
    double xy = x1 * y1, x2y = x2 * y1, xy2 = x1 * y2, x2y2 = x1 * xy2;
    double x3y = y1 * x3, xy3 = x1 * y3;
    double x2y3 = x2y2 * y1, x3y2 = x2y2 * x1, x4y = x4 * y1, xy4 = y4 * x1;
    (void)g;
    x[1] = x[1] + (B2 * x1 + A2 * y1) * cosangle / brho;
    x[3] = x[3] + (A2 * x1 - B2 * y1) * cosangle / brho;
    x[5] = x[5] + (B2 * (x2 - y2) * sinangle + A2 * xy * sinangle) * scale;
    x[1] = x[1] + (B3 * (x2 - y2) + 2 * A3 * xy) * cosangle / brho;
    x[3] = x[3] + (2 * B3 * xy + A3 * (x2 - y2)) * cosangle / brho;
    x[5] = x[5] + (B3 * (x3 - 3 * xy2) * sinangle + A3 * (3 * x2y - y3) * sinangle) * scale / 3.;
    x[1] = x[1] + (B4 * (x3 - 3 * xy2) - A4 * (y3 - 3 * x2y)) * cosangle / brho;
    x[3] = x[3] + (B4 * (y3 - 3 * x2y) + A4 * (x3 - 3 * xy2)) * cosangle / brho;
    x[5] = x[5] + (B4 * (x4 - 6 * x2 * y2 + y4) * sinangle + A4 * (4 * x3y - 4 * xy3) * sinangle) * scale / 4.;
    x[1] = x[1] + (B5 * (x4 - 6 * x2y2 + y4) + A5 * (4 * x3y - 4 * xy3)) * cosangle / brho;
    x[3] = x[3] + (B5 * (4 * xy3 - 4 * x3y) + A5 * (x4 - 6 * x2y2 + y4)) * cosangle / brho;
    x[5] = x[5] + (B5 * (x5 - 10 * x3y2 + 5 * xy4) * sinangle + A5 * (5 * x4y - 10 * x2y3 + y5) * sinangle) * scale / 5.;
}

/*
 * Apply one op: Element::Pass (SPIN = false) or Element::sPass (SPIN = true),
 * including the GtoL/LtoG frame changes.
 */
template <bool SPIN, bool HEXP, bool RAD>
ST_DEV void apply_op(St &s, Ptc &c, const st_op &op, st_pp par, st_gref g,
                     uint32_t pid, uint64_t &rctr, int turn)
{
    st_pp P = st_launder(par + op.po);
    if (op.pm >= 0) {
        mis_in(s, op.flags, st_launder(par + op.pm));
        if (op.flags & ST_F_MIS2) ptc_update(c, s.x[5], g);
    }
    double *x = s.x;
    switch (op.kind) {
    case K_DRIFT:
        drift<HEXP>(s, P[0], P[1], c);
        break;
    case K_DRIFT0:
        drift0<HEXP>(s, c);
        break;
    case K_QUAD:
        thick_multipole<K_QUAD, SPIN, HEXP>(s, c, op.nint, op.flags, P, g);
        break;
    case K_SEXT:
        thick_multipole<K_SEXT, SPIN, HEXP>(s, c, op.nint, op.flags, P, g);
        break;
    case K_OCT:
        thick_multipole<K_OCT, SPIN, HEXP>(s, c, op.nint, op.flags, P, g);
        break;
    case K_QUAD_THIN:
        thin_multipole<K_QUAD, SPIN>(s, c, P, g);
        break;
    case K_SEXT_THIN:
        thin_multipole<K_SEXT, SPIN>(s, c, P, g);
        break;
    case K_OCT_THIN:
        thin_multipole<K_OCT, SPIN>(s, c, P, g);
        break;
    case K_MULT:
        thick_mult<SPIN, HEXP>(s, c, op, P, g);
        break;
    case K_MULT_THIN:
        thin_mult<SPIN>(s, c, op, P, g);
        break;
    case K_SBEND_EXACT:
        edge(s, P[0]);
        sbend_exact(s, P[2], P[3], P[4], P[5], P[6], P[7], c, st_rcp_prep(P[3]));
        edge(s, P[1]);
        break;
    case K_SBEND_SLICED:
        sbend_sliced<SPIN, HEXP, RAD>(s, c, op.nint, P, g, pid, rctr);
        break;
    case K_GSBEND_EXACT:
    case K_GSBEND_YOSHIDA:
        gsbend<HEXP, RAD && !SPIN>(s, c, op.kind, op.nint, P, g, pid, rctr);
        break;
    case K_BENDMULT:
        bendmult<SPIN, HEXP, RAD>(s, c, op, P, g);
        break;
    case K_BENDMULT_THIN:
        bendmult_thin<SPIN>(s, c, op, P, g);
        break;
    case K_GSBENDMULT:
        gsbendmult<HEXP, RAD>(s, c, op, P, g);
        break;
    case K_KICK: /* KICK_Pass cpp:5887-5893, KICK_sPass cpp:8193-8208 */
        drift<HEXP>(s, P[0], P[1], c);
        if (SPIN) spin_rotate(s, x[0], x[1], x[3], 0, 0, P[3], -P[2], 0., c, g);
        x[1] = x[1] + P[2];
        x[3] = x[3] + P[3];
        drift<HEXP>(s, P[0], P[1], c);
        break;
    case K_RFCAV: /* RFCAV_Pass, cpp:6002-6012 */
        drift<HEXP>(s, P[0], P[1], c);
        x[5] = x[5] + P[2] * st_sin(P[3] * x[4] / 2.99792458e8 + 3.14159265358979323 - P[4]) / g.beta / g.energy;
        ptc_update(c, x[5], g);
        drift<HEXP>(s, P[0], P[1], c);
        break;
    case K_CRABRF: { /* CRABRF_Pass, cpp:6014-6025 */
        drift<HEXP>(s, P[0], P[1], c);
        double sn, cs;
        st_sincos(P[3] * x[4] / 2.99792458e8 + P[4], &sn, &cs);
        x[1] = x[1] - P[2] * sn / g.beta / g.energy;
        x[5] = x[5] - P[2] * cs * x[0] * P[5] / g.beta / g.energy;
        ptc_update(c, x[5], g);
        drift<HEXP>(s, P[0], P[1], c);
        break;
    }
    case K_CCMULT: /* CCMULT_Pass, cpp:6027-6067 */
        drift<HEXP>(s, P[0], P[1], c);
        ccmult_kick(s, P, g);
        ptc_update(c, x[5], g);
        drift<HEXP>(s, P[0], P[1], c);
        break;
    case K_LBT:
        lbt(s, P);
        ptc_update(c, x[5], g);
        break;
    case K_ILBT:
        ilbt(s, P);
        ptc_update(c, x[5], g);
        break;
    case K_BB4D:  /* BEAMBEAM_Pass TREATMENT != 6, cpp:7350-7353 */
    case K_LRBB:  /* LRBB_Pass, cpp:7358-7363, with the 6-argument BB4D (cpp:6384-6447) */
    case K_ELENS: /* ELENS_Pass, cpp:7370-7390 (Ne != 0) */
    {
        /* one inlined BB4D (with its complex error function) for the three */
        const bool lr = op.kind == K_LRBB, el = op.kind == K_ELENS;
        const int o = op.kind == K_BB4D ? 0 : 2;
        const int n = el ? op.nint : 1;
        for (int i = 0; i < n; i++) {
            if (el) drift<HEXP>(s, P[0], P[1], c);
            if (lr) {
                x[0] = x[0] + P[0];
                x[2] = x[2] + P[1];
            }
            double Dpx, Dpy;
            bb4d(x[0], x[2], P[o], P[o + 1], P[o + 2], Dpx, Dpy);
            if (el) {
                x[1] = x[1] - Dpx * P[5] * P[6];
                x[3] = x[3] - Dpy * P[5] * P[6];
                drift<HEXP>(s, P[0], P[1], c);
            } else {
                x[1] = x[1] + Dpx * P[o + 3];
                x[3] = x[3] + Dpy * P[o + 3];
            }
            if (lr) {
                x[0] = x[0] - P[0];
                x[2] = x[2] - P[1];
            }
        }
        break;
    }
    case K_BB6D:
        bb6d(s, op.norder, P);
        ptc_update(c, x[5], g);
        break;
    case K_BB6DA:
        bb6d_angle(s, op.norder, P);
        ptc_update(c, x[5], g);
        break;
    case K_MATRIX: { /* MATRIX_Pass, cpp:5873-5885 */
        double t[6];
#pragma unroll
        for (int i = 0; i < 6; i++) t[i] = x[i] - P[i];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double a = 0.0;
#pragma unroll
            for (int j = 0; j < 6; j++) a = a + P[6 + i * 6 + j] * t[j];
            x[i] = a;
        }
#pragma unroll
        for (int i = 0; i < 6; i++) x[i] = x[i] + P[42 + i];
        ptc_update(c, x[5], g);
        break;
    }
    case K_ROTAT: { /* ROTAT_sPass / SNAKE_sPass, cpp:8286-8348 */
        drift<HEXP>(s, P[0], P[1], c);
        double s0 = x[6], s1 = x[7], s2 = x[8];
        x[6] = ((0. + P[2] * s0) + P[3] * s1) + P[4] * s2;
        x[7] = ((0. + P[5] * s0) + P[6] * s1) + P[7] * s2;
        x[8] = ((0. + P[8] * s0) + P[9] * s1) + P[10] * s2;
        drift<HEXP>(s, P[0], P[1], c);
        break;
    }
    case K_PATCH: /* PATCH_Pass, cpp:6248-6256: TRANS_Pass (cpp:6140) then YROTAT_Pass (cpp:6192) */
        x[0] = x[0] - P[0];
        x[2] = x[2] - P[1];
        x[4] = x[4] - P[2];
        if (P[3] != 0.) {
            const double cosT = P[4], sinT = P[5], tanT = P[6];
            const double x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3], x4 = x[4];
            double pz = st_fsqrt(c.opd2 - x1 * x1 - x3 * x3);
            double temp = 1 - x1 * tanT / pz;
            x[0] = x0 / cosT / temp;
            x[1] = x1 * cosT + pz * sinT;
            x[2] = x2 + x3 * x0 * tanT / pz / temp;
            x[4] = x4 - c.opd * x0 * tanT / pz / temp / c.b1;
        }
        break;
    case K_SROTAT: { /* SROTAT_Pass, cpp:6153-6170 */
        const double cosT = P[0], sinT = P[1];
        const double x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3];
        x[0] = cosT * x0 + sinT * x2;
        x[1] = cosT * x1 + sinT * x3;
        x[2] = cosT * x2 - sinT * x0;
        x[3] = cosT * x3 - sinT * x1;
        break;
    }
    case K_XROTAT: { /* XROTAT_Pass, cpp:6220-6245 (delta1, beta1 from the cache) */
        const double cosT = P[0], sinT = P[1], tanT = P[2];
        const double x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3], x4 = x[4];
        double pz = st_fsqrt(c.opd2 - x1 * x1 - x3 * x3);
        double temp = 1 - x3 * tanT / pz;
        x[2] = x2 / cosT / temp;
        x[3] = x3 * cosT + pz * sinT;
        x[0] = x0 + x1 * x2 * tanT / pz / temp;
        x[4] = x4 - c.opd * x2 * tanT / pz / temp / c.b1;
        break;
    }
    case K_COOLING: /* COOLING_Pass, cpp:6132-6136 */
#pragma unroll
        for (int i = 0; i < 4; i++) x[i] = P[0] * x[i];
        break;
    case K_HELENS: /* HELENS_Pass, cpp:7392-7431 */
        for (int i = 0; i < op.nint; i++) {
            drift<HEXP>(s, P[0], P[1], c);
            double r = st_fsqrt(x[0] * x[0] + x[2] * x[2]);
            double scale;
            if (r <= P[3])
                scale = 0;
            else if (r > P[3] && r <= P[4])
                scale = (r * r - P[5]) * P[4] / r / P[6];
            else if (r > P[4])
                scale = P[4] / r;
            else
                scale = r; /* NaN radius: the reference leaves scale unset */
            x[1] = x[1] + scale * P[2] * x[0] / r;
            x[3] = x[3] + scale * P[2] * x[2] / r;
            drift<HEXP>(s, P[0], P[1], c);
        }
        break;
    case K_WIGGLER: { /* Wiggler_Pass_Forest cpp:4400-4452 (Pass), Wiggler_Pass_Wu cpp:4516-4572 (sPass) */
        const double h = P[0], kx = P[9], ky = P[10], kz = P[11], phiz0 = P[12], brho = P[13];
        if (!SPIN) {
            for (int i = 0; i < op.nint; i++) {
                const double z = h * i;
                double cz, sz, cx, sx, chy, shy;
                st_sincos(kz * z + phiz0, &sz, &cz);
                st_sincos(kx * x[0], &sx, &cx);
                st_coshsinh(ky * x[2], &chy, &shy);
                /* get_Axy_wiggler (cpp:4371-4397), then divided by brho */
                const double Ax0 = P[1] * cx * chy * sz / brho, Ay0 = P[2] * sx * shy * sz / brho;
                const double Ax1 = P[3] * sx * chy * sz / brho, Ay1 = P[4] * cx * shy * sz / brho;
                const double Ax2 = P[5] * cx * shy * sz / brho, Ay2 = P[6] * sx * chy * sz / brho;
                const double x00 = x[0], x10 = x[1], x20 = x[2], x30 = x[3];
                const double psi = 1.0 + x[5], hodp = h / psi;
                const double det = ((1 - Ax1 * hodp) * (1 - Ay2 * hodp) - Ax2 * Ay1 * hodp * hodp);
                const double x01 = Ax0 * Ax1, x02 = Ax0 * Ax2, y01 = Ay0 * Ay1, y02 = Ay0 * Ay2;
                x[1] = (x10 - (x01 + y01) * hodp) * (1 - Ay2 * hodp) + (x30 - (x02 + y02) * hodp) * Ay1 * hodp;
                x[1] = x[1] / det;
                x[3] = (x30 - (x02 + y02) * hodp) * (1 - Ax1 * hodp) + (x10 - (x01 + y01) * hodp) * Ax2 * hodp;
                x[3] = x[3] / det;
                x[0] = x00 + hodp * (x[1] - Ax0);
                x[2] = x20 + hodp * (x[3] - Ay0);
                const double t1 = (x[1] - Ax0) / psi, t2 = (x[3] - Ay0) / psi;
                x[4] = x[4] - h * (t1 * t1 + t2 * t2) / 2.0;
                if (RAD) {
                    const double Ax3 = P[7] * cx * chy * cz / brho, Ay3 = P[8] * sx * shy * cz / brho;
                    radiate(s, h, 0, -Ay3 * brho * h, Ax3 * brho * h, (Ay1 - Ax2) * brho * h, c, g);
                }
            }
        } else {
            /* Wu's splitting: six half kicks (y, y, x, x, y, y) around three
               drifts; one field evaluation in a loop over the kicks */
            double z = 0.0;
            for (int i = 0; i < op.nint; i++) {
                const double hd = h / (1.0 + x[5]);
                z += 0.5 * h;
                const double sz = st_sin(kz * z + phiz0);
#pragma unroll 1
                for (int k = 0; k < 6; k++) {
                    const bool ykick = k < 2 || k > 3;
                    const double sgn = (k & 1) ? 1.0 : -1.0;
                    double cx, sx, chy, shy;
                    st_sincos(kx * x[0], &sx, &cx);
                    st_coshsinh(ky * x[2], &chy, &shy);
                    double d1, d3;
                    if (ykick) { /* get_Ay_Wu: AyoBrho, IntpAypxdy */
                        d3 = P[2] * sx * shy * sz;
                        d1 = P[15] * cx * chy * sz;
                    } else {     /* get_Ax_Wu: AxoBrho, IntpAxpydx */
                        d1 = P[1] * cx * chy * sz;
                        d3 = P[14] * (kx == 0. ? x[0] : sx) * shy * sz;
                    }
                    /* x -+ d/brho written as x + (+-1)(d/brho): the same bits */
                    x[1] = x[1] + sgn * (d1 / brho);
                    x[3] = x[3] + sgn * (d3 / brho);
                    if (k == 0 || k == 4) {
                        x[2] = x[2] + 0.5 * hd * x[3];
                        x[4] = x[4] - hd * x[3] * x[3] / (1.0 + x[5]) / (1.0 + x[5]) / 4;
                    } else if (k == 2) {
                        x[0] = x[0] + hd * x[1];
                        x[4] = x[4] - 0.5 * hd * x[1] * x[1] / (1.0 + x[5]);
                    }
                }
                z += 0.5 * h;
                if (RAD) { /* get_wiggler_B (cpp:4497-4514) times the step */
                    double cx, sx, chy, shy, cz, szp;
                    st_sincos(kx * x[0], &sx, &cx);
                    st_coshsinh(ky * x[2], &chy, &shy);
                    st_sincos(kz * x[4] + phiz0, &szp, &cz);
                    radiate(s, h, 0.0, P[16] * sx * shy * cz * h, P[7] * cx * chy * cz * h,
                            P[17] * cx * shy * szp * h, c, g);
                }
            }
        }
        break;
    }
    case K_ACDIP:  /* ACDIP_Pass cpp:5930-5950, ACDIP_sPass cpp:8251-8284; GP.turn = turn */
    case K_ACMULT: /* ACMULT_Pass cpp:5895-5928, ACMULT_sPass cpp:8210-8249 */
    case K_SOLEN:  /* SOLEN_Pass cpp:5804-5857, SOLEN_sPass cpp:8118-8177 */
    {
        /* one inlined spin rotation for the three: after the AC kick's first
           half drift, or after each solenoid body step (delta1, gamma1,
           beta1 from the cache) */
        const bool dip = op.kind == K_ACDIP, sol = op.kind == K_SOLEN;
        if (dip && turn < P[8]) {
            drift<HEXP>(s, P[0], P[1], c);
            break;
        }
        const double hL = dip ? P[2] : P[0], hLb = dip ? P[3] : P[1];
        double dpx = 0., dpy = 0., B0 = 0., B1 = 0., B2 = 0.;
        double K = 0., vz = 0., sn = 0., cs = 0., m01 = 0., m03 = 0., m21 = 0.;
        int n = 1;
        if (sol) {
            const double L = P[0], Lint = P[1];
            double ps = st_fsqrt(c.opd2 - x[1] * x[1] - x[3] * x[3]);
            x[1] = x[1] / ps;
            x[3] = x[3] / ps;
            vz = 2.99792458e8 * c.b1 / st_fsqrt(1.0 + x[1] * x[1] + x[3] * x[3]);
            K = P[2] / c.g1 / P[3] / vz;
            x[1] = x[1] + x[2] * K;
            x[3] = x[3] - x[0] * K;
            const double theta = -2.0 * K * L / op.nint;
            sn = st_sin(theta);
            cs = st_cos(theta);
            m01 = Lint * sn / theta;
            m03 = -(1 - cs) * Lint / theta;
            m21 = (1 - cs) * Lint / theta;
            n = op.nint;
            B2 = P[6];
        } else {
            /* the kicks dpx, dpy (x - a is x + (-a) bit for bit) and the field */
            drift<HEXP>(s, hL, hLb, c);
            if (dip) {
                const double TS = P[8], TE = P[9], HK = P[4], VK = P[5];
                const double s1 = st_sin(P[6] * (turn - TS) + P[7]);
                if (turn >= TS && turn <= TE) {
                    dpx = ((turn - TS) * 1.0 * HK / (TE - TS)) * s1;
                    dpy = ((turn - TS) * 1.0 * VK / (TE - TS)) * s1;
                } else {
                    dpx = HK * s1;
                    dpy = VK * s1;
                }
                B0 = dpy;
                B1 = -dpx;
            } else {
                double By, Bx;
                if (op.norder == 0) {
                    By = P[2];
                    Bx = P[3];
                } else {
                    double Xn = 1., Yn = 0.;
                    int fac = 1;
                    for (int i = 1; i < op.norder + 1; i++) {
                        double Xn0 = Xn, Yn0 = Yn;
                        Xn = Xn0 * x[0] - Yn0 * x[2];
                        Yn = Xn0 * x[2] + Yn0 * x[0];
                        fac = fac * i;
                    }
                    By = 0. + (P[2] * Xn - P[3] * Yn) / fac;
                    Bx = 0. + (P[2] * Yn + P[3] * Xn) / fac;
                }
                const double s1 = st_sin(P[4] * turn / P[5] + P[6]);
                B0 = Bx * s1;
                B1 = By * s1;
                dpx = -(By * s1);
                dpy = Bx * s1;
            }
        }
        for (int i = 0; i < n; i++) {
            double xe = x[0], pxe = x[1], pye = x[3];
            if (sol) {
                const double t0 = x[0], t1 = x[1], t2 = x[2], t3 = x[3];
                /* the reference's 4x4 product term by term, its zero entries included */
                x[0] = (((0. + 1.0 * t0) + m01 * t1) + 0. * t2) + m03 * t3;
                x[1] = (((0. + 0. * t0) + cs * t1) + 0. * t2) + -sn * t3;
                x[2] = (((0. + 0. * t0) + m21 * t1) + 1. * t2) + m01 * t3;
                x[3] = (((0. + 0. * t0) + sn * t1) + 0. * t2) + cs * t3;
                xe = t0;
                pxe = c.g1 * P[3] * t1 * vz / P[4];
                pye = c.g1 * P[3] * t3 * vz / P[4];
            }
            if (SPIN) spin_rotate(s, xe, pxe, pye, 0, 0, B0, B1, B2, c, g);
        }
        if (sol) {
            x[1] = x[1] - x[2] * K;
            x[3] = x[3] + x[0] * K;
            x[1] = c.g1 * P[3] * x[1] * vz / P[4];
            x[3] = c.g1 * P[3] * x[3] * vz / P[4];
            x[4] = x[4] - 2.99792458e8 * (P[0] / vz - P[5]);
        } else {
            x[1] = x[1] + dpx;
            x[3] = x[3] + dpy;
            drift<HEXP>(s, hL, hLb, c);
        }
        break;
    }
    default:
        break;
    }
    if (op.pm >= 0) {
        mis_out(s, op.flags, st_launder(par + op.pm));
        if (op.flags & ST_F_MIS2) ptc_update(c, s.x[5], g);
    }
}

/* ------------------------------------------------------ tracking kernel */

/*
 * The turn loop of one wavefront: particle p (lane active if `in`) tracked
 * through turns [a.turn0, a.turn0 + a.nturn) of the table (ops, par, ap,
 * nelem), with Track's pre-check, loss test and loss records.  Used by the
 * one-table kernel and by the persistent multi-table kernel.
 */
template <bool SPIN, bool HEXP, bool RAD>
ST_DEV void track_turns(const st_track_args &a, const st_opp ops, const st_pp par, const st_pp ap, const st_pp x04,
                        const int nelem, const int64_t draws_per_turn, const int p, const bool in)
{
    constexpr int NC = SPIN ? 9 : 6;
    St s;
#pragma unroll
    for (int k = 0; k < 9; k++) s.x[k] = 0.;
    /* per-lane flags as integers: kept in VGPRs, they need no lane-mask
       bookkeeping at every element */
    int alive = 0;
    if (in) {
        const double *xp = a.x + (size_t)p * NC;
#pragma unroll
        for (int k = 0; k < NC; k++) s.x[k] = xp[k];
        alive = a.state[p] != 0 ? 1 : 0;
        if (a.precheck) {
            /* Track, simtrack.cpp:20795-20796 */
            if (__builtin_fabs(s.x[0]) > 0.1 || __builtin_fabs(s.x[2]) > 0.1 || !alive) {
                a.state[p] = 0;
                a.lost[2 * p] = 0;
                a.lost[2 * p + 1] = 0;
                alive = 0;
            }
        }
    }
    uint64_t rctr = 0;
    Ptc c;
    ptc_update(c, s.x[5], *st_launder_g((st_gp)a.gdev));
    int lt = 0, lp = 0;
    int died = 0;
    const int tend = a.turn0 + a.nturn;
    for (int turn = a.turn0; turn < tend; ++turn) {
        if (!__any(alive)) break;
        if (RAD) rctr = (uint64_t)turn * (uint64_t)draws_per_turn;
        if (a.xtbt && a.tbt_bpm < 0 && alive) {
            double *o = a.xtbt + ((size_t)p * a.tbt_nturn + (size_t)(turn - a.tbt_turn0)) * NC;
#pragma unroll
            for (int k = 0; k < NC; k++) o[k] = s.x[k];
        }
        for (int e = 0; e < nelem; ++e) {
            const st_op op = ops[e];
            if (alive) {
                /* drifts and zero-length drifts without misalignment are most
                   of a lattice's elements: dispatched ahead of the full switch
                   (orbit kernels only: in the spin kernels the extra path
                   costs the spin-tracked bends more than it saves) */
                if (!SPIN && op.kind == K_DRIFT && op.pm < 0) {
                    st_pp P = st_launder(par + op.po);
                    drift<HEXP>(s, P[0], P[1], c);
                } else if (!SPIN && op.kind == K_DRIFT0 && op.pm < 0) {
                    drift0<HEXP>(s, c);
                } else {
                    apply_op<SPIN, HEXP, RAD>(s, c, op, par, *st_launder_g((st_gp)a.gdev), (uint32_t)p, rctr, turn);
                }
                if (x04) s.x[4] = s.x[4] - x04[e];
                if (a.xtbt && e == a.tbt_bpm) {
                    double *o = a.xtbt + ((size_t)p * a.tbt_nturn + (size_t)(turn - a.tbt_turn0)) * NC;
#pragma unroll
                    for (int k = 0; k < NC; k++) o[k] = s.x[k];
                }
                /* loss test of Track, simtrack.cpp:20802 (none for a bare Element::Pass) */
                if (ap && (__builtin_fabs(s.x[0]) > ap[2 * e] || __builtin_fabs(s.x[2]) > ap[2 * e + 1] ||
                           s.x[0] != s.x[0] || s.x[2] != s.x[2])) {
                    alive = 0;
                    died = 1;
                    lt = turn;
                    lp = e;
                }
            }
        }
    }
    if (in) {
        double *xp = a.x + (size_t)p * NC;
#pragma unroll
        for (int k = 0; k < NC; k++) xp[k] = s.x[k];
        if (died) {
            a.state[p] = 0;
            a.lost[2 * p] = lt;
            a.lost[2 * p + 1] = lp;
        }
    }
}

/* minimum resident wavefronts per SIMD the register allocation must allow
   (1: the compiler's choice; 2: at most 256 VGPRs, 4: at most 128) */
#ifndef ST_TRACK_WAVES_PER_EU
#define ST_TRACK_WAVES_PER_EU 1
#endif

template <bool SPIN, bool HEXP, bool RAD>
__global__ __launch_bounds__(256, ST_TRACK_WAVES_PER_EU) void k_track(st_track_args a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = t < a.nact;
    const int p = in ? (a.idx ? a.idx[t] : t) : 0;
    track_turns<SPIN, HEXP, RAD>(a, (st_opp)a.ops, (st_pp)a.par, (st_pp)a.ap, (st_pp)a.x04, a.nelem,
                                 a.draws_per_turn, p, in);
}

/*
 * Persistent tracking of many tables (lattice seeds) in one launch.  Work
 * items are (table, up to 64 particles); every wavefront takes the next item
 * from a global counter until none is left, so each wavefront always exits
 * after at most nitems claims.  The host orders items longest table first.
 */
template <bool SPIN, bool HEXP, bool RAD>
__global__ __launch_bounds__(256, ST_TRACK_WAVES_PER_EU) void k_track_batch(st_batch_args b)
{
    const int lane = threadIdx.x & 63;
    st_track_args a;
    a.x = b.x;
    a.state = b.state;
    a.lost = b.lost;
    a.turn0 = b.turn0;
    a.nturn = b.nturn;
    a.precheck = b.precheck;
    a.xtbt = b.xtbt;
    a.tbt_bpm = -1;
    a.tbt_nturn = b.tbt_nturn;
    a.tbt_turn0 = b.tbt_turn0;
    a.g = b.g;
    a.gdev = b.gdev;
    const ST_KCONST st_seed_desc *seeds = (const ST_KCONST st_seed_desc *)b.seeds;
    const ST_KCONST int *items = (const ST_KCONST int *)b.items;
    for (;;) {
        int item = 0;
        if (lane == 0) item = atomicAdd(b.counter, 1);
        item = __builtin_amdgcn_readfirstlane(__shfl(item, 0));
        if (item >= b.nitems) break;
        const int sd = items[3 * item], start = items[3 * item + 1], cnt = items[3 * item + 2];
        const bool in = lane < cnt;
        const int p = in ? b.idx[start + lane] : 0;
        track_turns<SPIN, HEXP, RAD>(a, (st_opp)seeds[sd].ops, (st_pp)seeds[sd].par, (st_pp)seeds[sd].ap, nullptr,
                                     seeds[sd].nelem, seeds[sd].draws_per_turn, p, in);
    }
}

/* on-momentum z after each element: one particle from x = 0 (Track_Radial, cpp:21655-21661) */
template <bool HEXP>
__global__ void k_radial_x04(st_track_args a, double *x04)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    St s;
#pragma unroll
    for (int k = 0; k < 9; k++) s.x[k] = 0.;
    Ptc c;
    uint64_t rctr = 0;
    for (int e = 0; e < a.nelem; ++e) {
        s.x[4] = 0.;
        s.x[5] = 0.;
        ptc_update(c, s.x[5], *st_launder_g((st_gp)a.gdev));
        const st_op op = ((st_opp)a.ops)[e];
        apply_op<false, HEXP, false>(s, c, op, (st_pp)a.par, *st_launder_g((st_gp)a.gdev), 0u, rctr, 0);
        x04[e] = s.x[4];
    }
}

/* --------------------------------------------------- survivor compaction */

/* per-block count of alive particles among idx[0..n) */
__global__ __launch_bounds__(1024) void k_count_alive(const int *__restrict__ idx, int n, const int *__restrict__ state,
                                                      int *__restrict__ block_count)
{
    __shared__ int wsum[16];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int f = (t < n) ? (state[idx ? idx[t] : t] != 0) : 0;
    unsigned long long b = __ballot(f);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wsum[w] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += wsum[i];
        block_count[blockIdx.x] = s;
    }
}

/* exclusive scan of block counts (one block, loops for any count); total -> out_total */
__global__ __launch_bounds__(1024) void k_scan_blocks(int *__restrict__ block_count, int nb, int *__restrict__ out_total)
{
    __shared__ int tmp[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        int i = base + threadIdx.x;
        int v = (i < nb) ? block_count[i] : 0;
        tmp[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int add = (threadIdx.x >= off) ? tmp[threadIdx.x - off] : 0;
            __syncthreads();
            tmp[threadIdx.x] += add;
            __syncthreads();
        }
        int incl = tmp[threadIdx.x];
        if (i < nb) block_count[i] = carry + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) *out_total = carry;
}

/* order-preserving scatter of alive indices */
__global__ __launch_bounds__(1024) void k_compact(const int *__restrict__ idx, int n, const int *__restrict__ state,
                                                  const int *__restrict__ block_off, int *__restrict__ idx_out)
{
    __shared__ int woff[16];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int pid = (t < n) ? (idx ? idx[t] : t) : 0;
    int f = (t < n) ? (state[pid] != 0) : 0;
    unsigned long long b = __ballot(f);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) woff[w] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
            int v = woff[i];
            woff[i] = s;
            s += v;
        }
    }
    __syncthreads();
    if (f) {
        unsigned long long below = (lane == 0) ? 0ull : (b & ((~0ull) >> (64 - lane)));
        int pos = block_off[blockIdx.x] + woff[w] + __popcll(below);
        idx_out[pos] = pid;
    }
}

/* ------------------------------------------------------------- launchers */

#define ST_LAUNCH(S, H, R) \
    hipLaunchKernelGGL((k_track<S, H, R>), dim3(grid), dim3(256), 0, stream, a)

hipError_t st_launch_track(const st_track_args &a, bool spin, bool hexp, bool rad, hipStream_t stream)
{
    if (a.nact <= 0) return hipSuccess;
    unsigned grid = (unsigned)((a.nact + 255) / 256);
    if (!spin && !hexp && !rad) ST_LAUNCH(false, false, false);
    else if (!spin && !hexp && rad) ST_LAUNCH(false, false, true);
    else if (!spin && hexp && !rad) ST_LAUNCH(false, true, false);
    else if (!spin && hexp && rad) ST_LAUNCH(false, true, true);
    else if (spin && !hexp && !rad) ST_LAUNCH(true, false, false);
    else if (spin && !hexp && rad) ST_LAUNCH(true, false, true);
    else if (spin && hexp && !rad) ST_LAUNCH(true, true, false);
    else ST_LAUNCH(true, true, true);
    return hipGetLastError();
}

hipError_t st_launch_radial_x04(const st_track_args &a, double *x04, bool hexp, hipStream_t stream)
{
    if (hexp)
        hipLaunchKernelGGL((k_radial_x04<true>), dim3(1), dim3(64), 0, stream, a, x04);
    else
        hipLaunchKernelGGL((k_radial_x04<false>), dim3(1), dim3(64), 0, stream, a, x04);
    return hipGetLastError();
}

#define ST_LAUNCH_B(S, H, R)                                                                                   \
    hipLaunchKernelGGL((k_track_batch<S, H, R>), dim3(grid), dim3(block), 0, stream, b)

hipError_t st_launch_track_batch(const st_batch_args &b, int grid, int block, bool spin, bool hexp, bool rad,
                                 hipStream_t stream)
{
    if (b.nitems <= 0 || grid <= 0) return hipSuccess;
    if (!spin && !hexp && !rad) ST_LAUNCH_B(false, false, false);
    else if (!spin && !hexp && rad) ST_LAUNCH_B(false, false, true);
    else if (!spin && hexp && !rad) ST_LAUNCH_B(false, true, false);
    else if (!spin && hexp && rad) ST_LAUNCH_B(false, true, true);
    else if (spin && !hexp && !rad) ST_LAUNCH_B(true, false, false);
    else if (spin && !hexp && rad) ST_LAUNCH_B(true, false, true);
    else if (spin && hexp && !rad) ST_LAUNCH_B(true, true, false);
    else ST_LAUNCH_B(true, true, true);
    return hipGetLastError();
}

hipError_t st_launch_compact(const int *idx, int n, const int *state, int *block_buf, int *idx_out, int *d_total,
                             hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    int nb = (n + 1023) / 1024;
    hipLaunchKernelGGL(k_count_alive, dim3(nb), dim3(1024), 0, stream, idx, n, state, block_buf);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, stream, block_buf, nb, d_total);
    hipLaunchKernelGGL(k_compact, dim3(nb), dim3(1024), 0, stream, idx, n, state, block_buf, idx_out);
    return hipGetLastError();
}

#ifdef ST_PERF_MODE
} /* namespace st_perf */
#endif
