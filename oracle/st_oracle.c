/*
 * st_oracle.c -- CPU oracle for the SimTrack tracking hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests,
 * the smoke check and the CPU baseline leg of bench.py load it; the
 * product (simtrack_amd) never links it.
 *
 * It restates, in plain C99, the reference's element maps and tracking
 * drivers operation for operation (same expression trees, same evaluation
 * order, per-call recomputation of element constants exactly where the
 * reference recomputes them).  Each function cites the simtrack.cpp lines
 * it follows.  The reference itself cannot be built here: simtrack.h
 * includes GSL headers (gsl_vector.h, gsl_eigen.h, gsl_cdf.h, ...) that
 * the image does not have.
 *
 * Elementary functions: compiled with -DORACLE_GLIBC the oracle calls the
 * C library (what the reference calls); by default it calls the project
 * libm st_libm.h, the same implementation the GPU kernels use, so GPU and
 * oracle can be compared bit for bit.  The two builds are compared with
 * each other in the test suite.
 *
 * Build: oracle/build.sh (gcc -O2 -ffp-contract=off -mfma -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "simtrack_amd.h"

#define ST_FN static inline
#define ST_CONST_TABLE static const
#define ST_FMA(a, b, c) fma((a), (b), (c))
#define ST_RINT(x) nearbyint(x)
#define ST_SQRT(x) sqrt(x)
#include "st_libm.h"
#include "st_rng.h"

/* ORACLE_GLIBC_MASK picks the C library per function family (1: sin, cos,
   tan; 2: asin; 4: exp; 8: pow), so tests can tell which family moves a
   tracked particle away from the project-libm build */
#ifndef ORACLE_GLIBC_MASK
#ifdef ORACLE_GLIBC
#define ORACLE_GLIBC_MASK 15
#else
#define ORACLE_GLIBC_MASK 0
#endif
#endif
#if ORACLE_GLIBC_MASK & 1
#define OSIN(x) sin(x)
#define OCOS(x) cos(x)
#define OTAN(x) tan(x)
#else
#define OSIN(x) st_sin(x)
#define OCOS(x) st_cos(x)
#define OTAN(x) st_tan(x)
#endif
#if ORACLE_GLIBC_MASK & 2
#define OASIN(x) asin(x)
#else
#define OASIN(x) st_asin(x)
#endif
#if ORACLE_GLIBC_MASK & 4
#define OEXP(x) exp(x)
#define OCOSH(x) cosh(x)
#define OSINH(x) sinh(x)
#else
#define OEXP(x) st_exp(x)
#define OCOSH(x) st_cosh(x)
#define OSINH(x) st_sinh(x)
#endif
#if ORACLE_GLIBC_MASK & 8
#define OPOW(x, n) pow((x), (double)(n))
#else
#define OPOW(x, n) st_powi((x), (n))
#endif

#define EXPORT __attribute__((visibility("default")))
/* hooks of the op-counting build (oracle/opcount.cpp); no-ops here */
#ifndef ORC_UNCOUNTED_IF
#define ORC_UNCOUNTED_IF(cond)
#endif
#ifndef ORC_PT_CACHED
#define ORC_PT_CACHED(pt, v) (v)
#define ORC_PT_DERIVED(v) (v)
#endif
/* the plain value of an element field (the counting builds carry more) */
#ifndef ORC_VALUE
#define ORC_VALUE(v) (v)
#endif

enum { x_ = 0, px_ = 1, y_ = 2, py_ = 3, z_ = 4, pt_ = 5 };

/* --------------------------------------------------------- globals */
/* simtrack.cpp:23-32 */
static const double PI = 3.14159265358979323;
static const double speed_light = 299792458.;
static const double Cr = 8.84627632682e-5;
static double Cf;
static double Fdrift1, Fdrift2, Fkick1, Fkick2;
static st_globals GP;
static int g_init = 0;

static void orc_init_constants(void)
{
    /* simtrack.cpp:25-32 (pow/sqrt of constants, evaluated once at start-up) */
    Cf = (55. / 24. / sqrt(3.)) * 2.8179403e-15 * 1.054572e-34 / 9.1093822e-31;
    Fdrift1 = 1.e0 / (2.e0 * (2.e0 - pow(2.e0, 1.0 / 3.0)));
    Fdrift2 = 0.5e0 - Fdrift1;
    Fkick1 = 2.e0 * Fdrift1;
    Fkick2 = 1.e0 - 2e0 * Fkick1;
    g_init = 1;
}

EXPORT void orc_set_globals(const st_globals *gp)
{
    if (!g_init)
        orc_init_constants();
    GP = *gp;
}

EXPORT void orc_get_yoshida(double out[4])
{
    if (!g_init)
        orc_init_constants();
    out[0] = Fdrift1;
    out[1] = Fdrift2;
    out[2] = Fkick1;
    out[3] = Fkick2;
}

/* ---------------------------------------------------- misalignment */

/* GtoL1 / LtoG1, simtrack.cpp:3443-3481 */
static void GtoL1(double x[], double DX, double DY, double DT)
{
    double xtemp[6];
    if (fabs(DX) + fabs(DY) > 1.0e-10) {
        x[0] = x[0] - DX;
        x[2] = x[2] - DY;
    }
    if (fabs(DT) > 1.0e-10) {
        double cosT = OCOS(DT), sinT = OSIN(DT);
        for (int i = 0; i < 6; i++) xtemp[i] = x[i];
        x[0] = cosT * xtemp[0] + sinT * xtemp[2];
        x[1] = cosT * xtemp[1] + sinT * xtemp[3];
        x[2] = cosT * xtemp[2] - sinT * xtemp[0];
        x[3] = cosT * xtemp[3] - sinT * xtemp[1];
    }
}

static void LtoG1(double x[], double DX, double DY, double DT)
{
    double xtemp[6];
    if (fabs(DT) > 1.0e-10) {
        double cosT = OCOS(DT), sinT = OSIN(DT);
        for (int i = 0; i < 6; i++) xtemp[i] = x[i];
        x[0] = cosT * xtemp[0] - sinT * xtemp[2];
        x[1] = cosT * xtemp[1] - sinT * xtemp[3];
        x[2] = sinT * xtemp[0] + cosT * xtemp[2];
        x[3] = sinT * xtemp[1] + cosT * xtemp[3];
    }
    if (fabs(DX) + fabs(DY) > 1.0e-10) {
        x[0] = x[0] + DX;
        x[2] = x[2] + DY;
    }
}

/* Cal_W_Matrix, simtrack.cpp:3485-3503 */
static void Cal_W_Matrix(double theta, double phi, double psi, double W[])
{
    double costhe = OCOS(theta), cosphi = OCOS(phi), cospsi = OCOS(psi);
    double sinthe = OSIN(theta), sinphi = OSIN(phi), sinpsi = OSIN(psi);
    W[0] = +costhe * cospsi - sinthe * sinphi * sinpsi;
    W[1] = -costhe * sinpsi - sinthe * sinphi * cospsi;
    W[2] = sinthe * cosphi;
    W[3] = cosphi * sinpsi;
    W[4] = cosphi * cospsi;
    W[5] = sinphi;
    W[6] = -sinthe * cospsi - costhe * sinphi * sinpsi;
    W[7] = +sinthe * sinpsi - costhe * sinphi * cospsi;
    W[8] = costhe * cosphi;
}

/* Cal_VE_WE_Matrx, simtrack.cpp:3505-3565 */
static void Cal_VE_WE_Matrx(double length, double angle, double tilt, double VE[], double WE[])
{
    double temp, cospsi, costhe, sinpsi, sinthe;
    /* the reference tests < 1e-12 and > 1e-12, leaving VE/WE unset at
       |angle| == 1e-12 exactly; that case is taken as straight here and in
       the table builder */
    if (fabs(angle) <= 1.0e-12) {
        for (int i = 0; i < 3; i++) VE[i] = 0.;
        VE[2] = length;
        for (int i = 0; i < 9; i++) WE[i] = 0.;
        for (int i = 0; i < 3; i++) WE[i * 3 + i] = 1.0;
    }
    if (fabs(angle) > 1.0e-12) {
        tilt = 0.;
        cospsi = OCOS(tilt);
        sinpsi = OSIN(tilt);
        costhe = OCOS(angle);
        sinthe = OSIN(angle);
        temp = length * (costhe - 1.0) / angle;
        VE[0] = temp * cospsi;
        VE[1] = temp * sinpsi;
        VE[2] = length * OSIN(angle) / angle;
        WE[0 * 3 + 0] = costhe * cospsi * cospsi + sinpsi * sinpsi;
        WE[1 * 3 + 0] = (costhe - 1.0) * cospsi * sinpsi;
        WE[2 * 3 + 0] = sinthe * cospsi;
        WE[0 * 3 + 1] = WE[1 * 3 + 0];
        WE[1 * 3 + 1] = costhe * sinpsi * sinpsi + cospsi * cospsi;
        WE[2 * 3 + 1] = sinthe * sinpsi;
        WE[0 * 3 + 2] = -WE[2 * 3 + 0];
        WE[1 * 3 + 2] = -WE[2 * 3 + 1];
        WE[2 * 3 + 2] = costhe;
    }
}

/* mat_mult (simtrack.cpp:1351-1362), mat_transpose (cpp:1406-1416) */
static void mat_mult(const double a[], const double b[], double c[], int m, int n, int k)
{
    for (int i = 0; i <= m - 1; i++)
        for (int j = 0; j <= k - 1; j++) {
            int u = i * k + j;
            c[u] = 0.0;
            for (int l = 0; l <= n - 1; l++) c[u] = c[u] + a[i * n + l] * b[l * k + j];
        }
}

static void mat_transpose(const double a[], double at[], int m, int n)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) at[i * m + j] = a[j * n + i];
}

/* GtoL2, simtrack.cpp:3567-3626 (MAD-X tmali1) */
static void GtoL2(double x[], double length, double angle, double dx, double dy, double ds,
                  double theta, double phi, double psi)
{
    double s2, W[9], RM[36], xtemp[6];
    (void)length;
    (void)angle;
    Cal_W_Matrix(theta, phi, psi, W);
    s2 = (W[0 * 3 + 2] * dx + W[1 * 3 + 2] * dy + W[2 * 3 + 2] * ds) / W[2 * 3 + 2];
    for (int i = 0; i < 36; i++) RM[i] = 0.;
    for (int i = 0; i < 6; i++) RM[i * 6 + i] = 1.0;
    RM[1 * 6 + 1] = W[0 * 3 + 0];
    RM[1 * 6 + 3] = W[1 * 3 + 0];
    RM[1 * 6 + 5] = W[2 * 3 + 0] / GP.beta;
    RM[3 * 6 + 1] = W[0 * 3 + 1];
    RM[3 * 6 + 3] = W[1 * 3 + 1];
    RM[3 * 6 + 5] = W[2 * 3 + 1] / GP.beta;
    RM[0 * 6 + 0] = W[1 * 3 + 1] / W[2 * 3 + 2];
    RM[0 * 6 + 1] = RM[0 * 6 + 0] * s2;
    RM[0 * 6 + 2] = -W[0 * 3 + 1] / W[2 * 3 + 2];
    RM[0 * 6 + 3] = RM[0 * 6 + 2] * s2;
    RM[2 * 6 + 0] = -W[1 * 3 + 0] / W[2 * 3 + 2];
    RM[2 * 6 + 1] = RM[2 * 6 + 0] * s2;
    RM[2 * 6 + 2] = W[0 * 3 + 0] / W[2 * 3 + 2];
    RM[2 * 6 + 3] = RM[2 * 6 + 2] * s2;
    RM[4 * 6 + 0] = W[0 * 3 + 2] / (W[2 * 3 + 2] * GP.beta);
    RM[4 * 6 + 1] = RM[4 * 6 + 0] * s2;
    RM[4 * 6 + 2] = W[1 * 3 + 2] / (W[2 * 3 + 2] * GP.beta);
    RM[4 * 6 + 3] = RM[4 * 6 + 2] * s2;
    RM[4 * 6 + 5] = -s2 / (GP.beta * GP.gamma * GP.beta * GP.gamma);
    for (int i = 0; i < 6; i++) {
        xtemp[i] = 0.;
        for (int j = 0; j < 6; j++) xtemp[i] = xtemp[i] + RM[i * 6 + j] * x[j];
    }
    x[0] = xtemp[0] - (W[1 * 3 + 1] * dx - W[0 * 3 + 1] * dy) / W[2 * 3 + 2];
    x[1] = xtemp[1] + W[2 * 3 + 0];
    x[2] = xtemp[2] - (W[0 * 3 + 0] * dy - W[1 * 3 + 0] * dx) / W[2 * 3 + 2];
    x[3] = xtemp[3] + W[2 * 3 + 1];
    x[4] = xtemp[4] - s2 / GP.beta;
    x[5] = xtemp[5];
}

/* LtoG2, simtrack.cpp:3628-3699 (MAD-X tmali2) */
static void LtoG2(double x[], double length, double angle, double dx, double dy, double ds,
                  double theta, double phi, double psi)
{
    double V[3], W[9], VE[3], WE[9], V1[3], M1[9], M2[9], RM[36], s2, xtemp[6];
    double tilt = 0.;
    Cal_W_Matrix(theta, phi, psi, W);
    Cal_VE_WE_Matrx(length, angle, tilt, VE, WE);
    V[0] = dx + W[0 * 3 + 0] * VE[0] + W[0 * 3 + 1] * VE[1] + W[0 * 3 + 2] * VE[2] - VE[0];
    V[1] = dy + W[1 * 3 + 0] * VE[0] + W[1 * 3 + 1] * VE[1] + W[1 * 3 + 2] * VE[2] - VE[1];
    V[2] = ds + W[2 * 3 + 0] * VE[0] + W[2 * 3 + 1] * VE[1] + W[2 * 3 + 2] * VE[2] - VE[2];
    mat_transpose(WE, M1, 3, 3);
    for (int i = 0; i < 3; i++) {
        V1[i] = 0;
        for (int j = 0; j < 3; j++) V1[i] = V1[i] + M1[i * 3 + j] * V[j];
    }
    for (int i = 0; i < 3; i++) V[i] = V1[i];
    mat_mult(M1, W, M2, 3, 3, 3);
    mat_mult(M2, WE, W, 3, 3, 3);
    s2 = -(W[0 * 3 + 2] * V[0] + W[1 * 3 + 2] * V[1] + W[2 * 3 + 2] * V[2]) / W[2 * 3 + 2];
    for (int i = 0; i < 36; i++) RM[i] = 0.;
    for (int i = 0; i < 6; i++) RM[i * 6 + i] = 1.0;
    RM[0 * 6 + 0] = W[0 * 3 + 0];
    RM[2 * 6 + 0] = W[1 * 3 + 0];
    RM[4 * 6 + 0] = W[2 * 3 + 0] / GP.beta;
    RM[0 * 6 + 2] = W[0 * 3 + 1];
    RM[2 * 6 + 2] = W[1 * 3 + 1];
    RM[4 * 6 + 2] = W[2 * 3 + 1] / GP.beta;
    RM[1 * 6 + 1] = W[1 * 3 + 1] / W[2 * 3 + 2];
    RM[0 * 6 + 1] = RM[1 * 6 + 1] * s2;
    RM[3 * 6 + 1] = -W[0 * 3 + 1] / W[2 * 3 + 2];
    RM[2 * 6 + 1] = RM[3 * 6 + 1] * s2;
    RM[1 * 6 + 3] = -W[1 * 3 + 0] / W[2 * 3 + 2];
    RM[0 * 6 + 3] = RM[1 * 6 + 3] * s2;
    RM[3 * 6 + 3] = W[0 * 3 + 0] / W[2 * 3 + 2];
    RM[2 * 6 + 3] = RM[3 * 6 + 3] * s2;
    RM[1 * 6 + 5] = W[0 * 3 + 2] / (W[2 * 3 + 2] * GP.beta);
    RM[0 * 6 + 5] = RM[1 * 6 + 5] * s2;
    RM[3 * 6 + 5] = W[1 * 3 + 2] / (W[2 * 3 + 2] * GP.beta);
    RM[2 * 6 + 5] = RM[3 * 6 + 5] * s2;
    RM[4 * 6 + 5] = -s2 / (GP.beta * GP.gamma * GP.beta * GP.gamma);
    xtemp[0] = x[0] + (W[1 * 3 + 1] * V[0] - W[0 * 3 + 1] * V[1]) / W[2 * 3 + 2];
    xtemp[1] = x[1] - W[2 * 3 + 0];
    xtemp[2] = x[2] + (W[0 * 3 + 0] * V[1] - W[1 * 3 + 0] * V[0]) / W[2 * 3 + 2];
    xtemp[3] = x[3] - W[2 * 3 + 1];
    xtemp[4] = x[4] - s2 / GP.beta;
    xtemp[5] = x[5];
    for (int i = 0; i < 6; i++) {
        x[i] = 0.;
        for (int j = 0; j < 6; j++) x[i] = x[i] + RM[i * 6 + j] * xtemp[j];
    }
}

/* GtoL / LtoG, simtrack.cpp:3701-3741 */
static void GtoL(double x[], double length, double angle, double dx, double dy, double ds,
                 double theta, double phi, double psi)
{
    double temp = fabs(dx) + fabs(dy) + fabs(ds) + fabs(theta) + fabs(phi) + fabs(psi);
    if (temp < 1.0e-9) return;
    if (fabs(theta) + fabs(phi) < 1.0e-9)
        GtoL1(x, dx, dy, psi);
    else
        GtoL2(x, length, angle, dx, dy, ds, theta, phi, psi);
}

static void LtoG(double x[], double length, double angle, double dx, double dy, double ds,
                 double theta, double phi, double psi)
{
    double temp = fabs(dx) + fabs(dy) + fabs(ds) + fabs(theta) + fabs(phi) + fabs(psi);
    if (temp < 1.0e-9) return;
    if (fabs(theta) + fabs(phi) < 1.0e-9)
        LtoG1(x, dx, dy, psi);
    else
        LtoG2(x, length, angle, dx, dy, ds, theta, phi, psi);
}

/* -------------------------------------------------------- primitives */

/* DRIFT_Pass, simtrack.cpp:3743-3766 */
static void DRIFT_Pass(double x[], double L)
{
    ORC_UNCOUNTED_IF(L == 0.);      /* op-counting build only: a zero-length drift counts nothing */
    if (GP.H_expand) {
        double u = L / (1.0 + x[5]);
        x[x_] = x[x_] + x[px_] * u;
        x[y_] = x[y_] + x[py_] * u;
        x[z_] = x[z_] - (x[px_] * x[px_] + x[py_] * x[py_]) * u / 2.0 / (1 + x[5]);
    } else {
        double pz, betaz, delta1, gamma1, beta1;
        delta1 = ORC_PT_CACHED(x[pt_], sqrt(1.0 + 2 * x[pt_] / GP.beta + x[pt_] * x[pt_]) - 1.0);
        gamma1 = ORC_PT_CACHED(x[pt_], GP.gamma + sqrt(GP.gamma * GP.gamma - 1.) * x[pt_]);
        beta1 = ORC_PT_DERIVED(sqrt(1.0 - 1.0 / gamma1 / gamma1));
        pz = sqrt((1.0 + delta1) * (1.0 + delta1) - x[px_] * x[px_] - x[py_] * x[py_]);
        betaz = pz * beta1 / (1.0 + delta1);
        x[x_] = x[x_] + x[px_] * L / pz;
        x[y_] = x[y_] + x[py_] * L / pz;
        x[z_] = x[z_] - (L / betaz - L / GP.beta);
    }
}

/* YROTAT_Pass, simtrack.cpp:6192-6218: the frame y rotation of Forest's book */
static void YROTAT_Pass(double x[], double THETA)
{
    if (fabs(THETA) < 1.0e-9) return;
    double xtemp[6], temp, pz;
    double delta1 = ORC_PT_CACHED(x[pt_], sqrt(1.0 + 2 * x[pt_] / GP.beta + x[pt_] * x[pt_]) - 1.0);
    double gamma1 = ORC_PT_CACHED(x[pt_], GP.gamma + sqrt(GP.gamma * GP.gamma - 1.) * x[pt_]);
    double beta1 = ORC_PT_DERIVED(sqrt(1.0 - 1.0 / gamma1 / gamma1));
    double cosT = OCOS(THETA), sinT = OSIN(THETA), tanT = OTAN(THETA);
    for (int i = 0; i < 6; i++) xtemp[i] = x[i];
    pz = sqrt((1 + delta1) * (1 + delta1) - xtemp[1] * xtemp[1] - xtemp[3] * xtemp[3]);
    temp = 1 - xtemp[1] * tanT / pz;
    x[0] = xtemp[0] / cosT / temp;
    x[1] = xtemp[1] * cosT + pz * sinT;
    x[2] = xtemp[2] + xtemp[3] * xtemp[0] * tanT / pz / temp;
    x[3] = xtemp[3];
    x[4] = xtemp[4] - (1 + delta1) * xtemp[0] * tanT / pz / temp / beta1;
    x[5] = xtemp[5];
}

/* TRANS_Pass, simtrack.cpp:6140-6151 (its DRIFT_Pass by DS is commented out) */
static void TRANS_Pass(double x[], double DX, double DY, double DS)
{
    x[0] = x[0] - DX;
    x[2] = x[2] - DY;
    x[4] = x[4] - DS;
}

/* PATCH_Pass, simtrack.cpp:6248-6256: TRANS_Pass, then YROTAT_Pass */
static void PATCH_Pass(double x[], double DX, double DY, double DS, double THETA)
{
    TRANS_Pass(x, DX, DY, DS);
    YROTAT_Pass(x, THETA);
}

/* SROTAT_Pass, simtrack.cpp:6153-6170: frame rotation about s */
static void SROTAT_Pass(double x[], double PSI)
{
    if (PSI == 0.) return;
    double xtemp[6], cosT = OCOS(PSI), sinT = OSIN(PSI);
    for (int i = 0; i < 6; i++) xtemp[i] = x[i];
    x[0] = cosT * xtemp[0] + sinT * xtemp[2];
    x[1] = cosT * xtemp[1] + sinT * xtemp[3];
    x[2] = cosT * xtemp[2] - sinT * xtemp[0];
    x[3] = cosT * xtemp[3] - sinT * xtemp[1];
}

/* XROTAT_Pass, simtrack.cpp:6220-6245: frame rotation about x (Forest) */
static void XROTAT_Pass(double x[], double PHI)
{
    if (fabs(PHI) < 1.0e-9) return;
    double xtemp[6], temp, pz;
    double delta1 = ORC_PT_CACHED(x[pt_], sqrt(1.0 + 2 * x[pt_] / GP.beta + x[pt_] * x[pt_]) - 1.0);
    double gamma1 = ORC_PT_CACHED(x[pt_], GP.gamma + sqrt(GP.gamma * GP.gamma - 1.) * x[pt_]);
    double beta1 = ORC_PT_DERIVED(sqrt(1.0 - 1.0 / gamma1 / gamma1));
    double cosT = OCOS(PHI), sinT = OSIN(PHI), tanT = OTAN(PHI);
    for (int i = 0; i < 6; i++) xtemp[i] = x[i];
    pz = sqrt((1 + delta1) * (1 + delta1) - xtemp[1] * xtemp[1] - xtemp[3] * xtemp[3]);
    temp = 1 - xtemp[3] * tanT / pz;
    x[2] = xtemp[2] / cosT / temp;
    x[3] = xtemp[3] * cosT + pz * sinT;
    x[0] = xtemp[0] + xtemp[1] * xtemp[2] * tanT / pz / temp;
    x[1] = xtemp[1];
    x[4] = xtemp[4] - (1 + delta1) * xtemp[2] * tanT / pz / temp / beta1;
    x[5] = xtemp[5];
}

/* COOLING_Pass, simtrack.cpp:6132-6136 */
static void COOLING_Pass(double x[], double ALPHA)
{
    for (int i = 0; i < 4; i++) x[i] = (1. - ALPHA) * x[i];
}

/* The turn being tracked: GP.turn of the reference, which Track runs from 0
   (simtrack.cpp:20799) and the AC elements read. */
static __thread int orc_turn = 0;   /* per thread: particles are tracked in parallel */

/* ACDIP_Pass, simtrack.cpp:5930-5950 */
static void ACDIP_Pass(double x[], double L, double HKICKMAX, double VKICKMAX, double NUD, double TURNS,
                       double TURNE, double PHID)
{
    if (orc_turn < TURNS) {
        DRIFT_Pass(x, L);
    } else if (orc_turn >= TURNS && orc_turn <= TURNE) {
        DRIFT_Pass(x, L / 2.);
        x[1] += ((orc_turn - TURNS) * 1.0 * HKICKMAX / (TURNE - TURNS)) *
                OSIN(2.0 * 3.14159265 * NUD * (orc_turn - TURNS) + PHID);
        x[3] += ((orc_turn - TURNS) * 1.0 * VKICKMAX / (TURNE - TURNS)) *
                OSIN(2.0 * 3.14159265 * NUD * (orc_turn - TURNS) + PHID);
        DRIFT_Pass(x, L / 2.);
    } else {
        DRIFT_Pass(x, L / 2.);
        x[1] += HKICKMAX * OSIN(2.0 * 3.14159265 * NUD * (orc_turn - TURNS) + PHID);
        x[3] += VKICKMAX * OSIN(2.0 * 3.14159265 * NUD * (orc_turn - TURNS) + PHID);
        DRIFT_Pass(x, L / 2.);
    }
}

/* ACMULT_Pass, simtrack.cpp:5895-5928 (a single order, int factorial) */
static void ACMULT_Pass(double x[], double L, int Norder, double KL, double KSL, int TTURNS, double PHI0)
{
    int fac = 1;
    double Xn, Yn, Xn0, Yn0, KNL, KNSL, By, Bx;
    DRIFT_Pass(x, L / 2.0);
    if (Norder == 0) {
        By = KL;
        Bx = KSL;
    } else {
        KNL = KL;
        KNSL = KSL;
        By = 0.;
        Bx = 0.;
        Xn = 1.;
        Yn = 0.;
        for (int i = 1; i < Norder + 1; i++) {
            Xn0 = Xn;
            Yn0 = Yn;
            Xn = Xn0 * x[x_] - Yn0 * x[y_];
            Yn = Xn0 * x[y_] + Yn0 * x[x_];
            fac = fac * i;
        }
        By = By + (KNL * Xn - KNSL * Yn) / fac;
        Bx = Bx + (KNL * Yn + KNSL * Xn) / fac;
    }
    x[px_] = x[px_] - By * OSIN(2.0 * PI * orc_turn / TTURNS + PHI0);
    x[py_] = x[py_] + Bx * OSIN(2.0 * PI * orc_turn / TTURNS + PHI0);
    DRIFT_Pass(x, L / 2.0);
}

/* HELENS_Pass, simtrack.cpp:7392-7431 (speed_light = 299792458., cpp:24);
   for a NaN radius the reference leaves `scale` unset, here NaN */
static void HELENS_Pass(double x[], double Le, double Ne, double bbscale, double beta_e, int N_slice, double rout,
                        double rin)
{
    double Ne_slice = Ne * 1.0 / N_slice;
    double Le_slice = Le * 1.0 / N_slice;
    double RINNER = rin, ROUTER = rout, scale, I_e, Theta_max, r;
    (void)Ne_slice;
    (void)bbscale;
    I_e = 1.6021765e-19 * Ne * beta_e * 299792458. / Le;
    Theta_max = -2e-7 * Le_slice * I_e * (1 + beta_e * GP.beta) / beta_e / GP.beta / ROUTER / GP.brho;
    for (int i = 0; i < N_slice; i++) {
        DRIFT_Pass(x, Le_slice / 2.);
        r = sqrt(x[0] * x[0] + x[2] * x[2]);
        if (r <= RINNER)
            scale = 0;
        else if (r > RINNER && r <= ROUTER)
            scale = (r * r - RINNER * RINNER) * ROUTER / r / (ROUTER * ROUTER - RINNER * RINNER);
        else if (r > ROUTER)
            scale = ROUTER / r;
        else
            scale = r;
        x[1] = x[1] + scale * Theta_max * x[0] / r;
        x[3] = x[3] + scale * Theta_max * x[2] / r;
        DRIFT_Pass(x, Le_slice / 2.);
    }
}

/* SOLEN_Pass, simtrack.cpp:5804-5857: hard-edge solenoid in (x, x', y, y'):
   entrance kick, Nint times the body matrix, exit kick */
static void SOLEN_Pass(double x[], double L, int Nint, double KS)
{
    int i, j, k;
    double xtemp[6], ps, gamma1, delta1, beta1, vz;
    double Bz = KS * GP.brho, K, Lint = L / Nint, theta, sintheta, costheta, m44[16];
    if (KS == 0) {
        DRIFT_Pass(x, L);
        return;
    }
    delta1 = ORC_PT_CACHED(x[5], sqrt(1.0 + 2 * x[5] / GP.beta + x[5] * x[5]) - 1.0);
    gamma1 = ORC_PT_CACHED(x[5], GP.gamma + sqrt(GP.gamma * GP.gamma - 1.) * x[5]);
    beta1 = ORC_PT_DERIVED(sqrt(1.0 - 1.0 / gamma1 / gamma1));
    ps = sqrt((1 + delta1) * (1 + delta1) - x[1] * x[1] - x[3] * x[3]);
    x[1] = x[1] / ps;
    x[3] = x[3] / ps;
    vz = 2.99792458e8 * beta1 / sqrt(1.0 + x[1] * x[1] + x[3] * x[3]);
    K = GP.q * Bz / 2.0 / gamma1 / GP.m0 / vz;
    x[1] = x[1] + x[2] * K;
    x[3] = x[3] - x[0] * K;
    theta = -2.0 * K * L / Nint;
    sintheta = OSIN(theta);
    costheta = OCOS(theta);
    m44[0 * 4 + 0] = 1.0;  m44[0 * 4 + 1] = Lint * sintheta / theta;
    m44[0 * 4 + 2] = 0.;   m44[0 * 4 + 3] = -(1 - costheta) * Lint / theta;
    m44[1 * 4 + 0] = 0.;   m44[1 * 4 + 1] = costheta;
    m44[1 * 4 + 2] = 0.;   m44[1 * 4 + 3] = -sintheta;
    m44[2 * 4 + 0] = 0.;   m44[2 * 4 + 1] = (1 - costheta) * Lint / theta;
    m44[2 * 4 + 2] = 1.;   m44[2 * 4 + 3] = Lint * sintheta / theta;
    m44[3 * 4 + 0] = 0.;   m44[3 * 4 + 1] = sintheta;
    m44[3 * 4 + 2] = 0.;   m44[3 * 4 + 3] = costheta;
    for (i = 0; i < Nint; i++) {
        for (k = 0; k < 4; k++) xtemp[k] = x[k];
        for (k = 0; k < 4; k++) {
            x[k] = 0.;
            for (j = 0; j < 4; j++) x[k] = x[k] + m44[k * 4 + j] * xtemp[j];
        }
    }
    x[1] = x[1] - x[2] * K;
    x[3] = x[3] + x[0] * K;
    x[1] = gamma1 * GP.m0 * x[1] * vz / GP.p;
    x[3] = gamma1 * GP.m0 * x[3] * vz / GP.p;
    x[4] = x[4] - 2.99792458e8 * (L / vz - L / GP.beta / 2.99792458e8);
}

/* get_Axy_wiggler, simtrack.cpp:4371-4397: vector potential of the
   wiggler field and its derivatives (per unit B0 scaling as the reference) */
static void get_Axy_wiggler(double AxoBrho[], double AyoBrho[], double z, const double x[], double b0, double kx,
                            double kz, double phiz0)
{
    double ky = sqrt(kx * kx + kz * kz);
    double cx = OCOS(kx * x[0]), sx = OSIN(kx * x[0]);
    double chy = OCOSH(ky * x[2]), shy = OSINH(ky * x[2]);
    double cz = OCOS(kz * z + phiz0), sz = OSIN(kz * z + phiz0);
    AxoBrho[0] = (b0 / kz) * cx * chy * sz;
    AyoBrho[0] = (b0 * kx / ky / kz) * sx * shy * sz;
    AxoBrho[1] = -(b0 * kx / kz) * sx * chy * sz;
    AyoBrho[1] = (b0 * kx * kx / ky / kz) * cx * shy * sz;
    AxoBrho[2] = (b0 * ky / kz) * cx * shy * sz;
    AyoBrho[2] = (b0 * kx / kz) * sx * chy * sz;
    AxoBrho[3] = b0 * cx * chy * cz;
    AyoBrho[3] = (b0 * kx / ky) * sx * shy * cz;
}

static void radiate(double x[], double L, double href, const double BLbrho[]);

/* Wiggler_Pass_Forest, simtrack.cpp:4400-4452: implicit symplectic step per
   slice (WIGGLER::Pass) */
static void Wiggler_Pass_Forest(double x[], double l, int nint, double b0, double kx, double kz, double phiz0)
{
    double h = l / nint, z;
    double AxoBrho[4], AyoBrho[4], psi, hodp, det, t1, t2, x01, x02, y01, y02, x0[6], BL[3];
    for (int i = 0; i < nint; i++) {
        z = h * i;
        get_Axy_wiggler(AxoBrho, AyoBrho, z, x, b0, kx, kz, phiz0);
        for (int j = 0; j < 4; j++) {
            AxoBrho[j] = AxoBrho[j] / GP.brho;
            AyoBrho[j] = AyoBrho[j] / GP.brho;
        }
        for (int j = 0; j < 6; j++) x0[j] = x[j];
        psi = 1.0 + x0[5];
        hodp = h / psi;
        det = ((1 - AxoBrho[1] * hodp) * (1 - AyoBrho[2] * hodp) - AxoBrho[2] * AyoBrho[1] * hodp * hodp);
        x01 = AxoBrho[0] * AxoBrho[1];
        x02 = AxoBrho[0] * AxoBrho[2];
        y01 = AyoBrho[0] * AyoBrho[1];
        y02 = AyoBrho[0] * AyoBrho[2];
        x[1] = (x0[1] - (x01 + y01) * hodp) * (1 - AyoBrho[2] * hodp) +
               (x0[3] - (x02 + y02) * hodp) * AyoBrho[1] * hodp;
        x[1] = x[1] / det;
        x[3] = (x0[3] - (x02 + y02) * hodp) * (1 - AxoBrho[1] * hodp) +
               (x0[1] - (x01 + y01) * hodp) * AxoBrho[2] * hodp;
        x[3] = x[3] / det;
        x[0] = x0[0] + hodp * (x[1] - AxoBrho[0]);
        x[2] = x0[2] + hodp * (x[3] - AyoBrho[0]);
        t1 = (x[1] - AxoBrho[0]) / psi;
        t2 = (x[3] - AyoBrho[0]) / psi;
        x[4] = x[4] - h * (t1 * t1 + t2 * t2) / 2.0;
        if (GP.radiate) {
            BL[0] = -AyoBrho[3] * GP.brho * h;
            BL[1] = AxoBrho[3] * GP.brho * h;
            BL[2] = (AyoBrho[1] - AxoBrho[2]) * GP.brho * h;
            radiate(x, h, 0, BL);
        }
    }
}

/* get_Ax_Wu / get_Ay_Wu, simtrack.cpp:4456-4494 */
static void get_Ax_Wu(double *AxoBrho, double *IntpAxpydx, double z, const double x[], double b0, double kx,
                      double kz, double phiz0)
{
    double ky = sqrt(kx * kx + kz * kz);
    double cx = OCOS(kx * x[0]), sx = OSIN(kx * x[0]);
    double shy = OSINH(ky * x[2]), chy = OCOSH(ky * x[2]);
    double sz = OSIN(kz * z + phiz0);
    *AxoBrho = (b0 / kz) * cx * chy * sz;
    if (kx == 0.)
        *IntpAxpydx = (b0 * ky / kz) * x[0] * shy * sz;
    else
        *IntpAxpydx = (b0 * ky / kx / kz) * sx * shy * sz;
}
static void get_Ay_Wu(double *AyoBrho, double *IntpAypxdy, double z, const double x[], double b0, double kx,
                      double kz, double phiz0)
{
    double ky = sqrt(kx * kx + kz * kz);
    double cx = OCOS(kx * x[0]), sx = OSIN(kx * x[0]);
    double chy = OCOSH(ky * x[2]), shy = OSINH(ky * x[2]);
    double sz = OSIN(kz * z + phiz0);
    *AyoBrho = (b0 * kx / ky / kz) * sx * shy * sz;
    *IntpAypxdy = (b0 * kx * kx / ky / ky / kz) * cx * chy * sz;
}

/* get_wiggler_B, simtrack.cpp:4497-4514 */
static void get_wiggler_B(const double x[], double B[], double b0, double kx, double kz, double phiz0)
{
    double ky = sqrt(kx * kx + kz * kz);
    double cx = OCOS(kx * x[0]), sx = OSIN(kx * x[0]);
    double chy = OCOSH(ky * x[2]), shy = OSINH(ky * x[2]);
    double cz = OCOS(kz * x[4] + phiz0), sz = OSIN(kz * x[4] + phiz0);
    B[0] = (kx / ky) * b0 * sx * shy * cz;
    B[1] = b0 * cx * chy * cz;
    B[2] = -(kz * b0 / ky) * cx * shy * sz;
}

/* Wiggler_Pass_Wu, simtrack.cpp:4516-4572: second-order splitting per slice
   (WIGGLER::sPass; the spin is not rotated) */
static void Wiggler_Pass_Wu(double x[], double l, int nint, double b0, double kx, double kz, double phiz0)
{
    double h = l / nint, z = 0.0, hd, AxoBrho, AyoBrho, IntpAxpydx, IntpAypxdy, B[3];
    for (int i = 0; i < nint; i++) {
        hd = h / (1.0 + x[5]);
        z += 0.5 * h;
        get_Ay_Wu(&AyoBrho, &IntpAypxdy, z, x, b0, kx, kz, phiz0);
        x[1] = x[1] - IntpAypxdy / GP.brho;
        x[3] = x[3] - AyoBrho / GP.brho;
        x[2] = x[2] + 0.5 * hd * x[3];
        x[4] = x[4] - hd * x[3] * x[3] / (1.0 + x[5]) / (1.0 + x[5]) / 4;
        get_Ay_Wu(&AyoBrho, &IntpAypxdy, z, x, b0, kx, kz, phiz0);
        x[1] = x[1] + IntpAypxdy / GP.brho;
        x[3] = x[3] + AyoBrho / GP.brho;
        get_Ax_Wu(&AxoBrho, &IntpAxpydx, z, x, b0, kx, kz, phiz0);
        x[1] = x[1] - AxoBrho / GP.brho;
        x[3] = x[3] - IntpAxpydx / GP.brho;
        x[0] = x[0] + hd * x[1];
        x[4] = x[4] - 0.5 * hd * x[1] * x[1] / (1.0 + x[5]);
        get_Ax_Wu(&AxoBrho, &IntpAxpydx, z, x, b0, kx, kz, phiz0);
        x[1] = x[1] + AxoBrho / GP.brho;
        x[3] = x[3] + IntpAxpydx / GP.brho;
        get_Ay_Wu(&AyoBrho, &IntpAypxdy, z, x, b0, kx, kz, phiz0);
        x[1] = x[1] - IntpAypxdy / GP.brho;
        x[3] = x[3] - AyoBrho / GP.brho;
        x[2] = x[2] + 0.5 * hd * x[3];
        x[4] = x[4] - hd * x[3] * x[3] / (1.0 + x[5]) / (1.0 + x[5]) / 4;
        get_Ay_Wu(&AyoBrho, &IntpAypxdy, z, x, b0, kx, kz, phiz0);
        x[1] = x[1] + IntpAypxdy / GP.brho;
        x[3] = x[3] + AyoBrho / GP.brho;
        z += 0.5 * h;
        if (GP.radiate) {
            get_wiggler_B(x, B, b0, kx, kz, phiz0);
            for (int j = 0; j < 3; j++) B[j] = B[j] * h;
            radiate(x, h, 0.0, B);
        }
    }
}

/* bend_kick_pass, simtrack.cpp:3790-3794 */
static void bend_kick_pass(double x[], double L, double href)
{
    x[px_] = x[px_] + (href * x[pt_] - href * href * x[x_]) * L;
    x[z_] = x[z_] - href * x[x_] * L;
}

/* general_bend_kick_pass, simtrack.cpp:3796-3800 */
static void general_bend_kick_pass(double x[], double L, double href, double hreal)
{
    x[px_] = x[px_] + (-(hreal - href) + href * x[pt_] - href * hreal * x[x_]) * L;
    x[z_] = x[z_] - href * x[x_] * L;
}

/* sbend_exact_pass, simtrack.cpp:3833-3858 */
static void sbend_exact_pass(double x[], double L, double Angle, double cosAngle, double sinAngle)
{
    double href = Angle / L;
    double x0[6], pz0, pz1, delta, gamma, beta, derivative, temp, arcsin1, arcsin2;
    for (int i = 0; i < 6; i++) x0[i] = x[i];
    delta = ORC_PT_CACHED(x[pt_], sqrt(1.0 + 2 * x[pt_] / GP.beta + x[pt_] * x[pt_]) - 1.0);
    gamma = GP.gamma + sqrt(GP.gamma * GP.gamma - 1) * x[pt_];
    beta = sqrt(1.0 - 1.0 / gamma / gamma);
    pz0 = sqrt((1. + delta) * (1. + delta) - x0[1] * x0[1] - x0[3] * x0[3]);
    x[1] = x0[1] * cosAngle + (pz0 - 1. - href * x0[0]) * sinAngle;
    pz1 = sqrt((1. + delta) * (1. + delta) - x[1] * x[1] - x[3] * x[3]);
    derivative = -1.0 * x0[1] * sinAngle * href + (pz0 - 1. - x[0] * href) * cosAngle * href;
    x[0] = pz1 / href - derivative / href / href - 1.0 / href;
    temp = sqrt((1. + delta) * (1. + delta) - x0[3] * x0[3]);
    arcsin1 = OASIN(x0[1] / temp);
    arcsin2 = OASIN(x[1] / temp);
    x[2] = x0[2] + x0[3] * L + x0[3] * (arcsin1 - arcsin2) / href;
    temp = (1 + delta) * L + (1 + delta) * (arcsin1 - arcsin2) / href;
    x[4] = x[4] - (temp / beta - L / GP.beta);
}

/* general_sbend_exact_pass, simtrack.cpp:3974-4002 */
static void general_sbend_exact_pass(double x[], double L, double Angle, double K0L,
                                     double cosAngle, double sinAngle)
{
    double rhoc = L / Angle;
    double b0 = K0L / L;
    double x0[6], pz0, pz1, delta, gamma, beta, derivative, temp, arcsin1, arcsin2;
    for (int i = 0; i < 6; i++) x0[i] = x[i];
    delta = ORC_PT_CACHED(x[5], sqrt(1.0 + 2 * x[5] / GP.beta + x[5] * x[5]) - 1.0);
    gamma = GP.gamma + sqrt(GP.gamma * GP.gamma - 1) * x[5];
    beta = sqrt(1.0 - 1.0 / gamma / gamma);
    pz0 = sqrt((1. + delta) * (1. + delta) - x0[1] * x0[1] - x0[3] * x0[3]);
    x[1] = x0[1] * cosAngle + (pz0 - b0 * rhoc - b0 * x0[0]) * sinAngle;
    pz1 = sqrt((1. + delta) * (1. + delta) - x[1] * x[1] - x[3] * x[3]);
    derivative = -x0[1] * sinAngle / rhoc + (pz0 - b0 * rhoc - b0 * x0[0]) * cosAngle / rhoc;
    x[0] = pz1 / b0 - derivative * rhoc / b0 - rhoc;
    temp = sqrt((1. + delta) * (1. + delta) - x0[3] * x0[3]);
    arcsin1 = OASIN(x0[1] / temp);
    arcsin2 = OASIN(x[1] / temp);
    x[2] = x0[2] + x0[3] * L / b0 / rhoc + x0[3] * (arcsin1 - arcsin2) / b0;
    temp = (1 + delta) * L / b0 / rhoc + (1 + delta) * (arcsin1 - arcsin2) / b0;
    x[4] = x[4] - (temp / beta - L / GP.beta);
}

/* quad/sext/oct/mult kicks, simtrack.cpp:4046-4111 */
static void quad_kick_pass(double x[], double k1l, double k1sl)
{
    double preal = x[0], pimag = x[2];
    double Bx = k1sl * preal + k1l * pimag;
    double By = k1l * preal - k1sl * pimag;
    x[px_] = x[px_] - By;
    x[py_] = x[py_] + Bx;
}

static void sext_kick_pass(double x[], double k2l, double k2sl)
{
    double preal = x[0] * x[0] - x[2] * x[2];
    double pimag = 2. * x[0] * x[2];
    double Bx = (k2sl * preal + k2l * pimag) / 2.;
    double By = (k2l * preal - k2sl * pimag) / 2.;
    x[px_] = x[px_] - By;
    x[py_] = x[py_] + Bx;
}

static void oct_kick_pass(double x[], double k3l, double k3sl)
{
    double preal = x[0] * x[0] * x[0] - 3 * x[0] * x[2] * x[2];
    double pimag = 3 * x[0] * x[0] * x[2] - x[2] * x[2] * x[2];
    double Bx = (k3sl * preal + k3l * pimag) / 6.;
    double By = (k3l * preal - k3sl * pimag) / 6.;
    x[px_] = x[px_] - By;
    x[py_] = x[py_] + Bx;
}

static void mult_kick_pass(double x[], int Norder, const double KNL[11], const double KNSL[11])
{
    int fac = 1;
    double Xn, Yn, Xn0, Yn0, By, Bx;
    By = KNL[0];
    Bx = KNSL[0];
    Xn = 1.;
    Yn = 0.;
    for (int i = 1; i < Norder + 1; i++) {
        Xn0 = Xn;
        Yn0 = Yn;
        Xn = Xn0 * x[x_] - Yn0 * x[y_];
        Yn = Xn0 * x[y_] + Yn0 * x[x_];
        fac = fac * i;
        if (KNL[i] != 0. || KNSL[i] != 0.) {
            By = By + (KNL[i] * Xn - KNSL[i] * Yn) / fac;
            Bx = Bx + (KNL[i] * Yn + KNSL[i] * Xn) / fac;
        }
    }
    x[px_] = x[px_] - By;
    x[py_] = x[py_] + Bx;
}

/* bend_mult_kick_pass, simtrack.cpp:4113-4137 */
static void bend_mult_kick_pass(double x[], double L, double href, int Norder,
                                const double KNL[11], const double KNSL[11])
{
    int fac = 1;
    double Xn, Yn, Xn0, Yn0, By, Bx;
    By = KNL[0];
    Bx = KNSL[0];
    Xn = 1.;
    Yn = 0.;
    for (int i = 1; i < Norder + 1; i++) {
        Xn0 = Xn;
        Yn0 = Yn;
        Xn = Xn0 * x[x_] - Yn0 * x[y_];
        Yn = Xn0 * x[y_] + Yn0 * x[x_];
        fac = fac * i;
        if (KNL[i] != 0. || KNSL[i] != 0.) {
            By = By + (KNL[i] * Xn - KNSL[i] * Yn) / fac;
            Bx = Bx + (KNL[i] * Yn + KNSL[i] * Xn) / fac;
        }
    }
    x[px_] = x[px_] - By + (href * x[pt_] - href * href * x[x_]) * L;
    x[py_] = x[py_] + Bx;
    x[z_] = x[z_] - href * x[x_] * L;
}

/* general_bend_mult_kick_pass, simtrack.cpp:4139-4163 */
static void general_bend_mult_kick_pass(double x[], double L, double href, double hreal, int Norder,
                                        const double KNL[11], const double KNSL[11])
{
    int fac = 1;
    double Xn, Yn, Xn0, Yn0, By, Bx;
    By = KNL[0];
    Bx = KNSL[0];
    Xn = 1.;
    Yn = 0.;
    for (int i = 1; i < Norder + 1; i++) {
        Xn0 = Xn;
        Yn0 = Yn;
        Xn = Xn0 * x[x_] - Yn0 * x[y_];
        Yn = Xn0 * x[y_] + Yn0 * x[x_];
        fac = fac * i;
        if (KNL[i] != 0. || KNSL[i] != 0.) {
            By = By + (KNL[i] * Xn - KNSL[i] * Yn) / fac;
            Bx = Bx + (KNL[i] * Yn + KNSL[i] * Xn) / fac;
        }
    }
    x[px_] = x[px_] - By + (-(hreal - href) + href * x[pt_] - href * hreal * x[x_]) * L;
    x[py_] = x[py_] + Bx;
    x[z_] = x[z_] - href * x[x_] * L;
}

/* cal_Bfield, simtrack.cpp:4169-4196 */
static void cal_Bfield(const double x[], int Nint, int Norder, double Angle, const double KNL[11],
                       const double KNSL[11], double BLbrho[])
{
    int fac = 1;
    double Xn, Yn, Xn0, Yn0, By, Bx;
    By = KNL[0];
    Bx = KNSL[0];
    Xn = 1.;
    Yn = 0.;
    for (int i = 1; i < Norder + 1; i++) {
        Xn0 = Xn;
        Yn0 = Yn;
        Xn = Xn0 * x[x_] - Yn0 * x[y_];
        Yn = Xn0 * x[y_] + Yn0 * x[x_];
        fac = fac * i;
        if (KNL[i] != 0. || KNSL[i] != 0.) {
            By = By + (KNL[i] * Xn - KNSL[i] * Yn) / fac;
            Bx = Bx + (KNL[i] * Yn + KNSL[i] * Xn) / fac;
        }
    }
    BLbrho[0] = Bx / Nint;
    BLbrho[1] = (Angle + By) / Nint;
    BLbrho[2] = 0.;
}

/* vector_cross_product, simtrack.cpp (template near cal_Bfield) */
static void vector_cross_product(const double a[], const double b[], double c[])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = -1.0 * a[0] * b[2] + a[2] * b[0];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* ------------------------------------------------------- radiation */

/* Per-particle normal stream replacing gsl_ran_gaussian(gsl_r2, 1.0). */
typedef struct { uint64_t seed; uint32_t particle; uint64_t counter; } orc_rng;

/* radiate, simtrack.cpp:4198-4234 */
static void radiate(double x[], double L, double href, const double BLbrho[])
{
    double x0[6], delta1, gamma1, beta1, delta2, pz, xp, yp, temp1, temp2, ratio;
    double u[3], BLper[3], BLper2;
    for (int i = 0; i < 6; i++) x0[i] = x[i];
    delta1 = ORC_PT_CACHED(x[pt_], sqrt(1.0 + 2 * x[pt_] / GP.beta + x[pt_] * x[pt_]) - 1.0);
    gamma1 = ORC_PT_CACHED(x[pt_], GP.gamma + sqrt(GP.gamma * GP.gamma - 1) * x[pt_]);
    beta1 = ORC_PT_DERIVED(sqrt(1.0 - 1.0 / gamma1 / gamma1));
    (void)beta1;
    pz = sqrt((1 + delta1) * (1 + delta1) - x[px_] * x[px_] - x[py_] * x[py_]);
    xp = (1 + x[x_] * href) * x[px_] / pz;
    yp = (1 + x[x_] * href) * x[py_] / pz;
    temp1 = sqrt((1 + x[x_] * href) * (1 + x[x_] * href) + xp * xp + yp * yp);
    u[0] = xp / temp1;
    u[1] = yp / temp1;
    u[2] = (1 + x[x_] * href) / temp1;
    vector_cross_product(BLbrho, u, BLper);
    BLper2 = BLper[0] * BLper[0] + BLper[1] * BLper[1] + BLper[2] * BLper[2];
    temp2 = Cr * (1 + delta1) * (1 + delta1) * OPOW(GP.energy / 1000., 3) / 2 / PI;
    x[pt_] = x[pt_] - temp2 * BLper2 * temp1 / L;
    delta2 = ORC_PT_CACHED(x[pt_], sqrt(1.0 + 2 * x[pt_] / GP.beta + x[pt_] * x[pt_]) - 1.0);
    ratio = (1 + delta2) / (1 + delta1);
    x[px_] = x0[px_] * ratio;
    x[py_] = x0[py_] * ratio;
}

/* radiate1, simtrack.cpp:4236-4288 */
static void radiate1(double x[], double L, double href, const double BLbrho[], orc_rng *rng)
{
    double x0[6], delta1, gamma1, beta1, delta2, pz, xp, yp, temp1, temp2, ratio;
    double u[3], BLper[3], BLper2, g1, dtdl, ptloss, ptfluct, randgaus;
    for (int i = 0; i < 6; i++) x0[i] = x[i];
    delta1 = ORC_PT_CACHED(x[pt_], sqrt(1.0 + 2 * x[pt_] / GP.beta + x[pt_] * x[pt_]) - 1.0);
    gamma1 = ORC_PT_CACHED(x[pt_], GP.gamma + sqrt(GP.gamma * GP.gamma - 1) * x[pt_]);
    beta1 = ORC_PT_DERIVED(sqrt(1.0 - 1.0 / gamma1 / gamma1));
    pz = sqrt((1 + delta1) * (1 + delta1) - x[px_] * x[px_] - x[py_] * x[py_]);
    xp = (1 + x[x_] * href) * x[px_] / pz;
    yp = (1 + x[x_] * href) * x[py_] / pz;
    temp1 = sqrt((1 + x[x_] * href) * (1 + x[x_] * href) + xp * xp + yp * yp);
    u[0] = xp / temp1;
    u[1] = yp / temp1;
    u[2] = (1 + x[x_] * href) / temp1;
    vector_cross_product(BLbrho, u, BLper);
    BLper2 = BLper[0] * BLper[0] + BLper[1] * BLper[1] + BLper[2] * BLper[2];
    temp2 = Cr * (1 + delta1) * (1 + delta1) * OPOW(GP.energy / 1000., 3) / 2 / PI;
    ptloss = temp2 * BLper2 * temp1 / L;
    dtdl = temp1 / beta1 / speed_light;
    g1 = sqrt(BLper2 / L / L) / (1 + delta1);
    randgaus = st_rng_normal(rng->seed, rng->particle, rng->counter);
    rng->counter++;
    ptfluct = sqrt(Cf * gamma1 * gamma1 * gamma1 * gamma1 * gamma1 * g1 * g1 * g1 * dtdl * L) * randgaus;
    x[pt_] = x[pt_] - ptloss + ptfluct;
    delta2 = ORC_PT_CACHED(x[pt_], sqrt(1.0 + 2 * x[pt_] / GP.beta + x[pt_] * x[pt_]) - 1.0);
    ratio = (1.0 + delta2) / (1 + delta1);
    x[px_] = x0[px_] * ratio;
    x[py_] = x0[py_] * ratio;
}

/*
 * radiate2, simtrack.cpp:4290-4366.  gsl_ran_poisson(gsl_r1, Nr) and the
 * gsl_rng_uniform(gsl_r2) photon draws are replaced by st_rad2_photons: a
 * Poisson count by Knuth's product of uniforms (exact, in chunks of mean
 * <= 30) and one uniform per photon, all from the sub-stream of one draw
 * number, so that every particle's photons are independent of scheduling.
 * The photon-energy table and its linear interpolation (splint, cpp:333)
 * are the reference's.  Kept as the reference has them: yp from x[1]
 * (cpp:4331) and rho1 from 1 + x[5] (cpp:4339).
 */
static void radiate2(double x[], double L, double href, const double BLbrho[], orc_rng *rng)
{
    double alpha = 1.0 / 137.0359997;
    double x0[6], delta1, gamma1, beta1, delta2, pz, xp, yp, temp1, ratio;
    double u[3], BLper[3], BLper2, rho1, Nr, Etot, Ec;
    for (int i = 0; i < 6; i++) x0[i] = x[i];
    delta1 = ORC_PT_CACHED(x[5], sqrt(1.0 + 2 * x[5] / GP.beta + x[5] * x[5]) - 1.0);
    gamma1 = ORC_PT_CACHED(x[5], GP.gamma + sqrt(GP.gamma * GP.gamma - 1) * x[5]);
    beta1 = ORC_PT_DERIVED(sqrt(1.0 - 1.0 / gamma1 / gamma1));
    pz = sqrt((1 + delta1) * (1 + delta1) - x[1] * x[1] - x[3] * x[3]);
    xp = (1 + x[0] * href) * x[1] / pz;
    yp = (1 + x[0] * href) * x[1] / pz;
    temp1 = sqrt((1 + x[0] * href) * (1 + x[0] * href) + xp * xp + yp * yp);
    u[0] = xp / temp1;
    u[1] = yp / temp1;
    u[2] = (1 + x[0] * href) / temp1;
    vector_cross_product(BLbrho, u, BLper);
    BLper2 = BLper[0] * BLper[0] + BLper[1] * BLper[1] + BLper[2] * BLper[2];
    rho1 = (1 + x[5]) / sqrt(BLper2) * L;
    Ec = 1.5 * 6.582119e-16 * 2.99792458e8 * gamma1 * gamma1 * gamma1 / rho1 / 1.0e6;
    Nr = 5. * alpha * gamma1 / 2. / sqrt(3.) / rho1 * L * temp1 / beta1;
    Etot = st_rad2_photons(rng->seed, rng->particle, rng->counter, Nr);
    rng->counter++;
    Etot = Etot * Ec;
    x[5] = x[5] - Etot / GP.energy;
    delta2 = ORC_PT_CACHED(x[5], sqrt(1.0 + 2 * x[5] / GP.beta + x[5] * x[5]) - 1.0);
    ratio = (1 + delta2) / (1 + delta1);
    x[1] = x0[1] * ratio;
    x[3] = x0[3] * ratio;
}

/* ------------------------------------------------------------- spin */

/* spin(), simtrack.cpp:4574-4630 */
static void spin(double x[], double angle, double href, const double BLbrho[])
{
    double u[3], Bpar[3], udotB, temp;
    double delta1 = ORC_PT_CACHED(x[pt_], sqrt(1.0 + 2. * x[pt_] / GP.beta + x[pt_] * x[pt_]) - 1.0);
    double gamma1 = ORC_PT_CACHED(x[pt_], GP.gamma + sqrt(GP.gamma * GP.gamma - 1.) * x[pt_]);
    double Gr = GP.G * gamma1;
    double F[3], Fabs, uF[3], theta, costheta, sintheta, rot[9], spin0[3], sp[3];
    if (BLbrho[0] * BLbrho[0] + BLbrho[1] * BLbrho[1] + BLbrho[2] * BLbrho[2] != 0.) {
        for (int i = 0; i < 3; i++) spin0[i] = x[6 + i];
        temp = sqrt((1 + delta1) * (1 + delta1) - x[px_] * x[px_] - x[py_] * x[py_]);
        u[0] = x[px_] / temp;
        u[1] = x[py_] / temp;
        u[2] = 1.0;
        temp = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (int i = 0; i < 3; i++) u[i] = u[i] / temp;
        temp = (1 + href * x[0]) * temp;
        udotB = u[0] * BLbrho[0] + u[1] * BLbrho[1] + u[2] * BLbrho[2];
        for (int i = 0; i < 3; i++) Bpar[i] = udotB * u[i];
        for (int i = 0; i < 3; i++) F[i] = (1 + Gr) * BLbrho[i] + GP.G * (1 - gamma1) * Bpar[i];
        for (int i = 0; i < 3; i++) F[i] = F[i] * temp / (1 + delta1);
        F[1] = F[1] - angle;
        Fabs = sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
        for (int i = 0; i < 3; i++) uF[i] = F[i] / Fabs;
        theta = -Fabs;
        costheta = OCOS(theta);
        sintheta = OSIN(theta);
        rot[0] = uF[0] * uF[0] * (1 - costheta) + costheta;
        rot[1] = uF[0] * uF[1] * (1 - costheta) - uF[2] * sintheta;
        rot[2] = uF[0] * uF[2] * (1 - costheta) + uF[1] * sintheta;
        rot[3] = uF[0] * uF[1] * (1 - costheta) + uF[2] * sintheta;
        rot[4] = uF[1] * uF[1] * (1 - costheta) + costheta;
        rot[5] = uF[1] * uF[2] * (1 - costheta) - uF[0] * sintheta;
        rot[6] = uF[0] * uF[2] * (1 - costheta) - uF[1] * sintheta;
        rot[7] = uF[1] * uF[2] * (1 - costheta) + uF[0] * sintheta;
        rot[8] = uF[2] * uF[2] * (1 - costheta) + costheta;
        for (int i = 0; i < 3; i++) {
            sp[i] = 0.;
            for (int j = 0; j < 3; j++) sp[i] = sp[i] + rot[i * 3 + j] * spin0[j];
        }
        for (int i = 0; i < 3; i++) x[6 + i] = sp[i];
    }
}

/* ------------------------------------------------- thick elements */

static void edge(double x[], double E, double h)
{
    double edgefocus = OTAN(E) * h;
    x[1] = x[1] + edgefocus * x[0];
    x[3] = x[3] - edgefocus * x[2];
}

/* SBEND_Pass, simtrack.cpp:4762-4816 */
static void SBEND_Pass(double x[], double L, int Nint, double Angle, double E1, double E2, orc_rng *rng)
{
    double href = Angle / L, Lint = L / Nint, cosAngle, sinAngle, BLbrho[3];
    if (Angle == 0.) {
        DRIFT_Pass(x, L);
        return;
    }
    edge(x, E1, href);
    if (GP.H_expand && !GP.radiate) {
        for (int i = 0; i < Nint; i++) {
            DRIFT_Pass(x, Fdrift1 * Lint);
            bend_kick_pass(x, Fkick1 * Lint, href);
            DRIFT_Pass(x, Fdrift2 * Lint);
            bend_kick_pass(x, Fkick2 * Lint, href);
            DRIFT_Pass(x, Fdrift2 * Lint);
            bend_kick_pass(x, Fkick1 * Lint, href);
            DRIFT_Pass(x, Fdrift1 * Lint);
        }
    }
    if (GP.H_expand && GP.radiate) {
        BLbrho[0] = 0.;
        BLbrho[1] = Angle / Nint;
        BLbrho[2] = 0.;
        for (int i = 0; i < Nint; i++) {
            DRIFT_Pass(x, Fdrift1 * Lint);
            bend_kick_pass(x, Fkick1 * Lint, href);
            DRIFT_Pass(x, Fdrift2 * Lint);
            bend_kick_pass(x, Fkick2 * Lint, href);
            DRIFT_Pass(x, Fdrift2 * Lint);
            bend_kick_pass(x, Fkick1 * Lint, href);
            DRIFT_Pass(x, Fdrift1 * Lint);
            radiate1(x, Lint, href, BLbrho, rng);
        }
    }
    if (!GP.H_expand && !GP.radiate)
        sbend_exact_pass(x, L, Angle, OCOS(Angle), OSIN(Angle));
    if (!GP.H_expand && GP.radiate) {
        BLbrho[0] = 0.;
        BLbrho[1] = Angle / Nint;
        BLbrho[2] = 0.;
        cosAngle = OCOS(Angle / Nint);
        sinAngle = OSIN(Angle / Nint);
        for (int i = 0; i < Nint; i++) {
            sbend_exact_pass(x, Lint, Angle / Nint, cosAngle, sinAngle);
            radiate1(x, Lint, href, BLbrho, rng);
        }
    }
    edge(x, E2, href);
}

/* GSBEND_Pass, simtrack.cpp:4818-4880 */
static void GSBEND_Pass(double x[], double L, int Nint, double Angle, double K0L, double E1, double E2, orc_rng *rng)
{
    double href = Angle / L, hreal = K0L / L, Lint = L / Nint, BLbrho[3];
    if (Angle == 0.) {
        DRIFT_Pass(x, L);
        return;
    }
    edge(x, E1, hreal);
    if (GP.H_expand) {
        BLbrho[0] = 0.;
        BLbrho[1] = K0L / Nint;
        BLbrho[2] = 0.;
        for (int i = 0; i < Nint; i++) {
            DRIFT_Pass(x, Fdrift1 * Lint);
            general_bend_kick_pass(x, Fkick1 * Lint, href, hreal);
            DRIFT_Pass(x, Fdrift2 * Lint);
            general_bend_kick_pass(x, Fkick2 * Lint, href, hreal);
            DRIFT_Pass(x, Fdrift2 * Lint);
            general_bend_kick_pass(x, Fkick1 * Lint, href, hreal);
            DRIFT_Pass(x, Fdrift1 * Lint);
            if (GP.radiate) radiate2(x, Lint, href, BLbrho, rng);
        }
    } else if (!GP.radiate) {
        general_sbend_exact_pass(x, L, Angle, K0L, OCOS(Angle), OSIN(Angle));
    } else {
        BLbrho[0] = 0.;
        BLbrho[1] = K0L / Nint;
        BLbrho[2] = 0.;
        double cosAngle = OCOS(Angle / Nint), sinAngle = OSIN(Angle / Nint);
        for (int i = 0; i < Nint; i++) {
            general_sbend_exact_pass(x, Lint, Angle / Nint, K0L / Nint, cosAngle, sinAngle);
            radiate2(x, Lint, href, BLbrho, rng);
        }
    }
    edge(x, E2, hreal);
}

/* QUAD_Pass, simtrack.cpp:4882-4959 */
static void QUAD_Pass(double x[], double L, int Nint, double k1l, double k1sl)
{
    if (k1l == 0. && k1sl == 0.) {
        DRIFT_Pass(x, L);
        return;
    }
    if (L == 0.) {
        quad_kick_pass(x, k1l, k1sl);
        return;
    }
    if (L != 0. && GP.quad_fringe) {
        double x0[6], k = k1l / L;
        for (int i = 0; i < 6; i++) x0[i] = x[i];
        x[0] = x0[0] + k * (x0[0] * x0[0] * x0[0] + 3 * x0[0] * x0[2] * x0[2]) / 12 / (1 + x0[5]);
        x[1] = x0[1] - k * ((x0[0] * x0[0] + x0[2] * x0[2]) * x0[1] - 2 * x0[0] * x0[2] * x0[3]) / 4 / (1 + x0[5]);
        x[2] = x0[2] - k * (3 * x0[0] * x0[0] * x0[2] + x0[2] * x0[2] * x0[2]) / 12 / (1 + x0[5]);
        x[3] = x0[3] + k * ((x0[0] * x0[0] + x0[2] * x0[2]) * x0[3] - 2 * x0[0] * x0[2] * x0[1]) / 4 / (1 + x0[5]);
        x[4] = x0[4] + k * (x0[2] * x0[2] * x0[2] * x0[3] - x0[0] * x0[0] * x0[0] * x0[1] +
                            3 * x0[0] * x0[0] * x0[2] * x0[3] - 3 * x0[2] * x0[2] * x0[0] * x0[1]) /
                           12 / (1 + x0[5]) / (1 + x0[5]);
    }
    double Lint = L / Nint;
    double k1l_kick1 = Fkick1 * k1l / Nint, k1sl_kick1 = Fkick1 * k1sl / Nint;
    double k1l_kick2 = Fkick2 * k1l / Nint, k1sl_kick2 = Fkick2 * k1sl / Nint;
    for (int i = 0; i < Nint; i++) {
        DRIFT_Pass(x, Fdrift1 * Lint);
        quad_kick_pass(x, k1l_kick1, k1sl_kick1);
        DRIFT_Pass(x, Fdrift2 * Lint);
        quad_kick_pass(x, k1l_kick2, k1sl_kick2);
        DRIFT_Pass(x, Fdrift2 * Lint);
        quad_kick_pass(x, k1l_kick1, k1sl_kick1);
        DRIFT_Pass(x, Fdrift1 * Lint);
    }
    if (L != 0. && GP.quad_fringe) {
        double x0[6], k = k1l / L;
        for (int i = 0; i < 6; i++) x0[i] = x[i];
        x[0] = x0[0] - k * (x0[0] * x0[0] * x0[0] + 3 * x0[0] * x0[2] * x0[2]) / 12 / (1 + x0[5]);
        x[1] = x0[1] + k * ((x0[0] * x0[0] + x0[2] * x0[2]) * x0[1] - 2 * x0[0] * x0[2] * x0[3]) / 4 / (1 + x0[5]);
        x[2] = x0[2] + k * (3 * x0[0] * x0[0] * x0[2] + x0[2] * x0[2] * x0[2]) / 12 / (1 + x0[5]);
        x[3] = x0[3] - k * ((x0[0] * x0[0] + x0[2] * x0[2]) * x0[3] - 2 * x0[0] * x0[2] * x0[1]) / 4 / (1 + x0[5]);
        x[4] = x0[4] - k * (x0[2] * x0[2] * x0[2] * x0[3] - x0[0] * x0[0] * x0[0] * x0[1] +
                            3 * x0[0] * x0[0] * x0[2] * x0[3] - 3 * x0[2] * x0[2] * x0[0] * x0[1]) /
                           12 / (1 + x0[5]) / (1 + x0[5]);
    }
}

/* SEXT_Pass, simtrack.cpp:5004-5034 */
static void SEXT_Pass(double x[], double L, int Nint, double k2l, double k2sl)
{
    if (k2l == 0. && k2sl == 0.) {
        DRIFT_Pass(x, L);
        return;
    }
    if (L == 0.) {
        sext_kick_pass(x, k2l, k2sl);
        return;
    }
    double Lint = L / Nint;
    double k2l_kick1 = Fkick1 * k2l / Nint, k2sl_kick1 = Fkick1 * k2sl / Nint;
    double k2l_kick2 = Fkick2 * k2l / Nint, k2sl_kick2 = Fkick2 * k2sl / Nint;
    for (int i = 0; i < Nint; i++) {
        DRIFT_Pass(x, Fdrift1 * Lint);
        sext_kick_pass(x, k2l_kick1, k2sl_kick1);
        DRIFT_Pass(x, Fdrift2 * Lint);
        sext_kick_pass(x, k2l_kick2, k2sl_kick2);
        DRIFT_Pass(x, Fdrift2 * Lint);
        sext_kick_pass(x, k2l_kick1, k2sl_kick1);
        DRIFT_Pass(x, Fdrift1 * Lint);
    }
}

/* OCT_Pass, simtrack.cpp:5077-5107 */
static void OCT_Pass(double x[], double L, int Nint, double k3l, double k3sl)
{
    if (k3l == 0. && k3sl == 0.) {
        DRIFT_Pass(x, L);
        return;
    }
    if (L == 0.) {
        oct_kick_pass(x, k3l, k3sl);
        return;
    }
    double Lint = L / Nint;
    double k3l_kick1 = Fkick1 * k3l / Nint, k3sl_kick1 = Fkick1 * k3sl / Nint;
    double k3l_kick2 = Fkick2 * k3l / Nint, k3sl_kick2 = Fkick2 * k3sl / Nint;
    for (int i = 0; i < Nint; i++) {
        DRIFT_Pass(x, Fdrift1 * Lint);
        oct_kick_pass(x, k3l_kick1, k3sl_kick1);
        DRIFT_Pass(x, Fdrift2 * Lint);
        oct_kick_pass(x, k3l_kick2, k3sl_kick2);
        DRIFT_Pass(x, Fdrift2 * Lint);
        oct_kick_pass(x, k3l_kick1, k3sl_kick1);
        DRIFT_Pass(x, Fdrift1 * Lint);
    }
}

static void kick_coefs(const double KNL[11], const double KNSL[11], int Nint, double k1[11], double s1[11],
                       double k2[11], double s2[11])
{
    for (int i = 0; i < 11; i++) {
        k1[i] = Fkick1 * KNL[i] / Nint;
        s1[i] = Fkick1 * KNSL[i] / Nint;
    }
    for (int i = 0; i < 11; i++) {
        k2[i] = Fkick2 * KNL[i] / Nint;
        s2[i] = Fkick2 * KNSL[i] / Nint;
    }
}

/* MULT_Pass, simtrack.cpp:5152-5182 */
static void MULT_Pass(double x[], double L, int Nint, int Norder, const double KNL[11], const double KNSL[11])
{
    double k1[11], s1[11], k2[11], s2[11];
    if (L == 0.) {
        mult_kick_pass(x, Norder, KNL, KNSL);
        return;
    }
    double Lint = L / Nint;
    kick_coefs(KNL, KNSL, Nint, k1, s1, k2, s2);
    for (int i = 0; i < Nint; i++) {
        DRIFT_Pass(x, Fdrift1 * Lint);
        mult_kick_pass(x, Norder, k1, s1);
        DRIFT_Pass(x, Fdrift2 * Lint);
        mult_kick_pass(x, Norder, k2, s2);
        DRIFT_Pass(x, Fdrift2 * Lint);
        mult_kick_pass(x, Norder, k1, s1);
        DRIFT_Pass(x, Fdrift1 * Lint);
    }
}

/*
 * GMULT_Pass (simtrack.cpp:5223-5346), SBENDMULT_Pass (cpp:5348-5470) and
 * SMULT_Pass (cpp:5472-5595) have identical bodies.
 */
static void BENDMULT_Pass(double x[], double L, int Nint, int Norder, double Angle, double E1, double E2,
                          const double KNL[11], const double KNSL[11])
{
    double k1[11], s1[11], k2[11], s2[11], x1[6], BLbrho[3];
    double cosAngle, sinAngle, cosAngle1, sinAngle1;
    if (L == 0.) {
        mult_kick_pass(x, Norder, KNL, KNSL);
        return;
    }
    double href = Angle / L, Lint = L / Nint;
    edge(x, E1, href);
    kick_coefs(KNL, KNSL, Nint, k1, s1, k2, s2);
    if (GP.H_expand) {
        for (int i = 0; i < Nint; i++) {
            for (int j = 0; j < 6; j++) x1[j] = x[j];
            DRIFT_Pass(x, Fdrift1 * Lint);
            bend_mult_kick_pass(x, Fkick1 * Lint, href, Norder, k1, s1);
            DRIFT_Pass(x, Fdrift2 * Lint);
            bend_mult_kick_pass(x, Fkick2 * Lint, href, Norder, k2, s2);
            DRIFT_Pass(x, Fdrift2 * Lint);
            bend_mult_kick_pass(x, Fkick1 * Lint, href, Norder, k1, s1);
            DRIFT_Pass(x, Fdrift1 * Lint);
            if (GP.radiate) {
                for (int j = 0; j < 6; j++) x1[j] = (x[j] + x1[j]) / 2.;
                cal_Bfield(x1, Nint, Norder, Angle, KNL, KNSL, BLbrho);
                radiate(x, Lint, href, BLbrho);
            }
        }
    } else {
        cosAngle1 = OCOS(Fdrift1 * Angle / Nint);
        sinAngle1 = OSIN(Fdrift1 * Angle / Nint);
        cosAngle = OCOS(Fdrift2 * Angle / Nint);
        sinAngle = OSIN(Fdrift2 * Angle / Nint);
        for (int i = 0; i < Nint; i++) {
            for (int j = 0; j < 6; j++) x1[j] = x[j];
            sbend_exact_pass(x, Fdrift1 * Lint, Fdrift1 * Angle / Nint, cosAngle1, sinAngle1);
            mult_kick_pass(x, Norder, k1, s1);
            sbend_exact_pass(x, Fdrift2 * Lint, Fdrift2 * Angle / Nint, cosAngle, sinAngle);
            mult_kick_pass(x, Norder, k2, s2);
            sbend_exact_pass(x, Fdrift2 * Lint, Fdrift2 * Angle / Nint, cosAngle, sinAngle);
            mult_kick_pass(x, Norder, k1, s1);
            sbend_exact_pass(x, Fdrift1 * Lint, Fdrift1 * Angle / Nint, cosAngle1, sinAngle1);
            if (GP.radiate) {
                for (int j = 0; j < 6; j++) x1[j] = (x[j] + x1[j]) / 2.;
                cal_Bfield(x1, Nint, Norder, Angle, KNL, KNSL, BLbrho);
                radiate(x, Lint, href, BLbrho);
            }
        }
    }
    edge(x, E2, href);
}

/* GSBENDMULT_Pass, simtrack.cpp:5597-5723 */
static void GSBENDMULT_Pass(double x[], double L, int Nint, int Norder, double Angle, double K0L, double E1,
                            double E2, const double KNL[11], const double KNSL[11])
{
    double k1[11], s1[11], k2[11], s2[11], BLbrho[3];
    double cosAngle, sinAngle, cosAngle1, sinAngle1;
    if (L == 0.) {
        mult_kick_pass(x, Norder, KNL, KNSL);
        return;
    }
    double href = Angle / L, hreal = K0L / L, Lint = L / Nint;
    edge(x, E1, hreal);
    kick_coefs(KNL, KNSL, Nint, k1, s1, k2, s2);
    if (GP.H_expand) {
        for (int i = 0; i < Nint; i++) {
            DRIFT_Pass(x, Fdrift1 * Lint);
            general_bend_mult_kick_pass(x, Fkick1 * Lint, href, hreal, Norder, k1, s1);
            DRIFT_Pass(x, Fdrift2 * Lint);
            general_bend_mult_kick_pass(x, Fkick2 * Lint, href, hreal, Norder, k2, s2);
            DRIFT_Pass(x, Fdrift2 * Lint);
            general_bend_mult_kick_pass(x, Fkick1 * Lint, href, hreal, Norder, k1, s1);
            DRIFT_Pass(x, Fdrift1 * Lint);
            if (GP.radiate) {
                BLbrho[0] = 0.;
                BLbrho[1] = K0L / Nint;
                BLbrho[2] = 0.;
                radiate(x, Lint, href, BLbrho);
            }
        }
    } else {
        cosAngle1 = OCOS(Fdrift1 * Angle / Nint);
        sinAngle1 = OSIN(Fdrift1 * Angle / Nint);
        cosAngle = OCOS(Fdrift2 * Angle / Nint);
        sinAngle = OSIN(Fdrift2 * Angle / Nint);
        for (int i = 0; i < Nint; i++) {
            general_sbend_exact_pass(x, Fdrift1 * Lint, Fdrift1 * Angle / Nint, Fdrift1 * K0L / Nint, cosAngle1, sinAngle1);
            mult_kick_pass(x, Norder, k1, s1);
            general_sbend_exact_pass(x, Fdrift2 * Lint, Fdrift2 * Angle / Nint, Fdrift2 * K0L / Nint, cosAngle, sinAngle);
            mult_kick_pass(x, Norder, k2, s2);
            general_sbend_exact_pass(x, Fdrift2 * Lint, Fdrift2 * Angle / Nint, Fdrift2 * K0L / Nint, cosAngle, sinAngle);
            mult_kick_pass(x, Norder, k1, s1);
            general_sbend_exact_pass(x, Fdrift1 * Lint, Fdrift1 * Angle / Nint, Fdrift1 * K0L / Nint, cosAngle1, sinAngle1);
            if (GP.radiate) {
                BLbrho[0] = 0.;
                BLbrho[1] = K0L / Nint;
                BLbrho[2] = 0.;
                radiate(x, Lint, href, BLbrho);
            }
        }
    }
    edge(x, E2, hreal);
}

/* ------------------------------------------------------ thin maps */

/* MATRIX_Pass, simtrack.cpp:5873-5885 */
static void MATRIX_Pass(double x[], const double XCO_IN[6], const double XCO_OUT[6], const double M66[36])
{
    double x1[6];
    for (int i = 0; i < 6; i++) x[i] = x[i] - XCO_IN[i];
    for (int i = 0; i < 6; i++) {
        x1[i] = 0.0;
        for (int j = 0; j < 6; j++) x1[i] = x1[i] + M66[i * 6 + j] * x[j];
    }
    for (int i = 0; i < 6; i++) x[i] = x1[i];
    for (int i = 0; i < 6; i++) x[i] = x[i] + XCO_OUT[i];
}

/* KICK_Pass, simtrack.cpp:5887-5893 */
static void KICK_Pass(double x[], double L, double HKICK, double VKICK)
{
    DRIFT_Pass(x, L / 2.0);
    x[1] = x[1] + HKICK;
    x[3] = x[3] + VKICK;
    DRIFT_Pass(x, L / 2.0);
}

/* RFCAV_Pass, simtrack.cpp:6002-6012 */
static void RFCAV_Pass(double x[], double L, double VRF, double FRF, double PHASE0)
{
    if (VRF == 0.) {
        DRIFT_Pass(x, L);
        return;
    }
    DRIFT_Pass(x, L / 2.);
    x[pt_] = x[pt_] + fabs(VRF * GP.Q) * OSIN(-2.0 * PI * FRF * x[z_] / 2.99792458e8 + PI - PHASE0) / GP.beta / GP.energy;
    DRIFT_Pass(x, L / 2.);
}

/* CRABRF_Pass, simtrack.cpp:6014-6025 */
static void CRABRF_Pass(double x[], double L, double VRF, double FRF, double PHASE0)
{
    if (VRF == 0.) {
        DRIFT_Pass(x, L);
    } else {
        DRIFT_Pass(x, L / 2.);
        x[px_] = x[px_] - (VRF * GP.Q) * OSIN(2.0 * PI * FRF * x[z_] / 2.99792458e8 + PHASE0) / GP.beta / GP.energy;
        x[pt_] = x[pt_] - (VRF * GP.Q) * OCOS(2.0 * PI * FRF * x[z_] / 2.99792458e8 + PHASE0) * x[0] *
                              (2.0 * PI * FRF / 2.99792458e8) / GP.beta / GP.energy;
        DRIFT_Pass(x, L / 2.);
    }
}

/* CCMULT_Pass, simtrack.cpp:6027-6067 */
static void CCMULT_Pass(double x[], double L, double VRF, double FRF, double PHASE0, double B2, double A2,
                        double B3, double A3, double B4, double A4, double B5, double A5)
{
    if (VRF == 0.) {
        DRIFT_Pass(x, L);
    } else {
        DRIFT_Pass(x, L / 2.);
        double c = 2.99792458e8;
        double scale = 2 * PI * FRF / c / GP.brho;
        double sinangle = OSIN(2.0 * PI * FRF * x[z_] / 2.99792458e8 + PHASE0);
        double cosangle = OCOS(2.0 * PI * FRF * x[z_] / 2.99792458e8 + PHASE0);
        double x2 = x[x_] * x[x_], x3 = x2 * x[x_], x4 = x3 * x[x_], x5 = x4 * x[x_];
        double y2 = x[y_] * x[y_], y3 = y2 * x[y_], y4 = y3 * x[y_], y5 = y4 * x[y_];
        double xy = x[x_] * x[y_], x2y = x2 * x[y_], xy2 = x[x_] * y2, x2y2 = x[x_] * xy2;
        double x3y = x[y_] * x3, xy3 = x[x_] * y3;
        double x2y3 = x2y2 * x[y_], x3y2 = x2y2 * x[x_], x4y = x4 * x[y_], xy4 = y4 * x[x_];
        x[px_] = x[px_] + (B2 * x[x_] + A2 * x[y_]) * cosangle / GP.brho;
        x[py_] = x[py_] + (A2 * x[x_] - B2 * x[y_]) * cosangle / GP.brho;
        x[pt_] = x[pt_] + (B2 * (x2 - y2) * sinangle + A2 * xy * sinangle) * scale;
        x[px_] = x[px_] + (B3 * (x2 - y2) + 2 * A3 * xy) * cosangle / GP.brho;
        x[py_] = x[py_] + (2 * B3 * xy + A3 * (x2 - y2)) * cosangle / GP.brho;
        x[pt_] = x[pt_] + (B3 * (x3 - 3 * xy2) * sinangle + A3 * (3 * x2y - y3) * sinangle) * scale / 3.;
        x[px_] = x[px_] + (B4 * (x3 - 3 * xy2) - A4 * (y3 - 3 * x2y)) * cosangle / GP.brho;
        x[py_] = x[py_] + (B4 * (y3 - 3 * x2y) + A4 * (x3 - 3 * xy2)) * cosangle / GP.brho;
        x[pt_] = x[pt_] + (B4 * (x4 - 6 * x2 * y2 + y4) * sinangle + A4 * (4 * x3y - 4 * xy3) * sinangle) * scale / 4.;
        x[px_] = x[px_] + (B5 * (x4 - 6 * x2y2 + y4) + A5 * (4 * x3y - 4 * xy3)) * cosangle / GP.brho;
        x[py_] = x[py_] + (B5 * (4 * xy3 - 4 * x3y) + A5 * (x4 - 6 * x2y2 + y4)) * cosangle / GP.brho;
        x[pt_] = x[pt_] + (B5 * (x5 - 10 * x3y2 + 5 * xy4) * sinangle + A5 * (5 * x4y - 10 * x2y3 + y5) * sinangle) *
                              scale / 5.;
        DRIFT_Pass(x, L / 2.);
    }
}

/* LBT_Pass, simtrack.cpp:6069-6093 */
static void LBT_Pass(double x[], double theta)
{
    double sintheta = OSIN(theta), costheta = OCOS(theta), tantheta = sintheta / costheta;
    double x0 = x[0], px0 = x[1], y0 = x[2], py0 = x[3], z0 = x[4], pz0 = x[5];
    double ps0 = sqrt((1 + pz0) * (1 + pz0) - px0 * px0 - py0 * py0);
    double h0 = (1 + pz0) - ps0;
    double px_star = (px0 - tantheta * h0) / OCOS(theta);
    double py_star = py0 / costheta;
    double pz_star = pz0 - tantheta * px0 + tantheta * tantheta * h0;
    double ps_star = sqrt((1 + pz_star) * (1 + pz_star) - px_star * px_star - py_star * py_star);
    double h_star = (1 + pz_star) - ps_star;
    double x_star = tantheta * z0 + (1 + (px_star / ps_star) * sintheta) * x0;
    double y_star = y0 + sintheta * (py_star / ps_star) * x0;
    double z_star = z0 / costheta + (-h_star / ps_star) * sintheta * x0;
    x[0] = x_star;
    x[1] = px_star;
    x[2] = y_star;
    x[3] = py_star;
    x[4] = z_star;
    x[5] = pz_star;
}

/* ILBT_Pass, simtrack.cpp:6095-6123 */
static void ILBT_Pass(double x[], double theta)
{
    double sintheta = OSIN(theta), costheta = OCOS(theta);
    double x_star = x[0], px_star = x[1], y_star = x[2], py_star = x[3], z_star = x[4], pz_star = x[5];
    double ps_star = sqrt((1 + pz_star) * (1 + pz_star) - px_star * px_star - py_star * py_star);
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

/* ------------------------------------------------------ beam-beam */

/* cerrf, simtrack.cpp:6258-6323 (Gautschi's algorithm for w(z)) */
EXPORT void orc_cerrf(double xx, double yy, double *wx_out, double *wy_out)
{
    int n, nc, nu;
    double x, y, q, h, xl, xh, yh, tx, ty, tn, sx, sy, saux, wx, wy;
    double rx[33], ry[33];
    double cc = 1.12837916709551, zero = 0, one = 1, two = 2, half = .5, xlim = 5.33, ylim = 4.29, fac1 = 3.2,
           fac2 = 23, fac3 = 21;
    x = fabs(xx);
    y = fabs(yy);
    if (y < ylim && x < xlim) {
        q = (one - y / ylim) * sqrt(one - (x / xlim) * (x / xlim));
        h = one / (fac1 * q);
        nc = 7 + (int)(fac2 * q);
        xl = OPOW(h, (1 - nc));
        xh = y + half / h;
        yh = x;
        nu = 10 + (int)(fac3 * q);
        rx[nu + 1] = zero;
        ry[nu + 1] = zero;
        for (n = nu; n >= 1; n--) {
            tx = xh + n * rx[n + 1];
            ty = yh - n * ry[n + 1];
            tn = tx * tx + ty * ty;
            rx[n] = half * tx / tn;
            ry[n] = half * ty / tn;
        }
        sx = zero;
        sy = zero;
        for (n = nc; n >= 1; n--) {
            saux = sx + xl;
            sx = rx[n] * saux - ry[n] * sy;
            sy = rx[n] * sy + ry[n] * saux;
            xl = h * xl;
        }
        wx = cc * sx;
        wy = cc * sy;
    } else {
        xh = y;
        yh = x;
        rx[1] = zero;
        ry[1] = zero;
        for (n = 9; n >= 1; n--) {
            tx = xh + n * rx[1];
            ty = yh - n * ry[1];
            tn = tx * tx + ty * ty;
            rx[1] = half * tx / tn;
            ry[1] = half * ty / tn;
        }
        wx = cc * rx[1];
        wy = cc * ry[1];
    }
    if (yy < zero) {
        wx = two * OEXP(y * y - x * x) * OCOS(two * x * y) - wx;
        wy = -two * OEXP(y * y - x * x) * OSIN(two * x * y) - wy;
        if (xx > zero) wy = -wy;
    } else {
        if (xx < zero) wy = -wy;
    }
    *wx_out = wx;
    *wy_out = wy;
}

/* BB4D (scalar form), simtrack.cpp:6325-6382 */
EXPORT void orc_bb4d(double x, double y, double gamma, double N, double sigmax, double sigmay, double *Dpx,
                     double *Dpy)
{
    double rp = 1.53469826769e-18;
    double x1, y1, x2, y2;
    int signx = (x >= 0.0) ? 1 : -1;
    int signy = (y >= 0.0) ? 1 : -1;
    if (x == 0. && y == 0.) {
        *Dpx = 0.;
        *Dpy = 0.;
    } else {
        if (fabs(sigmax - sigmay) / sigmax < 1.0e-6) {
            double r2 = x * x + y * y;
            double temp1 = 2 * N * rp / gamma / r2;
            double temp2 = 1 - OEXP(-r2 / 2 / sigmax / sigmax);
            *Dpx = temp1 * x * temp2;
            *Dpy = temp1 * y * temp2;
        } else if (sigmax > sigmay) {
            double temp1 = (2. * N * rp / gamma) * sqrt(PI / 2. / (sigmax * sigmax - sigmay * sigmay));
            double temp2 = OEXP(-x * x / 2. / sigmax / sigmax - y * y / 2. / sigmay / sigmay);
            double temp3 = sqrt(2. * (sigmax * sigmax - sigmay * sigmay));
            double w1_real, w1_imag, w2_real, w2_imag;
            x1 = fabs(x) / temp3;
            y1 = fabs(y) / temp3;
            x2 = fabs(x) * sigmay / sigmax / temp3;
            y2 = fabs(y) * sigmax / sigmay / temp3;
            orc_cerrf(x1, y1, &w1_real, &w1_imag);
            orc_cerrf(x2, y2, &w2_real, &w2_imag);
            *Dpy = temp1 * (w1_real - temp2 * w2_real) * signy;
            *Dpx = temp1 * (w1_imag - temp2 * w2_imag) * signx;
        } else if (sigmax < sigmay) {
            double temp1 = (2 * N * rp / gamma) * sqrt(PI / 2 / (sigmay * sigmay - sigmax * sigmax));
            double temp2 = OEXP(-x * x / 2 / sigmax / sigmax - y * y / 2 / sigmay / sigmay);
            double temp3 = sqrt(2 * (sigmay * sigmay - sigmax * sigmax));
            double w1_real, w1_imag, w2_real, w2_imag;
            x1 = fabs(x) / temp3;
            y1 = fabs(y) / temp3;
            x2 = fabs(x) * sigmay / sigmax / temp3;
            y2 = fabs(y) * sigmax / sigmay / temp3;
            orc_cerrf(y1, x1, &w1_real, &w1_imag);
            orc_cerrf(y2, x2, &w2_real, &w2_imag);
            *Dpy = temp1 * (w1_imag - temp2 * w2_imag) * signy;
            *Dpx = temp1 * (w1_real - temp2 * w2_real) * signx;
        }
    }
}

/* gsl_cdf_ugaussian_Pinv: GSL not vendored; the project restatement of
   GSL's AS 241 algorithm is used on both sides (st_libm.h). */
EXPORT double orc_ugaussian_pinv(double p) { return st_ugaussian_pinv(p); }

/* slice centroids of BB6D (cpp:6594-6620) and BB6D_Angle (cpp:6937-6962) */
static void bb_slice_centers(int N_slice, double sigma_l, double z_star[])
{
    if (N_slice == 11) {
        z_star[0] = 1.7996765362153655;
        z_star[1] = 1.1049618149811977;
        z_star[2] = 0.75072324235797372;
        z_star[3] = 0.47407703558819131;
        z_star[4] = 0.23041113693387344;
        z_star[5] = 0.;
        z_star[6] = -0.23041113693387344;
        z_star[7] = -0.47407703558819131;
        z_star[8] = -0.75072324235797372;
        z_star[9] = -1.1049618149811977;
        z_star[10] = -1.7996765362153655;
        for (int i = 0; i < 11; i++) z_star[i] = z_star[i] * sigma_l;
    } else {
        double y[64];
        for (int i = 0; i < N_slice - 1; i++) y[i] = st_ugaussian_pinv((i + 1) * 1.0 / N_slice);
        z_star[0] = N_slice * (OEXP(-y[0] * y[0] / 2)) / sqrt(2.0 * 3.14159265);
        z_star[N_slice - 1] = N_slice * (0 - OEXP(-y[0] * y[0] / 2)) / sqrt(2.0 * 3.14159265);
        for (int i = 1; i < N_slice - 1; i++)
            z_star[i] = N_slice * (OEXP(-y[i] * y[i] / 2) - OEXP(-y[i - 1] * y[i - 1] / 2)) / sqrt(2.0 * 3.14159265);
        for (int i = 0; i < N_slice; i++) z_star[i] = z_star[i] * sigma_l;
    }
}

EXPORT void orc_bb_slice_centers(int N_slice, double sigma_l, double *z_star)
{
    bb_slice_centers(N_slice, sigma_l, z_star);
}

/* BB6D, simtrack.cpp:6579-6678 */
static void BB6D(double x[], double gamma, double Np, double bbscale, double sigma_l, int N_slice,
                 double emitx_rms, double betax_star, double alfx_star, double emity_rms, double betay_star,
                 double alfy_star)
{
    double rp = 1.534698e-18;
    double x0[6], Np_slice[64], z_star[64];
    double S, gx_star, gy_star, betax, betay, sigmax, sigmay, dsigmax2ds, dsigmay2ds, dUdsigmax2, dUdsigmay2;
    double X, Y, Dpx, Dpy, temp1, temp2, temp3, temp4;
    for (int i = 0; i < N_slice; i++) Np_slice[i] = 1.0 * Np / N_slice;
    bb_slice_centers(N_slice, sigma_l, z_star);
    for (int i = 0; i < N_slice; i++) {
        for (int j = 0; j < 6; j++) x0[j] = x[j];
        S = (x0[4] - z_star[i]) / 2.0;
        gx_star = (1.0 + alfx_star * alfx_star) / betax_star;
        gy_star = (1.0 + alfy_star * alfy_star) / betay_star;
        betax = betax_star + 2 * S * alfx_star + S * S * gx_star;
        betay = betay_star + 2 * S * alfy_star + S * S * gy_star;
        sigmax = sqrt(emitx_rms * betax);
        sigmay = sqrt(emity_rms * betay);
        dsigmax2ds = 2.0 * emitx_rms * (alfx_star + S * gx_star);
        dsigmay2ds = 2.0 * emity_rms * (alfy_star + S * gy_star);
        X = x0[0] + x0[1] * S;
        Y = x0[2] + x0[3] * S;
        orc_bb4d(X, Y, gamma, Np_slice[i], sigmax, sigmay, &Dpx, &Dpy);
        temp1 = 1.0 / 2. / (sigmax * sigmax - sigmay * sigmay);
        temp2 = X * Dpx + Y * Dpy;
        temp3 = 2. * Np_slice[i] * rp / gamma;
        temp4 = OEXP(-X * X / 2. / sigmax / sigmax - Y * Y / 2. / sigmay / sigmay);
        x[0] = x0[0] - S * Dpx * bbscale;
        x[1] = x0[1] + Dpx * bbscale;
        x[2] = x0[2] - S * Dpy * bbscale;
        x[3] = x0[3] + Dpy * bbscale;
        x[4] = x0[4];
        if (fabs(sigmax - sigmay) / sigmax < 1.0e-6) {
            x[5] = x0[5] + 0.5 * Dpx * bbscale * (x0[1] + 0.5 * Dpx * bbscale) +
                   0.5 * Dpy * bbscale * (x0[3] + 0.5 * Dpy * bbscale) +
                   (1.0 / sigmax / sigmax) * dsigmax2ds * (temp3 * bbscale / 4.) * temp4;
        } else if (sigmax > sigmay) {
            dUdsigmax2 = temp1 * bbscale * (temp2 + temp3 * (sigmay / sigmax * temp4 - 1.));
            dUdsigmay2 = -temp1 * bbscale * (temp2 + temp3 * (sigmax / sigmay * temp4 - 1.));
            x[5] = x0[5] + 0.5 * Dpx * bbscale * (x0[1] + 0.5 * Dpx * bbscale) +
                   0.5 * Dpy * bbscale * (x0[3] + 0.5 * Dpy * bbscale) -
                   (0.5 * dsigmax2ds * dUdsigmax2 + 0.5 * dsigmay2ds * dUdsigmay2);
        } else {
            dUdsigmay2 = -temp1 * bbscale * (temp2 + temp3 * (sigmax / sigmay * temp4 - 1.));
            dUdsigmax2 = temp1 * bbscale * (temp2 + temp3 * (sigmay / sigmax * temp4 - 1.));
            x[5] = x0[5] + 0.5 * Dpx * bbscale * (x0[1] + 0.5 * Dpx * bbscale) +
                   0.5 * Dpy * bbscale * (x0[3] + 0.5 * Dpy * bbscale) -
                   (0.5 * dsigmax2ds * dUdsigmax2 + 0.5 * dsigmay2ds * dUdsigmay2);
        }
    }
}

/* BB6D_Angle, simtrack.cpp:6917-7079 */
static void BB6D_Angle(double x[], double gamma, double Np, double bbscale, double sigma_l, double theta,
                       int N_slice, double emitx_rms, double betax_star, double alfx_star, double emity_rms,
                       double betay_star, double alfy_star)
{
    double rp = 1.534698e-18;
    double sintheta = OSIN(theta), costheta = OCOS(theta), tantheta = sintheta / costheta;
    double Np_slice[64], Z0[64], Z0_star[64];
    double S, gx_star, gy_star, betax, alfx, betay, alfy;
    double sigmax, sigmay, dsigmax2ds, dsigmay2ds, dUdsigmax2, dUdsigmay2;
    double X, Y, Dpx, Dpy, temp1, temp2, temp3, temp4;
    double x0, y0, z0, px0, py0, pz0, ps0, h0;
    double x_star, y_star, z_star, px_star, py_star, pz_star, hx_star, hy_star, hz_star, ps_star, h_star;
    double x_star_0, y_star_0, z_star_0, px_star_0, py_star_0, pz_star_0;
    for (int i = 0; i < N_slice; i++) Np_slice[i] = 1.0 * Np / N_slice;
    bb_slice_centers(N_slice, sigma_l, Z0);
    x0 = x[0];
    px0 = x[1];
    y0 = x[2];
    py0 = x[3];
    z0 = x[4];
    pz0 = x[5];
    ps0 = sqrt((1 + pz0) * (1 + pz0) - px0 * px0 - py0 * py0);
    h0 = (1 + pz0) - ps0;
    px_star = (px0 - tantheta * h0) / OCOS(theta);
    py_star = py0 / costheta;
    pz_star = pz0 - tantheta * px0 + tantheta * tantheta * h0;
    ps_star = sqrt((1 + pz_star) * (1 + pz_star) - px_star * px_star - py_star * py_star);
    h_star = (1 + pz_star) - ps_star;
    x_star = tantheta * z0 + (1 + (px_star / ps_star) * sintheta) * x0;
    y_star = y0 + sintheta * (py_star / ps_star) * x0;
    z_star = z0 / costheta + (-h_star / ps_star) * sintheta * x0;
    for (int i = 0; i < N_slice; i++) Z0_star[i] = Z0[i] / costheta;
    gx_star = (1.0 + alfx_star * alfx_star) / betax_star;
    gy_star = (1.0 + alfy_star * alfy_star) / betay_star;
    gx_star = gx_star / costheta;
    gy_star = gy_star / costheta;
    betax_star = betax_star * costheta;
    betay_star = betay_star * costheta;
    emitx_rms = emitx_rms / costheta;
    emity_rms = emity_rms / costheta;
    for (int i = 0; i < N_slice; i++) {
        x_star_0 = x_star;
        px_star_0 = px_star;
        y_star_0 = y_star;
        py_star_0 = py_star;
        z_star_0 = z_star;
        pz_star_0 = pz_star;
        S = (z_star_0 - Z0_star[i]) / 2.0;
        betax = betax_star - 2 * S * alfx_star + S * S * gx_star;
        alfx = (alfx_star - gx_star * S);
        betay = betay_star - 2 * S * alfy_star + S * S * gy_star;
        alfy = (alfy_star - gy_star * S);
        sigmax = sqrt(emitx_rms * betax);
        sigmay = sqrt(emity_rms * betay);
        dsigmax2ds = -2.0 * alfx * (emitx_rms);
        dsigmay2ds = -2.0 * alfy * (emity_rms);
        X = x_star_0 + px_star_0 * S - Z0_star[i] * sintheta;
        Y = y_star_0 + py_star_0 * S - 0;
        orc_bb4d(X, Y, gamma, Np_slice[i], sigmax, sigmay, &Dpx, &Dpy);
        temp1 = 1.0 / 2. / (sigmax * sigmax - sigmay * sigmay);
        temp2 = X * Dpx + Y * Dpy;
        temp3 = 2. * Np_slice[i] * rp / gamma;
        temp4 = OEXP(-X * X / 2. / sigmax / sigmax - Y * Y / 2. / sigmay / sigmay);
        x_star = x_star_0 - S * Dpx * bbscale;
        px_star = px_star_0 + Dpx * bbscale;
        y_star = y_star_0 - S * Dpy * bbscale;
        py_star = py_star_0 + Dpy * bbscale;
        z_star = z_star_0;
        if (fabs(sigmax - sigmay) / sigmax < 1.0e-6) {
            pz_star = pz_star_0 + 0.5 * Dpx * bbscale * (px_star_0 + 0.5 * Dpx * bbscale) +
                      0.5 * Dpy * bbscale * (py_star_0 + 0.5 * Dpy * bbscale) +
                      (1.0 / sigmax / sigmax) * dsigmax2ds * (temp3 * bbscale / 4.) * temp4;
        } else if (sigmax > sigmay) {
            dUdsigmax2 = temp1 * bbscale * (temp2 + temp3 * (sigmay / sigmax * temp4 - 1.));
            dUdsigmay2 = -temp1 * bbscale * (temp2 + temp3 * (sigmax / sigmay * temp4 - 1.));
            pz_star = pz_star_0 + 0.5 * Dpx * bbscale * (px_star_0 + 0.5 * Dpx * bbscale) +
                      0.5 * Dpy * bbscale * (py_star_0 + 0.5 * Dpy * bbscale) -
                      (0.5 * dsigmax2ds * dUdsigmax2 + 0.5 * dsigmay2ds * dUdsigmay2);
        } else {
            dUdsigmay2 = -temp1 * bbscale * (temp2 + temp3 * (sigmax / sigmay * temp4 - 1.));
            dUdsigmax2 = temp1 * bbscale * (temp2 + temp3 * (sigmay / sigmax * temp4 - 1.));
            pz_star = pz_star_0 + 0.5 * Dpx * bbscale * (px_star_0 + 0.5 * Dpx * bbscale) +
                      0.5 * Dpy * bbscale * (py_star_0 + 0.5 * Dpy * bbscale) -
                      (0.5 * dsigmax2ds * dUdsigmax2 + 0.5 * dsigmay2ds * dUdsigmay2);
        }
    }
    ps_star = sqrt((1 + pz_star) * (1 + pz_star) - px_star * px_star - py_star * py_star);
    h_star = 1 + pz_star - ps_star;
    hx_star = px_star / ps_star;
    hy_star = py_star / ps_star;
    hz_star = -h_star / ps_star;
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

/* BEAMBEAM_Pass, simtrack.cpp:7343-7356 */
static void BEAMBEAM_Pass(double x[], int TREATMENT, double NP, double BBSCALE, double SIGMAL, int NSLICE,
                          double EMITX, double BETAX, double ALFAX, double EMITY, double BETAY, double ALFAY)
{
    if (NP != 0.) {
        if (TREATMENT == 6) {
            BB6D(x, GP.gamma, NP, BBSCALE, SIGMAL, NSLICE, EMITX, BETAX, ALFAX, EMITY, BETAY, ALFAY);
        } else {
            double Dpx, Dpy;
            orc_bb4d(x[0], x[2], GP.gamma, NP, sqrt(EMITX * BETAX), sqrt(EMITY * BETAY), &Dpx, &Dpy);
            x[1] = x[1] + Dpx * BBSCALE;
            x[3] = x[3] + Dpy * BBSCALE;
        }
    }
}

/* LRBB_Pass, simtrack.cpp:7358-7363 (the 6-argument BB4D, cpp:6384-6447) */
static void LRBB_Pass(double x[], double NP, double BBSCALE, double SEPX, double SEPY, double SIGMAX, double SIGMAY)
{
    double Dpx, Dpy;
    x[0] = x[0] + SEPX;
    x[2] = x[2] + SEPY;
    orc_bb4d(x[0], x[2], GP.gamma, NP, SIGMAX, SIGMAY, &Dpx, &Dpy);
    x[1] = x[1] + Dpx * BBSCALE;
    x[3] = x[3] + Dpy * BBSCALE;
    x[0] = x[0] - SEPX;
    x[2] = x[2] - SEPY;
}

/* ELENS_Pass, simtrack.cpp:7370-7390 */
static void ELENS_Pass(double x[], double Le, double Ne, double bbscale, double beta_e, int N_slice, double sigmax,
                       double sigmay)
{
    double Le_slice = Le * 1.0 / N_slice;
    double Ne_slice = Ne * 1.0 / N_slice;
    double Dpx, Dpy;
    if (Ne != 0.) {
        for (int i = 0; i < N_slice; i++) {
            DRIFT_Pass(x, Le_slice / 2.);
            orc_bb4d(x[0], x[2], GP.gamma, Ne_slice, sigmax, sigmay, &Dpx, &Dpy);
            x[1] = x[1] - Dpx * (1. + beta_e) * bbscale;
            x[3] = x[3] - Dpy * (1. + beta_e) * bbscale;
            DRIFT_Pass(x, Le_slice / 2.);
        }
    } else {
        DRIFT_Pass(x, Le);
    }
}

/* ------------------------------------------------------ spin passes */

/* SBEND_sPass, simtrack.cpp:7486-7536 */
static void SBEND_sPass(double x[], double L, int Nint, double Angle, double E1, double E2)
{
    double href = Angle / L, Lint = L / Nint, BLbrho[3], x0[9];
    if (Angle == 0.) {
        DRIFT_Pass(x, L);
        return;
    }
    edge(x, E1, href);
    BLbrho[0] = 0.;
    BLbrho[1] = Angle / Nint;
    BLbrho[2] = 0.;
    if (GP.H_expand) {
        for (int i = 0; i < Nint; i++) {
            for (int j = 0; j < 9; j++) x0[j] = x[j];
            DRIFT_Pass(x, Fdrift1 * Lint);
            bend_kick_pass(x, Fkick1 * Lint, href);
            DRIFT_Pass(x, Fdrift2 * Lint);
            bend_kick_pass(x, Fkick2 * Lint, href);
            DRIFT_Pass(x, Fdrift2 * Lint);
            bend_kick_pass(x, Fkick1 * Lint, href);
            DRIFT_Pass(x, Fdrift1 * Lint);
            spin(x0, Angle / Nint, href, BLbrho);
            x[6] = x0[6];
            x[7] = x0[7];
            x[8] = x0[8];
            if (GP.radiate) radiate(x, Lint, href, BLbrho);
        }
    } else {
        for (int i = 0; i < Nint; i++) {
            for (int j = 0; j < 9; j++) x0[j] = x[j];
            sbend_exact_pass(x, Lint, Angle / Nint, OCOS(Angle / Nint), OSIN(Angle / Nint));
            spin(x0, Angle / Nint, href, BLbrho);
            x[6] = x0[6];
            x[7] = x0[7];
            x[8] = x0[8];
            if (GP.radiate) radiate(x, Lint, href, BLbrho);
        }
    }
    edge(x, E2, href);
}

/* QUAD_sPass, simtrack.cpp:7538-7590 */
static void QUAD_sPass(double x[], double L, int Nint, double k1l, double k1sl)
{
    double Lint = L / Nint, preal, pimag, BLbrho[3], x0[9], x1[9];
    if (k1l == 0. && k1sl == 0.) {
        DRIFT_Pass(x, L);
        return;
    }
    if (L == 0.) {
        for (int j = 0; j < 9; j++) x0[j] = x[j];
        quad_kick_pass(x, k1l, k1sl);
        for (int j = 0; j < 9; j++) x1[j] = (x[j] + x0[j]) / 2.;
        preal = x1[0];
        pimag = x1[2];
        BLbrho[0] = k1sl * preal + k1l * pimag;
        BLbrho[1] = k1l * preal - k1sl * pimag;
        BLbrho[2] = 0.;
        spin(x0, 0, 0, BLbrho);
        x[6] = x0[6];
        x[7] = x0[7];
        x[8] = x0[8];
    } else {
        double k1l_kick1 = Fkick1 * k1l / Nint, k1sl_kick1 = Fkick1 * k1sl / Nint;
        double k1l_kick2 = Fkick2 * k1l / Nint, k1sl_kick2 = Fkick2 * k1sl / Nint;
        for (int i = 0; i < Nint; i++) {
            for (int j = 0; j < 9; j++) x0[j] = x[j];
            DRIFT_Pass(x, Fdrift1 * Lint);
            quad_kick_pass(x, k1l_kick1, k1sl_kick1);
            DRIFT_Pass(x, Fdrift2 * Lint);
            quad_kick_pass(x, k1l_kick2, k1sl_kick2);
            DRIFT_Pass(x, Fdrift2 * Lint);
            quad_kick_pass(x, k1l_kick1, k1sl_kick1);
            DRIFT_Pass(x, Fdrift1 * Lint);
            for (int j = 0; j < 9; j++) x1[j] = (x[j] + x0[j]) / 2.;
            preal = x1[0];
            pimag = x1[2];
            BLbrho[0] = (k1sl * preal + k1l * pimag) / Nint;
            BLbrho[1] = (k1l * preal - k1sl * pimag) / Nint;
            BLbrho[2] = 0.;
            spin(x0, 0, 0, BLbrho);
            x[6] = x0[6];
            x[7] = x0[7];
            x[8] = x0[8];
        }
    }
}

/* SEXT_sPass, simtrack.cpp:7592-7645 */
static void SEXT_sPass(double x[], double L, int Nint, double k2l, double k2sl)
{
    double Lint = L / Nint, preal, pimag, BLbrho[3], x0[9], x1[9];
    if (k2l == 0. && k2sl == 0.) {
        DRIFT_Pass(x, L);
        return;
    }
    if (L == 0.) {
        for (int j = 0; j < 9; j++) x0[j] = x[j];
        sext_kick_pass(x, k2l, k2sl);
        for (int j = 0; j < 9; j++) x1[j] = (x[j] + x0[j]) / 2.;
        preal = x1[0] * x1[0] - x1[2] * x1[2];
        pimag = 2. * x1[0] * x1[2];
        BLbrho[0] = (k2sl * preal + k2l * pimag) / 2.;
        BLbrho[1] = (k2l * preal - k2sl * pimag) / 2.;
        BLbrho[2] = 0.;
        spin(x0, 0, 0, BLbrho);
        x[6] = x0[6];
        x[7] = x0[7];
        x[8] = x0[8];
    } else {
        double k2l_kick1 = Fkick1 * k2l / Nint, k2sl_kick1 = Fkick1 * k2sl / Nint;
        double k2l_kick2 = Fkick2 * k2l / Nint, k2sl_kick2 = Fkick2 * k2sl / Nint;
        for (int i = 0; i < Nint; i++) {
            for (int j = 0; j < 9; j++) x0[j] = x[j];
            DRIFT_Pass(x, Fdrift1 * Lint);
            sext_kick_pass(x, k2l_kick1, k2sl_kick1);
            DRIFT_Pass(x, Fdrift2 * Lint);
            sext_kick_pass(x, k2l_kick2, k2sl_kick2);
            DRIFT_Pass(x, Fdrift2 * Lint);
            sext_kick_pass(x, k2l_kick1, k2sl_kick1);
            DRIFT_Pass(x, Fdrift1 * Lint);
            for (int j = 0; j < 9; j++) x1[j] = (x[j] + x0[j]) / 2.;
            preal = x1[0] * x1[0] - x1[2] * x1[2];
            pimag = 2. * x1[0] * x1[2];
            BLbrho[0] = (k2sl * preal + k2l * pimag) / 2. / Nint;
            BLbrho[1] = (k2l * preal - k2sl * pimag) / 2. / Nint;
            BLbrho[2] = 0.;
            spin(x0, 0, 0, BLbrho);
            x[6] = x0[6];
            x[7] = x0[7];
            x[8] = x0[8];
        }
    }
}

/* OCT_sPass, simtrack.cpp:7647-7700 */
static void OCT_sPass(double x[], double L, int Nint, double k3l, double k3sl)
{
    double Lint = L / Nint, preal, pimag, BLbrho[3], x0[9], x1[9];
    if (k3l == 0. && k3sl == 0.) {
        DRIFT_Pass(x, L);
        return;
    }
    if (L == 0.) {
        for (int j = 0; j < 9; j++) x0[j] = x[j];
        oct_kick_pass(x, k3l, k3sl);
        for (int j = 0; j < 9; j++) x1[j] = (x[j] + x0[j]) / 2.;
        preal = x1[0] * x1[0] * x1[0] - 3 * x1[0] * x1[2] * x1[2];
        pimag = 3 * x1[0] * x1[0] * x1[2] - x1[2] * x1[2] * x1[2];
        BLbrho[0] = (k3sl * preal + k3l * pimag) / 6 / Nint;
        BLbrho[1] = (k3l * preal - k3sl * pimag) / 6 / Nint;
        BLbrho[2] = 0.;
        spin(x0, 0, 0, BLbrho);
        x[6] = x0[6];
        x[7] = x0[7];
        x[8] = x0[8];
    } else {
        double k3l_kick1 = Fkick1 * k3l / Nint, k3sl_kick1 = Fkick1 * k3sl / Nint;
        double k3l_kick2 = Fkick2 * k3l / Nint, k3sl_kick2 = Fkick2 * k3sl / Nint;
        for (int i = 0; i < Nint; i++) {
            for (int j = 0; j < 9; j++) x0[j] = x[j];
            DRIFT_Pass(x, Fdrift1 * Lint);
            oct_kick_pass(x, k3l_kick1, k3sl_kick1);
            DRIFT_Pass(x, Fdrift2 * Lint);
            oct_kick_pass(x, k3l_kick2, k3sl_kick2);
            DRIFT_Pass(x, Fdrift2 * Lint);
            oct_kick_pass(x, k3l_kick1, k3sl_kick1);
            DRIFT_Pass(x, Fdrift1 * Lint);
            for (int j = 0; j < 9; j++) x1[j] = (x[j] + x0[j]) / 2.;
            preal = x1[0] * x1[0] * x1[0] - 3 * x1[0] * x1[2] * x1[2];
            pimag = 3 * x1[0] * x1[0] * x1[2] - x1[2] * x1[2] * x1[2];
            BLbrho[0] = (k3sl * preal + k3l * pimag) / 6 / Nint;
            BLbrho[1] = (k3l * preal - k3sl * pimag) / 6 / Nint;
            BLbrho[2] = 0.;
            spin(x0, 0, 0, BLbrho);
            x[6] = x0[6];
            x[7] = x0[7];
            x[8] = x0[8];
        }
    }
}

/* MULT_sPass, simtrack.cpp:7702-7740 */
static void MULT_sPass(double x[], double L, int Nint, int Norder, const double KNL[11], const double KNSL[11])
{
    double Lint = L / Nint, BLbrho[3], x0[9], x1[9], k1[11], s1[11], k2[11], s2[11];
    if (L == 0.) {
        cal_Bfield(x, 1, Norder, 0, KNL, KNSL, BLbrho);
        spin(x, 0, 0, BLbrho);
        mult_kick_pass(x, Norder, KNL, KNSL);
    } else {
        kick_coefs(KNL, KNSL, Nint, k1, s1, k2, s2);
        for (int i = 0; i < Nint; i++) {
            for (int j = 0; j < 9; j++) x0[j] = x[j];
            DRIFT_Pass(x, Fdrift1 * Lint);
            mult_kick_pass(x, Norder, k1, s1);
            DRIFT_Pass(x, Fdrift2 * Lint);
            mult_kick_pass(x, Norder, k2, s2);
            DRIFT_Pass(x, Fdrift2 * Lint);
            mult_kick_pass(x, Norder, k1, s1);
            DRIFT_Pass(x, Fdrift1 * Lint);
            for (int j = 0; j < 9; j++) x1[j] = (x[j] + x0[j]) / 2.;
            cal_Bfield(x1, Nint, Norder, 0, KNL, KNSL, BLbrho);
            spin(x0, 0, 0, BLbrho);
            x[6] = x0[6];
            x[7] = x0[7];
            x[8] = x0[8];
        }
    }
}

/*
 * GMULT_sPass (simtrack.cpp:7742-7833), SBENDMULT_sPass (cpp:7835-7928),
 * SMULT_sPass (cpp:7930-8022) and GSBENDMULT_sPass (cpp:8024-8114, K0L
 * unused) have identical bodies.
 */
static void BENDMULT_sPass(double x[], double L, int Nint, int Norder, double Angle, double E1, double E2,
                           const double KNL[11], const double KNSL[11])
{
    double href = Angle / L, Lint = L / Nint, BLbrho[3], x0[9], x1[9], k1[11], s1[11], k2[11], s2[11];
    double cosAngle, sinAngle, cosAngle1, sinAngle1;
    if (L == 0.) {
        for (int j = 0; j < 9; j++) x0[j] = x[j];
        mult_kick_pass(x, Norder, KNL, KNSL);
        for (int j = 0; j < 9; j++) x1[j] = (x[j] + x0[j]) / 2.;
        cal_Bfield(x1, 1, Norder, Angle, KNL, KNSL, BLbrho);
        spin(x0, Angle, 0, BLbrho);
        x[6] = x0[6];
        x[7] = x0[7];
        x[8] = x0[8];
        return;
    }
    edge(x, E1, href);
    kick_coefs(KNL, KNSL, Nint, k1, s1, k2, s2);
    if (GP.H_expand) {
        for (int i = 0; i < Nint; i++) {
            for (int j = 0; j < 9; j++) x0[j] = x[j];
            DRIFT_Pass(x, Fdrift1 * Lint);
            bend_mult_kick_pass(x, Fkick1 * Lint, href, Norder, k1, s1);
            DRIFT_Pass(x, Fdrift2 * Lint);
            bend_mult_kick_pass(x, Fkick2 * Lint, href, Norder, k2, s2);
            DRIFT_Pass(x, Fdrift2 * Lint);
            bend_mult_kick_pass(x, Fkick1 * Lint, href, Norder, k1, s1);
            DRIFT_Pass(x, Fdrift1 * Lint);
            for (int j = 0; j < 9; j++) x1[j] = (x[j] + x0[j]) / 2.;
            cal_Bfield(x1, Nint, Norder, Angle, KNL, KNSL, BLbrho);
            spin(x0, Angle / Nint, href, BLbrho);
            x[6] = x0[6];
            x[7] = x0[7];
            x[8] = x0[8];
            if (GP.radiate) radiate(x, Lint, href, BLbrho);
        }
    } else {
        cosAngle1 = OCOS(Fdrift1 * Angle / Nint);
        sinAngle1 = OSIN(Fdrift1 * Angle / Nint);
        cosAngle = OCOS(Fdrift2 * Angle / Nint);
        sinAngle = OSIN(Fdrift2 * Angle / Nint);
        for (int i = 0; i < Nint; i++) {
            for (int j = 0; j < 9; j++) x0[j] = x[j];
            sbend_exact_pass(x, Fdrift1 * Lint, Fdrift1 * Angle / Nint, cosAngle1, sinAngle1);
            mult_kick_pass(x, Norder, k1, s1);
            sbend_exact_pass(x, Fdrift2 * Lint, Fdrift2 * Angle / Nint, cosAngle, sinAngle);
            mult_kick_pass(x, Norder, k2, s2);
            sbend_exact_pass(x, Fdrift2 * Lint, Fdrift2 * Angle / Nint, cosAngle, sinAngle);
            mult_kick_pass(x, Norder, k1, s1);
            sbend_exact_pass(x, Fdrift1 * Lint, Fdrift1 * Angle / Nint, cosAngle1, sinAngle1);
            for (int j = 0; j < 9; j++) x1[j] = (x[j] + x0[j]) / 2.;
            cal_Bfield(x1, Nint, Norder, Angle, KNL, KNSL, BLbrho);
            spin(x0, Angle / Nint, href, BLbrho);
            x[6] = x0[6];
            x[7] = x0[7];
            x[8] = x0[8];
            if (GP.radiate) radiate(x, Lint, href, BLbrho);
        }
    }
    edge(x, E2, href);
}

/* KICK_sPass, simtrack.cpp:8193-8208 */
static void KICK_sPass(double x[], double L, double HKICK, double VKICK)
{
    double BLbrho[3];
    DRIFT_Pass(x, L / 2.0);
    BLbrho[0] = VKICK;
    BLbrho[1] = -HKICK;
    BLbrho[2] = 0.;
    spin(x, 0, 0, BLbrho);
    x[1] = x[1] + HKICK;
    x[3] = x[3] + VKICK;
    DRIFT_Pass(x, L / 2.0);
}

/* SOLEN_sPass, simtrack.cpp:8118-8177: as SOLEN_Pass with Bz = KS Brho / L,
   and after each body step spin() of the step's entry state, its x', y'
   converted back to px, py, in the longitudinal field Bz Lint / Brho */
static void SOLEN_sPass(double x[], double L, int Nint, double KS)
{
    int i, j, k;
    double ps, gamma1, delta1, beta1, vz;
    double Bz = KS * GP.brho / L, K, Lint = L / Nint, theta, sintheta, costheta, m44[16];
    double BLbrho[3], xtemp[9];
    if (KS == 0) {
        DRIFT_Pass(x, L);
        return;
    }
    delta1 = ORC_PT_CACHED(x[5], sqrt(1.0 + 2 * x[5] / GP.beta + x[5] * x[5]) - 1.0);
    gamma1 = ORC_PT_CACHED(x[5], GP.gamma + sqrt(GP.gamma * GP.gamma - 1.) * x[5]);
    beta1 = ORC_PT_DERIVED(sqrt(1.0 - 1.0 / gamma1 / gamma1));
    ps = sqrt((1 + delta1) * (1 + delta1) - x[1] * x[1] - x[3] * x[3]);
    x[1] = x[1] / ps;
    x[3] = x[3] / ps;
    vz = 2.99792458e8 * beta1 / sqrt(1.0 + x[1] * x[1] + x[3] * x[3]);
    K = GP.q * Bz / 2.0 / gamma1 / GP.m0 / vz;
    x[1] = x[1] + x[2] * K;
    x[3] = x[3] - x[0] * K;
    theta = -2.0 * K * L / Nint;
    sintheta = OSIN(theta);
    costheta = OCOS(theta);
    m44[0 * 4 + 0] = 1.0;  m44[0 * 4 + 1] = Lint * sintheta / theta;
    m44[0 * 4 + 2] = 0.;   m44[0 * 4 + 3] = -(1 - costheta) * Lint / theta;
    m44[1 * 4 + 0] = 0.;   m44[1 * 4 + 1] = costheta;
    m44[1 * 4 + 2] = 0.;   m44[1 * 4 + 3] = -sintheta;
    m44[2 * 4 + 0] = 0.;   m44[2 * 4 + 1] = (1 - costheta) * Lint / theta;
    m44[2 * 4 + 2] = 1.;   m44[2 * 4 + 3] = Lint * sintheta / theta;
    m44[3 * 4 + 0] = 0.;   m44[3 * 4 + 1] = sintheta;
    m44[3 * 4 + 2] = 0.;   m44[3 * 4 + 3] = costheta;
    for (i = 0; i < Nint; i++) {
        for (k = 0; k < 9; k++) xtemp[k] = x[k];
        for (k = 0; k < 4; k++) {
            x[k] = 0.;
            for (j = 0; j < 4; j++) x[k] = x[k] + m44[k * 4 + j] * xtemp[j];
        }
        xtemp[1] = gamma1 * GP.m0 * xtemp[1] * vz / GP.p;
        xtemp[3] = gamma1 * GP.m0 * xtemp[3] * vz / GP.p;
        BLbrho[0] = 0.;
        BLbrho[1] = 0.;
        BLbrho[2] = Bz * Lint / GP.brho;
        spin(xtemp, 0, 0, BLbrho);
        x[6] = xtemp[6];
        x[7] = xtemp[7];
        x[8] = xtemp[8];
    }
    x[1] = x[1] - x[2] * K;
    x[3] = x[3] + x[0] * K;
    x[1] = gamma1 * GP.m0 * x[1] * vz / GP.p;
    x[3] = gamma1 * GP.m0 * x[3] * vz / GP.p;
    x[4] = x[4] - 2.99792458e8 * (L / vz - L / GP.beta / 2.99792458e8);
}

/* ACMULT_sPass, simtrack.cpp:8210-8249 */
static void ACMULT_sPass(double x[], double L, int Norder, double KL, double KSL, int TTURNS, double PHI0)
{
    int fac = 1;
    double Xn, Yn, Xn0, Yn0, KNL, KNSL, By, Bx, BLbrho[3];
    DRIFT_Pass(x, L / 2.0);
    if (Norder == 0) {
        By = KL;
        Bx = KSL;
    } else {
        KNL = KL;
        KNSL = KSL;
        By = 0.;
        Bx = 0.;
        Xn = 1.;
        Yn = 0.;
        for (int i = 1; i < Norder + 1; i++) {
            Xn0 = Xn;
            Yn0 = Yn;
            Xn = Xn0 * x[x_] - Yn0 * x[y_];
            Yn = Xn0 * x[y_] + Yn0 * x[x_];
            fac = fac * i;
        }
        By = By + (KNL * Xn - KNSL * Yn) / fac;
        Bx = Bx + (KNL * Yn + KNSL * Xn) / fac;
    }
    BLbrho[0] = Bx * OSIN(2.0 * PI * orc_turn / TTURNS + PHI0);
    BLbrho[1] = By * OSIN(2.0 * PI * orc_turn / TTURNS + PHI0);
    BLbrho[2] = 0.;
    spin(x, 0, 0, BLbrho);
    x[px_] = x[px_] - By * OSIN(2.0 * PI * orc_turn / TTURNS + PHI0);
    x[py_] = x[py_] + Bx * OSIN(2.0 * PI * orc_turn / TTURNS + PHI0);
    DRIFT_Pass(x, L / 2.0);
}

/* ACDIP_sPass, simtrack.cpp:8251-8284 */
static void ACDIP_sPass(double x[], double L, double HKICKMAX, double VKICKMAX, double NUD, double TURNS,
                        double TURNE, double PHID)
{
    double BLbrho[3];
    if (orc_turn < TURNS) {
        DRIFT_Pass(x, L);
    } else if (orc_turn >= TURNS && orc_turn <= TURNE) {
        DRIFT_Pass(x, L / 2.);
        BLbrho[0] = ((orc_turn - TURNS) * 1.0 * VKICKMAX / (TURNE - TURNS)) *
                    OSIN(2.0 * 3.14159265 * NUD * (orc_turn - TURNS) + PHID);
        BLbrho[1] = -((orc_turn - TURNS) * 1.0 * HKICKMAX / (TURNE - TURNS)) *
                    OSIN(2.0 * 3.14159265 * NUD * (orc_turn - TURNS) + PHID);
        BLbrho[2] = 0.;
        spin(x, 0, 0, BLbrho);
        x[1] += ((orc_turn - TURNS) * 1.0 * HKICKMAX / (TURNE - TURNS)) *
                OSIN(2.0 * 3.14159265 * NUD * (orc_turn - TURNS) + PHID);
        x[3] += ((orc_turn - TURNS) * 1.0 * VKICKMAX / (TURNE - TURNS)) *
                OSIN(2.0 * 3.14159265 * NUD * (orc_turn - TURNS) + PHID);
        DRIFT_Pass(x, L / 2.);
    } else {
        DRIFT_Pass(x, L / 2.);
        BLbrho[0] = VKICKMAX * OSIN(2.0 * 3.14159265 * NUD * (orc_turn - TURNS) + PHID);
        BLbrho[1] = -HKICKMAX * OSIN(2.0 * 3.14159265 * NUD * (orc_turn - TURNS) + PHID);
        BLbrho[2] = 0;
        spin(x, 0, 0, BLbrho);
        x[1] += HKICKMAX * OSIN(2.0 * 3.14159265 * NUD * (orc_turn - TURNS) + PHID);
        x[3] += VKICKMAX * OSIN(2.0 * 3.14159265 * NUD * (orc_turn - TURNS) + PHID);
        DRIFT_Pass(x, L / 2.);
    }
}

/* ROTAT_sPass / SNAKE_sPass, simtrack.cpp:8286-8348 (identical bodies) */
static void ROTAT_sPass(double x[], double L, const double n[3], double angle)
{
    double spin0[3], sp[3], F, uF[3], costheta = OCOS(-angle), sintheta = OSIN(-angle), rot[9];
    DRIFT_Pass(x, L / 2.);
    for (int i = 0; i < 3; i++) spin0[i] = x[6 + i];
    F = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int i = 0; i < 3; i++) uF[i] = n[i] / F;
    rot[0] = uF[0] * uF[0] * (1 - costheta) + costheta;
    rot[1] = uF[0] * uF[1] * (1 - costheta) - uF[2] * sintheta;
    rot[2] = uF[0] * uF[2] * (1 - costheta) + uF[1] * sintheta;
    rot[3] = uF[0] * uF[1] * (1 - costheta) + uF[2] * sintheta;
    rot[4] = uF[1] * uF[1] * (1 - costheta) + costheta;
    rot[5] = uF[1] * uF[2] * (1 - costheta) - uF[0] * sintheta;
    rot[6] = uF[0] * uF[2] * (1 - costheta) - uF[1] * sintheta;
    rot[7] = uF[1] * uF[2] * (1 - costheta) + uF[0] * sintheta;
    rot[8] = uF[2] * uF[2] * (1 - costheta) + costheta;
    for (int i = 0; i < 3; i++) {
        sp[i] = 0.;
        for (int j = 0; j < 3; j++) sp[i] = sp[i] + rot[i * 3 + j] * spin0[j];
    }
    for (int i = 0; i < 3; i++) x[6 + i] = sp[i];
    DRIFT_Pass(x, L / 2.);
}

/* ------------------------------------------------ Element::Pass/sPass */

#define MIS(e) (e)->dx, (e)->dy, (e)->ds, (e)->dtheta, (e)->dphi, (e)->dpsi

/* Element::Pass dispatch: the Pass() bodies of simtrack.cpp:8425-11470 */
static int elem_pass(const st_element *e, double x[], orc_rng *rng)
{
    switch (e->type) {
    case ST_DRIFT: case ST_MARKER: case ST_BPM: case ST_HBPM: case ST_VBPM:
    case ST_ROTAT: case ST_SNAKE:
        DRIFT_Pass(x, e->l);
        return 0;
    case ST_SBEND:
        GtoL(x, e->l, e->angle, MIS(e));
        SBEND_Pass(x, e->l, e->nint, e->angle, e->e1, e->e2, rng);
        LtoG(x, e->l, e->angle, MIS(e));
        return 0;
    case ST_GSBEND:
        GtoL(x, e->l, e->angle, MIS(e));
        GSBEND_Pass(x, e->l, e->nint, e->angle, e->k0l, e->e1, e->e2, rng);
        LtoG(x, e->l, e->angle, MIS(e));
        return 0;
    case ST_QUAD:
        GtoL(x, e->l, 0., MIS(e));
        QUAD_Pass(x, e->l, e->nint, e->knl[1], e->knsl[1]);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_SKEWQ:
        GtoL(x, e->l, 0., MIS(e));
        QUAD_Pass(x, e->l, e->nint, 0., e->knsl[1]);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_SEXT:
        GtoL(x, e->l, 0., MIS(e));
        SEXT_Pass(x, e->l, e->nint, e->knl[2], e->knsl[2]);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_OCT:
        GtoL(x, e->l, 0., MIS(e));
        OCT_Pass(x, e->l, e->nint, e->knl[3], e->knsl[3]);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_MULT:
        GtoL(x, e->l, 0., MIS(e));
        MULT_Pass(x, e->l, e->nint, e->norder, e->knl, e->knsl);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_GMULT: case ST_SBENDMULT: case ST_SMULT:
        GtoL(x, e->l, e->angle, MIS(e));
        BENDMULT_Pass(x, e->l, e->nint, e->norder, e->angle, e->e1, e->e2, e->knl, e->knsl);
        LtoG(x, e->l, e->angle, MIS(e));
        return 0;
    case ST_GSBENDMULT:
        GtoL(x, e->l, e->angle, MIS(e));
        GSBENDMULT_Pass(x, e->l, e->nint, e->norder, e->angle, e->k0l, e->e1, e->e2, e->knl, e->knsl);
        LtoG(x, e->l, e->angle, MIS(e));
        return 0;
    case ST_KICKER: case ST_HKICKER: case ST_VKICKER: {
        double hk = 0., vk = 0.;
        if (e->type != ST_VKICKER) hk = e->hkick;
        if (e->type != ST_HKICKER) vk = e->vkick;
        GtoL(x, e->l, 0., MIS(e));
        KICK_Pass(x, e->l, hk, vk);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    }
    case ST_RFCAV:
        GtoL(x, e->l, 0., MIS(e));
        RFCAV_Pass(x, e->l, e->vrf, e->frf, e->phase0);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_CRABRF:
        GtoL(x, e->l, 0., MIS(e));
        CRABRF_Pass(x, e->l, e->vrf, e->frf, e->phase0);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_CCMULT:
        GtoL(x, e->l, 0., MIS(e));
        CCMULT_Pass(x, e->l, e->vrf, e->frf, e->phase0, e->cc[0], e->cc[1], e->cc[2], e->cc[3], e->cc[4], e->cc[5],
                    e->cc[6], e->cc[7]);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_LBT:
        GtoL(x, e->l, 0., MIS(e));
        LBT_Pass(x, e->theta);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_ILBT:
        GtoL(x, e->l, 0., MIS(e));
        ILBT_Pass(x, e->theta);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_BEAMBEAM:
        GtoL(x, e->l, 0., MIS(e));
        if (e->np != 0.)
            BEAMBEAM_Pass(x, e->treatment, e->np, e->bbscale, e->sigmal, e->nslice, e->emitx, e->betax, e->alfax,
                          e->emity, e->betay, e->alfay);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_CRBB:
        GtoL(x, e->l, 0., MIS(e));
        BB6D_Angle(x, GP.gamma, e->np, e->bbscale, e->sigmal, e->theta, e->nslice, e->emitx, e->betax, e->alfax,
                   e->emity, e->betay, e->alfay);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_MATRIX:
        MATRIX_Pass(x, e->xco_in, e->xco_out, e->m66);
        return 0;
    case ST_PATCH:                          /* PATCH::Pass, cpp:12323-12325 */
        PATCH_Pass(x, e->dx, e->dy, e->ds, e->theta);
        return 0;
    case ST_SOLEN:                          /* SOLEN::Pass, cpp:9688-9692 */
        if (e->nint < 1) return -1;
        GtoL(x, e->l, 0., MIS(e));
        SOLEN_Pass(x, e->l, e->nint, e->ks);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_TRANS:                          /* TRANS::Pass, cpp:12108-12110 */
        TRANS_Pass(x, e->dx, e->dy, e->ds);
        return 0;
    case ST_YROTAT:                         /* YROTAT::Pass, cpp:12202-12204 */
        YROTAT_Pass(x, e->theta);
        return 0;
    case ST_SROTAT:                         /* SROTAT::Pass, cpp:12155-12157 */
        SROTAT_Pass(x, e->psi);
        return 0;
    case ST_XROTAT:                         /* XROTAT::Pass, cpp:12249-12251 */
        XROTAT_Pass(x, e->phi);
        return 0;
    case ST_COOLING:                        /* COOLING::Pass, cpp:12043-12045 */
        COOLING_Pass(x, e->alpha);
        return 0;
    case ST_LRBB:                           /* LRBB::Pass, cpp:11299-11303 */
        GtoL(x, e->l, 0., MIS(e));
        LRBB_Pass(x, e->np, e->bbscale, e->sepx, e->sepy, e->sigmax, e->sigmay);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_ELENS:                          /* ELENS::Pass, cpp:11649-11653 */
        if (e->nslice < 1) return -1;
        GtoL(x, e->l, 0., MIS(e));
        ELENS_Pass(x, e->l, e->ne, e->bbscale, e->betae, e->nslice, e->sigmax, e->sigmay);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_HELENS:                         /* HELENS::Pass, cpp:11765-11769 */
        if (e->nslice < 1) return -1;
        GtoL(x, e->l, 0., MIS(e));
        HELENS_Pass(x, e->l, e->ne, e->bbscale, e->betae, e->nslice, e->rout, e->rin);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_HACDIP: case ST_VACDIP: {       /* HACDIP::Pass cpp:10357-10361, VACDIP::Pass cpp:10453-10457 */
        double hk = 0., vk = 0.;
        if (e->type == ST_HACDIP) hk = e->hkick;
        if (e->type == ST_VACDIP) vk = e->vkick;
        GtoL(x, e->l, 0., MIS(e));
        ACDIP_Pass(x, e->l, hk, vk, e->nud, e->turns, e->turne, e->phid);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    }
    case ST_HACMULT: case ST_VACMULT: {     /* HACMULT::Pass cpp:10173-10177, VACMULT::Pass cpp:10260-10264 */
        double kl = 0., ksl = 0.;
        if (e->type == ST_HACMULT) kl = e->klmax;
        else ksl = e->klmax;
        GtoL(x, e->l, 0., MIS(e));
        ACMULT_Pass(x, e->l, e->norder, kl, ksl, (int)ORC_VALUE(e->tturns), e->phi0);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    }
    case ST_WIGGLER:                        /* WIGGLER::Pass, cpp:9755-9759 */
        if (e->nint < 1) return -1;
        GtoL(x, e->l, 0., MIS(e));
        Wiggler_Pass_Forest(x, e->l, e->nint, e->b0, e->kx, e->kz, e->phiz0);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    default:
        return -1;
    }
}

/* Element::sPass dispatch: the sPass() bodies of simtrack.cpp:8425-11470 */
static int elem_spass(const st_element *e, double x[])
{
    switch (e->type) {
    case ST_DRIFT: case ST_MARKER: case ST_BPM: case ST_HBPM: case ST_VBPM:
        DRIFT_Pass(x, e->l);
        return 0;
    case ST_ROTAT: case ST_SNAKE:
        ROTAT_sPass(x, e->l, e->rot_n, e->rot_angle);
        return 0;
    case ST_SBEND: case ST_GSBEND:
        GtoL(x, e->l, e->angle, MIS(e));
        SBEND_sPass(x, e->l, e->nint, e->angle, e->e1, e->e2);
        LtoG(x, e->l, e->angle, MIS(e));
        return 0;
    case ST_QUAD:
        GtoL(x, e->l, 0., MIS(e));
        QUAD_sPass(x, e->l, e->nint, e->knl[1], e->knsl[1]);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_SKEWQ:
        GtoL(x, e->l, 0., MIS(e));
        QUAD_sPass(x, e->l, e->nint, 0., e->knsl[1]);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_SEXT:
        GtoL(x, e->l, 0., MIS(e));
        SEXT_sPass(x, e->l, e->nint, e->knl[2], e->knsl[2]);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_OCT:
        GtoL(x, e->l, 0., MIS(e));
        OCT_sPass(x, e->l, e->nint, e->knl[3], e->knsl[3]);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_MULT:
        GtoL(x, e->l, 0., MIS(e));
        MULT_sPass(x, e->l, e->nint, e->norder, e->knl, e->knsl);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_GMULT: case ST_SBENDMULT: case ST_SMULT: case ST_GSBENDMULT:
        GtoL(x, e->l, e->angle, MIS(e));
        BENDMULT_sPass(x, e->l, e->nint, e->norder, e->angle, e->e1, e->e2, e->knl, e->knsl);
        LtoG(x, e->l, e->angle, MIS(e));
        return 0;
    case ST_KICKER: case ST_HKICKER: case ST_VKICKER: {
        double hk = 0., vk = 0.;
        if (e->type != ST_VKICKER) hk = e->hkick;
        if (e->type != ST_HKICKER) vk = e->vkick;
        GtoL(x, e->l, 0., MIS(e));
        KICK_sPass(x, e->l, hk, vk);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    }
    case ST_CRABRF:                         /* CRABRF::sPass is a drift, cpp:10808-10812 */
        GtoL(x, e->l, 0., MIS(e));
        DRIFT_Pass(x, e->l);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_LBT:                            /* no misalignment in sPass, cpp:11004-11006 */
        LBT_Pass(x, e->theta);
        return 0;
    case ST_ILBT:
        ILBT_Pass(x, e->theta);
        return 0;
    case ST_RFCAV: case ST_CCMULT: case ST_BEAMBEAM: case ST_CRBB: case ST_MATRIX:
        return elem_pass(e, x, NULL);       /* sPass == Pass (orbit only) */
    case ST_PATCH:                          /* PATCH::sPass, cpp:12331-12333: the orbit map */
        return elem_pass(e, x, NULL);
    case ST_WIGGLER:                        /* WIGGLER::sPass, cpp:9767-9771 (no spin rotation) */
        if (e->nint < 1) return -1;
        GtoL(x, e->l, 0., MIS(e));
        Wiggler_Pass_Wu(x, e->l, e->nint, e->b0, e->kx, e->kz, e->phiz0);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_SOLEN:                          /* SOLEN::sPass, cpp:9700-9704 */
        if (e->nint < 1) return -1;
        GtoL(x, e->l, 0., MIS(e));
        SOLEN_sPass(x, e->l, e->nint, e->ks);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    case ST_TRANS: case ST_YROTAT: case ST_SROTAT: case ST_XROTAT: case ST_COOLING:
    case ST_LRBB: case ST_ELENS: case ST_HELENS:
        return elem_pass(e, x, NULL);       /* sPass == Pass (cpp:12116, 12210, 12163, 12257, 12051, 11309, 11671, 11777) */
    case ST_HACDIP: case ST_VACDIP: {       /* HACDIP::sPass cpp:10367-10371, VACDIP::sPass cpp:10463-10467 */
        double hk = 0., vk = 0.;
        if (e->type == ST_HACDIP) hk = e->hkick;
        if (e->type == ST_VACDIP) vk = e->vkick;
        GtoL(x, e->l, 0., MIS(e));
        ACDIP_sPass(x, e->l, hk, vk, e->nud, e->turns, e->turne, e->phid);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    }
    case ST_HACMULT: case ST_VACMULT: {     /* HACMULT::sPass cpp:10183-10187, VACMULT::sPass cpp:10270-10274 */
        double kl = 0., ksl = 0.;
        if (e->type == ST_HACMULT) kl = e->klmax;
        else ksl = e->klmax;
        GtoL(x, e->l, 0., MIS(e));
        ACMULT_sPass(x, e->l, e->norder, kl, ksl, (int)ORC_VALUE(e->tturns), e->phi0);
        LtoG(x, e->l, 0., MIS(e));
        return 0;
    }
    default:
        return -1;
    }
}

/* ------------------------------------------------------- drivers */

EXPORT int orc_element_pass(const st_element *e, double *x, int spin_on, uint32_t pid)
{
    orc_rng rng = {GP.rng_seed, pid, 0u};
    orc_turn = 0;
    return spin_on ? elem_spass(e, x) : elem_pass(e, x, &rng);
}

static int lost_check(const st_element *e, const double x[])
{
    return fabs(x[0]) > e->apx || fabs(x[2]) > e->apy || isnan(x[0]) || isnan(x[2]);
}

/*
 * Track (simtrack.cpp:20791-20814), Track_spin (cpp:20816-20839),
 * Track_Radial (cpp:21641-21680) for one particle.  mode bit 1: spin,
 * bit 2: radial (on-momentum z subtraction with x04 from orc_radial_x04).
 * x_tbt (nullable): Track_tbt storage (cpp:20841-20861) at the start of each
 * turn, or after element bpm_index (cpp:20885-20907) when bpm_index >= 0.
 */
static int track_one(const st_element *line, int n, double *x, int nturn, int *stable, int *lost_turn,
                     int *lost_post, int mode, const double *x04, double *x_tbt, int bpm_index, uint32_t pid,
                     uint64_t draws_per_turn)
{
    int spin_on = mode & 1, radial = mode & 2;
    int ncol = spin_on ? 9 : 6;
    orc_rng rng = {GP.rng_seed, pid, 0u};
    if (fabs(x[0]) > 0.1 || fabs(x[2]) > 0.1 || *stable == 0) {
        *stable = 0;
        *lost_turn = 0;
        *lost_post = 0;
        return 0;
    }
    for (int turn = 0; turn < nturn; turn++) {
        orc_turn = turn;
        rng.counter = (uint64_t)turn * draws_per_turn;
        if (x_tbt && bpm_index < 0)
            for (int k = 0; k < ncol; k++) x_tbt[(size_t)turn * ncol + k] = x[k];
        for (int j = 0; j < n; j++) {
            int rc = spin_on ? elem_spass(&line[j], x) : elem_pass(&line[j], x, &rng);
            if (rc) return rc;
            if (radial) x[4] = x[4] - x04[j];
            if (x_tbt && j == bpm_index)
                for (int k = 0; k < ncol; k++) x_tbt[(size_t)turn * ncol + k] = x[k];
            if (lost_check(&line[j], x)) {
                *stable = 0;
                *lost_turn = turn;
                *lost_post = j;
                return 0;
            }
        }
    }
    return 0;
}

/*
 * Random draws per turn: radiate1 runs once per slice of every bending SBEND
 * (SBEND_Pass, cpp:4781-4813) and radiate2 once per slice of every bending
 * GSBEND (GSBEND_Pass, cpp:4848-4871) when GP.radiate; spin passes radiate
 * classically only.  Draw k of turn t has number t*D + k.
 */
EXPORT uint64_t orc_draws_per_turn(const st_element *line, int n, int spin_on)
{
    uint64_t d = 0;
    if (!GP.radiate || spin_on) return 0;
    for (int j = 0; j < n; j++)
        if ((line[j].type == ST_SBEND || line[j].type == ST_GSBEND) && line[j].angle != 0.)
            d += (uint64_t)(line[j].nint > 0 ? line[j].nint : 0);
    return d;
}

/* on-momentum z after every element (Track_Radial, cpp:21655-21661) */
EXPORT int orc_radial_x04(const st_element *line, int n, double *x04)
{
    double x0[6];
    orc_rng rng = {GP.rng_seed, 0u, 0u};
    for (int k = 0; k < 6; k++) x0[k] = 0.;
    for (int j = 0; j < n; j++) {
        x0[4] = 0.;
        x0[5] = 0.;
        int rc = elem_pass(&line[j], x0, &rng);
        if (rc) return rc;
        x04[j] = x0[4];
    }
    return 0;
}

/*
 * Batched driver over independent particles (the reference calls Track once
 * per particle).  nthreads > 1 spreads particles over OpenMP threads; each
 * particle's arithmetic is unchanged.
 */
EXPORT int orc_track(const st_element *line, int n, double *x, int np, int nturn, int *stable, int *lost_turn,
                     int *lost_post, int mode, double *x_tbt, int bpm_index, int nthreads)
{
    double *x04 = NULL;
    int ncol = (mode & 1) ? 9 : 6;
    int err = 0;
    uint64_t draws = orc_draws_per_turn(line, n, mode & 1);
    if (mode & 2) {
        x04 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        if (!x04) return -5;
        err = orc_radial_x04(line, n, x04);
        if (err) {
            free(x04);
            return err;
        }
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) reduction(| : err)
    for (int p = 0; p < np; p++) {
        double *tb = x_tbt ? x_tbt + (size_t)p * (size_t)nturn * ncol : NULL;
        err |= track_one(line, n, x + (size_t)p * ncol, nturn, stable + p, lost_turn + p, lost_post + p, mode, x04,
                         tb, bpm_index, (uint32_t)p, draws);
    }
    free(x04);
    return err ? -2 : 0;
}

/* Track_wrt_OnMomentumPart, simtrack.cpp:20942-20966: Track with dz0[j]
   subtracted from z after element j (the radial path with the caller's
   offsets) */
EXPORT int orc_track_dz(const st_element *line, int n, const double *dz0, double *x, int np, int nturn, int *stable,
                        int *lost_turn, int *lost_post)
{
    int err = 0;
    uint64_t draws = orc_draws_per_turn(line, n, 0);
    for (int p = 0; p < np; p++)
        err |= track_one(line, n, x + (size_t)p * 6, nturn, stable + p, lost_turn + p, lost_post + p, 2, dz0, NULL,
                         -1, (uint32_t)p, draws);
    return err ? -2 : 0;
}

/* FineTuneFinder, simtrack.cpp:2288-2319 */
EXPORT void orc_fine_tune(int n, const double *x, double tune_lower, double tune_upper, double *peaktune)
{
    double tune, tune_step = 0.001, sumr, sumi, amp, peakamp;
    do {
        tune = tune_lower;
        *peaktune = tune_lower;
        peakamp = 0;
        do {
            sumr = 0.;
            sumi = 0;
            for (int i = 0; i < n; i++) {
                sumr = sumr + x[i] * OCOS(-2 * PI * i * tune);
                sumi = sumi + x[i] * OSIN(-2 * PI * i * tune);
            }
            amp = sumr * sumr + sumi * sumi;
            if (amp > peakamp) {
                *peaktune = tune;
                peakamp = amp;
            }
            tune = tune + tune_step;
        } while (tune < tune_upper);
        tune_lower = *peaktune - tune_step;
        tune_upper = *peaktune + tune_step;
        tune_step = tune_step / 10.;
    } while (tune_step > 1.0e-12);
}

/* FineTuneFinder, x/px form, simtrack.cpp:2321-2352 */
EXPORT void orc_fine_tune_xpx(int n, const double *x, const double *px, double tune_lower, double tune_upper,
                              double *peaktune)
{
    double tune, tune_step = 0.001, sumr, sumi, amp, peakamp;
    do {
        tune = tune_lower;
        *peaktune = tune_lower;
        peakamp = 0;
        do {
            sumr = 0;
            sumi = 0.;
            for (int i = 0; i < n; i++) {
                sumr = sumr + x[i] * OCOS(-2 * PI * i * tune) - px[i] * OSIN(-2 * PI * i * tune);
                sumi = sumi + x[i] * OSIN(-2 * PI * i * tune) + px[i] * OCOS(-2 * PI * i * tune);
            }
            amp = sumr * sumr + sumi * sumi;
            if (amp > peakamp) {
                *peaktune = tune;
                peakamp = amp;
            }
            tune = tune + tune_step;
        } while (tune < tune_upper);
        tune_lower = *peaktune - tune_step;
        tune_upper = *peaktune + tune_step;
        tune_step = tune_step / 10.;
    } while (tune_step > 1.0e-12);
}

/* ------------------------------------------- elementary functions */
EXPORT double orc_sin(double x) { return OSIN(x); }
EXPORT double orc_cos(double x) { return OCOS(x); }
EXPORT double orc_tan(double x) { return OTAN(x); }
EXPORT double orc_asin(double x) { return OASIN(x); }
EXPORT double orc_exp(double x) { return OEXP(x); }
EXPORT double orc_cosh(double x) { return OCOSH(x); }
EXPORT double orc_sinh(double x) { return OSINH(x); }
#if !defined(ORACLE_GLIBC) && !defined(ORC_SCALAR_BUILD)
/* the branch-free common paths the GPU kernels evaluate first */
EXPORT double orc_exp_phase(double x, int phase) { return st_exp_phase(x, phase); }
EXPORT double orc_asin_small(double x) { return st_asin_small(x); }
EXPORT void orc_sincos_small(double x, double *s, double *c) { st_sincos_small(x, s, c); }
EXPORT void orc_sincos(double x, double *s, double *c) { st_sincos(x, s, c); }
#endif
EXPORT double orc_powi(double x, int n) { return OPOW(x, n); }
EXPORT double orc_log(double x) { return st_log(x); }
EXPORT double orc_atan(double x) { return st_atan(x); }
EXPORT double orc_rng_normal(uint64_t seed, uint32_t particle, uint64_t counter)
{
    return st_rng_normal(seed, particle, counter);
}
/* the photon sampler of radiate2 for particles 0..n-1 at one draw number:
   photon count and summed energy (units of the critical energy) */
EXPORT void orc_rad2_sample(uint64_t seed, int n, uint64_t draw, double nr, int *count, double *etot)
{
    for (int i = 0; i < n; i++) {
        uint32_t j;
        count[i] = st_rad2_count(seed, (uint32_t)i, draw, nr, &j);
        etot[i] = st_rad2_photons(seed, (uint32_t)i, draw, nr);
    }
}
EXPORT double orc_rad2_energy(double r) { return st_rad2_energy(r); }
EXPORT int orc_libm_kind(void)
{
#ifdef ORACLE_GLIBC
    return 1;
#else
    return 0;
#endif
}
EXPORT void orc_philox(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4])
{
    st_u32x4 c;
    for (int i = 0; i < 4; i++) c.v[i] = ctr[i];
    c = st_philox4x32_10(c, k0, k1);
    for (int i = 0; i < 4; i++) out[i] = c.v[i];
}

/* ------------------------------------------------- closed orbit, TBT scans */

/* mat_inv, simtrack.cpp:1440-1496 */
static int mat_inv(double a[], int n)
{
    int is[16], js[16], i, j, k, l, u, v;
    double d, p;
    for (k = 0; k <= n - 1; k++) {
        d = 0.0;
        for (i = k; i <= n - 1; i++)
            for (j = k; j <= n - 1; j++) {
                l = i * n + j;
                p = fabs(a[l]);
                if (p > d) {
                    d = p;
                    is[k] = i;
                    js[k] = j;
                }
            }
        if (d + 1.0 == 1.0) return 0;
        if (is[k] != k)
            for (j = 0; j <= n - 1; j++) {
                u = k * n + j;
                v = is[k] * n + j;
                p = a[u];
                a[u] = a[v];
                a[v] = p;
            }
        if (js[k] != k)
            for (i = 0; i <= n - 1; i++) {
                u = i * n + k;
                v = i * n + js[k];
                p = a[u];
                a[u] = a[v];
                a[v] = p;
            }
        l = k * n + k;
        a[l] = 1.0 / a[l];
        for (j = 0; j <= n - 1; j++)
            if (j != k) {
                u = k * n + j;
                a[u] = a[u] * a[l];
            }
        for (i = 0; i <= n - 1; i++)
            if (i != k)
                for (j = 0; j <= n - 1; j++)
                    if (j != k) {
                        u = i * n + j;
                        a[u] = a[u] - a[i * n + k] * a[k * n + j];
                    }
        for (i = 0; i <= n - 1; i++)
            if (i != k) {
                u = i * n + k;
                a[u] = -a[u] * a[l];
            }
    }
    for (k = n - 1; k >= 0; k--) {
        if (js[k] != k)
            for (j = 0; j <= n - 1; j++) {
                u = k * n + j;
                v = js[k] * n + j;
                p = a[u];
                a[u] = a[v];
                a[v] = p;
            }
        if (is[k] != k)
            for (i = 0; i <= n - 1; i++) {
                u = i * n + k;
                v = i * n + is[k];
                p = a[u];
                a[u] = a[v];
                a[v] = p;
            }
    }
    return 1;
}

/* one bare turn (Element::Pass of every element, no loss test) of particle pid */
static int pass_turn(const st_element *line, int n, double x[], uint32_t pid, uint64_t draws)
{
    orc_rng rng = {GP.rng_seed, pid, 0u};
    (void)draws;
    for (int j = 0; j < n; j++) {
        int rc = elem_pass(&line[j], x, &rng);
        if (rc) return rc;
    }
    return 0;
}

/* Cal_Orbit_Num, simtrack.cpp:14947-15030, on a copy of the line with the
   RF, crab and CCMULT voltages zeroed.  Returns 0, or -1 where the
   reference reports failure.  x_end: x_co passed once more. */
EXPORT int orc_closed_orbit(const st_element *line0, int n, double deltap, double x_co[6], double x_end[6],
                            int *iterations)
{
    st_element *line = (st_element *)malloc(sizeof(st_element) * (size_t)(n > 0 ? n : 1));
    if (!line) return -5;
    memcpy(line, line0, sizeof(st_element) * (size_t)n);
    for (int i = 0; i < n; i++)
        if (line[i].type == ST_RFCAV || line[i].type == ST_CRABRF || line[i].type == ST_CCMULT) line[i].vrf = 0.;
    int i, j, k, iter = 0, rc = 0;
    double x0[6], x01[6], x1[6], dx[4], d = 1.0e-09, mat[4][4], chi, codeps = 1e-12;
    int Max_iter = 20;
    for (i = 0; i < 6; i++) x0[i] = 0.;
    do {
        iter++;
        x0[4] = 0;
        x0[5] = deltap;
        for (j = 0; j < 6; j++) x01[j] = x0[j];
        if ((rc = pass_turn(line, n, x01, 0, 0))) goto out;
        for (k = 0; k < 4; k++) {
            for (j = 0; j < 6; j++) x1[j] = x0[j];
            x1[k] = x1[k] + d;
            if ((rc = pass_turn(line, n, x1, (uint32_t)(k + 1), 0))) goto out;
            for (j = 0; j < 4; j++) mat[j][k] = (x1[j] - x01[j]) / d;
        }
        for (j = 0; j < 6; j++) x1[j] = x0[j];
        if ((rc = pass_turn(line, n, x1, 0, 0))) goto out;
        for (i = 0; i < 4; i++) dx[i] = x0[i] - x1[i];
        chi = 0;
        for (i = 0; i < 4; i++) chi += dx[i] * dx[i];
        chi = sqrt(chi / 4.);
        if (chi > codeps) {
            for (i = 0; i < 4; i++) mat[i][i] = mat[i][i] - 1.0000001;
            if (mat_inv(&mat[0][0], 4) == 0) {
                rc = -1;
                goto out;
            }
            for (i = 0; i < 4; i++)
                for (j = 0; j < 4; j++) x0[i] += mat[i][j] * dx[j];
        }
    } while (chi > codeps && iter < Max_iter);
    if (iterations) *iterations = iter;
    if (iter == Max_iter - 1) {
        rc = -1;
        goto out;
    }
    x0[4] = 0.000;
    x0[5] = deltap;
    for (i = 0; i < 6; i++) x_co[i] = x0[i];
    if (x_end) {
        for (i = 0; i < 6; i++) x_end[i] = x0[i];
        rc = pass_turn(line, n, x_end, 0, 0);
    }
out:
    free(line);
    return rc;
}

/* sPass of every element with no loss test: the spin drivers walk
   linename.Cell[j]->sPass(x) directly (simtrack.cpp:20527, 20618, 20666) */
static int spass_turn(const st_element *line, int n, double x[])
{
    for (int j = 0; j < n; j++) {
        int rc = elem_spass(&line[j], x);
        if (rc) return rc;
    }
    return 0;
}

/*
 * Cal_Spin_Orbit, simtrack.cpp:20510-20572: Cal_Orbit_Num, then Newton on
 * the spin's (Sx, Ss) at the start with one-turn sPass maps (RF on),
 * difference 1e-6, tolerance 1e-12, at most 20 iterations, Sy re-normalised
 * after each step; mat_inv's result is not checked (cpp:20546).  x0: the
 * orbit Cell[Ncell-1]->X and the converged spin at the start; n0 (3 per
 * element, may be null): the normalised spin after every element of one
 * more pass.  Returns -1 where the reference reports failure.
 */
EXPORT int orc_spin_orbit(const st_element *line, int n, double deltap, double x0[9], double *n0, int *iterations)
{
    double x_co[6], x_end[6];
    int rc = orc_closed_orbit(line, n, deltap, x_co, x_end, NULL);
    if (rc) return rc;
    int i, j, iter = 0;
    double x01[9], x1[9], dx[3], d = 1.0e-06, mat[2][2], chi, codeps = 1e-12;
    int Max_iter = 20;
    for (i = 0; i < 6; i++) x0[i] = x_end[i];
    x0[6] = 0.;
    x0[7] = 1.0;
    x0[8] = 0.;
    do {
        iter++;
        for (j = 0; j < 9; j++) x01[j] = x0[j];
        if ((rc = spass_turn(line, n, x01))) return rc;
        for (j = 0; j < 9; j++) x1[j] = x0[j];
        x1[6] = x0[6] + d;
        if ((rc = spass_turn(line, n, x1))) return rc;
        mat[0][0] = (x1[6] - x01[6]) / d;
        mat[1][0] = (x1[8] - x01[8]) / d;
        for (j = 0; j < 9; j++) x1[j] = x0[j];
        x1[8] = x0[8] + d;
        if ((rc = spass_turn(line, n, x1))) return rc;
        mat[0][1] = (x1[6] - x01[6]) / d;
        mat[1][1] = (x1[8] - x01[8]) / d;
        dx[0] = x0[6] - x01[6];
        dx[1] = x0[8] - x01[8];
        chi = dx[0] * dx[0] + dx[1] * dx[1];
        chi = sqrt(chi / 2.);
        if (chi > codeps) {
            for (i = 0; i < 2; i++) mat[i][i] = mat[i][i] - 1.0000001;
            mat_inv(&mat[0][0], 2);
            x0[6] = x0[6] + mat[0][0] * dx[0] + mat[0][1] * dx[1];
            x0[8] = x0[8] + mat[1][0] * dx[0] + mat[1][1] * dx[1];
            x0[7] = sqrt(1. - x0[6] * x0[6] - x0[8] * x0[8]);
        }
    } while (chi > codeps && iter < Max_iter);
    if (iterations) *iterations = iter;
    if (iter == Max_iter - 1) return -1;
    if (n0) {
        double xs[9];
        for (j = 0; j < 9; j++) xs[j] = x0[j];
        for (j = 0; j < n; j++) {
            if ((rc = elem_spass(&line[j], xs))) return rc;
            chi = sqrt(xs[6] * xs[6] + xs[7] * xs[7] + xs[8] * xs[8]);
            for (i = 0; i < 3; i++) n0[3 * j + i] = xs[6 + i] / chi;
        }
    }
    return 0;
}

/*
 * Cal_Spin_Tune, simtrack.cpp:20574-20601: Cal_Spin_Orbit, a spin m
 * perpendicular to the last element's n0 (with the reference's grouping
 * n0^2 / n1 * n1, cpp:20587), 500 turns of sPass from the orbit
 * Cell[Ncell-1]->X recording Sx after each turn, and FineTuneFinder on
 * [0.001, 0.499].  n0_end (may be null): the last element's n0.
 */
EXPORT int orc_spin_tune(const st_element *line, int n, double deltap, double *tune, double n0_end[3])
{
    enum { NTURN = 500 };
    double x0[9], x[9], nv[3], m[3], sample[NTURN];
    double *n0 = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    if (!n0) return -5;
    int rc = orc_spin_orbit(line, n, deltap, x0, n0, NULL);
    if (rc) {
        free(n0);
        return rc;
    }
    for (int i = 0; i < 3; i++) nv[i] = n0[3 * (n - 1) + i];
    free(n0);
    if (n0_end)
        for (int i = 0; i < 3; i++) n0_end[i] = nv[i];
    m[2] = 0.;
    m[0] = 1.0 / sqrt(1. + nv[0] * nv[0] / nv[1] * nv[1]);
    m[1] = -nv[0] * m[0] / nv[1];
    for (int i = 0; i < 6; i++) x[i] = x0[i];
    x[6] = m[0];
    x[7] = m[1];
    x[8] = m[2];
    for (int j = 0; j < NTURN; j++) {
        if ((rc = spass_turn(line, n, x))) return rc;
        sample[j] = x[6];
    }
    orc_fine_tune(NTURN, sample, 0.001, 0.499, tune);
    return 0;
}

/*
 * Set_Initial_Spin, simtrack.cpp:20603-20635: the orbit of x with unit
 * spins along x, y and s, each tracked nturn turns of sPass; after every
 * turn its projection on n0 is summed; x's spin becomes the normalised
 * sums.
 */
EXPORT int orc_set_initial_spin(const st_element *line, int n, int nturn, const double n0[3], double x[9])
{
    int j, rc;
    double b = 0, b1 = 0., b2 = 0., b3 = 0.;
    double x1[9], x2[9], x3[9];
    for (j = 0; j < 6; j++) x1[j] = x[j];
    for (j = 0; j < 6; j++) x2[j] = x[j];
    for (j = 0; j < 6; j++) x3[j] = x[j];
    x1[6] = 1.;
    x1[7] = 0.;
    x1[8] = 0.;
    x2[6] = 0.;
    x2[7] = 1.;
    x2[8] = 0.;
    x3[6] = 0.;
    x3[7] = 0.;
    x3[8] = 1.;
    for (j = 0; j < nturn; j++) {
        if ((rc = spass_turn(line, n, x1))) return rc;
        b1 = b1 + x1[6] * n0[0] + x1[7] * n0[1] + x1[8] * n0[2];
    }
    for (j = 0; j < nturn; j++) {
        if ((rc = spass_turn(line, n, x2))) return rc;
        b2 = b2 + x2[6] * n0[0] + x2[7] * n0[1] + x2[8] * n0[2];
    }
    for (j = 0; j < nturn; j++) {
        if ((rc = spass_turn(line, n, x3))) return rc;
        b3 = b3 + x3[6] * n0[0] + x3[7] * n0[1] + x3[8] * n0[2];
    }
    b = sqrt(b1 * b1 + b2 * b2 + b3 * b3);
    b1 = b1 / b;
    b2 = b2 / b;
    b3 = b3 / b;
    x[6] = b1;
    x[7] = b2;
    x[8] = b3;
    return 0;
}

/* The one-turn matrix of Cal_OneTurnMap (simtrack.cpp:15122-15137: RFCAV
   voltages zeroed, crab cavities and CCMULT on) by symmetric differences of
   step h about x_co, one displaced particle after another (particle q moves
   coordinate q/2 by +h for even q, -h for odd q); m[6*i+k] = dx_i/dx_k. */
EXPORT int orc_one_turn_matrix(const st_element *line0, int n, const double x_co[6], double h, double m[36])
{
    st_element *line = (st_element *)malloc(sizeof(st_element) * (size_t)(n > 0 ? n : 1));
    if (!line) return -5;
    memcpy(line, line0, sizeof(st_element) * (size_t)n);
    for (int i = 0; i < n; i++)
        if (line[i].type == ST_RFCAV) line[i].vrf = 0.;
    double xp[6], xm[6];
    int rc = 0;
    for (int k = 0; k < 6 && !rc; k++) {
        for (int j = 0; j < 6; j++) xp[j] = xm[j] = x_co[j];
        xp[k] += h;
        xm[k] -= h;
        if ((rc = pass_turn(line, n, xp, (uint32_t)(2 * k), 0))) break;
        if ((rc = pass_turn(line, n, xm, (uint32_t)(2 * k + 1), 0))) break;
        for (int i = 0; i < 6; i++) m[6 * i + k] = (xp[i] - xm[i]) / (2 * h);
    }
    free(line);
    return rc;
}

/* Cal_Orbit_Num's last pass (simtrack.cpp:15020-15029) and Cal_ElementMap
   (cpp:15205-15253) as st_element_maps: orbit after every element, and every
   element's map by symmetric differences about the previous element's orbit
   with z = 0, x[5] = x5 */
EXPORT int orc_element_maps(const st_element *line0, int n, const double x_co[6], double x5, double h, double *T,
                            double *X)
{
    st_element *lo = (st_element *)malloc(sizeof(st_element) * (size_t)(n > 0 ? n : 1));
    st_element *lm = (st_element *)malloc(sizeof(st_element) * (size_t)(n > 0 ? n : 1));
    int rc = 0;
    if (!lo || !lm) {
        free(lo);
        free(lm);
        return -5;
    }
    memcpy(lo, line0, sizeof(st_element) * (size_t)n);
    memcpy(lm, line0, sizeof(st_element) * (size_t)n);
    for (int i = 0; i < n; i++) {
        if (lo[i].type == ST_RFCAV || lo[i].type == ST_CRABRF || lo[i].type == ST_CCMULT) lo[i].vrf = 0.;
        if (lm[i].type == ST_RFCAV) lm[i].vrf = 0.;
    }
    double xo[6], xp[6], xm[6];
    for (int j = 0; j < 6; j++) xo[j] = x_co[j];
    for (int k = 0; k < n && !rc; k++) {
        orc_rng rng = {GP.rng_seed, 0u, 0u};
        rc = elem_pass(&lo[k], xo, &rng);
        for (int j = 0; j < 6; j++) X[6 * k + j] = xo[j];
    }
    for (int k = 0; k < n && !rc; k++) {
        const double *x0 = X + 6 * (k > 0 ? k - 1 : n - 1);
        double *m = T + 36 * (size_t)k;
        for (int a = 0; a < 6 && !rc; a++) {
            for (int j = 0; j < 6; j++) xp[j] = xm[j] = x0[j];
            xp[4] = xm[4] = 0.;
            xp[5] = xm[5] = x5;
            xp[a] += h;
            xm[a] -= h;
            orc_rng rp = {GP.rng_seed, (uint32_t)(2 * a), 0u}, rmi = {GP.rng_seed, (uint32_t)(2 * a + 1), 0u};
            if ((rc = elem_pass(&lm[k], xp, &rp))) break;
            if ((rc = elem_pass(&lm[k], xm, &rmi))) break;
            for (int i = 0; i < 6; i++) m[6 * i + a] = (xp[i] - xm[i]) / (2 * h);
        }
    }
    free(lo);
    free(lm);
    return rc;
}

/* Track_tbt_FMA, simtrack.cpp:21179-21241, one particle after another;
   rows as st_fma_scan.  Returns the number of rows, or a negative code. */
EXPORT int orc_fma_scan(const st_element *line, int n, double deltap0, double sigmax0, double sigmay0,
                        const double x_end[4], double tune_low, double tune_high, int nturn, double sigma_step,
                        double sigma_max, double *rows, int max_rows)
{
    int nsigma = (int)((sigma_max - 0) * 1.0 / sigma_step) + 1, m = 0, w = nturn / 2, q = 0;
    double *x_tbt = (double *)malloc(sizeof(double) * (size_t)nturn * 6);
    double *xtemp = (double *)malloc(sizeof(double) * (size_t)w);
    uint64_t draws = orc_draws_per_turn(line, n, 0);
    if (!x_tbt || !xtemp) return -5;
    for (int i = 0; i < nsigma; i++)
        for (int j = 0; j < nsigma; j++) {
            if (!(sqrt(1.0 * i * i + 1.0 * j * j) <= (sigma_max / sigma_step))) continue;
            double x[6];
            x[0] = i * sigma_step * sigmax0 + 0.;
            x[1] = 0.;
            x[2] = j * sigma_step * sigmay0;
            x[3] = 0.;
            x[4] = 0.;
            x[5] = deltap0;
            x[0] = x[0] + x_end[0];
            x[1] = x[1] + x_end[1];
            x[2] = x[2] + x_end[2];
            x[3] = x[3] + x_end[3];
            if (x[0] < 1.0e-09) x[0] = x[0] + 1.0e-09;
            if (x[2] < 1.0e-09) x[2] = x[2] + 1.0e-09;
            int stable = 1, lt = 0, lp = 0;
            /* the GPU scan numbers the grid particles 0, 1, ...; radiation draws depend on it */
            int rc = track_one(line, n, x, nturn, &stable, &lt, &lp, 0, NULL, x_tbt, -1, (uint32_t)q++, draws);
            if (rc) return rc;
            if (stable != 1) continue;
            double t[4];
            for (int k = 0; k < w; k++) xtemp[k] = x_tbt[k * 6 + 0];
            orc_fine_tune(w, xtemp, tune_low, tune_high, &t[0]);
            for (int k = 0; k < w; k++) xtemp[k] = x_tbt[k * 6 + 2];
            orc_fine_tune(w, xtemp, tune_low, tune_high, &t[2]);
            for (int k = 0; k < w; k++) xtemp[k] = x_tbt[(k + w - 1) * 6 + 0];
            orc_fine_tune(w, xtemp, tune_low, tune_high, &t[1]);
            for (int k = 0; k < w; k++) xtemp[k] = x_tbt[(k + w - 1) * 6 + 2];
            orc_fine_tune(w, xtemp, tune_low, tune_high, &t[3]);
            if (m < max_rows) {
                double *o = rows + (size_t)m * 8;
                o[0] = i * sigma_step;
                o[1] = j * sigma_step;
                o[2] = t[0];
                o[3] = t[1];
                o[4] = t[2];
                o[5] = t[3];
                o[6] = t[0] - t[1];
                o[7] = t[2] - t[3];
            }
            m++;
        }
    free(x_tbt);
    free(xtemp);
    return m;
}

/* Track_tbt_Lyapunov, simtrack.cpp:21243-21303, one pair after another;
   rows as st_lyapunov_scan.  Returns the number of rows. */
EXPORT int orc_lyapunov_scan(const st_element *line, int n, double deltap0, double sigmax0, double sigmay0, int nturn,
                             double sigma_step, double sigma_max, double *rows, int max_rows)
{
    int nsigma = (int)((sigma_max - 0) * 1.0 / sigma_step), m = 0, q = 0;
    double distance0 = 1.0e-06;
    double *x1_tbt = (double *)malloc(sizeof(double) * (size_t)nturn * 6);
    double *x2_tbt = (double *)malloc(sizeof(double) * (size_t)nturn * 6);
    double *Lay = (double *)malloc(sizeof(double) * (size_t)nturn);
    uint64_t draws = orc_draws_per_turn(line, n, 0);
    if (!x1_tbt || !x2_tbt || !Lay) return -5;
    for (int i = 0; i < nsigma; i++)
        for (int j = 0; j < nsigma; j++) {
            if (!(sqrt(1.0 * i * i + 1.0 * j * j) <= (sigma_max / sigma_step))) continue;
            double x1[6], x2[6];
            x1[0] = i * sigma_step * sigmax0;
            x1[1] = 0.;
            x1[2] = j * sigma_step * sigmay0;
            x1[3] = 0.;
            x1[4] = 0.;
            x1[5] = deltap0;
            if (x1[0] == 0.) x1[0] = x1[0] + 1.0e-09;
            if (x1[2] == 0.) x1[2] = x1[2] + 1.0e-09;
            x2[0] = x1[0] + 0.707e-06;
            x2[1] = 0.;
            x2[2] = x1[2] + 0.707e-06;
            x2[3] = 0.;
            x2[4] = 0.;
            x2[5] = deltap0;
            int s1 = 1, s2 = 1, lt = 0, lp = 0;
            /* the GPU scan numbers the pair's particles 2q and 2q+1; radiation draws depend on it */
            int rc = track_one(line, n, x1, nturn, &s1, &lt, &lp, 0, NULL, x1_tbt, -1, (uint32_t)(2 * q), draws);
            if (!rc) rc = track_one(line, n, x2, nturn, &s2, &lt, &lp, 0, NULL, x2_tbt, -1, (uint32_t)(2 * q + 1), draws);
            q++;
            if (rc) return rc;
            if (s1 * s2 != 1) continue;
            for (int k = 0; k < nturn; k++) {
                double dd = sqrt((x1_tbt[k * 6 + 0] - x2_tbt[k * 6 + 0]) * (x1_tbt[k * 6 + 0] - x2_tbt[k * 6 + 0]) +
                                 (x1_tbt[k * 6 + 1] - x2_tbt[k * 6 + 1]) * (x1_tbt[k * 6 + 1] - x2_tbt[k * 6 + 1]) +
                                 (x1_tbt[k * 6 + 2] - x2_tbt[k * 6 + 2]) * (x1_tbt[k * 6 + 2] - x2_tbt[k * 6 + 2]) +
                                 (x1_tbt[k * 6 + 3] - x2_tbt[k * 6 + 3]) * (x1_tbt[k * 6 + 3] - x2_tbt[k * 6 + 3]));
                Lay[k] = st_log(dd / distance0) / (k + 1);
            }
            double sum = 0.0;
            for (int k = 0; k < 25; k++) sum = sum + Lay[nturn - 1 - k];
            if (m < max_rows) {
                double *o = rows + (size_t)m * 3;
                o[0] = i * sigma_step;
                o[1] = j * sigma_step;
                o[2] = sum / 25;
            }
            m++;
        }
    free(x1_tbt);
    free(x2_tbt);
    free(Lay);
    return m;
}

/* rnd, simtrack.cpp:283-291 */
static double rnd(double *r)
{
    int m;
    double s, u, v, p;
    s = 65536.0;
    u = 2053.0;
    v = 13849.0;
    m = (int)(*r / s);
    *r = *r - m * s;
    *r = u * *r + v;
    m = (int)(*r / s);
    *r = *r - m * s;
    p = *r / s;
    return (p);
}

/* Cal_Twiss_Emit, simtrack.cpp:771-815 (the emittance) */
static double Cal_Twiss_Emit(const double x[], const double px[], double n)
{
    int i;
    double sum, xmean, pxmean, sigma_xx, sigma_xpx, sigma_pxpx;
    sum = 0.;
    for (i = 0; i < n; i++) sum = sum + x[i];
    xmean = sum / n;
    sum = 0.;
    for (i = 0; i < n; i++) sum = sum + (x[i] - xmean) * (x[i] - xmean);
    sigma_xx = sum / n;
    sum = 0.;
    for (i = 0; i < n; i++) sum = sum + px[i];
    pxmean = sum / n;
    sum = 0.;
    for (i = 0; i < n; i++) sum = sum + (px[i] - pxmean) * (px[i] - pxmean);
    sigma_pxpx = sum / n;
    sum = 0.;
    for (i = 0; i < n; i++) sum = sum + (x[i] - xmean) * (px[i] - pxmean);
    sigma_xpx = sum / n;
    return sqrt(sigma_xx * sigma_pxpx - sigma_xpx * sigma_xpx);
}

/* Cal_Mean_RMS, simtrack.cpp:708-722 */
static void Cal_Mean_RMS(const double x[], int n, double *mean, double *rms)
{
    int i;
    *mean = 0;
    for (i = 0; i < n; i++) *mean += x[i];
    *mean = *mean * 1.0 / n;
    *rms = 0;
    for (i = 0; i < n; i++) *rms += (x[i] - *mean) * (x[i] - *mean);
    *rms = sqrt(*rms * 1.0 / (n - 1));
}

/* Track_tbt_ActionDiff, simtrack.cpp:21305-21383, one particle after
   another, start Twiss as arguments; rows as st_action_diff */
EXPORT int orc_action_diff(const st_element *line, int n, double betx, double alfx, double bety, double alfy,
                           double nsigma, double phase, double emitx0, double emity0, int nturn, int nstep, int npart,
                           double *seed, double *rows)
{
    const double PI = 3.14159265358979323;
    double nsigmax = nsigma * OCOS(phase), nsigmay = nsigma * OSIN(phase);
    double TwoJx0 = (nsigmax * nsigmax) * emitx0, TwoJy0 = (nsigmay * nsigmay) * emity0;
    double *x = (double *)malloc(sizeof(double) * (size_t)npart * 6);
    double *jx = (double *)calloc((size_t)npart, sizeof(double));
    double *jy = (double *)calloc((size_t)npart, sizeof(double));
    double *x_tbt = (double *)malloc(sizeof(double) * (size_t)nturn * 6);
    double *tx = (double *)malloc(sizeof(double) * (size_t)nturn);
    double *tp = (double *)malloc(sizeof(double) * (size_t)nturn);
    uint64_t draws = orc_draws_per_turn(line, n, 0);
    int rc = 0;
    if (!x || !jx || !jy || !x_tbt || !tx || !tp) return -5;
    for (int i = 0; i < npart; i++) {
        double phix0 = rnd(seed) * 360;
        double phiy0 = rnd(seed) * 360;
        double xn = sqrt(TwoJx0) * OCOS(phix0 * PI / 180), pxn = -sqrt(TwoJx0) * OSIN(phix0 * PI / 180);
        x[i * 6 + 0] = xn * sqrt(betx);
        x[i * 6 + 1] = -alfx * xn / sqrt(betx) + pxn / sqrt(betx);
        double yn = sqrt(TwoJy0) * OCOS(phiy0 * PI / 180), pyn = -sqrt(TwoJy0) * OSIN(phiy0 * PI / 180);
        x[i * 6 + 2] = yn * sqrt(bety);
        x[i * 6 + 3] = -alfy * yn / sqrt(bety) + pyn / sqrt(bety);
        x[i * 6 + 4] = 0.;
        x[i * 6 + 5] = 0.;
    }
    for (int j = 0; j < nstep && !rc; j++) {
        int ncount = 0;
        for (int i = 0; i < npart; i++) {
            int stable = 1, lt = 0, lp = 0;
            double *xt = x + (size_t)i * 6;
            if (xt[0] == 0.) xt[0] = xt[0] + 1.0e-09;
            if (xt[2] == 0.) xt[2] = xt[2] + 1.0e-09;
            rc = track_one(line, n, xt, nturn, &stable, &lt, &lp, 0, NULL, x_tbt, -1, (uint32_t)i, draws);
            if (rc) break;
            if (stable == 1) {
                ncount++;
                for (int k = 0; k < nturn; k++) {
                    tx[k] = x_tbt[k * 6 + 0];
                    tp[k] = x_tbt[k * 6 + 1];
                }
                jx[i] = Cal_Twiss_Emit(tx, tp, nturn);
                for (int k = 0; k < nturn; k++) {
                    tx[k] = x_tbt[k * 6 + 2];
                    tp[k] = x_tbt[k * 6 + 3];
                }
                jy[i] = Cal_Twiss_Emit(tx, tp, nturn);
            }
        }
        double *o = rows + (size_t)j * 6;
        o[0] = (double)((j + 1) * nturn);
        o[1] = ncount;
        Cal_Mean_RMS(jx, ncount, &o[2], &o[3]);
        Cal_Mean_RMS(jy, ncount, &o[4], &o[5]);
    }
    free(x);
    free(jx);
    free(jy);
    free(x_tbt);
    free(tx);
    free(tp);
    return rc;
}

/* ------------------------------------------- strong-strong beam-beam */

/* Cal_Mean_RMS(Npart, PartDist, mean[2], rms[2]), simtrack.cpp:724-742:
   sequential sums of x, y, x^2, y^2 over rows of 6 */
static void mean_rms_xy(int n, const double *p, double mean[2], double rms[2])
{
    double sumx = 0, sumy = 0, sumx2 = 0, sumy2 = 0;
    for (int i = 0; i < n; i++) {
        double tx = p[6 * i], ty = p[6 * i + 2];
        sumx = sumx + tx;
        sumy = sumy + ty;
        sumx2 = sumx2 + tx * tx;
        sumy2 = sumy2 + ty * ty;
    }
    mean[0] = sumx / n;
    mean[1] = sumy / n;
    rms[0] = sqrt(sumx2 / n - mean[0] * mean[0]);
    rms[1] = sqrt(sumy2 / n - mean[1] * mean[1]);
}

/* BeamBeam4D_SoftGaussian, simtrack.cpp:22846-22872: both beams' moments
   first, then every particle of each beam kicked by the other beam's
   Gaussian (BB4D about its centroid) */
EXPORT void orc_bb_soft_gaussian_4d(double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2, double gamma1,
                                    double gamma2, double bbscale1, double bbscale2)
{
    double m1[2], r1[2], m2[2], r2[2], dpx, dpy;
    mean_rms_xy(Mp1, x1, m1, r1);
    mean_rms_xy(Mp2, x2, m2, r2);
    for (int i = 0; i < Mp1; i++) {
        orc_bb4d(x1[6 * i] - m2[0], x1[6 * i + 2] - m2[1], gamma1, Np2, r2[0], r2[1], &dpx, &dpy);
        x1[6 * i + 1] += dpx * bbscale1;
        x1[6 * i + 3] += dpy * bbscale1;
    }
    for (int i = 0; i < Mp2; i++) {
        orc_bb4d(x2[6 * i] - m1[0], x2[6 * i + 2] - m1[1], gamma2, Np1, r1[0], r1[1], &dpx, &dpy);
        x2[6 * i + 1] += dpx * bbscale2;
        x2[6 * i + 3] += dpy * bbscale2;
    }
}

/* the collision order of both 6-D variants (cpp:22992-23015, 23192-23213):
   slice pairs (i, j) with S = (Z1[i] - Z2[j]) / 2, ordered by
   Z10 = -(Z1[i] + Z2[j]) / 2 ascending with the reference's exchange sort */
static void ss_schedule(int n1, const double *Z1, int n2, const double *Z2, int *I1, int *I2, double *S)
{
    int n = n1 * n2;
    double *Z10 = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) {
            I1[i * n2 + j] = i;
            I2[i * n2 + j] = j;
            S[i * n2 + j] = (Z1[i] - Z2[j]) / 2.;
            Z10[i * n2 + j] = -(Z1[i] + Z2[j]) / 2.;
        }
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (Z10[j] < Z10[i]) {
                double t = Z10[j];
                Z10[j] = Z10[i];
                Z10[i] = t;
                t = S[j];
                S[j] = S[i];
                S[i] = t;
                int it = I1[j];
                I1[j] = I1[i];
                I1[i] = it;
                it = I2[j];
                I2[j] = I2[i];
                I2[i] = it;
            }
    free(Z10);
}

/* the exchange sort of cpp:22895-22915: rows of 6 by z descending */
static void sort_z_desc(int n, double *x)
{
    double t[6];
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (x[j * 6 + 4] > x[i * 6 + 4]) {
                memcpy(t, x + j * 6, sizeof t);
                memcpy(x + j * 6, x + i * 6, sizeof t);
                memcpy(x + i * 6, t, sizeof t);
            }
}

/* drift by s, the moments, the kicks, the drift back (cpp:23035-23110) on
   slice copies c1 (m1 rows, drift +S) and c2 (m2 rows, drift -S) */
static void ss_collide_pair(int m1, double *c1, int m2, double *c2, double S, double N1, double N2, double gamma1,
                            double gamma2, double bbscale1, double bbscale2)
{
    double mean1[2], rms1[2], mean2[2], rms2[2], dpx, dpy;
    for (int j = 0; j < m1; j++) {
        c1[j * 6 + 0] = c1[j * 6 + 0] + c1[j * 6 + 1] * S;
        c1[j * 6 + 2] = c1[j * 6 + 2] + c1[j * 6 + 3] * S;
    }
    for (int j = 0; j < m2; j++) {
        c2[j * 6 + 0] = c2[j * 6 + 0] - c2[j * 6 + 1] * S;
        c2[j * 6 + 2] = c2[j * 6 + 2] - c2[j * 6 + 3] * S;
    }
    mean_rms_xy(m1, c1, mean1, rms1);
    mean_rms_xy(m2, c2, mean2, rms2);
    for (int j = 0; j < m1; j++) {
        orc_bb4d(c1[6 * j] - mean2[0], c1[6 * j + 2] - mean2[1], gamma1, N2, rms2[0], rms2[1], &dpx, &dpy);
        c1[6 * j + 1] = c1[6 * j + 1] + dpx * bbscale1;
        c1[6 * j + 3] = c1[6 * j + 3] + dpy * bbscale1;
    }
    for (int j = 0; j < m2; j++) {
        orc_bb4d(c2[6 * j] - mean1[0], c2[6 * j + 2] - mean1[1], gamma2, N1, rms1[0], rms1[1], &dpx, &dpy);
        c2[6 * j + 1] = c2[6 * j + 1] + dpx * bbscale2;
        c2[6 * j + 3] = c2[6 * j + 3] + dpy * bbscale2;
    }
    for (int j = 0; j < m1; j++) {
        c1[j * 6 + 0] = c1[j * 6 + 0] - c1[j * 6 + 1] * S;
        c1[j * 6 + 2] = c1[j * 6 + 2] - c1[j * 6 + 3] * S;
    }
    for (int j = 0; j < m2; j++) {
        c2[j * 6 + 0] = c2[j * 6 + 0] + c2[j * 6 + 1] * S;
        c2[j * 6 + 2] = c2[j * 6 + 2] + c2[j * 6 + 3] * S;
    }
}

/* BeamBeam6D_SoftGaussian, simtrack.cpp:22874-23132: both beams sorted by z
   (descending, left sorted), Mp/Nslice particles per slice (the last
   Mp % Nslice rows collide with nothing), slice pairs collided in order of
   their collision point.  Z1/Z2 (may be null) receive the slice centres the
   reference writes to bunch1_Z.dat / bunch2_Z.dat. */
EXPORT void orc_bb_soft_gaussian_6d(double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2, double gamma1,
                                    double gamma2, double bbscale1, double bbscale2, int Nslice1, int Nslice2,
                                    double *Z1o, double *Z2o)
{
    int n = Nslice1 * Nslice2;
    int mc1 = Mp1 / Nslice1, mc2 = Mp2 / Nslice2;
    double *Z1 = (double *)malloc(sizeof(double) * (size_t)Nslice1);
    double *Z2 = (double *)malloc(sizeof(double) * (size_t)Nslice2);
    int *I1 = (int *)malloc(sizeof(int) * (size_t)n), *I2 = (int *)malloc(sizeof(int) * (size_t)n);
    double *S = (double *)malloc(sizeof(double) * (size_t)n);
    sort_z_desc(Mp1, x1);
    sort_z_desc(Mp2, x2);
    for (int i = 0; i < Nslice1; i++) {
        double t = 0;
        for (int j = i * mc1; j < (i + 1) * mc1; j++) t = t + x1[6 * j + 4];
        Z1[i] = t / mc1;
    }
    for (int i = 0; i < Nslice2; i++) {
        double t = 0;
        for (int j = i * mc2; j < (i + 1) * mc2; j++) t = t + x2[6 * j + 4];
        Z2[i] = t / mc2;
    }
    if (Z1o) memcpy(Z1o, Z1, sizeof(double) * (size_t)Nslice1);
    if (Z2o) memcpy(Z2o, Z2, sizeof(double) * (size_t)Nslice2);
    ss_schedule(Nslice1, Z1, Nslice2, Z2, I1, I2, S);
    for (int i = 0; i < n; i++)
        ss_collide_pair(mc1, x1 + (size_t)I1[i] * mc1 * 6, mc2, x2 + (size_t)I2[i] * mc2 * 6, S[i], Np1 / Nslice1,
                        Np2 / Nslice2, gamma1, gamma2, bbscale1, bbscale2);
    free(Z1);
    free(Z2);
    free(I1);
    free(I2);
    free(S);
}

/* BeamBeam6D_SoftGaussian with slice boundaries, simtrack.cpp:23134-23284:
   slice i of a beam holds, in row order, the particles with
   zb[i+1] <= z < zb[i]; rows stay in place.  Each pair collides through
   BeamBeam4D_SoftGaussian with charges scaled by the slice populations.
   The reference clears the slice counters of both beams over Nslice1
   entries (cpp:23161-23165); here every counter starts at zero. */
EXPORT void orc_bb_soft_gaussian_6d_zb(double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2,
                                       double gamma1, double gamma2, double bbscale1, double bbscale2, int Nslice1,
                                       int Nslice2, const double *zb1, const double *zb2)
{
    int n = Nslice1 * Nslice2;
    double *Z1 = (double *)calloc((size_t)Nslice1, sizeof(double));
    double *Z2 = (double *)calloc((size_t)Nslice2, sizeof(double));
    int *I1 = (int *)malloc(sizeof(int) * (size_t)n), *I2 = (int *)malloc(sizeof(int) * (size_t)n);
    double *S = (double *)malloc(sizeof(double) * (size_t)n);
    double *c1 = (double *)malloc(sizeof(double) * 6 * (size_t)(Mp1 > 0 ? Mp1 : 1));
    double *c2 = (double *)malloc(sizeof(double) * 6 * (size_t)(Mp2 > 0 ? Mp2 : 1));
    for (int i = 0; i < Nslice1; i++) {
        double t = 0;
        int m = 0;
        for (int j = 0; j < Mp1; j++)
            if (x1[j * 6 + 4] < zb1[i] && x1[j * 6 + 4] >= zb1[i + 1]) {
                t = t + x1[6 * j + 4];
                m++;
            }
        Z1[i] = t / m;
    }
    for (int i = 0; i < Nslice2; i++) {
        double t = 0;
        int m = 0;
        for (int j = 0; j < Mp2; j++)
            if (x2[j * 6 + 4] < zb2[i] && x2[j * 6 + 4] >= zb2[i + 1]) {
                t = t + x2[6 * j + 4];
                m++;
            }
        Z2[i] = t / m;
    }
    ss_schedule(Nslice1, Z1, Nslice2, Z2, I1, I2, S);
    for (int i = 0; i < n; i++) {
        const double hi1 = zb1[I1[i]], lo1 = zb1[I1[i] + 1], hi2 = zb2[I2[i]], lo2 = zb2[I2[i] + 1];
        int m1 = 0, m2 = 0;
        for (int j = 0; j < Mp1; j++)
            if (x1[j * 6 + 4] < hi1 && x1[j * 6 + 4] >= lo1) memcpy(c1 + 6 * m1++, x1 + 6 * j, 6 * sizeof(double));
        for (int j = 0; j < Mp2; j++)
            if (x2[j * 6 + 4] < hi2 && x2[j * 6 + 4] >= lo2) memcpy(c2 + 6 * m2++, x2 + 6 * j, 6 * sizeof(double));
        for (int j = 0; j < m1; j++) {
            c1[j * 6 + 0] = c1[j * 6 + 0] + c1[j * 6 + 1] * S[i];
            c1[j * 6 + 2] = c1[j * 6 + 2] + c1[j * 6 + 3] * S[i];
        }
        for (int j = 0; j < m2; j++) {
            c2[j * 6 + 0] = c2[j * 6 + 0] - c2[j * 6 + 1] * S[i];
            c2[j * 6 + 2] = c2[j * 6 + 2] - c2[j * 6 + 3] * S[i];
        }
        orc_bb_soft_gaussian_4d(Np1 * m1 / Mp1, m1, c1, Np2 * m2 / Mp2, m2, c2, gamma1, gamma2, bbscale1, bbscale2);
        for (int j = 0; j < m1; j++) {
            c1[j * 6 + 0] = c1[j * 6 + 0] - c1[j * 6 + 1] * S[i];
            c1[j * 6 + 2] = c1[j * 6 + 2] - c1[j * 6 + 3] * S[i];
        }
        for (int j = 0; j < m2; j++) {
            c2[j * 6 + 0] = c2[j * 6 + 0] + c2[j * 6 + 1] * S[i];
            c2[j * 6 + 2] = c2[j * 6 + 2] + c2[j * 6 + 3] * S[i];
        }
        m1 = m2 = 0;
        for (int j = 0; j < Mp1; j++)
            if (x1[j * 6 + 4] < hi1 && x1[j * 6 + 4] >= lo1) memcpy(x1 + 6 * j, c1 + 6 * m1++, 6 * sizeof(double));
        for (int j = 0; j < Mp2; j++)
            if (x2[j * 6 + 4] < hi2 && x2[j * 6 + 4] >= lo2) memcpy(x2 + 6 * j, c2 + 6 * m2++, 6 * sizeof(double));
    }
    free(Z1);
    free(Z2);
    free(I1);
    free(I2);
    free(S);
    free(c1);
    free(c2);
}

/* Zboundary_slices_Gaussian, simtrack.cpp:3157-3168 */
EXPORT void orc_zboundary_slices_gaussian(double *zb, int Nslice, double sigmal)
{
    zb[0] = sigmal * 20;
    zb[Nslice] = -sigmal * 20;
    for (int i = 1; i <= Nslice - 1; i++) zb[i] = -st_ugaussian_pinv(i * 1.0 / Nslice) * sigmal;
}

/* ------------------------------------------- strong-strong PIC (4-D/6-D) */

#define PIC_NX 64
#define PIC_NY 64
#define PIC_NSIGMA 15

/* Cal_Grid_Weight, simtrack.cpp:2755-2791: the 3-point (TSC) weights */
static void pic_grid_weight(double x1, double y1, double x0, double y0, double hx, double hy, int xm[3], int ym[3],
                            double xw[3], double yw[3])
{
    double dx = x1 - x0, dy = y1 - y0, fx, fy, rx, rx2, ry, ry2;
    int ifx, ify;
    fx = dx / hx;
    ifx = (int)fx;
    fy = dy / hy;
    ify = (int)fy;
    if ((fx - ifx) <= 0.5) {
        xm[0] = ifx;
        xm[1] = ifx + 1;
        xm[2] = ifx - 1;
        rx = fx - ifx;
        rx2 = rx * rx;
    } else {
        xm[0] = ifx + 1;
        xm[1] = ifx;
        xm[2] = ifx + 2;
        rx = ifx + 1 - fx;
        rx2 = rx * rx;
    }
    xw[0] = 0.75 - rx2;
    xw[1] = 0.5 * (0.25 + rx + rx2);
    xw[2] = 0.5 * (0.25 - rx + rx2);
    if ((fy - ify) <= 0.5) {
        ym[0] = ify;
        ym[1] = ify + 1;
        ym[2] = ify - 1;
        ry = fy - ify;
        ry2 = ry * ry;
    } else {
        ym[0] = ify + 1;
        ym[1] = ify;
        ym[2] = ify + 2;
        ry = ify + 1 - fy;
        ry2 = ry * ry;
    }
    yw[0] = 0.75 - ry2;
    yw[1] = 0.5 * (0.25 + ry + ry2);
    yw[2] = 0.5 * (0.25 - ry + ry2);
}

/* int_green and calc_egreen, simtrack.cpp:2808-2879 */
static double pic_int_green(double x, double y)
{
    if (0 == x || 0 == y) return 0;
    double x2 = x * x, y2 = y * y;
    return -3 * x * y + x2 * st_atan(y / x) + y2 * st_atan(x / y) + x * y * st_log(x2 + y2);
}

static double pic_egreen(double x1, double x2, double y1, double y2, double hx, double hy)
{
    if (x1 * x2 * y1 * y2 != 0.) {
        int s11 = 1, s12 = 1, s21 = 1, s22 = 1;
        if (x1 > 0) {
            s11 *= -1;
            s12 *= -1;
        } else if (x1 < 0 && x2 < 0) {
            s21 *= -1;
            s22 *= -1;
        }
        if (y1 > 0) {
            s11 *= -1;
            s21 *= -1;
        } else if (y1 < 0 && y2 < 0) {
            s12 *= -1;
            s22 *= -1;
        }
        x1 = fabs(x1);
        x2 = fabs(x2);
        y1 = fabs(y1);
        y2 = fabs(y2);
        double v11 = pic_int_green(x1, y1), v12 = pic_int_green(x1, y2);
        double v21 = pic_int_green(x2, y1), v22 = pic_int_green(x2, y2);
        double r = s11 * v11 + s12 * v12 + s21 * v21 + s22 * v22;
        return -1.0 * r / hx / hy / 2;
    }
    double r = 0.;
    for (int i = 0; i < 20; i++)
        for (int j = 0; j < 20; j++) {
            double tx = x1 + (i - 1) * hx / 20 + hx / 40;
            double ty = y1 + (j - 1) * hy / 20 + hy / 40;
            if (tx != 0. && ty != 0.) r = r - 0.5 * st_log(tx * tx + ty * ty);
        }
    return r / 400.;
}

/* Cal_Effective_Green_Function, simtrack.cpp:2881-2921 (before its FFT) */
static void pic_green(double xoff, double yoff, double hx, double hy, double *g)
{
    const int Nx = PIC_NX, Ny = PIC_NY;
    for (int i = 0; i < 2 * Nx; i++)
        for (int j = 0; j < 2 * Ny; j++) {
            double x1 = (i <= Nx - 1) ? xoff + i * hx - 0.5 * hx : xoff - (2 * Nx - i) * hx - 0.5 * hx;
            double y1 = (j <= Ny - 1) ? yoff + j * hy - 0.5 * hy : yoff - (2 * Ny - j) * hy - 0.5 * hy;
            g[i * 2 * Nx + j] = pic_egreen(x1, x1 + hx, y1, y1 + hy, hx, hy);
        }
}

/* FFT1D / FFT2D, simtrack.cpp:2583-2698: radix 2, twiddles by recurrence,
   dir == 1 divides by the length */
static void pic_fft1d(double *x, double *y, int m, int dir)
{
    long nn = 1, i, i1, j, k, i2, l, l1, l2;
    double c1, c2, tx, ty, t1, t2, u1, u2, z;
    for (i = 0; i < m; i++) nn *= 2;
    i2 = nn >> 1;
    j = 0;
    for (i = 0; i < nn - 1; i++) {
        if (i < j) {
            tx = x[i];
            ty = y[i];
            x[i] = x[j];
            y[i] = y[j];
            x[j] = tx;
            y[j] = ty;
        }
        k = i2;
        while (k <= j) {
            j -= k;
            k >>= 1;
        }
        j += k;
    }
    c1 = -1.0;
    c2 = 0.0;
    l2 = 1;
    for (l = 0; l < m; l++) {
        l1 = l2;
        l2 <<= 1;
        u1 = 1.0;
        u2 = 0.0;
        for (j = 0; j < l1; j++) {
            for (i = j; i < nn; i += l2) {
                i1 = i + l1;
                t1 = u1 * x[i1] - u2 * y[i1];
                t2 = u1 * y[i1] + u2 * x[i1];
                x[i1] = x[i] - t1;
                y[i1] = y[i] - t2;
                x[i] += t1;
                y[i] += t2;
            }
            z = u1 * c1 - u2 * c2;
            u2 = u1 * c2 + u2 * c1;
            u1 = z;
        }
        c2 = sqrt((1.0 - c1) / 2.0);
        if (dir == 1) c2 = -c2;
        c1 = sqrt((1.0 + c1) / 2.0);
    }
    if (dir == 1)
        for (i = 0; i < nn; i++) {
            x[i] /= (double)nn;
            y[i] /= (double)nn;
        }
}

/* FFT2D for the square 2Nx x 2Ny grid (n = 128, m = 7) */
static void pic_fft2d(double *x, double *y, int n, int m, int dir)
{
    double re[128], im[128];
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < n; i++) {
            re[i] = x[i * n + j];
            im[i] = y[i * n + j];
        }
        pic_fft1d(re, im, m, dir);
        for (int i = 0; i < n; i++) {
            x[i * n + j] = re[i];
            y[i * n + j] = im[i];
        }
    }
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            re[j] = x[i * n + j];
            im[j] = y[i * n + j];
        }
        pic_fft1d(re, im, m, dir);
        for (int j = 0; j < n; j++) {
            x[i * n + j] = re[j];
            y[i * n + j] = im[j];
        }
    }
}

EXPORT void orc_fft2d_128(double *x, double *y, int dir) { pic_fft2d(x, y, 128, 7, dir); }

/* the field of one beam's charge (Charge2Grid, Cal_Electric_PIC with
   Cal_Potential_Convolution and Cal_Electric_Field, cpp:2793-3045) on the
   grid at (x0, y0), as seen at the other beam's offset */
static void pic_potential(int m, const double *xc, const double mean[2], double x0, double y0, double xoff,
                          double yoff, double hx, double hy, double *phi)
{
    const int Nx = PIC_NX, Ny = PIC_NY, N2 = 4 * PIC_NX * PIC_NY;
    const double eps0 = 8.854187817e-12;
    double *src = (double *)calloc((size_t)Nx * Ny, sizeof(double));
    double *gr = (double *)calloc((size_t)N2, sizeof(double)), *gi = (double *)calloc((size_t)N2, sizeof(double));
    double *sr = (double *)calloc((size_t)N2, sizeof(double)), *si = (double *)calloc((size_t)N2, sizeof(double));
    double *pr = (double *)calloc((size_t)N2, sizeof(double)), *pi = (double *)calloc((size_t)N2, sizeof(double));
    int xm[3], ym[3];
    double xw[3], yw[3];
    for (int k = 0; k < m; k++) {
        double xp = xc[6 * k], yp = xc[6 * k + 2];
        if (fabs(xp - mean[0]) < (Nx / 2 - 3) * hx && fabs(yp - mean[1]) < (Ny / 2 - 3) * hy) {
            pic_grid_weight(xp, yp, x0, y0, hx, hy, xm, ym, xw, yw);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) src[xm[i] * Nx + ym[j]] = src[xm[i] * Nx + ym[j]] + xw[i] * yw[j];
        }
    }
    pic_green(xoff, yoff, hx, hy, gr);
    pic_fft2d(gr, gi, 2 * Nx, 7, 1);
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++) sr[i * 2 * Nx + j] = src[i * Nx + j];
    pic_fft2d(sr, si, 2 * Nx, 7, 1);
    for (int i = 0; i < N2; i++) {
        pr[i] = sr[i] * gr[i] - si[i] * gi[i];
        pi[i] = sr[i] * gi[i] + si[i] * gr[i];
    }
    pic_fft2d(pr, pi, 2 * Nx, 7, -1);
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++) phi[i * Nx + j] = pr[i * 2 * Nx + j] * 2 * Nx * 2 * Ny / 2 / 3.14159265 / eps0;
    free(src);
    free(gr);
    free(gi);
    free(sr);
    free(si);
    free(pr);
    free(pi);
}

static void pic_field(int m, const double *xc, const double mean[2], double x0, double y0, double xoff, double yoff,
                      double hx, double hy, double *Ex, double *Ey)
{
    const int Nx = PIC_NX, Ny = PIC_NY;
    double *phi = (double *)calloc((size_t)Nx * Ny, sizeof(double));
    pic_potential(m, xc, mean, x0, y0, xoff, yoff, hx, hy, phi);
    for (int i = 0; i < Nx * Ny; i++) Ex[i] = Ey[i] = 0.;
    for (int i = 1; i < Nx - 1; i++)
        for (int j = 1; j < Ny - 1; j++) {
            Ex[i * Nx + j] = -((phi[(i + 1) * Nx + j + 1] - phi[(i - 1) * Nx + j + 1]) +
                               4 * (phi[(i + 1) * Nx + j] - phi[(i - 1) * Nx + j]) +
                               (phi[(i + 1) * Nx + j - 1] - phi[(i - 1) * Nx + j - 1])) /
                             12. / hx;
            Ey[i * Nx + j] = -((phi[(i + 1) * Nx + j + 1] - phi[(i + 1) * Nx + j - 1]) +
                               4 * (phi[i * Nx + j + 1] - phi[i * Nx + j - 1]) +
                               (phi[(i - 1) * Nx + j + 1] - phi[(i - 1) * Nx + j - 1])) /
                             12. / hy;
        }
    free(phi);
}

/* Interpolate_Electric_Field (cpp:3021-3036) and the kick of cpp:23337-23340 */
static void pic_kick(int m, double *xc, const double mean[2], double x0, double y0, double xoff, double yoff,
                     double hx, double hy, const double *Ex, const double *Ey, double Nsrc, int Msrc, double gamma,
                     double bbscale)
{
    const int Nx = PIC_NX, Ny = PIC_NY;
    int xm[3], ym[3];
    double xw[3], yw[3];
    for (int k = 0; k < m; k++) {
        double xp = xc[6 * k], yp = xc[6 * k + 2];
        if (fabs(xp - mean[0]) < (Nx / 2 - 3) * hx && fabs(yp - mean[1]) < (Ny / 2 - 3) * hy) {
            pic_grid_weight(xp - xoff, yp - yoff, x0, y0, hx, hy, xm, ym, xw, yw);
            double ex = 0., ey = 0.;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    ex = ex + Ex[xm[i] * Nx + ym[j]] * xw[i] * yw[j];
                    ey = ey + Ey[xm[i] * Nx + ym[j]] * xw[i] * yw[j];
                }
            xc[6 * k + 1] = xc[6 * k + 1] + ex * 1.60217662e-19 * bbscale / (gamma * 938.27201e6) * (Nsrc / Msrc);
            xc[6 * k + 3] = xc[6 * k + 3] + ey * 1.60217662e-19 * bbscale / (gamma * 938.27201e6) * (Nsrc / Msrc);
        }
    }
}

/* BeamBeam4D_PIC, simtrack.cpp:23286-23378: a 64 x 64 grid of nsigma = 15
   times the smaller rms size, centred on the source beam; beam 2 is kicked
   by beam 1's field, then beam 1 by beam 2's */
EXPORT void orc_bb_pic_4d(double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2, double gamma1,
                          double gamma2, double bbscale1, double bbscale2)
{
    const int Nx = PIC_NX, Ny = PIC_NY;
    double m1[2], r1[2], m2[2], r2[2];
    double *Ex = (double *)malloc(sizeof(double) * Nx * Ny), *Ey = (double *)malloc(sizeof(double) * Nx * Ny);
    mean_rms_xy(Mp1, x1, m1, r1);
    mean_rms_xy(Mp2, x2, m2, r2);
    double rmsx = r1[0] > r2[0] ? r2[0] : r1[0];
    double rmsy = r1[1] > r2[1] ? r2[1] : r1[1];
    double hx = PIC_NSIGMA * rmsx / Nx, hy = PIC_NSIGMA * rmsy / Ny;
    double x0 = m1[0] - Nx * hx / 2, y0 = m1[1] - Ny * hy / 2;
    double xoff = m2[0] - m1[0], yoff = m2[1] - m1[1];
    pic_field(Mp1, x1, m1, x0, y0, xoff, yoff, hx, hy, Ex, Ey);
    pic_kick(Mp2, x2, m2, x0, y0, xoff, yoff, hx, hy, Ex, Ey, Np1, Mp1, gamma2, bbscale2);
    x0 = m2[0] - Nx * hx / 2;
    y0 = m2[1] - Ny * hy / 2;
    xoff = m1[0] - m2[0];
    yoff = m1[1] - m2[1];
    pic_field(Mp2, x2, m2, x0, y0, xoff, yoff, hx, hy, Ex, Ey);
    pic_kick(Mp1, x1, m1, x0, y0, xoff, yoff, hx, hy, Ex, Ey, Np2, Mp2, gamma1, bbscale1);
    free(Ex);
    free(Ey);
}

/* Interpolate_Electric_Field_1Potential (cpp:3047-3100): the field at the
   3 x 3 weighted nodes by the same 4th-order stencil, taken from a 5 x 5
   patch of the potential around them, then the weighted sum */
static void pic_field_1potential(const double *phi, double x0, double y0, double hx, double hy, double x, double y,
                                 double *ex, double *ey)
{
    const int Nx = PIC_NX;
    int xm[3], ym[3];
    double xw[3], yw[3], p25[25], Ex25[25], Ey25[25];
    pic_grid_weight(x, y, x0, y0, hx, hy, xm, ym, xw, yw);
    int i0 = 10000, j0 = 10000;
    for (int i = 0; i < 3; i++) {
        if (xm[i] < i0) i0 = xm[i];
        if (ym[i] < j0) j0 = ym[i];
    }
    for (int i = 0; i < 25; i++) p25[i] = Ex25[i] = Ey25[i] = 0.;
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) p25[5 * i + j] = phi[Nx * (i0 - 1 + i) + (j0 - 1 + j)];
    for (int i = 1; i < 4; i++)
        for (int j = 1; j < 4; j++) {
            Ex25[i * 5 + j] = -((p25[(i + 1) * 5 + j + 1] - p25[(i - 1) * 5 + j + 1]) +
                                4 * (p25[(i + 1) * 5 + j] - p25[(i - 1) * 5 + j]) +
                                (p25[(i + 1) * 5 + j - 1] - p25[(i - 1) * 5 + j - 1])) /
                              12. / hx;
            Ey25[i * 5 + j] = -((p25[(i + 1) * 5 + j + 1] - p25[(i + 1) * 5 + j - 1]) +
                                4 * (p25[i * 5 + j + 1] - p25[i * 5 + j - 1]) +
                                (p25[(i - 1) * 5 + j + 1] - p25[(i - 1) * 5 + j - 1])) /
                              12. / hy;
        }
    double fx = 0., fy = 0.;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            fx = fx + Ex25[(xm[i] - i0 + 1) * 5 + (ym[j] - j0 + 1)] * xw[i] * yw[j];
            fy = fy + Ey25[(xm[i] - i0 + 1) * 5 + (ym[j] - j0 + 1)] * xw[i] * yw[j];
        }
    *ex = fx;
    *ey = fy;
}

/* BeamBeam4D_PIC_NonOffset, simtrack.cpp:24226-24280: a fixed 64 x 64 grid
   of cells hx x hy centred on the origin (no centroid offsets), the Green
   function of the unshifted grid (Cal_Effective_Green_Function with zero
   offsets, which the reference's caller computes once and passes in), the
   field interpolated from the potential; beam 2 kicked by beam 1, then
   beam 1 by beam 2; particles deposited or kicked only within 29 cells of
   the origin */
EXPORT void orc_bb_pic_4d_nonoffset(double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2, double gamma1,
                                    double gamma2, double bbscale1, double bbscale2, double hx, double hy)
{
    const int Nx = PIC_NX, Ny = PIC_NY;
    const double x0 = -Nx * hx / 2, y0 = -Ny * hy / 2, zero[2] = {0., 0.};
    double *phi = (double *)calloc((size_t)Nx * Ny, sizeof(double));
    for (int dir = 0; dir < 2; dir++) {
        const int ms = dir ? Mp2 : Mp1, mt = dir ? Mp1 : Mp2;
        const double *xs = dir ? x2 : x1, Ns = dir ? Np2 : Np1;
        double *xt = dir ? x1 : x2;
        const double gt = dir ? gamma1 : gamma2, kt = dir ? bbscale1 : bbscale2;
        pic_potential(ms, xs, zero, x0, y0, 0., 0., hx, hy, phi);
        for (int k = 0; k < mt; k++) {
            const double xp = xt[6 * k], yp = xt[6 * k + 2];
            if (fabs(xp) < (Nx / 2 - 3) * hx && fabs(yp) < (Ny / 2 - 3) * hy) {
                double ex, ey;
                pic_field_1potential(phi, x0, y0, hx, hy, xp, yp, &ex, &ey);
                xt[6 * k + 1] = xt[6 * k + 1] + ex * 1.60217662e-19 * kt / (gt * 938.27201e6) * (Ns / ms);
                xt[6 * k + 3] = xt[6 * k + 3] + ey * 1.60217662e-19 * kt / (gt * 938.27201e6) * (Ns / ms);
            }
        }
    }
    free(phi);
}

/* The same fixed-grid collision by the full-grid field of BeamBeam4D_PIC
   (Cal_Electric_Field, then Interpolate_Electric_Field): a check of the
   potential interpolation above, which must equal it wherever the window
   keeps the stencil inside the grid */
EXPORT void orc_bb_pic_4d_fixed_gridfield(double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2,
                                          double gamma1, double gamma2, double bbscale1, double bbscale2, double hx,
                                          double hy)
{
    const int Nx = PIC_NX, Ny = PIC_NY;
    const double x0 = -Nx * hx / 2, y0 = -Ny * hy / 2, zero[2] = {0., 0.};
    double *Ex = (double *)malloc(sizeof(double) * Nx * Ny), *Ey = (double *)malloc(sizeof(double) * Nx * Ny);
    pic_field(Mp1, x1, zero, x0, y0, 0., 0., hx, hy, Ex, Ey);
    pic_kick(Mp2, x2, zero, x0, y0, 0., 0., hx, hy, Ex, Ey, Np1, Mp1, gamma2, bbscale2);
    pic_field(Mp2, x2, zero, x0, y0, 0., 0., hx, hy, Ex, Ey);
    pic_kick(Mp1, x1, zero, x0, y0, 0., 0., hx, hy, Ex, Ey, Np2, Mp2, gamma1, bbscale1);
    free(Ex);
    free(Ey);
}

/* BeamBeam6D_PIC, simtrack.cpp:23380-23615: the z-sorted slicing and the
   collision order of BeamBeam6D_SoftGaussian, BeamBeam4D_PIC per pair on
   the drifted slices with charges Np/Nslice */
EXPORT void orc_bb_pic_6d(double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2, double gamma1,
                          double gamma2, double bbscale1, double bbscale2, int Nslice1, int Nslice2)
{
    int n = Nslice1 * Nslice2;
    int mc1 = Mp1 / Nslice1, mc2 = Mp2 / Nslice2;
    double *Z1 = (double *)malloc(sizeof(double) * (size_t)Nslice1);
    double *Z2 = (double *)malloc(sizeof(double) * (size_t)Nslice2);
    int *I1 = (int *)malloc(sizeof(int) * (size_t)n), *I2 = (int *)malloc(sizeof(int) * (size_t)n);
    double *S = (double *)malloc(sizeof(double) * (size_t)n);
    sort_z_desc(Mp1, x1);
    sort_z_desc(Mp2, x2);
    for (int i = 0; i < Nslice1; i++) {
        double t = 0;
        for (int j = i * mc1; j < (i + 1) * mc1; j++) t = t + x1[6 * j + 4];
        Z1[i] = t / mc1;
    }
    for (int i = 0; i < Nslice2; i++) {
        double t = 0;
        for (int j = i * mc2; j < (i + 1) * mc2; j++) t = t + x2[6 * j + 4];
        Z2[i] = t / mc2;
    }
    ss_schedule(Nslice1, Z1, Nslice2, Z2, I1, I2, S);
    for (int p = 0; p < n; p++) {
        double *c1 = x1 + (size_t)I1[p] * mc1 * 6, *c2 = x2 + (size_t)I2[p] * mc2 * 6;
        for (int j = 0; j < mc1; j++) {
            c1[j * 6 + 0] = c1[j * 6 + 0] + c1[j * 6 + 1] * S[p];
            c1[j * 6 + 2] = c1[j * 6 + 2] + c1[j * 6 + 3] * S[p];
        }
        for (int j = 0; j < mc2; j++) {
            c2[j * 6 + 0] = c2[j * 6 + 0] - c2[j * 6 + 1] * S[p];
            c2[j * 6 + 2] = c2[j * 6 + 2] - c2[j * 6 + 3] * S[p];
        }
        orc_bb_pic_4d(Np1 / Nslice1, mc1, c1, Np2 / Nslice2, mc2, c2, gamma1, gamma2, bbscale1, bbscale2);
        for (int j = 0; j < mc1; j++) {
            c1[j * 6 + 0] = c1[j * 6 + 0] - c1[j * 6 + 1] * S[p];
            c1[j * 6 + 2] = c1[j * 6 + 2] - c1[j * 6 + 3] * S[p];
        }
        for (int j = 0; j < mc2; j++) {
            c2[j * 6 + 0] = c2[j * 6 + 0] + c2[j * 6 + 1] * S[p];
            c2[j * 6 + 2] = c2[j * 6 + 2] + c2[j * 6 + 3] * S[p];
        }
    }
    free(Z1);
    free(Z2);
    free(I1);
    free(I2);
    free(S);
}
