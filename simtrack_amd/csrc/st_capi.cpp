/*
 * st_capi.cpp -- host side of the C ABI (include/simtrack_amd.h): context
 * and buffer management, lowering of st_element records to the packed device
 * table, and the tracking drivers (turn blocks with survivor compaction).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "simtrack_amd.h"

#define ST_FN static inline
#define ST_CONST_TABLE static const
#define ST_FMA(a, b, c) std::fma((a), (b), (c))
#define ST_RINT(x) std::nearbyint(x)
#define ST_SQRT(x) std::sqrt(x)
#include "st_libm.h"
#include "st_kernels.h"
#include "st_ssbb.h"
#include "st_tune.h"
#include "st_table_format.h"

/* ------------------------------------------------------------ errors */

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return fail(ST_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

extern "C" const char *st_last_error(void) { return g_err.c_str(); }

/* error reporting for the library's other translation units (not exported
   in the public header) */
extern "C" __attribute__((visibility("hidden"))) int st_internal_fail(int code, const char *msg)
{
    return fail(code, "%s", msg);
}

struct st_ctx;
/* the context's stream and device, for the library's other translation units */
extern "C" __attribute__((visibility("hidden"))) hipStream_t st_internal_stream(st_ctx *c, int *device);
extern "C" int st_abi_version(void) { return ST_ABI_VERSION; }
extern "C" void st_struct_sizes(int *element_bytes, int *globals_bytes)
{
    if (element_bytes) *element_bytes = (int)sizeof(st_element);
    if (globals_bytes) *globals_bytes = (int)sizeof(st_globals);
}

/* ------------------------------------------------------ reference GP */

/* 4th-order Yoshida coefficients, simtrack.cpp:29-32 (run-time pow, as at the
   reference's static initialisation) */
static double Fdrift1, Fdrift2, Fkick1, Fkick2;
static void init_yoshida()
{
    static bool done = false;
    if (done) return;
    volatile double two = 2.e0, third = 1.0 / 3.0;
    Fdrift1 = 1.e0 / (2.e0 * (2.e0 - std::pow(two, third)));
    Fdrift2 = 0.5e0 - Fdrift1;
    Fkick1 = 2.e0 * Fdrift1;
    Fkick2 = 1.e0 - 2e0 * Fkick1;
    done = true;
}

extern "C" void st_globals_default(st_globals *gp)
{
    /* GlobalVariables(), simtrack.h:31-58 */
    std::memset(gp, 0, sizeof *gp);
    gp->A = 1;
    gp->Q = 1;
    gp->m0 = 1.6726216e-27;
    gp->m = 1.6726216e-27;
    gp->q = 1.6021765e-19;
    gp->gamma = 106.;
    gp->beta = 0.999955;
    gp->energy0 = 938.27201;
    gp->energy = 99456.8;
    gp->p = 1.;
    gp->cp = 1.;
    gp->brho = 331.738;
    gp->G = 1.7928474;
    gp->H_expand = 0;
    gp->radiate = 0;
    gp->quad_fringe = 0;
    gp->rng_seed = 0;
}

extern "C" void st_globals_set_energy(st_globals *gp, double gamma, double E0, double Q, double A)
{
    /* Set_RefPartEnergy, simtrack.cpp:87-108 */
    gp->Q = (int32_t)Q;
    gp->A = (int32_t)A;
    gp->m0 = E0 * 1.0e6 * 1.6021765e-19 / 299792458. / 299792458.;
    gp->m = gamma * gp->m0;
    gp->q = Q * 1.6021765e-19;
    gp->gamma = gamma;
    gp->beta = std::sqrt(1.0 - 1.0 / gamma / gamma);
    gp->energy0 = E0;
    gp->energy = gamma * gp->energy0;
    gp->cp = gp->energy * gp->beta;
    gp->p = gp->energy * gp->beta * 1.0e6 * 1.6021765e-19 / 299792458.;
    gp->brho = 3.33564095198 * gp->beta * (gp->energy / 1000) / gp->Q;
}

/* ------------------------------------------------------- structures */

#define SS_LANES 8 /* slice pairs in flight at once */

struct st_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_gpu_s = 0.0;
    /* grow-only work buffers */
    double *d_x = nullptr;
    size_t cap_x = 0;
    int *d_state = nullptr, *d_lost = nullptr, *d_idx0 = nullptr, *d_idx1 = nullptr, *d_block = nullptr;
    int *d_total = nullptr, *h_total = nullptr;
    size_t cap_p = 0;
    double *d_tbt = nullptr;
    size_t cap_tbt = 0;
    double *d_x04 = nullptr;
    size_t cap_x04 = 0;
    /* per-launch kernel timing */
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> pending;  /* indices into ev_pool */
    std::vector<int64_t> pending_work;
    double k_seconds = 0.0;
    int64_t k_launches = 0, k_work = 0;
    /* strong-strong beam-beam work area (grow-only) */
    char *d_ss = nullptr;
    size_t cap_ss = 0;
    /* streams for slice pairs that touch disjoint slices (created on first
       use, default priority) and the events that fork and join them */
    hipStream_t lane[SS_LANES] = {};
    hipEvent_t lane_ev[SS_LANES] = {};
    hipEvent_t fork_ev = nullptr;
};

struct dev_table {
    st_op *d_ops = nullptr;
    double *d_par = nullptr;
    std::vector<st_op> ops;
    std::vector<double> par;
};

struct st_table {
    st_ctx *ctx = nullptr;
    st_globals gp;
    st_kglobals kg;
    int n = 0;
    dev_table orbit, spin;
    double *d_ap = nullptr;
    st_kglobals *d_kg = nullptr; /* kg in device memory, read by the kernels */
    int64_t draws_per_turn = 0;
    int radial_ok = 1;
};

/* ---------------------------------------------------------- lowering */

namespace {

struct Builder {
    const st_globals &gp;
    bool spin;
    dev_table &t;
    std::string err;

    Builder(const st_globals &g, bool s, dev_table &tab) : gp(g), spin(s), t(tab) {}

    int put(double v)
    {
        t.par.push_back(v);
        return (int)t.par.size() - 1;
    }

    st_op new_op(int kind)
    {
        st_op op;
        std::memset(&op, 0, sizeof op);
        op.kind = kind;
        op.po = (int)t.par.size();
        op.pm = -1;
        return op;
    }

    /* Cal_W_Matrix, simtrack.cpp:3485-3503 */
    static void w_matrix(double theta, double phi, double psi, double W[9])
    {
        double costhe = st_cos(theta), cosphi = st_cos(phi), cospsi = st_cos(psi);
        double sinthe = st_sin(theta), sinphi = st_sin(phi), sinpsi = st_sin(psi);
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

    /* mat_mult, simtrack.cpp:1351-1362 */
    static void mat_mult(const double *a, const double *b, double *c, int m, int n, int k)
    {
        for (int i = 0; i <= m - 1; i++)
            for (int j = 0; j <= k - 1; j++) {
                int u = i * k + j;
                c[u] = 0.0;
                for (int l = 0; l <= n - 1; l++) c[u] = c[u] + a[i * n + l] * b[l * k + j];
            }
    }

    /*
     * Misalignment block of GtoL/LtoG (simtrack.cpp:3701-3741) for an element of
     * length L and bend angle `angle`: offset/roll fast path (GtoL1/LtoG1) or
     * the general MAD-X frame change (GtoL2/LtoG2) with its element-constant
     * matrices evaluated here.
     */
    void misalign(st_op &op, const st_element &e, double L, double angle)
    {
        double dx = e.dx, dy = e.dy, ds = e.ds, theta = e.dtheta, phi = e.dphi, psi = e.dpsi;
        double temp = std::fabs(dx) + std::fabs(dy) + std::fabs(ds) + std::fabs(theta) + std::fabs(phi) +
                      std::fabs(psi);
        if (temp < 1.0e-9) return;
        if (std::fabs(theta) + std::fabs(phi) < 1.0e-9) {
            op.flags |= ST_F_MIS1;
            if (std::fabs(dx) + std::fabs(dy) > 1.0e-10) op.flags |= ST_F_MIS1_OFS;
            if (std::fabs(psi) > 1.0e-10) op.flags |= ST_F_MIS1_ROT;
            op.pm = put(dx);
            put(dy);
            put(st_cos(psi));
            put(st_sin(psi));
            return;
        }
        op.flags |= ST_F_MIS2;
        const double beta = gp.beta, gamma = gp.gamma;
        double W[9], RM[36];
        /* GtoL2, cpp:3567-3626 */
        w_matrix(theta, phi, psi, W);
        double s2 = (W[0 * 3 + 2] * dx + W[1 * 3 + 2] * dy + W[2 * 3 + 2] * ds) / W[2 * 3 + 2];
        for (int i = 0; i < 36; i++) RM[i] = 0.;
        for (int i = 0; i < 6; i++) RM[i * 6 + i] = 1.0;
        RM[1 * 6 + 1] = W[0 * 3 + 0];
        RM[1 * 6 + 3] = W[1 * 3 + 0];
        RM[1 * 6 + 5] = W[2 * 3 + 0] / beta;
        RM[3 * 6 + 1] = W[0 * 3 + 1];
        RM[3 * 6 + 3] = W[1 * 3 + 1];
        RM[3 * 6 + 5] = W[2 * 3 + 1] / beta;
        RM[0 * 6 + 0] = W[1 * 3 + 1] / W[2 * 3 + 2];
        RM[0 * 6 + 1] = RM[0 * 6 + 0] * s2;
        RM[0 * 6 + 2] = -W[0 * 3 + 1] / W[2 * 3 + 2];
        RM[0 * 6 + 3] = RM[0 * 6 + 2] * s2;
        RM[2 * 6 + 0] = -W[1 * 3 + 0] / W[2 * 3 + 2];
        RM[2 * 6 + 1] = RM[2 * 6 + 0] * s2;
        RM[2 * 6 + 2] = W[0 * 3 + 0] / W[2 * 3 + 2];
        RM[2 * 6 + 3] = RM[2 * 6 + 2] * s2;
        RM[4 * 6 + 0] = W[0 * 3 + 2] / (W[2 * 3 + 2] * beta);
        RM[4 * 6 + 1] = RM[4 * 6 + 0] * s2;
        RM[4 * 6 + 2] = W[1 * 3 + 2] / (W[2 * 3 + 2] * beta);
        RM[4 * 6 + 3] = RM[4 * 6 + 2] * s2;
        RM[4 * 6 + 5] = -s2 / (beta * gamma * beta * gamma);
        op.pm = (int)t.par.size();
        for (int i = 0; i < 36; i++) put(RM[i]);
        put(-((W[1 * 3 + 1] * dx - W[0 * 3 + 1] * dy) / W[2 * 3 + 2]));
        put(W[2 * 3 + 0]);
        put(-((W[0 * 3 + 0] * dy - W[1 * 3 + 0] * dx) / W[2 * 3 + 2]));
        put(W[2 * 3 + 1]);
        put(-(s2 / beta));
        /* LtoG2, cpp:3628-3699, with Cal_VE_WE_Matrx (cpp:3505-3565); an
           angle of exactly +-1e-12 (left uninitialised by the reference) is
           treated as straight */
        double VE[3], WE[9], V[3], V1[3], M1[9], M2[9];
        w_matrix(theta, phi, psi, W);
        if (!(std::fabs(angle) > 1.0e-12)) {
            VE[0] = VE[1] = 0.;
            VE[2] = L;
            for (int i = 0; i < 9; i++) WE[i] = 0.;
            for (int i = 0; i < 3; i++) WE[i * 3 + i] = 1.0;
        } else {
            double tilt = 0.;
            double cospsi = st_cos(tilt), sinpsi = st_sin(tilt);
            double costhe = st_cos(angle), sinthe = st_sin(angle);
            double tt = L * (costhe - 1.0) / angle;
            VE[0] = tt * cospsi;
            VE[1] = tt * sinpsi;
            VE[2] = L * st_sin(angle) / angle;
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
        V[0] = dx + W[0 * 3 + 0] * VE[0] + W[0 * 3 + 1] * VE[1] + W[0 * 3 + 2] * VE[2] - VE[0];
        V[1] = dy + W[1 * 3 + 0] * VE[0] + W[1 * 3 + 1] * VE[1] + W[1 * 3 + 2] * VE[2] - VE[1];
        V[2] = ds + W[2 * 3 + 0] * VE[0] + W[2 * 3 + 1] * VE[1] + W[2 * 3 + 2] * VE[2] - VE[2];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) M1[i * 3 + j] = WE[j * 3 + i];
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
        RM[4 * 6 + 0] = W[2 * 3 + 0] / beta;
        RM[0 * 6 + 2] = W[0 * 3 + 1];
        RM[2 * 6 + 2] = W[1 * 3 + 1];
        RM[4 * 6 + 2] = W[2 * 3 + 1] / beta;
        RM[1 * 6 + 1] = W[1 * 3 + 1] / W[2 * 3 + 2];
        RM[0 * 6 + 1] = RM[1 * 6 + 1] * s2;
        RM[3 * 6 + 1] = -W[0 * 3 + 1] / W[2 * 3 + 2];
        RM[2 * 6 + 1] = RM[3 * 6 + 1] * s2;
        RM[1 * 6 + 3] = -W[1 * 3 + 0] / W[2 * 3 + 2];
        RM[0 * 6 + 3] = RM[1 * 6 + 3] * s2;
        RM[3 * 6 + 3] = W[0 * 3 + 0] / W[2 * 3 + 2];
        RM[2 * 6 + 3] = RM[3 * 6 + 3] * s2;
        RM[1 * 6 + 5] = W[0 * 3 + 2] / (W[2 * 3 + 2] * beta);
        RM[0 * 6 + 5] = RM[1 * 6 + 5] * s2;
        RM[3 * 6 + 5] = W[1 * 3 + 2] / (W[2 * 3 + 2] * beta);
        RM[2 * 6 + 5] = RM[3 * 6 + 5] * s2;
        RM[4 * 6 + 5] = -s2 / (beta * gamma * beta * gamma);
        for (int i = 0; i < 36; i++) put(RM[i]);
        put((W[1 * 3 + 1] * V[0] - W[0 * 3 + 1] * V[1]) / W[2 * 3 + 2]);
        put(-W[2 * 3 + 0]);
        put((W[0 * 3 + 0] * V[1] - W[1 * 3 + 0] * V[0]) / W[2 * 3 + 2]);
        put(-W[2 * 3 + 1]);
        put(-(s2 / beta));
    }

    int drift_op(double L, const st_element *mis_e = nullptr, double mis_L = 0., double mis_angle = 0.)
    {
        st_op op = new_op(L == 0.0 && !std::signbit(L) ? K_DRIFT0 : K_DRIFT);
        if (op.kind == K_DRIFT) {
            put(L);
            put(L / gp.beta);
        }
        if (mis_e) misalign(op, *mis_e, mis_L, mis_angle);
        t.ops.push_back(op);
        return 0;
    }

    static int order_mask(const double *kn, const double *ks, int norder)
    {
        int m = 0;
        for (int i = 1; i < norder + 1; i++)
            if (kn[i] != 0. || ks[i] != 0.) m |= 1 << i;
        return m;
    }

    /* QUAD_Pass (cpp:4882) / SEXT_Pass (cpp:5004) / OCT_Pass (cpp:5077) and sPass twins */
    int multipole_op(const st_element &e, int thick_kind, int thin_kind, double kn, double ks)
    {
        double L = e.l;
        int Nint = e.nint;
        if (kn == 0. && ks == 0.) return drift_op(L, &e, L, 0.);
        if (L == 0.) {
            st_op op = new_op(thin_kind);
            put(kn);
            put(ks);
            put((double)Nint);
            misalign(op, e, L, 0.);
            t.ops.push_back(op);
            return 0;
        }
        if (Nint < 1) return -1;
        st_op op = new_op(thick_kind);
        op.nint = Nint;
        double Lint = L / Nint;
        put(Fdrift1 * Lint);
        put((Fdrift1 * Lint) / gp.beta);
        put(Fdrift2 * Lint);
        put((Fdrift2 * Lint) / gp.beta);
        put(Fkick1 * kn / Nint);
        put(Fkick1 * ks / Nint);
        put(Fkick2 * kn / Nint);
        put(Fkick2 * ks / Nint);
        put(kn);
        put(ks);
        put((double)Nint);
        put(kn / L);
        if (!spin && thick_kind == K_QUAD && gp.quad_fringe) op.flags |= ST_F_FRINGE;
        misalign(op, e, L, 0.);
        t.ops.push_back(op);
        return 0;
    }

    void kick_arrays(const double *KNL, const double *KNSL, int Nint, int norder, st_op &op)
    {
        double k1[11], s1[11], k2[11], s2[11];
        for (int i = 0; i < 11; i++) {
            k1[i] = Fkick1 * KNL[i] / Nint;
            s1[i] = Fkick1 * KNSL[i] / Nint;
            k2[i] = Fkick2 * KNL[i] / Nint;
            s2[i] = Fkick2 * KNSL[i] / Nint;
        }
        for (int i = 0; i <= norder; i++) put(k1[i]);
        for (int i = 0; i <= norder; i++) put(s1[i]);
        for (int i = 0; i <= norder; i++) put(k2[i]);
        for (int i = 0; i <= norder; i++) put(s2[i]);
        op.mask1 = order_mask(k1, s1, norder);
        op.mask2 = order_mask(k2, s2, norder);
    }

    void plain_arrays(const double *KNL, const double *KNSL, int norder)
    {
        for (int i = 0; i <= norder; i++) put(KNL[i]);
        for (int i = 0; i <= norder; i++) put(KNSL[i]);
    }

    /* MULT_Pass (cpp:5152) / MULT_sPass (cpp:7702).  mis_angle: the ANGLE the
       owning element passes to GtoL/LtoG (0 for MULT, ANGLE for the bend
       multipoles, whose L = 0 branch is a plain mult_kick_pass). */
    int mult_op(const st_element &e, double mis_angle = 0.)
    {
        double L = e.l;
        int Nint = e.nint, norder = e.norder;
        if (L == 0.) {
            st_op op = new_op(K_MULT_THIN);
            op.norder = norder;
            op.mask1 = order_mask(e.knl, e.knsl, norder);
            plain_arrays(e.knl, e.knsl, norder);
            misalign(op, e, L, mis_angle);
            t.ops.push_back(op);
            return 0;
        }
        if (Nint < 1) return -1;
        st_op op = new_op(K_MULT);
        op.nint = Nint;
        op.norder = norder;
        double Lint = L / Nint;
        put(Fdrift1 * Lint);
        put((Fdrift1 * Lint) / gp.beta);
        put(Fdrift2 * Lint);
        put((Fdrift2 * Lint) / gp.beta);
        put((double)Nint);
        put((double)order_mask(e.knl, e.knsl, norder));
        kick_arrays(e.knl, e.knsl, Nint, norder, op);
        plain_arrays(e.knl, e.knsl, norder);
        misalign(op, e, L, 0.);
        t.ops.push_back(op);
        return 0;
    }

    /* sbend_exact_pass piece (L, href, 1/href, cos, sin, L/beta) */
    void exact_piece(double L, double A, double c, double s)
    {
        double href = A / L;
        put(L);
        put(href);
        put(1.0 / href);
        put(c);
        put(s);
        put(L / gp.beta);
    }

    /* general_sbend_exact_pass piece (L, rhoc, b0, b0*rhoc, cos, sin, L/beta) */
    void general_piece(double L, double A, double K0L, double c, double s)
    {
        double rhoc = L / A, b0 = K0L / L;
        put(L);
        put(rhoc);
        put(b0);
        put(b0 * rhoc);
        put(c);
        put(s);
        put(L / gp.beta);
    }

    void yoshida_pieces(double Lint)
    {
        put(Fdrift1 * Lint);
        put((Fdrift1 * Lint) / gp.beta);
        put(Fdrift2 * Lint);
        put((Fdrift2 * Lint) / gp.beta);
        put(Fkick1 * Lint);
        put(Fkick2 * Lint);
    }

    /* SBEND_Pass (cpp:4762) / SBEND_sPass (cpp:7486) */
    int sbend_op(const st_element &e)
    {
        double L = e.l, Angle = e.angle;
        int Nint = e.nint;
        if (Angle == 0.) return drift_op(L, &e, L, Angle);
        double href = Angle / L;
        bool hexp = gp.H_expand != 0, rad = gp.radiate != 0;
        if (!spin && !hexp && !rad) {
            st_op op = new_op(K_SBEND_EXACT);
            put(st_tan(e.e1) * href);
            put(st_tan(e.e2) * href);
            put(L);
            put(href);
            put(1.0 / href);
            put(st_cos(Angle));
            put(st_sin(Angle));
            put(L / gp.beta);
            misalign(op, e, L, Angle);
            t.ops.push_back(op);
            return 0;
        }
        if (Nint < 1) return -1;
        st_op op = new_op(K_SBEND_SLICED);
        op.nint = Nint;
        double Lint = L / Nint, An = Angle / Nint;
        put(st_tan(e.e1) * href);
        put(st_tan(e.e2) * href);
        put(href);
        put(Lint);
        put(An);
        put((double)Nint);
        if (hexp) {
            yoshida_pieces(Lint);
        } else {
            double hs = An / Lint;
            put(hs);
            put(1.0 / hs);
            put(st_cos(An));
            put(st_sin(An));
            put(Lint / gp.beta);
        }
        if (rad) op.flags |= ST_F_RAD;
        misalign(op, e, L, Angle);
        t.ops.push_back(op);
        return 0;
    }

    /* GSBEND_Pass (cpp:4818); with GP.radiate, radiate2 after every slice */
    int gsbend_op(const st_element &e)
    {
        double L = e.l, Angle = e.angle, K0L = e.k0l;
        int Nint = e.nint;
        if (Angle == 0.) return drift_op(L, &e, L, Angle);
        double href = Angle / L, hreal = K0L / L;
        bool rad = gp.radiate != 0;
        if ((gp.H_expand || rad) && Nint < 1) return -1;
        if (gp.H_expand) {
            st_op op = new_op(K_GSBEND_YOSHIDA);
            op.nint = Nint;
            double Lint = L / Nint;
            put(st_tan(e.e1) * hreal);
            put(st_tan(e.e2) * hreal);
            put(Fdrift1 * Lint);
            put((Fdrift1 * Lint) / gp.beta);
            put(Fdrift2 * Lint);
            put((Fdrift2 * Lint) / gp.beta);
            put(Fkick1 * Lint);
            put(Fkick2 * Lint);
            put(href);
            put(hreal);
            if (rad) {
                put(K0L / Nint);
                put(Lint);
                op.flags |= ST_F_RAD;
            }
            misalign(op, e, L, Angle);
            t.ops.push_back(op);
            return 0;
        }
        st_op op = new_op(K_GSBEND_EXACT);
        put(st_tan(e.e1) * hreal);
        put(st_tan(e.e2) * hreal);
        if (rad) {
            double Lint = L / Nint, An = Angle / Nint;
            op.nint = Nint;
            general_piece(Lint, An, K0L / Nint, st_cos(An), st_sin(An));
            put(href);
            put(K0L / Nint);
            put(Lint);
            op.flags |= ST_F_RAD;
        } else {
            general_piece(L, Angle, K0L, st_cos(Angle), st_sin(Angle));
        }
        misalign(op, e, L, Angle);
        t.ops.push_back(op);
        return 0;
    }

    /* GMULT/SBENDMULT/SMULT_Pass (cpp:5223-5595) and their sPass twins */
    int bendmult_op(const st_element &e)
    {
        double L = e.l, Angle = e.angle;
        int Nint = e.nint, norder = e.norder;
        if (L == 0.) {
            if (!spin) return mult_op(e, Angle);
            st_op op = new_op(K_BENDMULT_THIN);
            op.norder = norder;
            op.mask1 = order_mask(e.knl, e.knsl, norder);
            put(Angle);
            plain_arrays(e.knl, e.knsl, norder);
            misalign(op, e, L, Angle);
            t.ops.push_back(op);
            return 0;
        }
        if (Nint < 1) return -1;
        st_op op = new_op(K_BENDMULT);
        op.nint = Nint;
        op.norder = norder;
        double href = Angle / L, Lint = L / Nint;
        put(st_tan(e.e1) * href);
        put(st_tan(e.e2) * href);
        put(href);
        put(Lint);
        put((double)Nint);
        put(Angle);
        put(Angle / Nint);
        put((double)order_mask(e.knl, e.knsl, norder));
        if (gp.H_expand) {
            yoshida_pieces(Lint);
            for (int i = 0; i < 6; i++) put(0.);
        } else {
            double A1 = Fdrift1 * Angle / Nint, A2 = Fdrift2 * Angle / Nint;
            exact_piece(Fdrift1 * Lint, A1, st_cos(A1), st_sin(A1));
            exact_piece(Fdrift2 * Lint, A2, st_cos(A2), st_sin(A2));
        }
        kick_arrays(e.knl, e.knsl, Nint, norder, op);
        plain_arrays(e.knl, e.knsl, norder);
        if (gp.radiate) op.flags |= ST_F_RAD;
        misalign(op, e, L, Angle);
        t.ops.push_back(op);
        return 0;
    }

    /* GSBENDMULT_Pass (cpp:5597) */
    int gsbendmult_op(const st_element &e)
    {
        double L = e.l, Angle = e.angle, K0L = e.k0l;
        int Nint = e.nint, norder = e.norder;
        if (spin) return bendmult_op(e); /* GSBENDMULT_sPass == SBENDMULT_sPass (cpp:8024) */
        if (L == 0.) return mult_op(e, Angle);
        if (Nint < 1) return -1;
        st_op op = new_op(K_GSBENDMULT);
        op.nint = Nint;
        op.norder = norder;
        double href = Angle / L, hreal = K0L / L, Lint = L / Nint;
        put(st_tan(e.e1) * hreal);
        put(st_tan(e.e2) * hreal);
        put(href);
        put(hreal);
        put(Lint);
        put((double)Nint);
        put(K0L / Nint);
        put(0.);
        if (gp.H_expand) {
            yoshida_pieces(Lint);
            for (int i = 0; i < 8; i++) put(0.);
        } else {
            double A1 = Fdrift1 * Angle / Nint, A2 = Fdrift2 * Angle / Nint;
            general_piece(Fdrift1 * Lint, A1, Fdrift1 * K0L / Nint, st_cos(A1), st_sin(A1));
            general_piece(Fdrift2 * Lint, A2, Fdrift2 * K0L / Nint, st_cos(A2), st_sin(A2));
        }
        kick_arrays(e.knl, e.knsl, Nint, norder, op);
        if (gp.radiate) op.flags |= ST_F_RAD;
        misalign(op, e, L, Angle);
        t.ops.push_back(op);
        return 0;
    }

    /* slice centroids of BB6D / BB6D_Angle (cpp:6594-6620, 6937-6962) */
    static void slice_centers(int N_slice, double sigma_l, double *z)
    {
        if (N_slice == 11) {
            static const double tab[11] = {1.7996765362153655,   1.1049618149811977,   0.75072324235797372,
                                           0.47407703558819131,  0.23041113693387344,  0.,
                                           -0.23041113693387344, -0.47407703558819131, -0.75072324235797372,
                                           -1.1049618149811977,  -1.7996765362153655};
            for (int i = 0; i < 11; i++) z[i] = tab[i] * sigma_l;
            return;
        }
        double y[64];
        for (int i = 0; i < N_slice - 1; i++) y[i] = st_ugaussian_pinv((i + 1) * 1.0 / N_slice);
        z[0] = N_slice * (st_exp(-y[0] * y[0] / 2)) / std::sqrt(2.0 * 3.14159265);
        z[N_slice - 1] = N_slice * (0 - st_exp(-y[0] * y[0] / 2)) / std::sqrt(2.0 * 3.14159265);
        for (int i = 1; i < N_slice - 1; i++)
            z[i] = N_slice * (st_exp(-y[i] * y[i] / 2) - st_exp(-y[i - 1] * y[i - 1] / 2)) / std::sqrt(2.0 * 3.14159265);
        for (int i = 0; i < N_slice; i++) z[i] = z[i] * sigma_l;
    }

    int lower(const st_element &e)
    {
        const double PI = 3.14159265358979323;
        switch (e.type) {
        case ST_DRIFT: case ST_MARKER: case ST_BPM: case ST_HBPM: case ST_VBPM:
            return drift_op(e.l);
        case ST_ROTAT: case ST_SNAKE: {
            if (!spin) return drift_op(e.l);
            /* ROTAT_sPass / SNAKE_sPass, cpp:8286-8348 */
            st_op op = new_op(K_ROTAT);
            double costheta = st_cos(-e.rot_angle), sintheta = st_sin(-e.rot_angle);
            const double *n = e.rot_n;
            double F = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            double uF[3] = {n[0] / F, n[1] / F, n[2] / F};
            double rot[9];
            rot[0] = uF[0] * uF[0] * (1 - costheta) + costheta;
            rot[1] = uF[0] * uF[1] * (1 - costheta) - uF[2] * sintheta;
            rot[2] = uF[0] * uF[2] * (1 - costheta) + uF[1] * sintheta;
            rot[3] = uF[0] * uF[1] * (1 - costheta) + uF[2] * sintheta;
            rot[4] = uF[1] * uF[1] * (1 - costheta) + costheta;
            rot[5] = uF[1] * uF[2] * (1 - costheta) - uF[0] * sintheta;
            rot[6] = uF[0] * uF[2] * (1 - costheta) - uF[1] * sintheta;
            rot[7] = uF[1] * uF[2] * (1 - costheta) + uF[0] * sintheta;
            rot[8] = uF[2] * uF[2] * (1 - costheta) + costheta;
            put(e.l / 2.);
            put((e.l / 2.) / gp.beta);
            for (int i = 0; i < 9; i++) put(rot[i]);
            t.ops.push_back(op);
            return 0;
        }
        case ST_SBEND:
            return sbend_op(e);
        case ST_GSBEND:
            if (spin) return sbend_op(e); /* GSBEND::sPass calls SBEND_sPass, cpp:8619-8623 */
            return gsbend_op(e);
        case ST_QUAD:
            return multipole_op(e, K_QUAD, K_QUAD_THIN, e.knl[1], e.knsl[1]);
        case ST_SKEWQ:
            return multipole_op(e, K_QUAD, K_QUAD_THIN, 0., e.knsl[1]);
        case ST_SEXT:
            return multipole_op(e, K_SEXT, K_SEXT_THIN, e.knl[2], e.knsl[2]);
        case ST_OCT:
            return multipole_op(e, K_OCT, K_OCT_THIN, e.knl[3], e.knsl[3]);
        case ST_MULT:
            return mult_op(e);
        case ST_GMULT: case ST_SBENDMULT: case ST_SMULT:
            return bendmult_op(e);
        case ST_GSBENDMULT:
            return gsbendmult_op(e);
        case ST_KICKER: case ST_HKICKER: case ST_VKICKER: {
            /* KICK_Pass cpp:5887, KICK_sPass cpp:8193 */
            st_op op = new_op(K_KICK);
            double hk = e.type == ST_VKICKER ? 0. : e.hkick;
            double vk = e.type == ST_HKICKER ? 0. : e.vkick;
            put(e.l / 2.0);
            put((e.l / 2.0) / gp.beta);
            put(hk);
            put(vk);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_RFCAV: {
            /* RFCAV_Pass cpp:6002 (sPass identical, cpp:10712) */
            if (e.vrf == 0.) return drift_op(e.l, &e, e.l, 0.);
            st_op op = new_op(K_RFCAV);
            put(e.l / 2.);
            put((e.l / 2.) / gp.beta);
            put(std::fabs(e.vrf * gp.Q));
            put(-2.0 * PI * e.frf);
            put(e.phase0);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_CRABRF: {
            /* CRABRF_Pass cpp:6014; CRABRF::sPass is a drift (cpp:10808-10812) */
            if (e.vrf == 0. || spin) return drift_op(e.l, &e, e.l, 0.);
            st_op op = new_op(K_CRABRF);
            put(e.l / 2.);
            put((e.l / 2.) / gp.beta);
            put(e.vrf * gp.Q);
            put(2.0 * PI * e.frf);
            put(e.phase0);
            put(2.0 * PI * e.frf / 2.99792458e8);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_CCMULT: {
            /* CCMULT_Pass cpp:6027 (B1..A4 of the element feed B2..A5, cpp:10943-10947) */
            if (e.vrf == 0.) return drift_op(e.l, &e, e.l, 0.);
            st_op op = new_op(K_CCMULT);
            double c = 2.99792458e8;
            put(e.l / 2.);
            put((e.l / 2.) / gp.beta);
            put(2.0 * PI * e.frf);
            put(e.phase0);
            put(2 * PI * e.frf / c / gp.brho);
            put(gp.brho);
            for (int i = 0; i < 8; i++) put(e.cc[i]);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_LBT: case ST_ILBT: {
            /* LBT_Pass cpp:6069 / ILBT_Pass cpp:6095; sPass without misalignment */
            st_op op = new_op(e.type == ST_LBT ? K_LBT : K_ILBT);
            double sn = st_sin(e.theta), cs = st_cos(e.theta);
            put(sn);
            put(cs);
            if (e.type == ST_LBT) {
                put(sn / cs);
                put(st_cos(e.theta));
            }
            if (!spin) misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_BEAMBEAM: {
            /* BEAMBEAM::Pass cpp:11191 -> BEAMBEAM_Pass cpp:7343 */
            st_op op = new_op(K_NOP);
            if (e.np != 0.) {
                if (e.treatment == 6) {
                    if (e.nslice < 2 || e.nslice > 64) {
                        err = "BEAMBEAM NSLICE must be in [2, 64]";
                        return -3;
                    }
                    op.kind = K_BB6D;
                    op.norder = e.nslice;
                    double Nps = 1.0 * e.np / e.nslice;
                    put(2 * Nps * 1.53469826769e-18 / gp.gamma);
                    put(2. * Nps * 1.534698e-18 / gp.gamma);
                    put(e.bbscale);
                    put((1.0 + e.alfax * e.alfax) / e.betax);
                    put((1.0 + e.alfay * e.alfay) / e.betay);
                    put(e.betax);
                    put(e.alfax);
                    put(e.betay);
                    put(e.alfay);
                    put(e.emitx);
                    put(e.emity);
                    put(2.0 * e.emitx);
                    put(2.0 * e.emity);
                    double z[64];
                    slice_centers(e.nslice, e.sigmal, z);
                    for (int i = 0; i < e.nslice; i++) put(z[i]);
                } else {
                    op.kind = K_BB4D;
                    put(2 * e.np * 1.53469826769e-18 / gp.gamma);
                    put(std::sqrt(e.emitx * e.betax));
                    put(std::sqrt(e.emity * e.betay));
                    put(e.bbscale);
                }
            }
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_CRBB: {
            /* CRBB::Pass cpp:11442 -> BB6D_Angle cpp:6917 */
            if (e.nslice < 2 || e.nslice > 64) {
                err = "CRBB NSLICE must be in [2, 64]";
                return -3;
            }
            st_op op = new_op(K_BB6DA);
            op.norder = e.nslice;
            double sn = st_sin(e.theta), cs = st_cos(e.theta), tn = sn / cs;
            double Nps = 1.0 * e.np / e.nslice;
            put(2 * Nps * 1.53469826769e-18 / gp.gamma);
            put(2. * Nps * 1.534698e-18 / gp.gamma);
            put(e.bbscale);
            put(sn);
            put(cs);
            put(tn);
            put(st_cos(e.theta));
            double gx = (1.0 + e.alfax * e.alfax) / e.betax;
            double gy = (1.0 + e.alfay * e.alfay) / e.betay;
            put(gx / cs);
            put(gy / cs);
            put(e.betax * cs);
            put(e.alfax);
            put(e.betay * cs);
            put(e.alfay);
            put(e.emitx / cs);
            put(e.emity / cs);
            double z[64];
            slice_centers(e.nslice, e.sigmal, z);
            for (int i = 0; i < e.nslice; i++) put(z[i] / cs);
            misalign(op, e, e.l, 0.);
            t.ops.push_back(op);
            return 0;
        }
        case ST_PATCH: {
            /* PATCH::Pass / sPass = PATCH_Pass (cpp:12323-12333, 6248-6256):
               TRANS_Pass by DX, DY, DS, then YROTAT_Pass by THETA (skipped
               for |THETA| < 1e-9, cpp:6195); no GtoL/LtoG */
            st_op op = new_op(K_PATCH);
            put(e.dx);
            put(e.dy);
            put(e.ds);
            put(std::fabs(e.theta) < 1.0e-9 ? 0. : 1.);
            put(st_cos(e.theta));
            put(st_sin(e.theta));
            put(st_tan(e.theta));
            t.ops.push_back(op);
            return 0;
        }
        case ST_SOLEN:
            /* SOLEN::Pass / sPass (cpp:9688-9704): GtoL, SOLEN_Pass, LtoG; with
               KS = 0 SOLEN_Pass and SOLEN_sPass are DRIFT_Pass (cpp:5811, 8125) */
            if (e.ks != 0.) {
                err = "SOLEN with KS != 0 is not supported (only KS = 0, a drift)";
                return -2;
            }
            return drift_op(e.l, &e, e.l, 0.);
        case ST_MATRIX: {
            st_op op = new_op(K_MATRIX);
            for (int i = 0; i < 6; i++) put(e.xco_in[i]);
            for (int i = 0; i < 36; i++) put(e.m66[i]);
            for (int i = 0; i < 6; i++) put(e.xco_out[i]);
            t.ops.push_back(op);
            return 0;
        }
        default:
            return -2;
        }
    }
};

} // namespace

/* ----------------------------------------------------------- context */

extern "C" int st_ctx_create(int device, st_ctx **out)
{
    if (!out) return fail(ST_ERR_ARG, "st_ctx_create: out is null");
    *out = nullptr;
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(ST_ERR_ARG, "st_ctx_create: device %d not in [0, %d)", device, n);
    HIPCHK(hipSetDevice(device));
    st_ctx *c = new st_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipHostMalloc((void **)&c->h_total, sizeof(int)) != hipSuccess) {
        st_ctx_destroy(c);
        return fail(ST_ERR_DEVICE, "st_ctx_create: stream/event/pinned allocation failed");
    }
    init_yoshida();
    *out = c;
    return ST_OK;
}

/* fold finished timed launches into the counters (stream must be idle) */
static int drain_timing(st_ctx *c)
{
    for (size_t i = 0; i < c->pending.size(); ++i) {
        float ms = 0.f;
        hipError_t e = hipEventElapsedTime(&ms, c->ev_pool[c->pending[i].first], c->ev_pool[c->pending[i].second]);
        if (e != hipSuccess) return fail(ST_ERR_DEVICE, "kernel timing: %s", hipGetErrorString(e));
        c->k_seconds += ms * 1e-3;
        c->k_launches += 1;
        c->k_work += c->pending_work[i];
    }
    c->pending.clear();
    c->pending_work.clear();
    return ST_OK;
}

static int timed_event(st_ctx *c, int *idx)
{
    size_t used = 2 * c->pending.size();
    if (used + 1 >= c->ev_pool.size()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c->ev_pool.push_back(e);
        HIPCHK(hipEventCreate(&e));
        c->ev_pool.push_back(e);
    }
    *idx = (int)used;
    return ST_OK;
}

extern "C" int st_ctx_set_timing(st_ctx *c, int enable)
{
    if (!c) return fail(ST_ERR_ARG, "st_ctx_set_timing: null context");
    c->timing = enable != 0;
    return ST_OK;
}

extern "C" int st_ctx_kernel_stats(st_ctx *c, double *sec, int64_t *launches, int64_t *work)
{
    if (!c) return fail(ST_ERR_ARG, "st_ctx_kernel_stats: null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    int rc = drain_timing(c);
    if (rc) return rc;
    if (sec) *sec = c->k_seconds;
    if (launches) *launches = c->k_launches;
    if (work) *work = c->k_work;
    c->k_seconds = 0.0;
    c->k_launches = c->k_work = 0;
    return ST_OK;
}

static void free_bufs(st_ctx *c)
{
    hipFree(c->d_x);
    hipFree(c->d_state);
    hipFree(c->d_lost);
    hipFree(c->d_idx0);
    hipFree(c->d_idx1);
    hipFree(c->d_block);
    hipFree(c->d_total);
    hipFree(c->d_tbt);
    hipFree(c->d_x04);
    hipFree(c->d_ss);
    c->d_ss = nullptr;
    c->cap_ss = 0;
    c->d_x = nullptr;
    c->d_state = c->d_lost = c->d_idx0 = c->d_idx1 = c->d_block = c->d_total = nullptr;
    c->d_tbt = c->d_x04 = nullptr;
    c->cap_x = c->cap_p = c->cap_tbt = c->cap_x04 = 0;
}

extern "C" int st_ctx_destroy(st_ctx *c)
{
    if (!c) return ST_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    free_bufs(c);
    if (c->h_total) hipHostFree(c->h_total);
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    for (int l = 0; l < SS_LANES; ++l) {
        if (c->lane[l]) {
            hipStreamSynchronize(c->lane[l]);
            hipStreamDestroy(c->lane[l]);
        }
        if (c->lane_ev[l]) hipEventDestroy(c->lane_ev[l]);
    }
    if (c->fork_ev) hipEventDestroy(c->fork_ev);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return ST_OK;
}

extern "C" int st_ctx_sync(st_ctx *c)
{
    if (!c) return fail(ST_ERR_ARG, "st_ctx_sync: null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ST_OK;
}

extern "C" int st_last_gpu_seconds(st_ctx *c, double *s)
{
    if (!c || !s) return fail(ST_ERR_ARG, "st_last_gpu_seconds: null argument");
    *s = c->last_gpu_s;
    return ST_OK;
}

static int ensure_particles(st_ctx *c, size_t np, int ncol)
{
    size_t need_x = np * (size_t)ncol;
    if (need_x > c->cap_x) {
        hipFree(c->d_x);
        c->d_x = nullptr;
        c->cap_x = 0;
        HIPCHK(hipMalloc(&c->d_x, need_x * sizeof(double)));
        c->cap_x = need_x;
    }
    if (np > c->cap_p) {
        hipFree(c->d_state);
        hipFree(c->d_lost);
        hipFree(c->d_idx0);
        hipFree(c->d_idx1);
        hipFree(c->d_block);
        hipFree(c->d_total);
        c->d_state = c->d_lost = c->d_idx0 = c->d_idx1 = c->d_block = c->d_total = nullptr;
        c->cap_p = 0;
        size_t nb = (np + 1023) / 1024 + 1;
        HIPCHK(hipMalloc(&c->d_state, np * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_lost, 2 * np * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_idx0, np * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_idx1, np * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_block, nb * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_total, sizeof(int)));
        c->cap_p = np;
    }
    return ST_OK;
}

/* ------------------------------------------------------------- table */

static int upload(dev_table &t)
{
    size_t no = t.ops.size() ? t.ops.size() : 1, np = t.par.size() ? t.par.size() : 1;
    HIPCHK(hipMalloc(&t.d_ops, no * sizeof(st_op)));
    HIPCHK(hipMalloc(&t.d_par, np * sizeof(double)));
    if (!t.ops.empty()) HIPCHK(hipMemcpy(t.d_ops, t.ops.data(), t.ops.size() * sizeof(st_op), hipMemcpyHostToDevice));
    if (!t.par.empty()) HIPCHK(hipMemcpy(t.d_par, t.par.data(), t.par.size() * sizeof(double), hipMemcpyHostToDevice));
    return ST_OK;
}

extern "C" int st_table_destroy(st_table *t)
{
    if (!t) return ST_OK;
    if (t->ctx) hipSetDevice(t->ctx->device);
    hipFree(t->orbit.d_ops);
    hipFree(t->orbit.d_par);
    hipFree(t->spin.d_ops);
    hipFree(t->spin.d_par);
    hipFree(t->d_ap);
    hipFree(t->d_kg);
    delete t;
    return ST_OK;
}

static int validate(const st_element &e, int j)
{
    switch (e.type) {
    case ST_MULT: case ST_GMULT: case ST_SBENDMULT: case ST_SMULT: case ST_GSBENDMULT:
        if (e.norder < 0 || e.norder > 10)
            return fail(ST_ERR_PARAM, "element %d: Norder %d outside [0, 10] (KNL has 11 entries)", j, e.norder);
        break;
    default:
        break;
    }
    /* a PATCH's L is its DS, which may be negative (cpp:12267); no map uses it */
    if (e.l < 0. && e.type != ST_PATCH) return fail(ST_ERR_PARAM, "element %d: negative length", j);
    return ST_OK;
}

extern "C" int st_table_build(st_ctx *ctx, const st_globals *gp, const st_element *elems, int n, st_table **out)
{
    if (!ctx || !gp || !out || (n > 0 && !elems) || n < 0) return fail(ST_ERR_ARG, "st_table_build: bad argument");
    *out = nullptr;
    HIPCHK(hipSetDevice(ctx->device));
    init_yoshida();
    st_table *t = new st_table();
    t->ctx = ctx;
    t->gp = *gp;
    t->n = n;
    st_kglobals &k = t->kg;
    std::memset(&k, 0, sizeof k);
    k.beta = gp->beta;
    k.gamma = gp->gamma;
    k.sg = std::sqrt(gp->gamma * gp->gamma - 1.);
    k.G = gp->G;
    k.energy = gp->energy;
    k.brho = gp->brho;
    k.bg2 = gp->beta * gp->gamma * gp->beta * gp->gamma;
    k.e3 = st_powi(gp->energy / 1000., 3);
    k.Cr = 8.84627632682e-5;
    k.Cf = (55. / 24. / std::sqrt(3.)) * 2.8179403e-15 * 1.054572e-34 / 9.1093822e-31;
    k.rng_seed = gp->rng_seed;
    k.H_expand = gp->H_expand;
    k.radiate = gp->radiate;
    std::vector<double> ap(2 * (size_t)(n > 0 ? n : 1));
    for (int pass = 0; pass < 2; ++pass) {
        Builder b(*gp, pass == 1, pass == 1 ? t->spin : t->orbit);
        for (int j = 0; j < n; ++j) {
            int rc = validate(elems[j], j);
            if (rc) {
                st_table_destroy(t);
                return rc;
            }
            rc = b.lower(elems[j]);
            if (rc) {
                std::string msg = b.err.empty() ? "unsupported element type or parameters" : b.err;
                st_table_destroy(t);
                return fail(rc == -2 ? ST_ERR_TYPE : ST_ERR_PARAM, "element %d (type %d): %s", j, elems[j].type,
                            msg.c_str());
            }
        }
    }
    for (int j = 0; j < n; ++j) {
        ap[2 * j] = elems[j].apx;
        ap[2 * j + 1] = elems[j].apy;
        if (gp->radiate && (elems[j].type == ST_SBEND || elems[j].type == ST_GSBEND) && elems[j].angle != 0.)
            t->draws_per_turn += elems[j].nint > 0 ? elems[j].nint : 0;
    }
    int rc = upload(t->orbit);
    if (!rc) rc = upload(t->spin);
    if (!rc) {
        hipError_t e = hipMalloc(&t->d_ap, ap.size() * sizeof(double));
        if (e == hipSuccess) e = hipMemcpy(t->d_ap, ap.data(), ap.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&t->d_kg, sizeof(st_kglobals));
        if (e == hipSuccess) e = hipMemcpy(t->d_kg, &t->kg, sizeof(st_kglobals), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(ST_ERR_DEVICE, "table upload: %s", hipGetErrorString(e));
    }
    if (rc) {
        st_table_destroy(t);
        return rc;
    }
    *out = t;
    return ST_OK;
}

extern "C" int st_table_size(const st_table *t, int *n_elems, int64_t *bytes)
{
    if (!t) return fail(ST_ERR_ARG, "st_table_size: null table");
    if (n_elems) *n_elems = t->n;
    if (bytes)
        *bytes = (int64_t)((t->orbit.ops.size() + t->spin.ops.size()) * sizeof(st_op) +
                           (t->orbit.par.size() + t->spin.par.size()) * sizeof(double) + 2 * t->n * sizeof(double));
    return ST_OK;
}

/* ---------------------------------------------------------- tracking */

static st_track_args base_args(st_table *t, bool spin)
{
    st_track_args a;
    std::memset(&a, 0, sizeof a);
    dev_table &d = spin ? t->spin : t->orbit;
    a.ops = d.d_ops;
    a.par = d.d_par;
    a.ap = t->d_ap;
    a.nelem = t->n;
    a.tbt_bpm = -1;
    a.g = t->kg;
    a.gdev = t->d_kg;
    a.draws_per_turn = spin ? 0 : t->draws_per_turn;
    return a;
}

/* turns per launch: long enough to amortise launch and compaction, short
   enough that compaction removes lost particles early */
static int turn_block(int n_elems, int np)
{
    (void)np;
    long work = (long)n_elems > 0 ? (long)n_elems : 1;
    long tb = 200000 / work;
    if (tb < 1) tb = 1;
    if (tb > 256) tb = 256;
    return (int)tb;
}

/* device-side core: d_x/d_state/d_lost already on the device */
static int track_device_core(st_table *t, double *d_x, int np, int turn0, int nturn, int *d_state, int *d_lost,
                             bool spin, bool radial, double *d_tbt, int tbt_bpm, int tbt_nturn, bool bare = false)
{
    st_ctx *c = t->ctx;
    if (ensure_particles(c, (size_t)np, 1)) return ST_ERR_NOMEM;
    st_track_args a = base_args(t, spin);
    if (bare) a.ap = nullptr; /* Element::Pass turn after turn: no loss test */
    a.x = d_x;
    a.state = d_state;
    a.lost = d_lost;
    a.xtbt = d_tbt;
    a.tbt_bpm = tbt_bpm;
    a.tbt_nturn = tbt_nturn;
    a.tbt_turn0 = 0;
    if (radial) {
        if ((size_t)t->n > c->cap_x04) {
            hipFree(c->d_x04);
            c->d_x04 = nullptr;
            c->cap_x04 = 0;
            HIPCHK(hipMalloc(&c->d_x04, (size_t)(t->n > 0 ? t->n : 1) * sizeof(double)));
            c->cap_x04 = (size_t)t->n;
        }
        st_track_args ar = base_args(t, false);
        HIPCHK(st_launch_radial_x04(ar, c->d_x04, t->gp.H_expand != 0, c->stream));
        a.x04 = c->d_x04;
    }
    const bool hexp = t->gp.H_expand != 0, rad = t->gp.radiate != 0;
    int tb = turn_block(t->n, np);
    int nact = np;
    const int *idx = nullptr;
    int *bufs[2] = {c->d_idx0, c->d_idx1};
    int which = 0;
    bool first = true;
    for (int t0 = turn0; t0 < turn0 + nturn && nact > 0;) {
        int nt = std::min(tb, turn0 + nturn - t0);
        a.turn0 = t0;
        a.nturn = nt;
        a.precheck = (first && turn0 == 0 && !bare) ? 1 : 0;
        a.idx = idx;
        a.nact = nact;
        int ev = -1;
        if (c->timing) {
            int rc2 = timed_event(c, &ev);
            if (rc2) return rc2;
            HIPCHK(hipEventRecord(c->ev_pool[ev], c->stream));
        }
        HIPCHK(st_launch_track(a, spin, hexp, rad, c->stream));
        if (c->timing) {
            HIPCHK(hipEventRecord(c->ev_pool[ev + 1], c->stream));
            c->pending.push_back({ev, ev + 1});
            c->pending_work.push_back((int64_t)nact * nt);
        }
        first = false;
        t0 += nt;
        if (t0 < turn0 + nturn) {
            int *out = bufs[which];
            HIPCHK(st_launch_compact(idx, nact, d_state, c->d_block, out, c->d_total, c->stream));
            HIPCHK(hipMemcpyAsync(c->h_total, c->d_total, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (c->timing) {
                int rc2 = drain_timing(c);
                if (rc2) return rc2;
            }
            nact = *c->h_total;
            idx = out;
            which ^= 1;
        }
    }
    if (nturn == 0 && turn0 == 0 && !bare) {
        /* the reference's pre-check runs even for nturn == 0 */
        a.turn0 = 0;
        a.nturn = 0;
        a.precheck = 1;
        a.idx = nullptr;
        a.nact = np;
        HIPCHK(st_launch_track(a, spin, hexp, rad, c->stream));
    }
    return ST_OK;
}

static int track_host(st_table *t, double *x, int np, int nturn, int *stable, int *lost_turn, int *lost_post,
                      bool spin, bool radial, double *x_tbt, int tbt_bpm, bool bare = false)
{
    if (!t || (np > 0 && (!x || !stable || !lost_turn || !lost_post)) || np < 0 || nturn < 0)
        return fail(ST_ERR_ARG, "track: bad argument");
    if (np == 0) return ST_OK;
    st_ctx *c = t->ctx;
    HIPCHK(hipSetDevice(c->device));
    const int ncol = spin ? 9 : 6;
    int rc = ensure_particles(c, (size_t)np, ncol);
    if (rc) return rc;
    std::vector<int> lost(2 * (size_t)np);
    for (int p = 0; p < np; ++p) {
        lost[2 * p] = lost_turn[p];
        lost[2 * p + 1] = lost_post[p];
    }
    double *d_tbt = nullptr;
    size_t tbt_n = 0;
    if (x_tbt) {
        tbt_n = (size_t)np * (size_t)nturn * ncol;
        if (tbt_n > c->cap_tbt) {
            hipFree(c->d_tbt);
            c->d_tbt = nullptr;
            c->cap_tbt = 0;
            HIPCHK(hipMalloc(&c->d_tbt, (tbt_n ? tbt_n : 1) * sizeof(double)));
            c->cap_tbt = tbt_n;
        }
        d_tbt = c->d_tbt;
        if (tbt_n) HIPCHK(hipMemcpyAsync(d_tbt, x_tbt, tbt_n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipMemcpyAsync(c->d_x, x, (size_t)np * ncol * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_state, stable, (size_t)np * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_lost, lost.data(), 2 * (size_t)np * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    rc = track_device_core(t, c->d_x, np, 0, nturn, c->d_state, c->d_lost, spin, radial, d_tbt, tbt_bpm, nturn,
                           bare);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipMemcpyAsync(x, c->d_x, (size_t)np * ncol * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(stable, c->d_state, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(lost.data(), c->d_lost, 2 * (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (x_tbt && tbt_n)
        HIPCHK(hipMemcpyAsync(x_tbt, d_tbt, tbt_n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_gpu_s = ms * 1e-3;
    for (int p = 0; p < np; ++p) {
        lost_turn[p] = lost[2 * p];
        lost_post[p] = lost[2 * p + 1];
    }
    return ST_OK;
}

extern "C" int st_track(st_table *t, double *x, int np, int nturn, int *stable, int *lost_turn, int *lost_post)
{
    return track_host(t, x, np, nturn, stable, lost_turn, lost_post, false, false, nullptr, -1);
}

extern "C" int st_track_spin(st_table *t, double *x, int np, int nturn, int *stable, int *lost_turn, int *lost_post)
{
    return track_host(t, x, np, nturn, stable, lost_turn, lost_post, true, false, nullptr, -1);
}

extern "C" int st_track_tbt(st_table *t, double *x, int np, int nturn, double *x_tbt, int bpm_index, int spin,
                            int *stable, int *lost_turn, int *lost_post)
{
    if (!x_tbt && np > 0) return fail(ST_ERR_ARG, "st_track_tbt: x_tbt is null");
    if (t && (bpm_index >= t->n)) return fail(ST_ERR_ARG, "st_track_tbt: bpm_index %d >= %d", bpm_index, t->n);
    return track_host(t, x, np, nturn, stable, lost_turn, lost_post, spin != 0, false, x_tbt,
                      bpm_index < 0 ? -1 : bpm_index);
}

extern "C" int st_track_radial(st_table *t, double *x, int np, int nturn, int *stable, int *lost_turn,
                               int *lost_post)
{
    return track_host(t, x, np, nturn, stable, lost_turn, lost_post, false, true, nullptr, -1);
}

extern "C" int st_track_device(st_table *t, double *d_x, int np, int turn0, int nturn, int *d_state, int *d_lost,
                               int flags)
{
    if (!t || np < 0 || nturn < 0 || turn0 < 0 || (np > 0 && (!d_x || !d_state || !d_lost)))
        return fail(ST_ERR_ARG, "st_track_device: bad argument");
    if (np == 0) return ST_OK;
    HIPCHK(hipSetDevice(t->ctx->device));
    return track_device_core(t, d_x, np, turn0, nturn, d_state, d_lost, (flags & ST_TRACK_SPIN) != 0,
                             (flags & ST_TRACK_RADIAL) != 0, nullptr, -1, 0, (flags & ST_TRACK_BARE) != 0);
}

extern "C" int st_element_pass(st_table *t, int elem_index, double *x, int np, int spin)
{
    if (!t || elem_index < 0 || elem_index >= t->n || np < 0 || (np > 0 && !x))
        return fail(ST_ERR_ARG, "st_element_pass: bad argument");
    if (np == 0) return ST_OK;
    st_ctx *c = t->ctx;
    HIPCHK(hipSetDevice(c->device));
    const int ncol = spin ? 9 : 6;
    int rc = ensure_particles(c, (size_t)np, ncol);
    if (rc) return rc;
    std::vector<int> ones((size_t)np, 1);
    HIPCHK(hipMemcpyAsync(c->d_x, x, (size_t)np * ncol * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_state, ones.data(), (size_t)np * sizeof(int), hipMemcpyHostToDevice, c->stream));
    st_track_args a = base_args(t, spin != 0);
    a.ops += elem_index;
    a.nelem = 1;
    a.ap = nullptr;
    a.x = c->d_x;
    a.state = c->d_state;
    a.lost = c->d_lost;
    a.nact = np;
    a.turn0 = 0;
    a.nturn = 1;
    a.draws_per_turn = 0;
    HIPCHK(st_launch_track(a, spin != 0, t->gp.H_expand != 0, t->gp.radiate != 0, c->stream));
    HIPCHK(hipMemcpyAsync(x, c->d_x, (size_t)np * ncol * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ST_OK;
}

/* ------------------------------------------------------ tune finder */

static int tune_check(const st_ctx *c, int n, int nsig, double lo, double hi, bool px, int *max_cand)
{
    if (!c || n <= 0 || nsig < 0 || !(hi > lo) || !std::isfinite(lo) || !std::isfinite(hi))
        return fail(ST_ERR_ARG, "st_fine_tune: bad argument");
    double lvl0 = std::ceil((hi - lo) / 0.001) + 2.0;
    if (lvl0 > 4096.) return fail(ST_ERR_ARG, "st_fine_tune: tune range too wide (more than 4096 steps of 1e-3)");
    int mc = std::max((int)lvl0, 24);
    size_t lds = sizeof(double) * ((size_t)n * (px ? 2 : 1) + (size_t)mc);
    if (lds > 60 * 1024)
        return fail(ST_ERR_ARG, "st_fine_tune: %d samples%s and %d candidates exceed the 60 KiB LDS budget", n,
                    px ? " (x and px)" : "", mc);
    *max_cand = mc;
    return ST_OK;
}

extern "C" int st_fine_tune_window_device(st_ctx *c, const double *d_x, const double *d_px, int64_t stride_sig,
                                          int64_t stride, int n, int nsig, double tune_lower, double tune_upper,
                                          const double *d_window, double *d_tune);

extern "C" int st_fine_tune_device(st_ctx *c, const double *d_x, const double *d_px, int64_t stride_sig,
                                   int64_t stride, int n, int nsig, double tune_lower, double tune_upper,
                                   double *d_tune)
{
    return st_fine_tune_window_device(c, d_x, d_px, stride_sig, stride, n, nsig, tune_lower, tune_upper, nullptr,
                                      d_tune);
}

extern "C" int st_fine_tune_window_device(st_ctx *c, const double *d_x, const double *d_px, int64_t stride_sig,
                                          int64_t stride, int n, int nsig, double tune_lower, double tune_upper,
                                          const double *d_window, double *d_tune)
{
    int mc = 0;
    int rc = tune_check(c, n, nsig, tune_lower, tune_upper, d_px != nullptr, &mc);
    if (rc) return rc;
    if (nsig == 0) return ST_OK;
    if (!d_x || !d_tune || stride < 1 || stride_sig < 0) return fail(ST_ERR_ARG, "st_fine_tune_device: bad argument");
    HIPCHK(hipSetDevice(c->device));
    st_tune_args a;
    a.x = d_x;
    a.px = d_px;
    a.stride_sig = stride_sig;
    a.stride = stride;
    a.n = n;
    a.nsig = nsig;
    a.max_cand = mc;
    a.lower = tune_lower;
    a.upper = tune_upper;
    a.tune = d_tune;
    a.window = d_window;
    HIPCHK(st_launch_fine_tune(a, c->stream));
    return ST_OK;
}

extern "C" int st_fine_tune(st_ctx *c, const double *x, const double *px, int n, int nsig, double tune_lower,
                            double tune_upper, double *tune)
{
    int mc = 0;
    int rc = tune_check(c, n, nsig, tune_lower, tune_upper, px != nullptr, &mc);
    if (rc) return rc;
    if (nsig == 0) return ST_OK;
    if (!x || !tune) return fail(ST_ERR_ARG, "st_fine_tune: bad argument");
    HIPCHK(hipSetDevice(c->device));
    size_t cnt = (size_t)n * (size_t)nsig;
    double *d = nullptr, *dt = nullptr;
    HIPCHK(hipMalloc(&d, cnt * (px ? 2 : 1) * sizeof(double)));
    hipError_t e = hipMalloc(&dt, (size_t)nsig * sizeof(double));
    if (e == hipSuccess) e = hipMemcpyAsync(d, x, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && px)
        e = hipMemcpyAsync(d + cnt, px, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        rc = st_fine_tune_device(c, d, px ? d + cnt : nullptr, n, 1, n, nsig, tune_lower, tune_upper, dt);
        if (rc == ST_OK) e = hipMemcpyAsync(tune, dt, (size_t)nsig * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (rc == ST_OK && e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    hipFree(d);
    if (dt) hipFree(dt);
    if (rc) return rc;
    if (e != hipSuccess) return fail(ST_ERR_DEVICE, "st_fine_tune: %s", hipGetErrorString(e));
    return ST_OK;
}

/* --------------------------------------------------- closed orbit, TBT scans */

/* mat_inv, simtrack.cpp:1440-1496: Gauss-Jordan with full pivoting, in place;
   0 when the matrix is singular */
static int mat_inv_ref(double *a, int n)
{
    std::vector<int> is(n), js(n);
    int l, u, v;
    double d, p;
    for (int k = 0; k <= n - 1; k++) {
        d = 0.0;
        for (int i = k; i <= n - 1; i++)
            for (int j = k; j <= n - 1; j++) {
                l = i * n + j;
                p = std::fabs(a[l]);
                if (p > d) {
                    d = p;
                    is[k] = i;
                    js[k] = j;
                }
            }
        if (d + 1.0 == 1.0) return 0;
        if (is[k] != k)
            for (int j = 0; j <= n - 1; j++) {
                u = k * n + j;
                v = is[k] * n + j;
                p = a[u];
                a[u] = a[v];
                a[v] = p;
            }
        if (js[k] != k)
            for (int i = 0; i <= n - 1; i++) {
                u = i * n + k;
                v = i * n + js[k];
                p = a[u];
                a[u] = a[v];
                a[v] = p;
            }
        l = k * n + k;
        a[l] = 1.0 / a[l];
        for (int j = 0; j <= n - 1; j++)
            if (j != k) {
                u = k * n + j;
                a[u] = a[u] * a[l];
            }
        for (int i = 0; i <= n - 1; i++)
            if (i != k)
                for (int j = 0; j <= n - 1; j++)
                    if (j != k) {
                        u = i * n + j;
                        a[u] = a[u] - a[i * n + k] * a[k * n + j];
                    }
        for (int i = 0; i <= n - 1; i++)
            if (i != k) {
                u = i * n + k;
                a[u] = -a[u] * a[l];
            }
    }
    for (int k = n - 1; k >= 0; k--) {
        if (js[k] != k)
            for (int j = 0; j <= n - 1; j++) {
                u = k * n + j;
                v = js[k] * n + j;
                p = a[u];
                a[u] = a[v];
                a[v] = p;
            }
        if (is[k] != k)
            for (int i = 0; i <= n - 1; i++) {
                u = i * n + k;
                v = i * n + is[k];
                p = a[u];
                a[u] = a[v];
                a[v] = p;
            }
    }
    return 1;
}

/* Cal_Orbit_Num, simtrack.cpp:14947-15030: Newton iteration on the 4-D
   one-turn map with RF, crab and CCMULT voltages off, finite differences
   of 1e-9, bare Element::Pass (no loss test).  The five particles of an
   iteration (x0 and four displaced copies) are one batch on the GPU. */
extern "C" int st_closed_orbit(st_ctx *c, const st_globals *gp, const st_element *rec, int n, double deltap,
                               double x_co[6], double x_end[6], int *iterations)
{
    if (!c || !gp || (n > 0 && !rec) || n < 0 || !x_co) return fail(ST_ERR_ARG, "st_closed_orbit: bad argument");
    std::vector<st_element> r(rec, rec + n);
    for (auto &e : r)
        if (e.type == ST_RFCAV || e.type == ST_CRABRF || e.type == ST_CCMULT) e.vrf = 0.;
    st_table *t = nullptr;
    int rc = st_table_build(c, gp, r.data(), n, &t);
    if (rc) return rc;
    std::unique_ptr<st_table, int (*)(st_table *)> guard(t, st_table_destroy);
    const double d = 1.0e-09, codeps = 1e-12;
    const int Max_iter = 20;
    double x0[6] = {0., 0., 0., 0., 0., 0.}, mat[4][4], dx[4], chi;
    double xb[5 * 6];
    int st[5], lt[5], lp[5];
    int iter = 0;
    do {
        iter++;
        x0[4] = 0;
        x0[5] = deltap;
        for (int q = 0; q < 5; ++q) {
            for (int j = 0; j < 6; j++) xb[q * 6 + j] = x0[j];
            if (q > 0) xb[q * 6 + q - 1] = xb[q * 6 + q - 1] + d;
            st[q] = 1;
            lt[q] = lp[q] = 0;
        }
        rc = track_host(t, xb, 5, 1, st, lt, lp, false, false, nullptr, -1, true);
        if (rc) return rc;
        const double *x01 = xb;
        for (int k = 0; k < 4; k++)
            for (int j = 0; j < 4; j++) mat[j][k] = (xb[(k + 1) * 6 + j] - x01[j]) / d;
        /* the reference passes x0 once more; the map is deterministic, so that is x01 */
        for (int i = 0; i < 4; i++) dx[i] = x0[i] - x01[i];
        chi = 0;
        for (int i = 0; i < 4; i++) chi += dx[i] * dx[i];
        chi = std::sqrt(chi / 4.);
        if (chi > codeps) {
            for (int i = 0; i < 4; i++) mat[i][i] = mat[i][i] - 1.0000001;
            if (mat_inv_ref(&mat[0][0], 4) == 0) return fail(ST_ERR_PARAM, "closed orbit: singular matrix");
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) x0[i] += mat[i][j] * dx[j];
        }
    } while (chi > codeps && iter < Max_iter);
    if (iterations) *iterations = iter;
    /* the reference's test, cpp:15008 */
    if (iter == Max_iter - 1) return fail(ST_ERR_PARAM, "closed orbit: failed to find COD");
    x0[4] = 0.000;
    x0[5] = deltap;
    for (int j = 0; j < 6; j++) x_co[j] = x0[j];
    if (x_end) {
        /* Cell[Ncell-1]->X: x0 passed once more through every element (cpp:15015-15018) */
        for (int j = 0; j < 6; j++) xb[j] = x0[j];
        st[0] = 1;
        lt[0] = lp[0] = 0;
        rc = track_host(t, xb, 1, 1, st, lt, lp, false, false, nullptr, -1, true);
        if (rc) return rc;
        for (int j = 0; j < 6; j++) x_end[j] = xb[j];
    }
    return ST_OK;
}

/* The 6x6 one-turn matrix about x_co: symmetric differences of step h of
   bare passes (no loss test) with the RFCAV voltages off (crab cavities and
   CCMULT stay on), the map Cal_OneTurnMap (simtrack.cpp:15122-15137)
   expands.  The twelve displaced particles are one batch; m is row-major,
   m[6*i + k] = d x_i / d x_k. */
extern "C" int st_one_turn_matrix(st_ctx *c, const st_globals *gp, const st_element *rec, int n,
                                  const double x_co[6], double h, double m[36])
{
    if (!c || !gp || (n > 0 && !rec) || n < 0 || !x_co || !m || !(h > 0.))
        return fail(ST_ERR_ARG, "st_one_turn_matrix: bad argument");
    std::vector<st_element> r(rec, rec + n);
    for (auto &e : r)
        if (e.type == ST_RFCAV) e.vrf = 0.;
    st_table *t = nullptr;
    int rc = st_table_build(c, gp, r.data(), n, &t);
    if (rc) return rc;
    std::unique_ptr<st_table, int (*)(st_table *)> guard(t, st_table_destroy);
    double xb[12 * 6];
    int st[12], lt[12], lp[12];
    for (int q = 0; q < 12; ++q) {
        for (int j = 0; j < 6; j++) xb[q * 6 + j] = x_co[j];
        xb[q * 6 + q / 2] += (q & 1) ? -h : h;
        st[q] = 1;
        lt[q] = lp[q] = 0;
    }
    rc = track_host(t, xb, 12, 1, st, lt, lp, false, false, nullptr, -1, true);
    if (rc) return rc;
    for (int k = 0; k < 6; k++)
        for (int i = 0; i < 6; i++) m[6 * i + k] = (xb[2 * k * 6 + i] - xb[(2 * k + 1) * 6 + i]) / (2 * h);
    return ST_OK;
}

/*
 * The closed orbit after every element and every element's linear map, as
 * Cal_Orbit_Num's last pass (simtrack.cpp:15020-15029) and Cal_ElementMap
 * (cpp:15205-15253) leave them for Trace_A: X[k] is x_co passed through
 * elements 0..k with every RF, crab and CCMULT voltage off; T[k] is element
 * k's 6x6 map about X[k-1] (X[n-1] for k = 0) with z = 0 and x[5] = x5,
 * through the line with the RFCAV voltages off, as symmetric differences of
 * step h of GPU element passes (the reference differentiates with TPS).
 */
extern "C" int st_element_maps(st_ctx *c, const st_globals *gp, const st_element *rec, int n, const double x_co[6],
                               double x5, double h, double *T, double *X)
{
    if (!c || !gp || n <= 0 || !rec || !x_co || !T || !X || !(h > 0.))
        return fail(ST_ERR_ARG, "st_element_maps: bad argument");
    std::vector<st_element> ro(rec, rec + n), rm(rec, rec + n);
    for (auto &e : ro)
        if (e.type == ST_RFCAV || e.type == ST_CRABRF || e.type == ST_CCMULT) e.vrf = 0.;
    for (auto &e : rm)
        if (e.type == ST_RFCAV) e.vrf = 0.;
    st_table *to = nullptr, *tm = nullptr;
    int rc = st_table_build(c, gp, ro.data(), n, &to);
    if (rc) return rc;
    std::unique_ptr<st_table, int (*)(st_table *)> go(to, st_table_destroy);
    rc = st_table_build(c, gp, rm.data(), n, &tm);
    if (rc) return rc;
    std::unique_ptr<st_table, int (*)(st_table *)> gm(tm, st_table_destroy);
    double xo[6];
    for (int j = 0; j < 6; j++) xo[j] = x_co[j];
    for (int k = 0; k < n; ++k) {
        rc = st_element_pass(to, k, xo, 1, 0);
        if (rc) return rc;
        for (int j = 0; j < 6; j++) X[6 * k + j] = xo[j];
    }
    double xb[12 * 6];
    for (int k = 0; k < n; ++k) {
        const double *x0 = X + 6 * (k > 0 ? k - 1 : n - 1);
        for (int q = 0; q < 12; ++q) {
            for (int j = 0; j < 6; j++) xb[q * 6 + j] = x0[j];
            xb[q * 6 + 4] = 0.;
            xb[q * 6 + 5] = x5;
            xb[q * 6 + q / 2] += (q & 1) ? -h : h;
        }
        rc = st_element_pass(tm, k, xb, 12, 0);
        if (rc) return rc;
        double *m = T + 36 * (size_t)k;
        for (int a = 0; a < 6; a++)
            for (int i = 0; i < 6; i++) m[6 * i + a] = (xb[2 * a * 6 + i] - xb[(2 * a + 1) * 6 + i]) / (2 * h);
    }
    return ST_OK;
}

/* TBT of np particles on the device; the buffer stays in the context */
static int tbt_device(st_table *t, const std::vector<double> &x, int nturn, std::vector<int> &stable,
                      double **d_tbt_out)
{
    st_ctx *c = t->ctx;
    const int np = (int)(x.size() / 6);
    int rc = ensure_particles(c, (size_t)np, 6);
    if (rc) return rc;
    size_t tbt_n = (size_t)np * (size_t)nturn * 6;
    if (tbt_n > c->cap_tbt) {
        hipFree(c->d_tbt);
        c->d_tbt = nullptr;
        c->cap_tbt = 0;
        HIPCHK(hipMalloc(&c->d_tbt, (tbt_n ? tbt_n : 1) * sizeof(double)));
        c->cap_tbt = tbt_n;
    }
    std::vector<int> zeros(2 * (size_t)np, 0);
    stable.assign(np, 1);
    HIPCHK(hipMemcpyAsync(c->d_x, x.data(), x.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_state, stable.data(), (size_t)np * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_lost, zeros.data(), zeros.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    rc = track_device_core(t, c->d_x, np, 0, nturn, c->d_state, c->d_lost, false, false, c->d_tbt, -1, nturn);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(stable.data(), c->d_state, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *d_tbt_out = c->d_tbt;
    return ST_OK;
}

/* Track_tbt_FMA, simtrack.cpp:21179-21241: the (i, j) grid of amplitude
   steps sigma_step within sigma_max (the reference: 0.1 and 6, nturn 2048),
   offset by the closed orbit Cell[Ncell-1]->X (x_end[0..3]); every stable
   particle's tunes from its x and y turn-by-turn data in turns [0, w) and
   [w-1, 2w-1), w = nturn/2, by FineTuneFinder on the GPU.  rows: 8 doubles
   each, i*step, j*step, tunex1, tunex2, tuney1, tuney2, tunex1-tunex2,
   tuney1-tuney2, for the stable particles in the reference's order. */
extern "C" int st_fma_scan(st_table *t, double deltap0, double sigmax0, double sigmay0, const double x_end[4],
                           double tune_low, double tune_high, int nturn, double sigma_step, double sigma_max,
                           double *rows, int max_rows, int *nrows)
{
    if (!t || !x_end || !nrows || nturn < 4 || (nturn & 1) || !(sigma_step > 0.) || !(sigma_max >= 0.))
        return fail(ST_ERR_ARG, "st_fma_scan: bad argument");
    st_ctx *c = t->ctx;
    HIPCHK(hipSetDevice(c->device));
    const int w = nturn / 2;
    const int nsigma = int((sigma_max - 0) * 1.0 / sigma_step) + 1;
    std::vector<double> x;
    std::vector<std::pair<int, int>> ij;
    for (int i = 0; i < nsigma; i++)
        for (int j = 0; j < nsigma; j++)
            if (std::sqrt(1.0 * i * i + 1.0 * j * j) <= (sigma_max / sigma_step)) {
                double p[6] = {i * sigma_step * sigmax0 + 0., 0., j * sigma_step * sigmay0, 0., 0., deltap0};
                p[0] = p[0] + x_end[0];
                p[1] = p[1] + x_end[1];
                p[2] = p[2] + x_end[2];
                p[3] = p[3] + x_end[3];
                if (p[0] < 1.0e-09) p[0] = p[0] + 1.0e-09;
                if (p[2] < 1.0e-09) p[2] = p[2] + 1.0e-09;
                x.insert(x.end(), p, p + 6);
                ij.push_back({i, j});
            }
    const int np = (int)ij.size();
    *nrows = 0;
    if (np == 0) return ST_OK;
    std::vector<int> stable;
    double *d_tbt = nullptr;
    int rc = tbt_device(t, x, nturn, stable, &d_tbt);
    if (rc) return rc;
    double *d_q = nullptr;
    HIPCHK(hipMalloc(&d_q, 4 * (size_t)np * sizeof(double)));
    std::unique_ptr<double, hipError_t (*)(void *)> gq(d_q, hipFree);
    const int64_t sig = (int64_t)nturn * 6;
    const int64_t offs[4] = {0, (int64_t)(w - 1) * 6, 2, (int64_t)(w - 1) * 6 + 2}; /* x1, x2, y1, y2 */
    for (int k = 0; k < 4; ++k) {
        rc = st_fine_tune_device(c, d_tbt + offs[k], nullptr, sig, 6, w, np, tune_low, tune_high, d_q + (size_t)k * np);
        if (rc) return rc;
    }
    std::vector<double> q(4 * (size_t)np);
    HIPCHK(hipMemcpyAsync(q.data(), d_q, q.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int m = 0;
    for (int p = 0; p < np; ++p) {
        if (stable[p] != 1) continue;
        if (rows && m < max_rows) {
            double tx1 = q[p], tx2 = q[np + p], ty1 = q[2 * (size_t)np + p], ty2 = q[3 * (size_t)np + p];
            double *o = rows + (size_t)m * 8;
            o[0] = ij[p].first * sigma_step;
            o[1] = ij[p].second * sigma_step;
            o[2] = tx1;
            o[3] = tx2;
            o[4] = ty1;
            o[5] = ty2;
            o[6] = tx1 - tx2;
            o[7] = ty1 - ty2;
        }
        ++m;
    }
    *nrows = m;
    if (rows && m > max_rows) return fail(ST_ERR_ARG, "st_fma_scan: %d rows do not fit in %d", m, max_rows);
    return ST_OK;
}

/* Track_tbt_Lyapunov, simtrack.cpp:21243-21303: twin particles 0.707e-6
   apart in x and y on the (i, j) grid (the reference: step 0.05, 6 sigma,
   2048 turns); for pairs that both survive, the mean over the last 25 turns
   of log(d_k / 1e-6) / (k + 1), d_k the 4-D distance at the start of turn k.
   rows: i*step, j*step, the mean. */
extern "C" int st_lyapunov_scan(st_table *t, double deltap0, double sigmax0, double sigmay0, int nturn,
                                double sigma_step, double sigma_max, double *rows, int max_rows, int *nrows)
{
    if (!t || !nrows || nturn < 25 || !(sigma_step > 0.) || !(sigma_max >= 0.))
        return fail(ST_ERR_ARG, "st_lyapunov_scan: bad argument");
    st_ctx *c = t->ctx;
    HIPCHK(hipSetDevice(c->device));
    const int nsigma = int((sigma_max - 0) * 1.0 / sigma_step);
    const double distance0 = 1.0e-06;
    std::vector<double> x;
    std::vector<std::pair<int, int>> ij;
    for (int i = 0; i < nsigma; i++)
        for (int j = 0; j < nsigma; j++)
            if (std::sqrt(1.0 * i * i + 1.0 * j * j) <= (sigma_max / sigma_step)) {
                double x1[6] = {i * sigma_step * sigmax0, 0., j * sigma_step * sigmay0, 0., 0., deltap0};
                if (x1[0] == 0.) x1[0] = x1[0] + 1.0e-09;
                if (x1[2] == 0.) x1[2] = x1[2] + 1.0e-09;
                double x2[6] = {x1[0] + 0.707e-06, 0., x1[2] + 0.707e-06, 0., 0., deltap0};
                x.insert(x.end(), x1, x1 + 6);
                x.insert(x.end(), x2, x2 + 6);
                ij.push_back({i, j});
            }
    const int npair = (int)ij.size(), np = 2 * npair;
    *nrows = 0;
    if (npair == 0) return ST_OK;
    std::vector<int> stable;
    double *d_tbt = nullptr;
    int rc = tbt_device(t, x, nturn, stable, &d_tbt);
    if (rc) return rc;
    /* only the last 25 turns enter the result */
    const int K = 25;
    std::vector<double> last((size_t)np * K * 6);
    HIPCHK(hipMemcpy2DAsync(last.data(), K * 6 * sizeof(double), d_tbt + (size_t)(nturn - K) * 6,
                            (size_t)nturn * 6 * sizeof(double), K * 6 * sizeof(double), np, hipMemcpyDeviceToHost,
                            c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int m = 0;
    for (int q = 0; q < npair; ++q) {
        if (stable[2 * q] * stable[2 * q + 1] != 1) continue;
        const double *a = &last[(size_t)(2 * q) * K * 6], *b = &last[(size_t)(2 * q + 1) * K * 6];
        double sum = 0.0;
        for (int k = 0; k < K; k++) {
            const int kk = nturn - 1 - k, r = K - 1 - k; /* turn kk is row r of the copied block */
            const double *u = a + r * 6, *v = b + r * 6;
            double distance = std::sqrt((u[0] - v[0]) * (u[0] - v[0]) + (u[1] - v[1]) * (u[1] - v[1]) +
                                        (u[2] - v[2]) * (u[2] - v[2]) + (u[3] - v[3]) * (u[3] - v[3]));
            double lay = st_log(distance / distance0) / (kk + 1);
            sum = sum + lay;
        }
        if (rows && m < max_rows) {
            double *o = rows + (size_t)m * 3;
            o[0] = ij[q].first * sigma_step;
            o[1] = ij[q].second * sigma_step;
            o[2] = sum / 25;
        }
        ++m;
    }
    *nrows = m;
    if (rows && m > max_rows) return fail(ST_ERR_ARG, "st_lyapunov_scan: %d rows do not fit in %d", m, max_rows);
    return ST_OK;
}

/* rnd, simtrack.cpp:283-291: the reference's LCG; r is its state (global `seed`) */
static double rnd_ref(double &r)
{
    int m;
    double s, u, v, p;
    s = 65536.0;
    u = 2053.0;
    v = 13849.0;
    m = int(r / s);
    r = r - m * s;
    r = u * r + v;
    m = (int)(r / s);
    r = r - m * s;
    p = r / s;
    return (p);
}

/* Cal_Twiss_Emit (cpp:771-815): rms emittance of n (x, px) samples */
static double twiss_emit_ref(const double *x, const double *px, int stride, double n)
{
    double sum = 0., xmean, pxmean, sxx, sxp, spp;
    for (int i = 0; i < n; i++) sum = sum + x[(size_t)i * stride];
    xmean = sum / n;
    sum = 0.;
    for (int i = 0; i < n; i++) sum = sum + (x[(size_t)i * stride] - xmean) * (x[(size_t)i * stride] - xmean);
    sxx = sum / n;
    sum = 0.;
    for (int i = 0; i < n; i++) sum = sum + px[(size_t)i * stride];
    pxmean = sum / n;
    sum = 0.;
    for (int i = 0; i < n; i++) sum = sum + (px[(size_t)i * stride] - pxmean) * (px[(size_t)i * stride] - pxmean);
    spp = sum / n;
    sum = 0.;
    for (int i = 0; i < n; i++) sum = sum + (x[(size_t)i * stride] - xmean) * (px[(size_t)i * stride] - pxmean);
    sxp = sum / n;
    return std::sqrt(sxx * spp - sxp * sxp);
}

/* Cal_Mean_RMS (cpp:708-722) */
static void mean_rms_ref(const double *x, int n, double &mean, double &rms)
{
    mean = 0;
    for (int i = 0; i < n; i++) mean += x[i];
    mean = mean * 1.0 / n;
    rms = 0;
    for (int i = 0; i < n; i++) rms += (x[i] - mean) * (x[i] - mean);
    rms = std::sqrt(rms * 1.0 / (n - 1));
}

/* Track_tbt_ActionDiff (simtrack.cpp:21305-21383) with the Twiss functions at
   the start as arguments (the reference takes them from Cal_Twiss).  npart
   particles at nsigma sigma (emittance split by `phase`), phases from the
   reference's LCG whose state *seed is updated; nstep steps of nturn turns,
   each particle continuing from where the last step left it with its stable
   flag reset; per step the row ((j+1)*nturn, ncount, mean and rms of 2Jx,
   mean and rms of 2Jy) with the reference's statistics, including that the
   mean runs over the first ncount entries of the per-particle array, whose
   entries of lost particles keep earlier (here initially zero) values. */
extern "C" int st_action_diff(st_table *t, double betx, double alfx, double bety, double alfy, double nsigma,
                              double phase, double emitx0, double emity0, int nturn, int nstep, int npart,
                              double *seed, double *rows)
{
    if (!t || !seed || !rows || nturn < 1 || nstep < 0 || npart < 1 || !(betx > 0.) || !(bety > 0.))
        return fail(ST_ERR_ARG, "st_action_diff: bad argument");
    const double PI = 3.14159265358979323;
    double nsigmax = nsigma * st_cos(phase), nsigmay = nsigma * st_sin(phase);
    double TwoJx0 = (nsigmax * nsigmax) * emitx0, TwoJy0 = (nsigmay * nsigmay) * emity0;
    std::vector<double> x((size_t)npart * 6), jx(npart, 0.), jy(npart, 0.);
    for (int i = 0; i < npart; i++) {
        double phix0 = rnd_ref(*seed) * 360;
        double phiy0 = rnd_ref(*seed) * 360;
        double xn = std::sqrt(TwoJx0) * st_cos(phix0 * PI / 180), pxn = -std::sqrt(TwoJx0) * st_sin(phix0 * PI / 180);
        double *p = &x[(size_t)i * 6];
        p[0] = xn * std::sqrt(betx);
        p[1] = -alfx * xn / std::sqrt(betx) + pxn / std::sqrt(betx);
        double yn = std::sqrt(TwoJy0) * st_cos(phiy0 * PI / 180), pyn = -std::sqrt(TwoJy0) * st_sin(phiy0 * PI / 180);
        p[2] = yn * std::sqrt(bety);
        p[3] = -alfy * yn / std::sqrt(bety) + pyn / std::sqrt(bety);
        p[4] = 0.;
        p[5] = 0.;
    }
    std::vector<double> tbt((size_t)npart * nturn * 6);
    std::vector<int> st(npart), lt(npart), lp(npart);
    for (int j = 0; j < nstep; j++) {
        for (int i = 0; i < npart; i++) {
            st[i] = 1;
            lt[i] = lp[i] = 0;
            if (x[(size_t)i * 6] == 0.) x[(size_t)i * 6] = x[(size_t)i * 6] + 1.0e-09;
            if (x[(size_t)i * 6 + 2] == 0.) x[(size_t)i * 6 + 2] = x[(size_t)i * 6 + 2] + 1.0e-09;
        }
        int rc = track_host(t, x.data(), npart, nturn, st.data(), lt.data(), lp.data(), false, false, tbt.data(), -1);
        if (rc) return rc;
        int ncount = 0;
        for (int i = 0; i < npart; i++) {
            if (st[i] != 1) continue;
            ncount++;
            const double *b = &tbt[(size_t)i * nturn * 6];
            jx[i] = twiss_emit_ref(b + 0, b + 1, 6, nturn);
            jy[i] = twiss_emit_ref(b + 2, b + 3, 6, nturn);
        }
        double *o = rows + (size_t)j * 6;
        o[0] = (double)((j + 1) * nturn);
        o[1] = ncount;
        mean_rms_ref(jx.data(), ncount, o[2], o[3]);
        mean_rms_ref(jy.data(), ncount, o[4], o[5]);
    }
    return ST_OK;
}

/* --------------------------------------------- many tables, one launch */

struct st_batch {
    st_ctx *c = nullptr;
    std::vector<st_table *> tabs;
    st_seed_desc *d_seeds[2] = {nullptr, nullptr}; /* orbit, spin */
    int *d_items = nullptr;
    size_t cap_items = 0;
    int *d_counter = nullptr;
    int n_cu = 0;
    bool hexp = false, rad = false;
    st_kglobals kg;
    ~st_batch()
    {
        if (!c) return;
        hipSetDevice(c->device);
        hipFree(d_seeds[0]);
        hipFree(d_seeds[1]);
        hipFree(d_items);
        hipFree(d_counter);
    }
};

extern "C" int st_batch_create(st_ctx *c, st_table *const *tables, int ntab, st_batch **out)
{
    if (!c || !tables || ntab <= 0 || !out) return fail(ST_ERR_ARG, "st_batch_create: bad argument");
    *out = nullptr;
    for (int k = 0; k < ntab; ++k) {
        if (!tables[k] || tables[k]->ctx != c) return fail(ST_ERR_ARG, "st_batch_create: table %d not on this context", k);
        const st_table *t0 = tables[0], *t = tables[k];
        if (std::memcmp(&t->kg, &t0->kg, sizeof(st_kglobals)) != 0 || t->gp.H_expand != t0->gp.H_expand ||
            t->gp.radiate != t0->gp.radiate)
            return fail(ST_ERR_ARG, "st_batch_create: table %d has other globals than table 0 (one launch shares them)", k);
    }
    HIPCHK(hipSetDevice(c->device));
    std::unique_ptr<st_batch> b(new st_batch);
    b->c = c;
    b->tabs.assign(tables, tables + ntab);
    b->hexp = tables[0]->gp.H_expand != 0;
    b->rad = tables[0]->gp.radiate != 0;
    b->kg = tables[0]->kg;
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    b->n_cu = ncu > 0 ? ncu : 1;
    for (int s = 0; s < 2; ++s) {
        std::vector<st_seed_desc> d((size_t)ntab);
        for (int k = 0; k < ntab; ++k) {
            st_table *t = tables[k];
            dev_table &dt = s ? t->spin : t->orbit;
            d[k].ops = dt.d_ops;
            d[k].par = dt.d_par;
            d[k].ap = t->d_ap;
            d[k].nelem = t->n;
            d[k].reserved = 0;
            d[k].draws_per_turn = s ? 0 : t->draws_per_turn;
        }
        HIPCHK(hipMalloc(&b->d_seeds[s], (size_t)ntab * sizeof(st_seed_desc)));
        HIPCHK(hipMemcpy(b->d_seeds[s], d.data(), (size_t)ntab * sizeof(st_seed_desc), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&b->d_counter, sizeof(int)));
    *out = b.release();
    return ST_OK;
}

extern "C" int st_batch_destroy(st_batch *b)
{
    delete b;
    return ST_OK;
}

/* items of the alive particles, 64 per item, never mixing tables; the
   longest tables first so that the tail of the launch is short */
static void build_items(const st_batch *b, const std::vector<int> &state, const int *offsets, std::vector<int> &idx,
                        std::vector<int> &items)
{
    const int ntab = (int)b->tabs.size();
    std::vector<int> order((size_t)ntab);
    for (int k = 0; k < ntab; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return b->tabs[u]->n > b->tabs[v]->n; });
    idx.clear();
    items.clear();
    for (int k : order) {
        int start = (int)idx.size();
        for (int p = offsets[k]; p < offsets[k + 1]; ++p)
            if (state[p] != 0) idx.push_back(p);
        for (int s0 = start; s0 < (int)idx.size(); s0 += 64) {
            items.push_back(k);
            items.push_back(s0);
            items.push_back(std::min(64, (int)idx.size() - s0));
        }
    }
}

static int batch_core(st_batch *b, double *d_x, const int *offsets, int turn0, int nturn, int *d_state, int *d_lost,
                      bool spin, const st_batch_config &cfg, double *d_tbt = nullptr)
{
    st_ctx *c = b->c;
    const int ntab = (int)b->tabs.size();
    const int np = offsets[ntab];
    if (np <= 0 || nturn <= 0) return ST_OK;
    if (cfg.wg_waves < 1 || cfg.wg_waves > 4 || cfg.waves_per_simd < 1 || cfg.waves_per_simd > 8 || cfg.turn_block < 1)
        return fail(ST_ERR_ARG, "st_batch: bad config (wg_waves %d, waves_per_simd %d, turn_block %d)", cfg.wg_waves,
                    cfg.waves_per_simd, cfg.turn_block);
    int rc = ensure_particles(c, (size_t)np, 1);
    if (rc) return rc;
    std::vector<int> state((size_t)np), idx, items;
    HIPCHK(hipMemcpyAsync(state.data(), d_state, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int grid = std::max(1, b->n_cu * 4 * cfg.waves_per_simd / cfg.wg_waves);
    const int block = 64 * cfg.wg_waves;
    bool first = true;
    for (int t0 = turn0; t0 < turn0 + nturn;) {
        const int nt = std::min(cfg.turn_block, turn0 + nturn - t0);
        /* the pre-check runs on every particle, alive or not, before turn 0 */
        if (first && turn0 == 0) {
            std::vector<int> all((size_t)np, 1);
            build_items(b, all, offsets, idx, items);
        } else {
            build_items(b, state, offsets, idx, items);
        }
        const int nitems = (int)items.size() / 3;
        if (nitems == 0) break;
        if ((size_t)items.size() > b->cap_items) {
            hipFree(b->d_items);
            b->d_items = nullptr;
            b->cap_items = 0;
            HIPCHK(hipMalloc(&b->d_items, items.size() * sizeof(int)));
            b->cap_items = items.size();
        }
        HIPCHK(hipMemcpyAsync(c->d_idx0, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(b->d_items, items.data(), items.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemsetAsync(b->d_counter, 0, sizeof(int), c->stream));
        st_batch_args a;
        std::memset(&a, 0, sizeof a);
        a.seeds = b->d_seeds[spin ? 1 : 0];
        a.items = b->d_items;
        a.nitems = nitems;
        a.counter = b->d_counter;
        a.idx = c->d_idx0;
        a.x = d_x;
        a.state = d_state;
        a.lost = d_lost;
        a.turn0 = t0;
        a.nturn = nt;
        a.precheck = (first && turn0 == 0) ? 1 : 0;
        a.xtbt = d_tbt;
        a.tbt_nturn = nturn;
        a.tbt_turn0 = turn0;
        a.g = b->kg;
        a.gdev = b->tabs[0]->d_kg;
        int ev = -1;
        if (c->timing) {
            rc = timed_event(c, &ev);
            if (rc) return rc;
            HIPCHK(hipEventRecord(c->ev_pool[ev], c->stream));
        }
        HIPCHK(st_launch_track_batch(a, grid, block, spin, b->hexp, b->rad, c->stream));
        if (c->timing) {
            HIPCHK(hipEventRecord(c->ev_pool[ev + 1], c->stream));
            c->pending.push_back({ev, ev + 1});
            c->pending_work.push_back((int64_t)idx.size() * nt);
        }
        HIPCHK(hipMemcpyAsync(state.data(), d_state, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->timing) {
            rc = drain_timing(c);
            if (rc) return rc;
        }
        first = false;
        t0 += nt;
    }
    return ST_OK;
}

static int batch_check(const st_batch *b, const int *offsets)
{
    if (!b || !offsets) return fail(ST_ERR_ARG, "st_batch: bad argument");
    const int ntab = (int)b->tabs.size();
    if (offsets[0] != 0) return fail(ST_ERR_ARG, "st_batch: offsets[0] must be 0");
    for (int k = 0; k < ntab; ++k)
        if (offsets[k + 1] < offsets[k]) return fail(ST_ERR_ARG, "st_batch: offsets must not decrease");
    return ST_OK;
}

extern "C" int st_batch_default_config(st_batch_config *cfg)
{
    if (!cfg) return fail(ST_ERR_ARG, "st_batch_default_config: null");
    cfg->wg_waves = 1;
    cfg->waves_per_simd = 2;
    cfg->turn_block = 256;
    return ST_OK;
}

extern "C" int st_batch_track_device(st_batch *b, double *d_x, const int *offsets, int turn0, int nturn,
                                     int *d_state, int *d_lost, int flags, const st_batch_config *cfg)
{
    int rc = batch_check(b, offsets);
    if (rc) return rc;
    if (!d_x || !d_state || !d_lost || turn0 < 0 || nturn < 0) return fail(ST_ERR_ARG, "st_batch_track_device: bad argument");
    st_batch_config dc;
    st_batch_default_config(&dc);
    HIPCHK(hipSetDevice(b->c->device));
    return batch_core(b, d_x, offsets, turn0, nturn, d_state, d_lost, (flags & ST_TRACK_SPIN) != 0, cfg ? *cfg : dc);
}

extern "C" int st_batch_track_tbt_device(st_batch *b, double *d_x, const int *offsets, int turn0, int nturn,
                                         double *d_tbt, int *d_state, int *d_lost, int flags,
                                         const st_batch_config *cfg)
{
    int rc = batch_check(b, offsets);
    if (rc) return rc;
    if (!d_x || !d_tbt || !d_state || !d_lost || turn0 < 0 || nturn < 0)
        return fail(ST_ERR_ARG, "st_batch_track_tbt_device: bad argument");
    st_batch_config dc;
    st_batch_default_config(&dc);
    HIPCHK(hipSetDevice(b->c->device));
    return batch_core(b, d_x, offsets, turn0, nturn, d_state, d_lost, (flags & ST_TRACK_SPIN) != 0, cfg ? *cfg : dc,
                      d_tbt);
}

extern "C" int st_batch_track(st_batch *b, double *x, const int *offsets, int nturn, int *stable, int *lost_turn,
                              int *lost_post, int flags, const st_batch_config *cfg)
{
    int rc = batch_check(b, offsets);
    if (rc) return rc;
    const int np = offsets[b->tabs.size()];
    if (np < 0 || nturn < 0 || (np > 0 && (!x || !stable || !lost_turn || !lost_post)))
        return fail(ST_ERR_ARG, "st_batch_track: bad argument");
    if (np == 0) return ST_OK;
    st_ctx *c = b->c;
    HIPCHK(hipSetDevice(c->device));
    const bool spin = (flags & ST_TRACK_SPIN) != 0;
    const int ncol = spin ? 9 : 6;
    rc = ensure_particles(c, (size_t)np, ncol);
    if (rc) return rc;
    std::vector<int> lost(2 * (size_t)np);
    for (int p = 0; p < np; ++p) {
        lost[2 * p] = lost_turn[p];
        lost[2 * p + 1] = lost_post[p];
    }
    HIPCHK(hipMemcpyAsync(c->d_x, x, (size_t)np * ncol * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_state, stable, (size_t)np * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_lost, lost.data(), 2 * (size_t)np * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    st_batch_config dc;
    st_batch_default_config(&dc);
    rc = batch_core(b, c->d_x, offsets, 0, nturn, c->d_state, c->d_lost, spin, cfg ? *cfg : dc);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipMemcpyAsync(x, c->d_x, (size_t)np * ncol * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(stable, c->d_state, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(lost.data(), c->d_lost, 2 * (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_gpu_s = ms * 1e-3;
    for (int p = 0; p < np; ++p) {
        lost_turn[p] = lost[2 * p];
        lost_post[p] = lost[2 * p + 1];
    }
    return ST_OK;
}

/*
 * Autotuner: times every candidate launch shape on the same calibration run
 * (calib_turns turns of the given particles, re-uploaded for each candidate)
 * and returns the fastest.  Candidates: 1, 2 or 4 wavefronts per workgroup;
 * 1 or 2 resident wavefronts per SIMD; turn blocks of 16, 64 or 256 turns.
 */
extern "C" int st_batch_autotune(st_batch *b, const double *x, const int *offsets, int calib_turns, int flags,
                                 st_batch_config *best, double *seconds_out, int *n_candidates)
{
    int rc = batch_check(b, offsets);
    if (rc) return rc;
    const int np = offsets[b->tabs.size()];
    if (np <= 0 || !x || !best || calib_turns <= 0) return fail(ST_ERR_ARG, "st_batch_autotune: bad argument");
    st_ctx *c = b->c;
    HIPCHK(hipSetDevice(c->device));
    const bool spin = (flags & ST_TRACK_SPIN) != 0;
    const int ncol = spin ? 9 : 6;
    rc = ensure_particles(c, (size_t)np, ncol);
    if (rc) return rc;
    const int wgs[3] = {1, 2, 4}, wps[2] = {1, 2}, tbs[3] = {16, 64, 256};
    std::vector<int> ones((size_t)np, 1), zeros(2 * (size_t)np, 0);
    double tbest = 1e300;
    int k = 0;
    for (int w : wgs)
        for (int o : wps)
            for (int tb : tbs) {
                st_batch_config cfg{w, o, tb};
                HIPCHK(hipMemcpyAsync(c->d_x, x, (size_t)np * ncol * sizeof(double), hipMemcpyHostToDevice, c->stream));
                HIPCHK(hipMemcpyAsync(c->d_state, ones.data(), (size_t)np * sizeof(int), hipMemcpyHostToDevice, c->stream));
                HIPCHK(hipMemcpyAsync(c->d_lost, zeros.data(), 2 * (size_t)np * sizeof(int), hipMemcpyHostToDevice,
                                      c->stream));
                HIPCHK(hipEventRecord(c->ev0, c->stream));
                rc = batch_core(b, c->d_x, offsets, 0, calib_turns, c->d_state, c->d_lost, spin, cfg);
                if (rc) return rc;
                HIPCHK(hipEventRecord(c->ev1, c->stream));
                HIPCHK(hipEventSynchronize(c->ev1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
                if (seconds_out) seconds_out[k] = ms * 1e-3;
                if (ms * 1e-3 < tbest) {
                    tbest = ms * 1e-3;
                    *best = cfg;
                }
                ++k;
            }
    if (n_candidates) *n_candidates = k;
    return ST_OK;
}

extern "C" __attribute__((visibility("hidden"))) hipStream_t st_internal_stream(st_ctx *c, int *device)
{
    if (device) *device = c ? c->device : 0;
    return c ? c->stream : nullptr;
}

extern "C" int st_surrogate_features(st_ctx *c, const double *d_x0, int ncol, const double *d_tunes, int n,
                                     double beta_x, double alpha_x, double beta_y, double alpha_y, float *d_feat)
{
    if (!c || n < 0 || (n > 0 && (!d_x0 || !d_tunes || !d_feat)) || ncol < 4 || !(beta_x > 0.) || !(beta_y > 0.))
        return fail(ST_ERR_ARG, "st_surrogate_features: bad argument");
    if (n == 0) return ST_OK;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(st_launch_features(d_x0, ncol, d_tunes, n, beta_x, alpha_x, beta_y, alpha_y, d_feat, c->stream));
    return ST_OK;
}

/* ------------------------------------------ strong-strong beam-beam */

namespace {

const double SS_RP = 1.53469826769e-18; /* BB4D's classical radius (simtrack.cpp:6328) */

/* carve 256-byte aligned pieces out of the context's work area */
struct ss_arena {
    size_t off = 0;
    char *base = nullptr;
    template <class T> T *take(size_t n)
    {
        size_t at = off;
        off = (off + n * sizeof(T) + 255) & ~(size_t)255;
        return base ? (T *)(base + at) : nullptr;
    }
};

struct ss_work {
    ss_sort_buffers s;
    int *rows[2];
    double *part, *Z, *zb[2];
    int *seg_i, *cnt;
    double *hx[2];
    size_t cap_part, cap_seg;
    /* particle-in-cell: one set of buffers per lane */
    struct pic_lane {
        int *pkey_in, *pkey_out, *pval_in, *pval_out;
        double *pw, *pws, *grid, *params, *part;
        void *ptemp;
    } pk[SS_LANES];
    int lanes;
    double *tw;
    size_t ptemp_bytes;
};

/* lay out (base == null: measure) the work area for beams of n1, n2 rows;
   pic_rows > 0 adds the particle-in-cell buffers for slices of that size */
size_t ss_layout(char *base, int n1, int n2, int ns1, int ns2, bool host_rows, size_t temp_bytes, ss_work &w,
                 int pic_rows = 0, size_t pic_temp = 0, int lanes = 1)
{
    ss_arena a;
    a.base = base;
    const int nmax = std::max(n1, n2) > 0 ? std::max(n1, n2) : 1;
    w.s.dkey_in = a.take<double>(nmax);
    w.s.dkey_out = a.take<double>(nmax);
    w.s.ikey_in = a.take<int>(nmax);
    w.s.ikey_out = a.take<int>(nmax);
    w.s.row_in = a.take<int>(nmax);
    w.s.row_out = a.take<int>(nmax);
    w.s.rows = a.take<double>((size_t)nmax * 6);
    w.s.counts = a.take<int>((size_t)std::max(ns1, ns2) + 1);
    w.s.temp = a.take<char>(temp_bytes);
    w.s.temp_bytes = temp_bytes;
    w.rows[0] = a.take<int>(n1 > 0 ? n1 : 1);
    w.rows[1] = a.take<int>(n2 > 0 ? n2 : 1);
    w.cap_part = (size_t)(st_ss_partial_blocks(n1) + st_ss_partial_blocks(n2) + ns1 + ns2 + 2) * 4;
    w.part = a.take<double>(w.cap_part);
    w.Z = a.take<double>((size_t)ns1 + ns2);
    w.zb[0] = a.take<double>((size_t)ns1 + 1);
    w.zb[1] = a.take<double>((size_t)ns2 + 1);
    w.cap_seg = (size_t)(std::max(n1, n2) / SS_CHUNK + std::max(ns1, ns2) + 2);
    w.seg_i = a.take<int>(w.cap_seg * 6);
    w.cnt = a.take<int>((size_t)std::max(ns1, ns2) + 1);
    w.hx[0] = host_rows ? a.take<double>((size_t)(n1 > 0 ? n1 : 1) * 6) : nullptr;
    w.hx[1] = host_rows ? a.take<double>((size_t)(n2 > 0 ? n2 : 1) * 6) : nullptr;
    std::memset(w.pk, 0, sizeof w.pk);
    w.tw = nullptr;
    w.ptemp_bytes = pic_temp;
    w.lanes = lanes;
    for (int l = 0; l < lanes; ++l) w.pk[l].part = l == 0 ? w.part : a.take<double>(w.cap_part);
    if (pic_rows > 0) {
        const size_t nc = (size_t)pic_rows * 9;
        for (int l = 0; l < lanes; ++l) {
            ss_work::pic_lane &k = w.pk[l];
            k.pkey_in = a.take<int>(nc);
            k.pkey_out = a.take<int>(nc);
            k.pval_in = a.take<int>(nc);
            k.pval_out = a.take<int>(nc);
            k.pw = a.take<double>(nc);
            k.pws = a.take<double>(nc);
            k.ptemp = a.take<char>(pic_temp);
            k.grid = a.take<double>((size_t)4096 * 4 + (size_t)16384 * 6);
            k.params = a.take<double>(16);
        }
        w.tw = a.take<double>(2 * 254);
    }
    return a.off;
}

} // namespace

static int ss_prepare(st_ctx *c, int n1, int n2, int ns1, int ns2, bool host_rows, ss_work &w, int pic_rows = 0)
{
    size_t temp = 0, ptemp = 0;
    HIPCHK(st_ss_sort_temp_bytes(std::max(std::max(n1, n2), 1), &temp));
    if (pic_rows > 0) HIPCHK(st_sort_cells_bytes(pic_rows * 9, &ptemp));
    /* as many lanes as the smaller beam has slices */
    const int lanes = std::max(1, std::min(SS_LANES, std::min(ns1, ns2)));
    size_t need = ss_layout(nullptr, n1, n2, ns1, ns2, host_rows, temp, w, pic_rows, ptemp, lanes);
    if (need > c->cap_ss) {
        hipFree(c->d_ss);
        c->d_ss = nullptr;
        c->cap_ss = 0;
        HIPCHK(hipMalloc((void **)&c->d_ss, need));
        c->cap_ss = need;
    }
    ss_layout(c->d_ss, n1, n2, ns1, ns2, host_rows, temp, w, pic_rows, ptemp, lanes);
    for (int l = 1; l < lanes; ++l) {
        if (!c->lane[l]) HIPCHK(hipStreamCreateWithFlags(&c->lane[l], hipStreamNonBlocking));
        if (!c->lane_ev[l]) HIPCHK(hipEventCreateWithFlags(&c->lane_ev[l], hipEventDisableTiming));
    }
    if (lanes > 1 && !c->fork_ev) HIPCHK(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
    if (pic_rows > 0) {
        /* FFT1D's twiddles (simtrack.cpp:2614-2637), both directions, by its
           recurrence; built once (thread-safe static initialisation) */
        static const std::vector<double> tw = [] {
            std::vector<double> tw(2 * 254, 0.);
            for (int d = 0; d < 2; ++d) {
                const int dir = d == 0 ? 1 : -1;
                double *t = tw.data() + 254 * d;
                double c1 = -1.0, c2 = 0.0;
                for (int l = 0; l < 7; l++) {
                    const int l1 = 1 << l;
                    double u1 = 1.0, u2 = 0.0;
                    for (int j = 0; j < l1; j++) {
                        t[2 * (l1 - 1 + j)] = u1;
                        t[2 * (l1 - 1 + j) + 1] = u2;
                        const double z = u1 * c1 - u2 * c2;
                        u2 = u1 * c2 + u2 * c1;
                        u1 = z;
                    }
                    c2 = std::sqrt((1.0 - c1) / 2.0);
                    if (dir == 1) c2 = -c2;
                    c1 = std::sqrt((1.0 + c1) / 2.0);
                }
            }
            return tw;
        }();
        HIPCHK(hipMemcpyAsync(w.tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    return ST_OK;
}

/* slice centres of one beam: Z[s] = (sum of z over the slice's rows) / count;
   segments s are [start[s], start[s] + count[s]) of the row list idx */
static int ss_slice_centres(st_ctx *c, ss_work &w, const double *d_x, const int *d_idx, const std::vector<int> &start,
                            const std::vector<int> &count, double *d_Z)
{
    const int ns = (int)start.size();
    std::vector<int> blk_seg, blk_chunk, first(ns), nblk(ns);
    for (int s = 0; s < ns; ++s) {
        first[s] = (int)blk_seg.size();
        nblk[s] = st_ss_partial_blocks(count[s]);
        for (int k = 0; k < nblk[s]; ++k) {
            blk_seg.push_back(s);
            blk_chunk.push_back(k);
        }
    }
    const int nb = (int)blk_seg.size();
    if ((size_t)nb > w.cap_seg || (size_t)ns > w.cap_seg || (size_t)nb > w.cap_part)
        return fail(ST_ERR_PARAM, "strong-strong: slice table larger than its work area");
    int *d = w.seg_i;
    std::vector<int> h(w.cap_seg * 6, 0);
    std::copy(start.begin(), start.end(), h.begin());
    std::copy(count.begin(), count.end(), h.begin() + w.cap_seg);
    std::copy(blk_seg.begin(), blk_seg.end(), h.begin() + 2 * w.cap_seg);
    std::copy(blk_chunk.begin(), blk_chunk.end(), h.begin() + 3 * w.cap_seg);
    std::copy(first.begin(), first.end(), h.begin() + 4 * w.cap_seg);
    std::copy(nblk.begin(), nblk.end(), h.begin() + 5 * w.cap_seg);
    HIPCHK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(st_ss_zsums(d_x, d_idx, d, d + w.cap_seg, d + 2 * w.cap_seg, d + 3 * w.cap_seg, nb, d + 4 * w.cap_seg,
                       d + 5 * w.cap_seg, ns, w.part, d_Z, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream)); /* h is freed on return */
    return ST_OK;
}

/* the collision order (cpp:22992-23015): pairs by -(Z1 + Z2)/2 ascending,
   the reference's exchange sort */
static void ss_schedule(const std::vector<double> &Z1, const std::vector<double> &Z2, std::vector<int> &I1,
                        std::vector<int> &I2, std::vector<double> &S)
{
    const int n1 = (int)Z1.size(), n2 = (int)Z2.size(), n = n1 * n2;
    std::vector<double> Z10(n);
    I1.assign(n, 0);
    I2.assign(n, 0);
    S.assign(n, 0.);
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
                std::swap(Z10[i], Z10[j]);
                std::swap(S[i], S[j]);
                std::swap(I1[i], I1[j]);
                std::swap(I2[i], I2[j]);
            }
}

/*
 * The slice pairs of a schedule in runs: consecutive pairs that touch
 * disjoint slices of both beams commute, so each run goes out at once, pair
 * k of the run on lane k (lane 0 is the context stream); pairs that share a
 * slice fall in different runs and keep their order, so the result is the
 * sequential one bit for bit.  pair(i, lane, stream) enqueues pair i.
 */
template <class F>
static int ss_runs(st_ctx *c, const ss_work &w, const std::vector<int> &I1, const std::vector<int> &I2, int Ns1,
                   int Ns2, F &&pair)
{
    const size_t np = I1.size();
    const int lanes = std::max(1, w.lanes);
    size_t i0 = 0;
    while (i0 < np) {
        size_t i1 = i0;
        std::vector<char> u1(Ns1, 0), u2(Ns2, 0);
        while (i1 < np && (int)(i1 - i0) < lanes && !u1[I1[i1]] && !u2[I2[i1]]) {
            u1[I1[i1]] = u2[I2[i1]] = 1;
            ++i1;
        }
        if (i1 - i0 > 1) HIPCHK(hipEventRecord(c->fork_ev, c->stream));
        for (size_t i = i0; i < i1; ++i) {
            const int l = (int)(i - i0);
            hipStream_t st = l == 0 ? c->stream : c->lane[l];
            if (l > 0) HIPCHK(hipStreamWaitEvent(st, c->fork_ev, 0));
            int rc = pair(i, l, st);
            if (rc) return rc;
            if (l > 0) HIPCHK(hipEventRecord(c->lane_ev[l], st));
        }
        for (size_t i = i0 + 1; i < i1; ++i) HIPCHK(hipStreamWaitEvent(c->stream, c->lane_ev[i - i0], 0));
        i0 = i1;
    }
    return ST_OK;
}

static int ss_pair(st_ctx *c, ss_work &w, double *d_x1, const int *idx1, int start1, int m1, double *d_x2,
                   const int *idx2, int start2, int m2, double S, int drift, double N1, double N2, double gamma1,
                   double gamma2, double bbscale1, double bbscale2, int lane = 0)
{
    hipStream_t st = lane == 0 ? c->stream : c->lane[lane];
    ss_pair_args a{};
    a.x[0] = d_x1;
    a.x[1] = d_x2;
    a.idx[0] = idx1;
    a.idx[1] = idx2;
    a.start[0] = start1;
    a.start[1] = start2;
    a.count[0] = m1;
    a.count[1] = m2;
    a.S = S;
    a.drift = drift;
    /* BB4D's 2 N rp / gamma with the other beam's charge and this beam's energy */
    a.c2Nrg[0] = 2 * N2 * SS_RP / gamma1;
    a.c2Nrg[1] = 2 * N1 * SS_RP / gamma2;
    a.bbscale[0] = bbscale1;
    a.bbscale[1] = bbscale2;
    a.part = w.pk[lane].part;
    int ev = -1;
    if (c->timing) {
        int rc = timed_event(c, &ev);
        if (rc) return rc;
        HIPCHK(hipEventRecord(c->ev_pool[ev], st));
    }
    HIPCHK(st_ss_collide(a, st));
    if (c->timing) {
        HIPCHK(hipEventRecord(c->ev_pool[ev + 1], st));
        c->pending.push_back({ev, ev + 1});
        c->pending_work.push_back((int64_t)m1 + m2);
    }
    return ST_OK;
}

static int ss_check(st_ctx *c, int Mp1, const double *x1, int Mp2, const double *x2, const char *what)
{
    if (!c) return fail(ST_ERR_ARG, "%s: null context", what);
    if (Mp1 < 0 || Mp2 < 0 || (Mp1 > 0 && !x1) || (Mp2 > 0 && !x2)) return fail(ST_ERR_ARG, "%s: bad beam", what);
    if ((size_t)Mp1 * 6 > (size_t)INT32_MAX || (size_t)Mp2 * 6 > (size_t)INT32_MAX)
        return fail(ST_ERR_ARG, "%s: beam too large", what);
    return ST_OK;
}

/* BeamBeam4D_SoftGaussian, simtrack.cpp:22846-22872 */
static int ss_4d(st_ctx *c, ss_work &w, double Np1, int Mp1, double *d_x1, double Np2, int Mp2, double *d_x2,
                 double gamma1, double gamma2, double bbscale1, double bbscale2)
{
    return ss_pair(c, w, d_x1, nullptr, 0, Mp1, d_x2, nullptr, 0, Mp2, 0., 0, Np1, Np2, gamma1, gamma2, bbscale1,
                   bbscale2);
}

/* BeamBeam6D_SoftGaussian, simtrack.cpp:22874-23132 */
static int ss_6d(st_ctx *c, ss_work &w, double Np1, int Mp1, double *d_x1, double Np2, int Mp2, double *d_x2,
                 double gamma1, double gamma2, double bbscale1, double bbscale2, int Ns1, int Ns2, double *Z1o,
                 double *Z2o)
{
    HIPCHK(st_ss_sort_by_z(d_x1, Mp1, w.s, c->stream));
    HIPCHK(st_ss_sort_by_z(d_x2, Mp2, w.s, c->stream));
    const int mc1 = Mp1 / Ns1, mc2 = Mp2 / Ns2;
    std::vector<int> st1(Ns1), ct1(Ns1, mc1), st2(Ns2), ct2(Ns2, mc2);
    for (int i = 0; i < Ns1; i++) st1[i] = i * mc1;
    for (int i = 0; i < Ns2; i++) st2[i] = i * mc2;
    int rc = ss_slice_centres(c, w, d_x1, nullptr, st1, ct1, w.Z);
    if (!rc) rc = ss_slice_centres(c, w, d_x2, nullptr, st2, ct2, w.Z + Ns1);
    if (rc) return rc;
    std::vector<double> Z1(Ns1), Z2(Ns2);
    HIPCHK(hipMemcpyAsync(Z1.data(), w.Z, Ns1 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(Z2.data(), w.Z + Ns1, Ns2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (Z1o) std::copy(Z1.begin(), Z1.end(), Z1o);
    if (Z2o) std::copy(Z2.begin(), Z2.end(), Z2o);
    std::vector<int> I1, I2;
    std::vector<double> S;
    ss_schedule(Z1, Z2, I1, I2, S);
    return ss_runs(c, w, I1, I2, Ns1, Ns2, [&](size_t i, int l, hipStream_t) {
        return ss_pair(c, w, d_x1, nullptr, I1[i] * mc1, mc1, d_x2, nullptr, I2[i] * mc2, mc2, S[i], 1, Np1 / Ns1,
                       Np2 / Ns2, gamma1, gamma2, bbscale1, bbscale2, l);
    });
}

/* BeamBeam6D_SoftGaussian with slice boundaries, simtrack.cpp:23134-23284 */
static int ss_6d_zb(st_ctx *c, ss_work &w, double Np1, int Mp1, double *d_x1, double Np2, int Mp2, double *d_x2,
                    double gamma1, double gamma2, double bbscale1, double bbscale2, int Ns1, int Ns2,
                    const double *zb1, const double *zb2)
{
    double *dx[2] = {d_x1, d_x2};
    const int Mp[2] = {Mp1, Mp2}, Ns[2] = {Ns1, Ns2};
    const double *zb[2] = {zb1, zb2};
    std::vector<int> st[2], ct[2];
    for (int b = 0; b < 2; ++b) {
        HIPCHK(hipMemcpyAsync(w.zb[b], zb[b], (Ns[b] + 1) * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(st_ss_group_by_slice(dx[b], Mp[b], w.zb[b], Ns[b], w.s, c->stream));
        if (Mp[b] > 0)
            HIPCHK(hipMemcpyAsync(w.rows[b], w.s.row_out, Mp[b] * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
        std::vector<int> cnt(Ns[b] + 1, 0);
        HIPCHK(hipMemcpyAsync(cnt.data(), w.s.counts, (Ns[b] + 1) * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (Mp[b] == 0) std::fill(cnt.begin(), cnt.end(), 0);
        st[b].assign(Ns[b], 0);
        ct[b].assign(Ns[b], 0);
        int at = 0;
        for (int i = 0; i < Ns[b]; ++i) {
            st[b][i] = at;
            ct[b][i] = cnt[i];
            at += cnt[i];
        }
        int rc = ss_slice_centres(c, w, dx[b], w.rows[b], st[b], ct[b], w.Z + (b ? Ns1 : 0));
        if (rc) return rc;
    }
    std::vector<double> Z1(Ns1), Z2(Ns2);
    HIPCHK(hipMemcpyAsync(Z1.data(), w.Z, Ns1 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(Z2.data(), w.Z + Ns1, Ns2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<int> I1, I2;
    std::vector<double> S;
    ss_schedule(Z1, Z2, I1, I2, S);
    return ss_runs(c, w, I1, I2, Ns1, Ns2, [&](size_t i, int l, hipStream_t) {
        const int a = I1[i], b = I2[i];
        const int m1 = ct[0][a], m2 = ct[1][b];
        /* BeamBeam4D_SoftGaussian(Np1*Mp1_coll/Mp1, ...) on the drifted slices (cpp:23252) */
        return ss_pair(c, w, d_x1, w.rows[0], st[0][a], m1, d_x2, w.rows[1], st[1][b], m2, S[i], 1,
                       Np1 * m1 / Mp1, Np2 * m2 / Mp2, gamma1, gamma2, bbscale1, bbscale2, l);
    });
}

static int ss_check_zb(int Ns, const double *zb, const char *what)
{
    if (!zb) return fail(ST_ERR_ARG, "%s: null slice boundaries", what);
    for (int i = 0; i < Ns; ++i)
        if (!(zb[i] > zb[i + 1])) return fail(ST_ERR_PARAM, "%s: slice boundaries must decrease strictly", what);
    return ST_OK;
}

extern "C" int st_bb_soft_gaussian_4d(st_ctx *c, double Np1, int Mp1, double *d_x1, double Np2, int Mp2,
                                      double *d_x2, double gamma1, double gamma2, double bbscale1, double bbscale2)
{
    int rc = ss_check(c, Mp1, d_x1, Mp2, d_x2, "st_bb_soft_gaussian_4d");
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    ss_work w;
    if ((rc = ss_prepare(c, Mp1, Mp2, 1, 1, false, w))) return rc;
    if ((rc = ss_4d(c, w, Np1, Mp1, d_x1, Np2, Mp2, d_x2, gamma1, gamma2, bbscale1, bbscale2))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return ST_OK;
}

extern "C" int st_bb_soft_gaussian_6d(st_ctx *c, double Np1, int Mp1, double *d_x1, double Np2, int Mp2,
                                      double *d_x2, double gamma1, double gamma2, double bbscale1, double bbscale2,
                                      int Nslice1, int Nslice2, double *Z1, double *Z2)
{
    int rc = ss_check(c, Mp1, d_x1, Mp2, d_x2, "st_bb_soft_gaussian_6d");
    if (rc) return rc;
    if (Nslice1 < 1 || Nslice2 < 1) return fail(ST_ERR_PARAM, "st_bb_soft_gaussian_6d: Nslice must be >= 1");
    HIPCHK(hipSetDevice(c->device));
    ss_work w;
    if ((rc = ss_prepare(c, Mp1, Mp2, Nslice1, Nslice2, false, w))) return rc;
    if ((rc = ss_6d(c, w, Np1, Mp1, d_x1, Np2, Mp2, d_x2, gamma1, gamma2, bbscale1, bbscale2, Nslice1, Nslice2, Z1,
                    Z2)))
        return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return ST_OK;
}

extern "C" int st_bb_soft_gaussian_6d_zb(st_ctx *c, double Np1, int Mp1, double *d_x1, double Np2, int Mp2,
                                         double *d_x2, double gamma1, double gamma2, double bbscale1,
                                         double bbscale2, int Nslice1, int Nslice2, const double *zb1,
                                         const double *zb2)
{
    int rc = ss_check(c, Mp1, d_x1, Mp2, d_x2, "st_bb_soft_gaussian_6d_zb");
    if (rc) return rc;
    if (Nslice1 < 1 || Nslice2 < 1) return fail(ST_ERR_PARAM, "st_bb_soft_gaussian_6d_zb: Nslice must be >= 1");
    if ((rc = ss_check_zb(Nslice1, zb1, "st_bb_soft_gaussian_6d_zb")) ||
        (rc = ss_check_zb(Nslice2, zb2, "st_bb_soft_gaussian_6d_zb")))
        return rc;
    HIPCHK(hipSetDevice(c->device));
    ss_work w;
    if ((rc = ss_prepare(c, Mp1, Mp2, Nslice1, Nslice2, false, w))) return rc;
    if ((rc = ss_6d_zb(c, w, Np1, Mp1, d_x1, Np2, Mp2, d_x2, gamma1, gamma2, bbscale1, bbscale2, Nslice1, Nslice2,
                       zb1, zb2)))
        return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return ST_OK;
}

/* host-array forms: the beams are copied to the context's work area and back */
static int ss_host_in(st_ctx *c, ss_work &w, int Mp1, const double *x1, int Mp2, const double *x2)
{
    if (Mp1 > 0) HIPCHK(hipMemcpyAsync(w.hx[0], x1, (size_t)Mp1 * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (Mp2 > 0) HIPCHK(hipMemcpyAsync(w.hx[1], x2, (size_t)Mp2 * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return ST_OK;
}

static int ss_host_out(st_ctx *c, ss_work &w, int Mp1, double *x1, int Mp2, double *x2)
{
    if (Mp1 > 0) HIPCHK(hipMemcpyAsync(x1, w.hx[0], (size_t)Mp1 * 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (Mp2 > 0) HIPCHK(hipMemcpyAsync(x2, w.hx[1], (size_t)Mp2 * 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ST_OK;
}

extern "C" int st_bb_soft_gaussian_4d_host(st_ctx *c, double Np1, int Mp1, double *x1, double Np2, int Mp2,
                                           double *x2, double gamma1, double gamma2, double bbscale1,
                                           double bbscale2)
{
    int rc = ss_check(c, Mp1, x1, Mp2, x2, "st_bb_soft_gaussian_4d_host");
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    ss_work w;
    if ((rc = ss_prepare(c, Mp1, Mp2, 1, 1, true, w)) || (rc = ss_host_in(c, w, Mp1, x1, Mp2, x2)) ||
        (rc = ss_4d(c, w, Np1, Mp1, w.hx[0], Np2, Mp2, w.hx[1], gamma1, gamma2, bbscale1, bbscale2)))
        return rc;
    return ss_host_out(c, w, Mp1, x1, Mp2, x2);
}

extern "C" int st_bb_soft_gaussian_6d_host(st_ctx *c, double Np1, int Mp1, double *x1, double Np2, int Mp2,
                                           double *x2, double gamma1, double gamma2, double bbscale1,
                                           double bbscale2, int Nslice1, int Nslice2, double *Z1, double *Z2)
{
    int rc = ss_check(c, Mp1, x1, Mp2, x2, "st_bb_soft_gaussian_6d_host");
    if (rc) return rc;
    if (Nslice1 < 1 || Nslice2 < 1) return fail(ST_ERR_PARAM, "st_bb_soft_gaussian_6d_host: Nslice must be >= 1");
    HIPCHK(hipSetDevice(c->device));
    ss_work w;
    if ((rc = ss_prepare(c, Mp1, Mp2, Nslice1, Nslice2, true, w)) || (rc = ss_host_in(c, w, Mp1, x1, Mp2, x2)) ||
        (rc = ss_6d(c, w, Np1, Mp1, w.hx[0], Np2, Mp2, w.hx[1], gamma1, gamma2, bbscale1, bbscale2, Nslice1,
                    Nslice2, Z1, Z2)))
        return rc;
    return ss_host_out(c, w, Mp1, x1, Mp2, x2);
}

extern "C" int st_bb_soft_gaussian_6d_zb_host(st_ctx *c, double Np1, int Mp1, double *x1, double Np2, int Mp2,
                                              double *x2, double gamma1, double gamma2, double bbscale1,
                                              double bbscale2, int Nslice1, int Nslice2, const double *zb1,
                                              const double *zb2)
{
    int rc = ss_check(c, Mp1, x1, Mp2, x2, "st_bb_soft_gaussian_6d_zb_host");
    if (rc) return rc;
    if (Nslice1 < 1 || Nslice2 < 1) return fail(ST_ERR_PARAM, "st_bb_soft_gaussian_6d_zb_host: Nslice must be >= 1");
    if ((rc = ss_check_zb(Nslice1, zb1, "st_bb_soft_gaussian_6d_zb_host")) ||
        (rc = ss_check_zb(Nslice2, zb2, "st_bb_soft_gaussian_6d_zb_host")))
        return rc;
    HIPCHK(hipSetDevice(c->device));
    ss_work w;
    if ((rc = ss_prepare(c, Mp1, Mp2, Nslice1, Nslice2, true, w)) || (rc = ss_host_in(c, w, Mp1, x1, Mp2, x2)) ||
        (rc = ss_6d_zb(c, w, Np1, Mp1, w.hx[0], Np2, Mp2, w.hx[1], gamma1, gamma2, bbscale1, bbscale2, Nslice1,
                       Nslice2, zb1, zb2)))
        return rc;
    return ss_host_out(c, w, Mp1, x1, Mp2, x2);
}

/* BeamBeam4D_PIC on rows [s1, s1 + m1) of beam 1 and [s2, s2 + m2) of beam 2
   (simtrack.cpp:23286-23378): beam 1's charge kicks beam 2, then beam 2's
   charge kicks beam 1 */
static int pic_pair(st_ctx *c, ss_work &w, double *d_x1, int s1, int m1, double *d_x2, int s2, int m2, double Np1,
                    double Np2, double gamma1, double gamma2, double bbscale1, double bbscale2, int lane = 0)
{
    const ss_work::pic_lane &k = w.pk[lane];
    hipStream_t st = lane == 0 ? c->stream : c->lane[lane];
    int ev = -1;
    if (c->timing) {
        int rc = timed_event(c, &ev);
        if (rc) return rc;
        HIPCHK(hipEventRecord(c->ev_pool[ev], st));
    }
    ss_pair_args a{};
    a.x[0] = d_x1;
    a.x[1] = d_x2;
    a.start[0] = s1;
    a.start[1] = s2;
    a.count[0] = m1;
    a.count[1] = m2;
    a.part = k.part;
    HIPCHK(st_ss_partials(a, st));
    HIPCHK(st_pic_params(k.part, a.nb[0], a.nb[1], m1, m2, k.params, st));
    pic_args p{};
    p.x[0] = d_x1;
    p.x[1] = d_x2;
    p.start[0] = s1;
    p.start[1] = s2;
    p.count[0] = m1;
    p.count[1] = m2;
    p.params = k.params;
    p.key_in = k.pkey_in;
    p.key_out = k.pkey_out;
    p.val_in = k.pval_in;
    p.val_out = k.pval_out;
    p.w = k.pw;
    p.ws = k.pws;
    p.temp = k.ptemp;
    p.temp_bytes = w.ptemp_bytes;
    double *g = k.grid;
    p.src = g;
    p.phi = g + 4096;
    p.Ex = g + 2 * 4096;
    p.Ey = g + 3 * 4096;
    p.gr = g + 4 * 4096;
    p.gi = p.gr + 16384;
    p.sr = p.gi + 16384;
    p.si = p.sr + 16384;
    p.pr = p.si + 16384;
    p.pi = p.pr + 16384;
    p.tw_fwd = w.tw;
    p.tw_inv = w.tw + 254;
    p.nsm[0] = Np1 / m1; /* (Np1/Mp1), cpp:23338 */
    p.nsm[1] = Np2 / m2;
    p.gamma[0] = gamma1;
    p.gamma[1] = gamma2;
    p.bbscale[0] = bbscale1;
    p.bbscale[1] = bbscale2;
    HIPCHK(st_pic_solve_and_kick(p, 0, st));
    HIPCHK(st_pic_solve_and_kick(p, 1, st));
    if (c->timing) {
        HIPCHK(hipEventRecord(c->ev_pool[ev + 1], st));
        c->pending.push_back({ev, ev + 1});
        c->pending_work.push_back((int64_t)m1 + m2);
    }
    return ST_OK;
}

/* BeamBeam6D_PIC, simtrack.cpp:23380-23615 */
static int pic_6d(st_ctx *c, ss_work &w, double Np1, int Mp1, double *d_x1, double Np2, int Mp2, double *d_x2,
                  double gamma1, double gamma2, double bbscale1, double bbscale2, int Ns1, int Ns2, double *Z1o,
                  double *Z2o)
{
    HIPCHK(st_ss_sort_by_z(d_x1, Mp1, w.s, c->stream));
    HIPCHK(st_ss_sort_by_z(d_x2, Mp2, w.s, c->stream));
    const int mc1 = Mp1 / Ns1, mc2 = Mp2 / Ns2;
    std::vector<int> st1(Ns1), ct1(Ns1, mc1), st2(Ns2), ct2(Ns2, mc2);
    for (int i = 0; i < Ns1; i++) st1[i] = i * mc1;
    for (int i = 0; i < Ns2; i++) st2[i] = i * mc2;
    int rc = ss_slice_centres(c, w, d_x1, nullptr, st1, ct1, w.Z);
    if (!rc) rc = ss_slice_centres(c, w, d_x2, nullptr, st2, ct2, w.Z + Ns1);
    if (rc) return rc;
    std::vector<double> Z1(Ns1), Z2(Ns2);
    HIPCHK(hipMemcpyAsync(Z1.data(), w.Z, Ns1 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(Z2.data(), w.Z + Ns1, Ns2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (Z1o) std::copy(Z1.begin(), Z1.end(), Z1o);
    if (Z2o) std::copy(Z2.begin(), Z2.end(), Z2o);
    std::vector<int> I1, I2;
    std::vector<double> S;
    ss_schedule(Z1, Z2, I1, I2, S);
    return ss_runs(c, w, I1, I2, Ns1, Ns2, [&](size_t i, int l, hipStream_t st) {
        const int a = I1[i] * mc1, b = I2[i] * mc2;
        HIPCHK(st_pic_drift(d_x1, a, mc1, S[i], 1, st));
        HIPCHK(st_pic_drift(d_x2, b, mc2, S[i], -1, st));
        int rc2 = pic_pair(c, w, d_x1, a, mc1, d_x2, b, mc2, Np1 / Ns1, Np2 / Ns2, gamma1, gamma2, bbscale1,
                           bbscale2, l);
        if (rc2) return rc2;
        HIPCHK(st_pic_drift(d_x1, a, mc1, S[i], -1, st));
        HIPCHK(st_pic_drift(d_x2, b, mc2, S[i], 1, st));
        return (int)ST_OK;
    });
}

extern "C" int st_bb_pic_4d(st_ctx *c, double Np1, int Mp1, double *d_x1, double Np2, int Mp2, double *d_x2,
                            double gamma1, double gamma2, double bbscale1, double bbscale2)
{
    int rc = ss_check(c, Mp1, d_x1, Mp2, d_x2, "st_bb_pic_4d");
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    ss_work w;
    if ((rc = ss_prepare(c, Mp1, Mp2, 1, 1, false, w, std::max(std::max(Mp1, Mp2), 1)))) return rc;
    if ((rc = pic_pair(c, w, d_x1, 0, Mp1, d_x2, 0, Mp2, Np1, Np2, gamma1, gamma2, bbscale1, bbscale2))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return ST_OK;
}

extern "C" int st_bb_pic_6d(st_ctx *c, double Np1, int Mp1, double *d_x1, double Np2, int Mp2, double *d_x2,
                            double gamma1, double gamma2, double bbscale1, double bbscale2, int Nslice1, int Nslice2,
                            double *Z1, double *Z2)
{
    int rc = ss_check(c, Mp1, d_x1, Mp2, d_x2, "st_bb_pic_6d");
    if (rc) return rc;
    if (Nslice1 < 1 || Nslice2 < 1) return fail(ST_ERR_PARAM, "st_bb_pic_6d: Nslice must be >= 1");
    HIPCHK(hipSetDevice(c->device));
    ss_work w;
    const int pr = std::max(std::max(Mp1 / Nslice1, Mp2 / Nslice2), 1);
    if ((rc = ss_prepare(c, Mp1, Mp2, Nslice1, Nslice2, false, w, pr))) return rc;
    if ((rc = pic_6d(c, w, Np1, Mp1, d_x1, Np2, Mp2, d_x2, gamma1, gamma2, bbscale1, bbscale2, Nslice1, Nslice2, Z1,
                     Z2)))
        return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return ST_OK;
}

extern "C" int st_bb_pic_4d_host(st_ctx *c, double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2,
                                 double gamma1, double gamma2, double bbscale1, double bbscale2)
{
    int rc = ss_check(c, Mp1, x1, Mp2, x2, "st_bb_pic_4d_host");
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    ss_work w;
    if ((rc = ss_prepare(c, Mp1, Mp2, 1, 1, true, w, std::max(std::max(Mp1, Mp2), 1))) ||
        (rc = ss_host_in(c, w, Mp1, x1, Mp2, x2)) ||
        (rc = pic_pair(c, w, w.hx[0], 0, Mp1, w.hx[1], 0, Mp2, Np1, Np2, gamma1, gamma2, bbscale1, bbscale2)))
        return rc;
    return ss_host_out(c, w, Mp1, x1, Mp2, x2);
}

extern "C" int st_bb_pic_6d_host(st_ctx *c, double Np1, int Mp1, double *x1, double Np2, int Mp2, double *x2,
                                 double gamma1, double gamma2, double bbscale1, double bbscale2, int Nslice1,
                                 int Nslice2, double *Z1, double *Z2)
{
    int rc = ss_check(c, Mp1, x1, Mp2, x2, "st_bb_pic_6d_host");
    if (rc) return rc;
    if (Nslice1 < 1 || Nslice2 < 1) return fail(ST_ERR_PARAM, "st_bb_pic_6d_host: Nslice must be >= 1");
    HIPCHK(hipSetDevice(c->device));
    ss_work w;
    const int pr = std::max(std::max(Mp1 / Nslice1, Mp2 / Nslice2), 1);
    if ((rc = ss_prepare(c, Mp1, Mp2, Nslice1, Nslice2, true, w, pr)) || (rc = ss_host_in(c, w, Mp1, x1, Mp2, x2)) ||
        (rc = pic_6d(c, w, Np1, Mp1, w.hx[0], Np2, Mp2, w.hx[1], gamma1, gamma2, bbscale1, bbscale2, Nslice1,
                     Nslice2, Z1, Z2)))
        return rc;
    return ss_host_out(c, w, Mp1, x1, Mp2, x2);
}

/* Zboundary_slices_Gaussian, simtrack.cpp:3157-3168 (zb: Nslice + 1 doubles) */
extern "C" int st_zboundary_slices_gaussian(double *zb, int Nslice, double sigmal)
{
    if (!zb || Nslice < 1) return fail(ST_ERR_ARG, "st_zboundary_slices_gaussian: bad argument");
    zb[0] = sigmal * 20;
    zb[Nslice] = -sigmal * 20;
    for (int i = 1; i <= Nslice - 1; i++) zb[i] = -st_ugaussian_pinv(i * 1.0 / Nslice) * sigmal;
    return ST_OK;
}
