/*
 * st_scan.cpp -- drivers built on tracking: FineTuneFinder, Cal_Orbit_Num,
 * the one-turn and element maps, and the turn-by-turn scans (frequency map,
 * Lyapunov, action diffusion).
 */
#include "st_capi_internal.h"

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

/* ----------------------------------------------------- spin drivers */

/* Cal_Spin_Orbit (simtrack.cpp:20510-20572): Cal_Orbit_Num, then Newton on
   the spin's (Sx, Ss) at the start with one-turn sPass maps of the line as
   given (RF on), no loss test; the three particles of an iteration (x0 and
   the two displaced spins) are one batch.  n0: one more pass, the
   normalised spin after every element. */
extern "C" int st_spin_orbit(st_ctx *c, const st_globals *gp, const st_element *rec, int n, double deltap,
                             double x0[9], double *n0, int *iterations)
{
    if (!c || !gp || n <= 0 || !rec || !x0) return fail(ST_ERR_ARG, "st_spin_orbit: bad argument");
    double x_co[6], x_end[6];
    int rc = st_closed_orbit(c, gp, rec, n, deltap, x_co, x_end, nullptr);
    if (rc) return rc;
    st_table *t = nullptr;
    rc = st_table_build(c, gp, rec, n, &t);
    if (rc) return rc;
    std::unique_ptr<st_table, int (*)(st_table *)> guard(t, st_table_destroy);
    const double d = 1.0e-06, codeps = 1e-12;
    const int Max_iter = 20;
    double mat[2][2], dx[2], chi, xb[3 * 9];
    int st[3], lt[3], lp[3];
    for (int i = 0; i < 6; i++) x0[i] = x_end[i];
    x0[6] = 0.;
    x0[7] = 1.0;
    x0[8] = 0.;
    int iter = 0;
    do {
        iter++;
        for (int q = 0; q < 3; ++q) {
            for (int j = 0; j < 9; j++) xb[q * 9 + j] = x0[j];
            st[q] = 1;
            lt[q] = lp[q] = 0;
        }
        xb[9 + 6] = x0[6] + d;
        xb[18 + 8] = x0[8] + d;
        rc = track_host(t, xb, 3, 1, st, lt, lp, true, false, nullptr, -1, true);
        if (rc) return rc;
        const double *x01 = xb, *x1 = xb + 9, *x2 = xb + 18;
        mat[0][0] = (x1[6] - x01[6]) / d;
        mat[1][0] = (x1[8] - x01[8]) / d;
        mat[0][1] = (x2[6] - x01[6]) / d;
        mat[1][1] = (x2[8] - x01[8]) / d;
        dx[0] = x0[6] - x01[6];
        dx[1] = x0[8] - x01[8];
        chi = dx[0] * dx[0] + dx[1] * dx[1];
        chi = std::sqrt(chi / 2.);
        if (chi > codeps) {
            for (int i = 0; i < 2; i++) mat[i][i] = mat[i][i] - 1.0000001;
            mat_inv_ref(&mat[0][0], 2); /* unchecked, as the reference's (cpp:20546) */
            x0[6] = x0[6] + mat[0][0] * dx[0] + mat[0][1] * dx[1];
            x0[8] = x0[8] + mat[1][0] * dx[0] + mat[1][1] * dx[1];
            x0[7] = std::sqrt(1. - x0[6] * x0[6] - x0[8] * x0[8]);
        }
    } while (chi > codeps && iter < Max_iter);
    if (iterations) *iterations = iter;
    /* the reference's test, cpp:20554 */
    if (iter == Max_iter - 1) return fail(ST_ERR_PARAM, "spin closed orbit: failed to find spin COD");
    if (n0) {
        double xs[9];
        for (int j = 0; j < 9; j++) xs[j] = x0[j];
        for (int k = 0; k < n; ++k) {
            rc = st_element_pass(t, k, xs, 1, 1);
            if (rc) return rc;
            const double a = std::sqrt(xs[6] * xs[6] + xs[7] * xs[7] + xs[8] * xs[8]);
            for (int i = 0; i < 3; i++) n0[3 * k + i] = xs[6 + i] / a;
        }
    }
    return ST_OK;
}

/* `nturn` turns of bare spin tracking (no loss test) of np particles,
   out[(p*nturn + j)*9 + k] = coordinate k of particle p after turn j + 1 */
static int spin_turns(st_table *t, double *x, int np, int nturn, std::vector<double> &out)
{
    std::vector<int> st(np, 1), lt(np, 0), lp(np, 0);
    std::vector<double> tbt((size_t)np * (size_t)nturn * 9, 0.);
    int rc = track_host(t, x, np, nturn, st.data(), lt.data(), lp.data(), true, false, tbt.data(), -1, true);
    if (rc) return rc;
    /* the capture holds the state at the start of each turn: shift by one */
    out.assign(tbt.size(), 0.);
    for (int p = 0; p < np; ++p) {
        const size_t b = (size_t)p * (size_t)nturn * 9;
        for (int j = 0; j + 1 < nturn; ++j)
            for (int k = 0; k < 9; ++k) out[b + (size_t)j * 9 + k] = tbt[b + (size_t)(j + 1) * 9 + k];
        for (int k = 0; k < 9; ++k) out[b + (size_t)(nturn - 1) * 9 + k] = x[(size_t)p * 9 + k];
    }
    return ST_OK;
}

/* Cal_Spin_Tune (simtrack.cpp:20574-20601): Cal_Spin_Orbit; a spin m
   perpendicular to the last element's n0, with the reference's grouping
   n0^2 / n1 * n1 (cpp:20587); 500 turns of sPass from Cell[Ncell-1]->X;
   FineTuneFinder of Sx after each turn on [0.001, 0.499]. */
extern "C" int st_spin_tune(st_ctx *c, const st_globals *gp, const st_element *rec, int n, double deltap,
                            double *tune, double n0_end[3])
{
    if (!c || !gp || n <= 0 || !rec || !tune) return fail(ST_ERR_ARG, "st_spin_tune: bad argument");
    const int Nturn = 500;
    double x0[9];
    std::vector<double> n0((size_t)n * 3);
    int rc = st_spin_orbit(c, gp, rec, n, deltap, x0, n0.data(), nullptr);
    if (rc) return rc;
    double nv[3], m[3];
    for (int i = 0; i < 3; i++) nv[i] = n0[3 * (size_t)(n - 1) + i];
    if (n0_end)
        for (int i = 0; i < 3; i++) n0_end[i] = nv[i];
    m[2] = 0.;
    m[0] = 1.0 / std::sqrt(1. + nv[0] * nv[0] / nv[1] * nv[1]);
    m[1] = -nv[0] * m[0] / nv[1];
    double x[9];
    for (int i = 0; i < 6; i++) x[i] = x0[i];
    x[6] = m[0];
    x[7] = m[1];
    x[8] = m[2];
    st_table *t = nullptr;
    rc = st_table_build(c, gp, rec, n, &t);
    if (rc) return rc;
    std::unique_ptr<st_table, int (*)(st_table *)> guard(t, st_table_destroy);
    std::vector<double> turns;
    rc = spin_turns(t, x, 1, Nturn, turns);
    if (rc) return rc;
    std::vector<double> sample(Nturn);
    for (int j = 0; j < Nturn; j++) sample[j] = turns[(size_t)j * 9 + 6];
    return st_fine_tune(c, sample.data(), nullptr, Nturn, 1, 0.001, 0.499, tune);
}

/* Set_Initial_Spin (simtrack.cpp:20603-20635): the orbit of x with unit
   spins along x, y and s (one batch), nturn turns of sPass each, their
   projections on n0 summed after every turn in the reference's order; x's
   spin becomes the normalised sums. */
extern "C" int st_set_initial_spin(st_ctx *c, const st_globals *gp, const st_element *rec, int n, int nturn,
                                   const double n0[3], double x[9])
{
    if (!c || !gp || n <= 0 || !rec || nturn < 0 || !n0 || !x)
        return fail(ST_ERR_ARG, "st_set_initial_spin: bad argument");
    st_table *t = nullptr;
    int rc = st_table_build(c, gp, rec, n, &t);
    if (rc) return rc;
    std::unique_ptr<st_table, int (*)(st_table *)> guard(t, st_table_destroy);
    double xb[3 * 9];
    for (int q = 0; q < 3; ++q) {
        for (int j = 0; j < 6; j++) xb[q * 9 + j] = x[j];
        for (int k = 0; k < 3; k++) xb[q * 9 + 6 + k] = (k == q) ? 1. : 0.;
    }
    double bq[3] = {0., 0., 0.};
    if (nturn > 0) {
        std::vector<double> turns;
        rc = spin_turns(t, xb, 3, nturn, turns);
        if (rc) return rc;
        for (int q = 0; q < 3; ++q)
            for (int j = 0; j < nturn; j++) {
                const double *s = turns.data() + ((size_t)q * (size_t)nturn + (size_t)j) * 9 + 6;
                bq[q] = bq[q] + s[0] * n0[0] + s[1] * n0[1] + s[2] * n0[2];
            }
    }
    double b = std::sqrt(bq[0] * bq[0] + bq[1] * bq[1] + bq[2] * bq[2]);
    double b1 = bq[0] / b, b2 = bq[1] / b, b3 = bq[2] / b;
    x[6] = b1;
    x[7] = b2;
    x[8] = b3;
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
