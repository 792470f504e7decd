/*
 * st_da.cpp -- dynamic-aperture drivers on top of the batched tracker.
 *
 * The reference's drivers call Track() once per particle and amplitude:
 *   Track_DA (simtrack.cpp:21389), Track_DA1 (cpp:21443) and
 *   Track_DA_Radial (cpp:21682) bisect the amplitude with twin particles,
 *   Track_DA2 (cpp:21497) and Track_DA3 (cpp:21535) with one particle;
 *   Track_DA_Uniform (cpp:21574), Track_DA_Radial_Uniform (cpp:21736) and
 *   Track_DA_Radial_Uniform_Test (Example2/DA_cal.cpp:61) step it uniformly
 *   and stop at the first loss.
 * Here every particle those loops could track is launched in one batch and
 * the host replays the loops on the outcomes.  Particles are independent, so
 * each outcome is exactly the one the sequential loop would have obtained;
 * only the order of work changes.
 *
 * Bisection: the loop visits the midpoints of dyadic sub-intervals of
 * [0, sigma_max] (all exactly representable), at most `depth` of them per
 * angle, where depth is the number of halvings until the interval is
 * <= 0.1.  The whole tree (2^depth - 1 midpoints, each with its twin at
 * +0.05 for the twin drivers) is tracked speculatively -- enough parallel
 * work to fill the GPU,
 * in one launch, instead of depth dependent launches of two particles.
 *
 * Initial conditions use the C library's cos/sin, as the reference drivers
 * do on the same host.
 */
#include <cmath>
#include <cstring>
#include <vector>

#include "simtrack_amd.h"

extern "C" int st_internal_fail(int code, const char *msg);

namespace {

const double REF_PI = 3.14159265358979323; /* PI, simtrack.cpp:23 */

void set_particle(double *x, double sigma, double angle, double sigmax0, double sigmay0, double deltap0)
{
    x[0] = sigma * std::cos(angle) * sigmax0;
    x[1] = 0.;
    x[2] = sigma * std::sin(angle) * sigmay0;
    x[3] = 0.;
    x[4] = 0.;
    x[5] = deltap0;
}

/* number of midpoints the reference loop tests for a given sigma_max */
int bisection_depth(double sigma_max)
{
    double s = 0., e = sigma_max;
    int d = 0;
    do {
        double sigma = (s + e) / 2.0;
        e = sigma; /* the width evolves identically on either branch */
        ++d;
    } while (std::fabs(e - s) > 0.1 && d < 60);
    return d;
}

int track_batch(st_table *t, std::vector<double> &x, int np, int nturn, int radial, std::vector<int> &stable)
{
    stable.assign((size_t)np, 1);
    std::vector<int> lt((size_t)np, 0), lp((size_t)np, 0);
    if (radial) return st_track_radial(t, x.data(), np, nturn, stable.data(), lt.data(), lp.data());
    return st_track(t, x.data(), np, nturn, stable.data(), lt.data(), lp.data());
}

} // namespace

extern "C" int st_da_bisect(st_table *t, int nturn, double deltap0, double sigmax0, double sigmay0,
                            const double *angles, int nangle, double sigma_max, int radial, double *sigma_out)
{
    if (!t || nturn < 0 || nangle < 0 || (nangle > 0 && (!angles || !sigma_out)) || !(sigma_max > 0.1))
        return st_internal_fail(ST_ERR_ARG, "st_da_bisect: bad argument");
    if (nangle == 0) return ST_OK;
    const int depth = bisection_depth(sigma_max);
    if (depth > 20) return st_internal_fail(ST_ERR_ARG, "st_da_bisect: sigma_max too large");
    const int nodes = (1 << depth) - 1; /* heap order: node k has children 2k+1, 2k+2 */
    std::vector<double> mid((size_t)nodes);
    {
        /* midpoint of each tree node, built by the loop's own arithmetic */
        std::vector<double> lo((size_t)nodes), hi((size_t)nodes);
        lo[0] = 0.;
        hi[0] = sigma_max;
        for (int k = 0; k < nodes; ++k) {
            mid[k] = (lo[k] + hi[k]) / 2.0;
            if (2 * k + 2 < nodes) {
                lo[2 * k + 1] = lo[k];    /* unstable: sigma_e = sigma */
                hi[2 * k + 1] = mid[k];
                lo[2 * k + 2] = mid[k];   /* stable: sigma_s = sigma   */
                hi[2 * k + 2] = hi[k];
            }
        }
    }
    /* twins at sigma and sigma + 0.05 (Track_DA, cpp:21410-21426), or one
       particle at sigma (Track_DA2, cpp:21513-21522) */
    const int per = (radial & ST_DA_SINGLE) ? 1 : 2;
    const int rad = radial & ST_DA_RADIAL;
    const size_t np = (size_t)nangle * (size_t)nodes * per;
    if (np > (size_t)1 << 30) return st_internal_fail(ST_ERR_ARG, "st_da_bisect: too many particles");
    std::vector<double> x(np * 6);
    for (int a = 0; a < nangle; ++a)
        for (int k = 0; k < nodes; ++k) {
            size_t p = ((size_t)a * nodes + k) * per;
            double sigma1 = mid[k], sigma2 = mid[k] + 0.05;
            set_particle(&x[p * 6], sigma1, angles[a], sigmax0, sigmay0, deltap0);
            if (per == 2) set_particle(&x[(p + 1) * 6], sigma2, angles[a], sigmax0, sigmay0, deltap0);
        }
    std::vector<int> stable;
    int rc = track_batch(t, x, (int)np, nturn, rad, stable);
    if (rc) return rc;
    for (int a = 0; a < nangle; ++a) {
        /* replay of the do-while loop of Track_DA (cpp:21403-21435) */
        double sigma_s = 0., sigma_e = sigma_max, sigma = 0.;
        int k = 0;
        do {
            sigma = (sigma_s + sigma_e) / 2.0;
            size_t p = ((size_t)a * nodes + k) * per;
            int st = per == 1 ? stable[p] : ((stable[p] * stable[p + 1] == 1) ? 1 : 0);
            if (st == 0) sigma_e = sigma;
            if (st == 1) sigma_s = sigma;
            k = 2 * k + (st ? 2 : 1);
        } while (std::fabs(sigma_e - sigma_s) > 0.1 && k < nodes);
        sigma_out[a] = sigma;
    }
    return ST_OK;
}

extern "C" int st_da_uniform(st_table *t, int nturn, double deltap0, double angle, double sigmax0, double sigmay0,
                             double sigma0, double dsigma, int n, int radial, double *sigma_out, int *stable_out,
                             int *n_scanned)
{
    if (!t || nturn < 0 || n < 0 || (n > 0 && (!sigma_out || !stable_out)) || !n_scanned)
        return st_internal_fail(ST_ERR_ARG, "st_da_uniform: bad argument");
    *n_scanned = 0;
    if (n == 0) return ST_OK;
    std::vector<double> x((size_t)n * 6);
    for (int i = 0; i < n; ++i) {
        double sigma = sigma0 + i * dsigma; /* e.g. 1. + i*0.2, cpp:21587 */
        sigma_out[i] = sigma;
        set_particle(&x[(size_t)i * 6], sigma, angle, sigmax0, sigmay0, deltap0);
    }
    std::vector<int> stable;
    int rc = track_batch(t, x, n, nturn, radial, stable);
    if (rc) return rc;
    int m = n;
    for (int i = 0; i < n; ++i) {
        stable_out[i] = stable[i];
        if (stable[i] == 0 && m == n) m = i + 1; /* the reference stops after this line */
    }
    *n_scanned = m;
    return ST_OK;
}

extern "C" int st_da_angles(int variant, double *angles, int *nangle, double *sigma_max)
{
    /* the angle sets and bisection ranges of the reference's drivers */
    if (!angles || !nangle || !sigma_max) return st_internal_fail(ST_ERR_ARG, "st_da_angles: bad argument");
    switch (variant) {
    case ST_DA_TRACK_DA: /* cpp:21399-21400 */
    case ST_DA_TRACK_DA_RADIAL: /* cpp:21692-21693 */
        *nangle = 10;
        for (int i = 0; i < 10; ++i) angles[i] = (90 / 11.) * (i + 1) * REF_PI / 180;
        *sigma_max = variant == ST_DA_TRACK_DA ? 50.0 : 15.0;
        return ST_OK;
    case ST_DA_TRACK_DA1: /* cpp:21453-21455 */
        *nangle = 5;
        for (int i = 0; i < 5; ++i) angles[i] = 15 * (i + 1) * REF_PI / 180;
        *sigma_max = 50.0;
        return ST_OK;
    case ST_DA_TRACK_DA2: /* cpp:21507-21510 */
        *nangle = 10;
        for (int i = 0; i < 10; ++i) angles[i] = (90 / 11.) * (i + 1) * REF_PI / 180;
        *sigma_max = 50.0;
        return ST_OK;
    case ST_DA_TRACK_DA3: /* cpp:21545-21550 */
        *nangle = 5;
        for (int i = 0; i < 5; ++i) angles[i] = 15 * (i + 1) * REF_PI / 180;
        *sigma_max = 30.0;
        return ST_OK;
    default:
        return st_internal_fail(ST_ERR_ARG, "st_da_angles: unknown variant");
    }
}
