/*
 * st_tune.hip -- FineTuneFinder on the GPU: the peak of the unwindowed DFT
 * power of many turn-by-turn signals at once.
 *
 * FineTuneFinder (simtrack.cpp:2288-2319, and the x/px form cpp:2321-2352)
 * scans tunes from tune_lower in steps of 1e-3 while tune < tune_upper,
 * keeps the first tune of maximal power sum_r^2 + sum_i^2, then rescans
 * +-step around it with step/10, until the step drops to 1e-12.
 *
 * One wavefront per signal.  The samples are staged once in LDS; within a
 * level the candidate tunes (generated by the reference's own repeated
 * additions, so every candidate carries the same bits) are spread over the
 * 64 lanes, each lane summing its candidates over all samples in the
 * reference's order; the level's peak is the first candidate of maximal
 * power, found by a cross-lane reduction.  An optional window w[i]
 * multiplies the samples as they are staged (a Hann window makes the peak
 * search NAFF-like: leakage falls from O(1/N) to O(1/N^3)).  Elementary functions are the
 * project libm, as in the tracking kernels.
 */
#include <hip/hip_runtime.h>

#include <stdint.h>

#define ST_FN __device__ __forceinline__
#define ST_FN_COLD __device__ __noinline__
#define ST_CONST_TABLE static __device__ __constant__ const
#define ST_FMA(a, b, c) __builtin_fma((a), (b), (c))
#define ST_RINT(x) __builtin_rint(x)
#define ST_SQRT(x) __builtin_sqrt(x)
#include "st_libm.h"
#include "st_tune.h"

namespace {

const double REF_PI = 3.14159265358979323; /* PI, simtrack.cpp:23 */

__global__ __launch_bounds__(64) void k_fine_tune(st_tune_args a)
{
    extern __shared__ double lds[];
    const int j = blockIdx.x;
    const int lane = threadIdx.x;
    const int n = a.n;
    double *xs = lds;                     /* n samples                  */
    double *ps = lds + n;                 /* n samples of px (x/px form) */
    double *cand = lds + (a.px ? 2 * n : n); /* candidate tunes           */
    __shared__ double s_amp[64];
    __shared__ int s_idx[64];
    __shared__ int s_ncand;
    const double *src = a.x + (int64_t)j * a.stride_sig;
    const double *psrc = a.px ? a.px + (int64_t)j * a.stride_sig : nullptr;
    for (int i = lane; i < n; i += 64) {
        xs[i] = src[(int64_t)i * a.stride];
        if (psrc) ps[i] = psrc[(int64_t)i * a.stride];
        if (a.window) {
            /* the windowed form: the same search on w[i] * x[i] */
            xs[i] = a.window[i] * xs[i];
            if (psrc) ps[i] = a.window[i] * ps[i];
        }
    }
    double lower = a.lower, upper = a.upper, step = 0.001, peak = lower;
    do {
        if (lane == 0) {
            /* tune = tune_lower; do { ...; tune = tune + tune_step; } while (tune < tune_upper) */
            int m = 0;
            double t = lower;
            do {
                if (m < a.max_cand) cand[m] = t;
                ++m;
                t = t + step;
            } while (t < upper);
            s_ncand = m;
        }
        __syncthreads();
        const int m = s_ncand;
        if (m > a.max_cand) {
            /* the host sizes max_cand from the range; flag and stop */
            if (lane == 0) a.tune[j] = __builtin_nan("");
            return;
        }
        double best = 0.;
        int bidx = -1;
        for (int c = lane; c < m; c += 64) {
            const double tune = cand[c];
            double sumr = 0., sumi = 0.;
            if (psrc) {
                for (int i = 0; i < n; i++) {
                    double sn, cs;
                    st_sincos(-2 * REF_PI * i * tune, &sn, &cs);
                    sumr = sumr + xs[i] * cs - ps[i] * sn;
                    sumi = sumi + xs[i] * sn + ps[i] * cs;
                }
            } else {
                for (int i = 0; i < n; i++) {
                    double sn, cs;
                    st_sincos(-2 * REF_PI * i * tune, &sn, &cs);
                    sumr = sumr + xs[i] * cs;
                    sumi = sumi + xs[i] * sn;
                }
            }
            double amp = sumr * sumr + sumi * sumi;
            if (amp > best) { /* first maximum along this lane's candidates */
                best = amp;
                bidx = c;
            }
        }
        s_amp[lane] = best;
        s_idx[lane] = bidx;
        __syncthreads();
        if (lane == 0) {
            /* the first candidate of maximal power; peakamp starts at 0, so
               candidates with amp <= 0 (or NaN) never win */
            double pa = 0.;
            int pi = -1;
            for (int l = 0; l < 64; ++l) {
                int ci = s_idx[l];
                if (ci < 0) continue;
                double v = s_amp[l];
                if (v > pa || (v == pa && pi >= 0 && ci < pi)) {
                    pa = v;
                    pi = ci;
                }
            }
            s_idx[0] = pi;
        }
        __syncthreads();
        const int pi = s_idx[0];
        peak = pi >= 0 ? cand[pi] : lower;
        __syncthreads();
        lower = peak - step;
        upper = peak + step;
        step = step / 10.;
    } while (step > 1.0e-12);
    if (lane == 0) a.tune[j] = peak;
}

/*
 * Surrogate features of one particle (FP32): Courant-Snyder actions Jx, Jy
 * and angles (cos, sin) of the initial coordinates with the Twiss
 * parameters of the start point, the tunes of two turn-by-turn windows
 * (nux1, nuy1, nux2, nuy2) and log10 |tune change| of both planes.
 */
__global__ void k_features(const double *x0, int ncol, const double *tunes, int n, double bx, double ax, double by,
                           double ay, float *feat)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *x = x0 + (int64_t)i * ncol;
    const double *q = tunes + (int64_t)i * 4;
    double sbx = sqrt(bx), sby = sqrt(by);
    double X = x[0] / sbx, PX = (ax * x[0] + bx * x[1]) / sbx;
    double Y = x[2] / sby, PY = (ay * x[2] + by * x[3]) / sby;
    double jx = 0.5 * (X * X + PX * PX), jy = 0.5 * (Y * Y + PY * PY);
    double fx = atan2(-PX, X), fy = atan2(-PY, Y);
    float *o = feat + (int64_t)i * 12;
    o[0] = (float)jx;
    o[1] = (float)jy;
    o[2] = (float)cos(fx);
    o[3] = (float)sin(fx);
    o[4] = (float)cos(fy);
    o[5] = (float)sin(fy);
    o[6] = (float)q[0];
    o[7] = (float)q[1];
    o[8] = (float)q[2];
    o[9] = (float)q[3];
    o[10] = (float)log10(fabs(q[2] - q[0]) + 1e-12);
    o[11] = (float)log10(fabs(q[3] - q[1]) + 1e-12);
}

} // namespace

hipError_t st_launch_features(const double *x0, int ncol, const double *tunes, int n, double bx, double ax,
                              double by, double ay, float *feat, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, x0, ncol, tunes, n, bx,
                       ax, by, ay, feat);
    return hipGetLastError();
}

hipError_t st_launch_fine_tune(const st_tune_args &a, hipStream_t stream)
{
    if (a.nsig <= 0) return hipSuccess;
    size_t lds = sizeof(double) * ((size_t)a.n * (a.px ? 2 : 1) + (size_t)a.max_cand);
    hipLaunchKernelGGL(k_fine_tune, dim3((unsigned)a.nsig), dim3(64), lds, stream, a);
    return hipGetLastError();
}
