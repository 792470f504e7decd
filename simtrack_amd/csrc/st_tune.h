/*
 * st_tune.h -- launch interface of the FineTuneFinder kernel (st_tune.hip).
 */
#ifndef ST_TUNE_H
#define ST_TUNE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

struct st_tune_args {
    const double *x;    /* signal j, sample i: x[j*stride_sig + i*stride]   */
    const double *px;   /* same layout, or null for the one-signal form      */
    int64_t stride_sig;
    int64_t stride;
    int n;              /* samples per signal                                */
    int nsig;           /* signals                                           */
    int max_cand;       /* LDS slots for one level's candidate tunes         */
    double lower, upper;
    double *tune;       /* nsig results                                      */
    const double *window; /* n weights applied to every signal, or null     */
};

hipError_t st_launch_fine_tune(const st_tune_args &a, hipStream_t stream);
hipError_t st_launch_features(const double *x0, int ncol, const double *tunes, int n, double bx, double ax,
                              double by, double ay, float *feat, hipStream_t stream);

#endif
