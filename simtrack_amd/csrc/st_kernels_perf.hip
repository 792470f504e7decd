/*
 * st_kernels_perf.hip -- the perf-mode tracking kernels: st_kernels.hip
 * compiled a second time with ST_PERF_MODE, in namespace st_perf, and with
 * floating-point contraction on (build.py).  The maps are the same source;
 * ST_PERF_MODE changes their arithmetic policy (st_maps.h): a division by a
 * prepared divisor is one multiplication by its refined reciprocal, square
 * roots skip the range bookkeeping, and a * b + c may become one FMA.  The
 * results are therefore not bit-equal to the reference's operations; they
 * differ by rounding and are reached only through ST_TRACK_PERF.
 */
#define ST_PERF_MODE 1
#include "st_kernels.hip"
