/*
 * st_ssbb.hip -- strong-strong soft-Gaussian beam-beam on gfx950.
 *
 * Two macroparticle beams (rows of 6 doubles, the reference's layout) meet
 * slice pair after slice pair.  A pair is two launches:
 *   k_ss_partial  deterministic partial sums (x, y, x^2, y^2) of both
 *                 slices' drifted coordinates, one block per fixed chunk;
 *   k_ss_kick     every block folds the other beam's partials in a fixed
 *                 order (mean, rms), then each row is drifted by S, kicked
 *                 by BB4D about the other slice's centroid and drifted back
 *                 with its new momentum (simtrack.cpp:23035-23110).
 * Sums are formed in a fixed order that does not depend on timing, so a
 * run repeats bit for bit; the order is not the reference's sequential one,
 * so moments agree with it to rounding.
 *
 * Rows of a slice are given by a row list (idx) and a range in it; the
 * z-sorted variant sorts rows by z (hipcub radix sort, descending) and
 * gathers them, the boundary variant sorts row numbers by slice number
 * (stable, so each slice keeps row order).
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <stdint.h>

#define ST_FN __device__ __forceinline__
#define ST_FN_COLD __device__ __noinline__
#define ST_CONST_TABLE static __device__ __constant__ const
#define ST_FMA(a, b, c) __builtin_fma((a), (b), (c))
#define ST_RINT(x) __builtin_rint(x)
#include "st_fastdiv.h"
#define ST_SQRT(x) st_fsqrt(x)
#include "st_libm.h"
#include "st_rng.h"
#include "st_maps.h"
#include "st_ssbb.h"

#define SS_BLK 256

/* row r of segment (idx, start) */
ST_DEV int ss_row(const int *idx, int start, int k) { return idx ? idx[start + k] : start + k; }

/* x, y after the drift of beam b (beam 0: + p S, beam 1: - p S) */
ST_DEV void ss_drifted(const double *row, int b, int drift, double S, double &x, double &y)
{
    if (!drift) {
        x = row[0];
        y = row[2];
    } else if (b == 0) {
        x = row[0] + row[1] * S;
        y = row[2] + row[3] * S;
    } else {
        x = row[0] - row[1] * S;
        y = row[2] - row[3] * S;
    }
}

/* fixed-order block reduction of 4 sums held by each thread */
ST_DEV void ss_block_sum4(double v[4], double (*sh)[SS_BLK])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; k++) sh[k][t] = v[k];
    __syncthreads();
    for (int w = SS_BLK / 2; w > 0; w >>= 1) {
        if (t < w) {
#pragma unroll
            for (int k = 0; k < 4; k++) sh[k][t] = sh[k][t] + sh[k][t + w];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = sh[k][0];
    __syncthreads();
}

__global__ __launch_bounds__(SS_BLK) void k_ss_partial(ss_pair_args a)
{
    __shared__ double sh[4][SS_BLK];
    const int b = blockIdx.x < (unsigned)a.nb[0] ? 0 : 1;
    const int chunk = blockIdx.x - (b ? a.nb[0] : 0);
    const int lo = chunk * SS_CHUNK;
    const int hi = min(lo + SS_CHUNK, a.count[b]);
    double v[4] = {0., 0., 0., 0.};
    for (int k = lo + (int)threadIdx.x; k < hi; k += SS_BLK) {
        const double *row = a.x[b] + (size_t)ss_row(a.idx[b], a.start[b], k) * 6;
        double x, y;
        ss_drifted(row, b, a.drift, a.S, x, y);
        v[0] = v[0] + x;
        v[1] = v[1] + y;
        v[2] = v[2] + x * x;
        v[3] = v[3] + y * y;
    }
    ss_block_sum4(v, sh);
    if (threadIdx.x < 4) a.part[(size_t)blockIdx.x * 4 + threadIdx.x] = v[threadIdx.x];
}

/* Cal_Mean_RMS of beam b from its partials (cpp:724-742 formulas) */
ST_DEV void ss_moments(const ss_pair_args &a, int b, double (*sh)[SS_BLK], double &mx, double &my, double &sx,
                       double &sy)
{
    const int base = b ? a.nb[0] : 0, nb = a.nb[b];
    double v[4] = {0., 0., 0., 0.};
    for (int k = threadIdx.x; k < nb; k += SS_BLK) {
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = v[q] + a.part[(size_t)(base + k) * 4 + q];
    }
    ss_block_sum4(v, sh);
    const int n = a.count[b];
    mx = v[0] / n;
    my = v[1] / n;
    sx = __builtin_sqrt(v[2] / n - mx * mx);
    sy = __builtin_sqrt(v[3] / n - my * my);
}

__global__ __launch_bounds__(SS_BLK) void k_ss_kick(ss_pair_args a, int kb0)
{
    __shared__ double sh[4][SS_BLK];
    const int b = blockIdx.x < (unsigned)kb0 ? 0 : 1;
    const int o = 1 - b;
    double mx, my, sx, sy;
    ss_moments(a, o, sh, mx, my, sx, sy);
    const int k = (blockIdx.x - (b ? kb0 : 0)) * SS_BLK + threadIdx.x;
    if (k >= a.count[b]) return;
    double *row = a.x[b] + (size_t)ss_row(a.idx[b], a.start[b], k) * 6;
    double x, y;
    ss_drifted(row, b, a.drift, a.S, x, y);
    double dpx, dpy;
    bb4d(x - mx, y - my, a.c2Nrg[b], sx, sy, dpx, dpy);
    const double px = row[1] + dpx * a.bbscale[b];
    const double py = row[3] + dpy * a.bbscale[b];
    if (!a.drift) {
        row[1] = px;
        row[3] = py;
    } else if (b == 0) {
        row[0] = x - px * a.S;
        row[1] = px;
        row[2] = y - py * a.S;
        row[3] = py;
    } else {
        row[0] = x + px * a.S;
        row[1] = px;
        row[2] = y + py * a.S;
        row[3] = py;
    }
}

/* per-slice sums of z over row ranges (nseg segments, fixed chunking) */
__global__ __launch_bounds__(SS_BLK) void k_ss_zpartial(const double *x, const int *idx, const int *seg_start,
                                                        const int *seg_count, const int *blk_seg,
                                                        const int *blk_chunk, double *part)
{
    __shared__ double sh[4][SS_BLK];
    const int s = blk_seg[blockIdx.x], chunk = blk_chunk[blockIdx.x];
    const int lo = chunk * SS_CHUNK, hi = min(lo + SS_CHUNK, seg_count[s]);
    double v[4] = {0., 0., 0., 0.};
    for (int k = lo + (int)threadIdx.x; k < hi; k += SS_BLK)
        v[0] = v[0] + x[(size_t)ss_row(idx, seg_start[s], k) * 6 + 4];
    ss_block_sum4(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ void k_ss_zfinal(const double *part, const int *seg_first_blk, const int *seg_nblk, const int *seg_count,
                            int nseg, double *Z)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    double t = 0.;
    for (int k = 0; k < seg_nblk[s]; k++) t = t + part[seg_first_blk[s] + k];
    Z[s] = t / seg_count[s];
}

/* z keys and row numbers of a beam */
__global__ void k_ss_zkeys(const double *x, int n, double *key, int *row)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = x[(size_t)i * 6 + 4];
    row[i] = i;
}

/* slice number of each row: i with zb[i+1] <= z < zb[i], else nslice */
__global__ void k_ss_slice_id(const double *x, int n, const double *zb, int nslice, int *key, int *row)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double z = x[(size_t)i * 6 + 4];
    int s = nslice;
    for (int k = 0; k < nslice; k++)
        if (z < zb[k] && z >= zb[k + 1]) {
            s = k;
            break;
        }
    key[i] = s;
    row[i] = i;
}

/* rows per slice number from the sorted slice numbers: the first row of
   slice s is the lower bound of s (binary search, one thread per slice) */
__global__ void k_ss_count(const int *key, int n, int nslice, int *count)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nslice) return;
    int lo[2];
    for (int q = 0; q < 2; q++) {
        int a = 0, b = n, v = s + q;
        while (a < b) {
            const int m = a + (b - a) / 2;
            if (key[m] < v)
                a = m + 1;
            else
                b = m;
        }
        lo[q] = a;
    }
    count[s] = lo[1] - lo[0];
}

__global__ void k_ss_gather(const double *x, const int *row, int n, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * 6) return;
    const int r = i / 6, k = i - r * 6;
    out[i] = x[(size_t)row[r] * 6 + k];
}

/* ------------------------------------------------------------ launch */

hipError_t st_ss_collide(const ss_pair_args &a0, hipStream_t stream)
{
    ss_pair_args a = a0;
    a.nb[0] = (a.count[0] + SS_CHUNK - 1) / SS_CHUNK;
    a.nb[1] = (a.count[1] + SS_CHUNK - 1) / SS_CHUNK;
    if (a.nb[0] + a.nb[1] == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ss_partial, dim3(a.nb[0] + a.nb[1]), dim3(SS_BLK), 0, stream, a);
    const int kb0 = (a.count[0] + SS_BLK - 1) / SS_BLK, kb1 = (a.count[1] + SS_BLK - 1) / SS_BLK;
    if (kb0 + kb1 > 0) hipLaunchKernelGGL(k_ss_kick, dim3(kb0 + kb1), dim3(SS_BLK), 0, stream, a, kb0);
    return hipGetLastError();
}

hipError_t st_ss_partials(ss_pair_args &a, hipStream_t stream)
{
    a.nb[0] = (a.count[0] + SS_CHUNK - 1) / SS_CHUNK;
    a.nb[1] = (a.count[1] + SS_CHUNK - 1) / SS_CHUNK;
    if (a.nb[0] + a.nb[1] == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ss_partial, dim3(a.nb[0] + a.nb[1]), dim3(SS_BLK), 0, stream, a);
    return hipGetLastError();
}

/* stable sort of (cell, contribution) pairs, cells in [0, 4096] */
hipError_t st_sort_cells_bytes(int n, size_t *bytes)
{
    return hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, (const int *)nullptr, (int *)nullptr,
                                              (const int *)nullptr, (int *)nullptr, n, 0, 13);
}

hipError_t st_sort_cells(void *temp, size_t bytes, const int *kin, int *kout, const int *vin, int *vout, int n,
                         hipStream_t stream)
{
    size_t need = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, vin, vout, n, 0, 13, stream);
    if (e != hipSuccess) return e;
    if (need > bytes) return hipErrorInvalidValue;
    return hipcub::DeviceRadixSort::SortPairs(temp, need, kin, kout, vin, vout, n, 0, 13, stream);
}

int st_ss_partial_blocks(int count) { return (count + SS_CHUNK - 1) / SS_CHUNK; }

hipError_t st_ss_zsums(const double *x, const int *idx, const int *seg_start, const int *seg_count,
                       const int *blk_seg, const int *blk_chunk, int nblk, const int *seg_first_blk,
                       const int *seg_nblk, int nseg, double *part, double *Z, hipStream_t stream)
{
    if (nblk > 0)
        hipLaunchKernelGGL(k_ss_zpartial, dim3(nblk), dim3(SS_BLK), 0, stream, x, idx, seg_start, seg_count, blk_seg,
                           blk_chunk, part);
    if (nseg > 0)
        hipLaunchKernelGGL(k_ss_zfinal, dim3((nseg + 63) / 64), dim3(64), 0, stream, part, seg_first_blk, seg_nblk,
                           seg_count, nseg, Z);
    return hipGetLastError();
}

hipError_t st_ss_sort_by_z(double *x, int n, ss_sort_buffers &s, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const int g = (n + 255) / 256;
    hipLaunchKernelGGL(k_ss_zkeys, dim3(g), dim3(256), 0, stream, x, n, s.dkey_in, s.row_in);
    size_t need = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, need, s.dkey_in, s.dkey_out, s.row_in,
                                                                 s.row_out, n, 0, 64, stream);
    if (e != hipSuccess) return e;
    if (need > s.temp_bytes) return hipErrorInvalidValue;
    e = hipcub::DeviceRadixSort::SortPairsDescending(s.temp, need, s.dkey_in, s.dkey_out, s.row_in, s.row_out, n, 0,
                                                     64, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_ss_gather, dim3((n * 6 + 255) / 256), dim3(256), 0, stream, x, s.row_out, n, s.rows);
    return hipMemcpyAsync(x, s.rows, (size_t)n * 6 * sizeof(double), hipMemcpyDeviceToDevice, stream);
}

hipError_t st_ss_group_by_slice(const double *x, int n, const double *d_zb, int nslice, ss_sort_buffers &s,
                                hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const int g = (n + 255) / 256;
    hipLaunchKernelGGL(k_ss_slice_id, dim3(g), dim3(256), 0, stream, x, n, d_zb, nslice, s.ikey_in, s.row_in);
    size_t need = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, need, s.ikey_in, s.ikey_out, s.row_in, s.row_out, n,
                                                      0, 32, stream);
    if (e != hipSuccess) return e;
    if (need > s.temp_bytes) return hipErrorInvalidValue;
    e = hipcub::DeviceRadixSort::SortPairs(s.temp, need, s.ikey_in, s.ikey_out, s.row_in, s.row_out, n, 0, 32,
                                           stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_ss_count, dim3((nslice + 64) / 64), dim3(64), 0, stream, s.ikey_out, n, nslice, s.counts);
    return hipGetLastError();
}

hipError_t st_ss_sort_temp_bytes(int n, size_t *bytes)
{
    size_t a = 0, b = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, a, (const double *)nullptr,
                                                                 (double *)nullptr, (const int *)nullptr,
                                                                 (int *)nullptr, n, 0, 64);
    if (e != hipSuccess) return e;
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const int *)nullptr, (int *)nullptr, (const int *)nullptr,
                                           (int *)nullptr, n, 0, 32);
    *bytes = a > b ? a : b;
    return e;
}
