/*
 * st_pic.hip -- particle-in-cell strong-strong beam-beam on gfx950
 * (BeamBeam4D_PIC / BeamBeam6D_PIC, simtrack.cpp:23286-23615).
 *
 * One field solve (the charge of one beam, felt by the other):
 *   k_pic_contrib    each source row's 3 x 3 weights (Cal_Grid_Weight) as
 *                    (cell, weight) contributions, row-major;
 *   (radix sort of the contributions by cell, stable)
 *   k_pic_gather_w, k_pic_cell_sum
 *                    each cell's charge, its contributions added in row order
 *                    from zero: the reference's Charge2Grid sums, bit for bit;
 *   k_pic_green      the integrated Green function on the doubled grid
 *                    (Cal_Effective_Green_Function, before its FFT);
 *   k_pic_fft_lines  the reference's radix-2 FFT1D on 128-point lines, one
 *                    wavefront per line: the same bit-reversal, the same
 *                    butterflies, the twiddles of its recurrence from a table;
 *   k_pic_pad, k_pic_mul, k_pic_phi, k_pic_efield
 *                    Cal_Potential_Convolution and Cal_Electric_Field;
 *   k_pic_kick       Interpolate_Electric_Field and the kick (cpp:23337-23340).
 * Every step after the moments repeats the reference's operations in its
 * order; the moments (grid size and origin) are the fixed-order sums of the
 * soft-Gaussian kernels.
 */
#include <hip/hip_runtime.h>

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

#define PIC_NX 64
#define PIC_NY 64
#define PIC_N2 128
#define PIC_NCELL (PIC_NX * PIC_NY)

/* Cal_Grid_Weight, simtrack.cpp:2755-2791 */
ST_DEV void pic_grid_weight(double x1, double y1, double x0, double y0, double hx, double hy, int xm[3], int ym[3],
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

ST_DEV bool pic_inside(double xp, double yp, const double *mean, double hx, double hy)
{
    return __builtin_fabs(xp - mean[0]) < (PIC_NX / 2 - 3) * hx && __builtin_fabs(yp - mean[1]) < (PIC_NY / 2 - 3) * hy;
}

/* the grid of one pair from both slices' partial sums (BeamBeam4D_PIC,
   cpp:23300-23326): p[0..3] mean1, rms1; p[4..7] mean2, rms2; p[8], p[9]
   hx, hy */
__global__ void k_pic_params(const double *part, int nb0, int nb1, int m0, int m1, double *p)
{
    __shared__ double sh[8][256];
    const int t = threadIdx.x;
    double v[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
    for (int b = 0; b < 2; ++b) {
        const int base = b ? nb0 : 0, nb = b ? nb1 : nb0;
        for (int k = t; k < nb; k += 256)
            for (int q = 0; q < 4; q++) v[4 * b + q] = v[4 * b + q] + part[(size_t)(base + k) * 4 + q];
    }
    for (int q = 0; q < 8; q++) sh[q][t] = v[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int q = 0; q < 8; q++) sh[q][t] = sh[q][t] + sh[q][t + w];
        __syncthreads();
    }
    if (t != 0) return;
    for (int b = 0; b < 2; ++b) {
        const int n = b ? m1 : m0;
        const double mx = sh[4 * b][0] / n, my = sh[4 * b + 1][0] / n;
        p[4 * b + 0] = mx;
        p[4 * b + 1] = my;
        p[4 * b + 2] = __builtin_sqrt(sh[4 * b + 2][0] / n - mx * mx);
        p[4 * b + 3] = __builtin_sqrt(sh[4 * b + 3][0] / n - my * my);
    }
    const double rmsx = p[2] > p[6] ? p[6] : p[2];
    const double rmsy = p[3] > p[7] ? p[7] : p[3];
    p[8] = 15 * rmsx / PIC_NX;
    p[9] = 15 * rmsy / PIC_NY;
}

/* the source beam s (0 or 1) of a solve: grid origin and offset
   (cpp:23319-23322, 23345-23348) */
ST_DEV void pic_frame(const double *p, int s, double &x0, double &y0, double &xoff, double &yoff)
{
    const double *ms = p + 4 * s, *mt = p + 4 * (1 - s);
    x0 = ms[0] - PIC_NX * p[8] / 2;
    y0 = ms[1] - PIC_NY * p[9] / 2;
    xoff = mt[0] - ms[0];
    yoff = mt[1] - ms[1];
}

__global__ void k_pic_contrib(const double *x, int start, int m, const double *p, int s, int *key, int *val,
                              double *w)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const double *row = x + (size_t)(start + k) * 6;
    const double xp = row[0], yp = row[2];
    const double hx = p[8], hy = p[9];
    double x0, y0, xoff, yoff;
    pic_frame(p, s, x0, y0, xoff, yoff);
    int xm[3], ym[3];
    double xw[3], yw[3];
    const bool in = pic_inside(xp, yp, p + 4 * s, hx, hy);
    if (in) pic_grid_weight(xp, yp, x0, y0, hx, hy, xm, ym, xw, yw);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const size_t q = (size_t)k * 9 + 3 * i + j;
            int cell = PIC_NCELL;
            if (in) {
                const int c = xm[i] * PIC_NX + ym[j];
                cell = (c >= 0 && c < PIC_NCELL) ? c : PIC_NCELL;
            }
            key[q] = cell;
            val[q] = (int)q;
            w[q] = in ? xw[i] * yw[j] : 0.;
        }
}

ST_DEV int pic_lower_bound(const int *key, int n, int v)
{
    int a = 0, b = n;
    while (a < b) {
        const int mid = a + (b - a) / 2;
        if (key[mid] < v)
            a = mid + 1;
        else
            b = mid;
    }
    return a;
}

/* the weights in cell order (ws[q] = w[val[q]]), so that each cell's sum
   below reads consecutive memory */
__global__ void k_pic_gather_w(const int *val, int n, const double *w, double *ws)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) ws[q] = w[val[q]];
}

/* lane j's double, read by every lane (two v_readlane) */
ST_DEV double pic_readlane(double v, int j)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, j);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

#define PIC_CS_AHEAD 8 /* blocks of 64 weights in flight per wavefront */

/*
 * One wavefront per cell.  The lanes load the cell's weights 64 at a time,
 * coalesced and PIC_CS_AHEAD blocks ahead of the additions; each block is
 * then added to one running sum in row order, every weight broadcast from
 * its lane, so the sum is the sequential one of the deposition loop.
 */
__global__ __launch_bounds__(256) void k_pic_cell_sum(const int *key, int n, const double *ws, double *src)
{
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (c >= PIC_NCELL) return;
    const int lo = pic_lower_bound(key, n, c), hi = pic_lower_bound(key, n, c + 1);
    const int nb = (hi - lo + 63) >> 6;
    double buf[PIC_CS_AHEAD];
#pragma unroll
    for (int d = 0; d < PIC_CS_AHEAD; ++d) {
        const int q = lo + d * 64 + lane;
        buf[d] = q < hi ? ws[q] : 0.;
    }
    double t = 0;
    for (int b = 0; b < nb; b += PIC_CS_AHEAD) {
#pragma unroll
        for (int d = 0; d < PIC_CS_AHEAD; ++d) {
            if (b + d < nb) {
                const double v = buf[d];
                const int q = lo + (b + d + PIC_CS_AHEAD) * 64 + lane;
                buf[d] = q < hi ? ws[q] : 0.;
                const int m = min(64, hi - lo - (b + d) * 64);
                if (m == 64) {
#pragma unroll
                    for (int u = 0; u < 64; ++u) t = t + pic_readlane(v, u);
                } else {
                    for (int u = 0; u < m; ++u) t = t + pic_readlane(v, u);
                }
            }
        }
    }
    if (lane == 0) src[c] = t;
}

/* int_green and calc_egreen, simtrack.cpp:2808-2879 */
ST_DEV double pic_int_green(double x, double y)
{
    if (0 == x || 0 == y) return 0;
    double x2 = x * x, y2 = y * y;
    return -3 * x * y + x2 * st_atan(y / x) + y2 * st_atan(x / y) + x * y * st_log(x2 + y2);
}

ST_DEV double pic_egreen(double x1, double x2, double y1, double y2, double hx, double hy)
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
        x1 = __builtin_fabs(x1);
        x2 = __builtin_fabs(x2);
        y1 = __builtin_fabs(y1);
        y2 = __builtin_fabs(y2);
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

/* Cal_Effective_Green_Function, simtrack.cpp:2881-2921 */
__global__ void k_pic_green(const double *p, int s, double *gr, double *gi)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= PIC_N2 * PIC_N2) return;
    const int i = q / PIC_N2, j = q - i * PIC_N2;
    const double hx = p[8], hy = p[9];
    double x0, y0, xoff, yoff;
    pic_frame(p, s, x0, y0, xoff, yoff);
    double x1 = (i <= PIC_NX - 1) ? xoff + i * hx - 0.5 * hx : xoff - (2 * PIC_NX - i) * hx - 0.5 * hx;
    double y1 = (j <= PIC_NY - 1) ? yoff + j * hy - 0.5 * hy : yoff - (2 * PIC_NY - j) * hy - 0.5 * hy;
    gr[q] = pic_egreen(x1, x1 + hx, y1, y1 + hy, hx, hy);
    gi[q] = 0.;
}

/* FFT1D (simtrack.cpp:2583-2647) on 128-point lines: line l holds the
   elements base(l) + k * step, k = 0..127; one 64-thread block per line.
   tw: the 127 twiddles (u1, u2) of the reference's recurrence, stage l at
   offset 2^l - 1 */
__global__ __launch_bounds__(64) void k_pic_fft_lines(double *x, double *y, int line_stride, int step,
                                                      const double *tw, int dir)
{
    __shared__ double re[PIC_N2], im[PIC_N2];
    const int t = threadIdx.x;
    double *bx = x + (size_t)blockIdx.x * line_stride, *by = y + (size_t)blockIdx.x * line_stride;
    for (int k = t; k < PIC_N2; k += 64) {
        const int r = (int)(__builtin_bitreverse32((uint32_t)k) >> 25);
        re[k] = bx[(size_t)r * step];
        im[k] = by[(size_t)r * step];
    }
    __syncthreads();
    for (int l = 0; l < 7; ++l) {
        const int l1 = 1 << l, l2 = l1 << 1;
        const int j = t & (l1 - 1), i = (t >> l) * l2 + j, i1 = i + l1;
        const double u1 = tw[2 * (l1 - 1 + j)], u2 = tw[2 * (l1 - 1 + j) + 1];
        const double t1 = u1 * re[i1] - u2 * im[i1];
        const double t2 = u1 * im[i1] + u2 * re[i1];
        const double xi = re[i], yi = im[i];
        re[i1] = xi - t1;
        im[i1] = yi - t2;
        re[i] = xi + t1;
        im[i] = yi + t2;
        __syncthreads();
    }
    for (int k = t; k < PIC_N2; k += 64) {
        double a = re[k], b = im[k];
        if (dir == 1) {
            a /= (double)PIC_N2;
            b /= (double)PIC_N2;
        }
        bx[(size_t)k * step] = a;
        by[(size_t)k * step] = b;
    }
}

__global__ void k_pic_pad(const double *src, double *sr, double *si)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= PIC_N2 * PIC_N2) return;
    const int i = q / PIC_N2, j = q - i * PIC_N2;
    sr[q] = (i < PIC_NX && j < PIC_NY) ? src[i * PIC_NX + j] : 0.;
    si[q] = 0.;
}

__global__ void k_pic_mul(const double *sr, const double *si, const double *gr, const double *gi, double *pr,
                          double *pi)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= PIC_N2 * PIC_N2) return;
    pr[q] = sr[q] * gr[q] - si[q] * gi[q];
    pi[q] = sr[q] * gi[q] + si[q] * gr[q];
}

__global__ void k_pic_phi(const double *pr, double *phi)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= PIC_NCELL) return;
    const int i = q / PIC_NX, j = q - i * PIC_NX;
    const double eps0 = 8.854187817e-12;
    phi[q] = pr[i * 2 * PIC_NX + j] * 2 * PIC_NX * 2 * PIC_NY / 2 / 3.14159265 / eps0;
}

/* Cal_Electric_Field, simtrack.cpp:2998-3019 */
__global__ void k_pic_efield(const double *phi, const double *p, double *Ex, double *Ey)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= PIC_NCELL) return;
    const int Nx = PIC_NX, i = q / Nx, j = q - i * Nx;
    const double hx = p[8], hy = p[9];
    if (i < 1 || i > Nx - 2 || j < 1 || j > PIC_NY - 2) {
        Ex[q] = 0.;
        Ey[q] = 0.;
        return;
    }
    Ex[q] = -((phi[(i + 1) * Nx + j + 1] - phi[(i - 1) * Nx + j + 1]) + 4 * (phi[(i + 1) * Nx + j] - phi[(i - 1) * Nx + j]) +
              (phi[(i + 1) * Nx + j - 1] - phi[(i - 1) * Nx + j - 1])) /
            12. / hx;
    Ey[q] = -((phi[(i + 1) * Nx + j + 1] - phi[(i + 1) * Nx + j - 1]) + 4 * (phi[i * Nx + j + 1] - phi[i * Nx + j - 1]) +
              (phi[(i - 1) * Nx + j + 1] - phi[(i - 1) * Nx + j - 1])) /
            12. / hy;
}

/* Interpolate_Electric_Field (cpp:3021-3036) and the kick of cpp:23337-23340
   on the target beam t = 1 - s; nsm = (source charge) / (source rows) */
__global__ void k_pic_kick(double *x, int start, int m, const double *p, int s, const double *Ex, const double *Ey,
                           double nsm, double gamma, double bbscale)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    double *row = x + (size_t)(start + k) * 6;
    const double xp = row[0], yp = row[2];
    const double hx = p[8], hy = p[9];
    if (!pic_inside(xp, yp, p + 4 * (1 - s), hx, hy)) return;
    double x0, y0, xoff, yoff;
    pic_frame(p, s, x0, y0, xoff, yoff);
    int xm[3], ym[3];
    double xw[3], yw[3];
    pic_grid_weight(xp - xoff, yp - yoff, x0, y0, hx, hy, xm, ym, xw, yw);
    double ex = 0., ey = 0.;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const int c = xm[i] * PIC_NX + ym[j];
            const double fx = (c >= 0 && c < PIC_NCELL) ? Ex[c] : 0., fy = (c >= 0 && c < PIC_NCELL) ? Ey[c] : 0.;
            ex = ex + fx * xw[i] * yw[j];
            ey = ey + fy * xw[i] * yw[j];
        }
    row[1] = row[1] + ex * 1.60217662e-19 * bbscale / (gamma * 938.27201e6) * nsm;
    row[3] = row[3] + ey * 1.60217662e-19 * bbscale / (gamma * 938.27201e6) * nsm;
}

/* x += sgn p S (sgn = +1 / -1) on rows [start, start + m) (cpp:23556-23563) */
__global__ void k_pic_drift(double *x, int start, int m, double S, int sgn)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    double *row = x + (size_t)(start + k) * 6;
    if (sgn > 0) {
        row[0] = row[0] + row[1] * S;
        row[2] = row[2] + row[3] * S;
    } else {
        row[0] = row[0] - row[1] * S;
        row[2] = row[2] - row[3] * S;
    }
}

/* ------------------------------------------------------------ launch */

static hipError_t pic_fft2d(double *x, double *y, const double *tw, int dir, hipStream_t stream)
{
    /* FFT2D: first along i for every j (stride 128), then along j for every i */
    hipLaunchKernelGGL(k_pic_fft_lines, dim3(PIC_N2), dim3(64), 0, stream, x, y, 1, PIC_N2, tw, dir);
    hipLaunchKernelGGL(k_pic_fft_lines, dim3(PIC_N2), dim3(64), 0, stream, x, y, PIC_N2, 1, tw, dir);
    return hipGetLastError();
}

hipError_t st_pic_solve_and_kick(const pic_args &a, int s, hipStream_t stream)
{
    const int ms = a.count[s], mt = a.count[1 - s];
    const int n = ms * 9;
    if (ms > 0) {
        hipLaunchKernelGGL(k_pic_contrib, dim3((ms + 255) / 256), dim3(256), 0, stream, a.x[s], a.start[s], ms,
                           a.params, s, a.key_in, a.val_in, a.w);
        hipError_t e = st_sort_cells(a.temp, a.temp_bytes, a.key_in, a.key_out, a.val_in, a.val_out, n, stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_pic_gather_w, dim3((n + 255) / 256), dim3(256), 0, stream, a.val_out, n, a.w, a.ws);
    }
    hipLaunchKernelGGL(k_pic_cell_sum, dim3(PIC_NCELL / 4), dim3(256), 0, stream, a.key_out, n, a.ws, a.src);
    const int g2 = PIC_N2 * PIC_N2 / 256;
    hipLaunchKernelGGL(k_pic_green, dim3(g2), dim3(256), 0, stream, a.params, s, a.gr, a.gi);
    hipError_t e = pic_fft2d(a.gr, a.gi, a.tw_fwd, 1, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pic_pad, dim3(g2), dim3(256), 0, stream, a.src, a.sr, a.si);
    if ((e = pic_fft2d(a.sr, a.si, a.tw_fwd, 1, stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pic_mul, dim3(g2), dim3(256), 0, stream, a.sr, a.si, a.gr, a.gi, a.pr, a.pi);
    if ((e = pic_fft2d(a.pr, a.pi, a.tw_inv, -1, stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pic_phi, dim3(PIC_NCELL / 256), dim3(256), 0, stream, a.pr, a.phi);
    hipLaunchKernelGGL(k_pic_efield, dim3(PIC_NCELL / 256), dim3(256), 0, stream, a.phi, a.params, a.Ex, a.Ey);
    if (mt > 0)
        hipLaunchKernelGGL(k_pic_kick, dim3((mt + 255) / 256), dim3(256), 0, stream, a.x[1 - s], a.start[1 - s], mt,
                           a.params, s, a.Ex, a.Ey, a.nsm[s], a.gamma[1 - s], a.bbscale[1 - s]);
    return hipGetLastError();
}

/* the fixed grid of BeamBeam4D_PIC_NonOffset (cpp:24236-24237): both
   centroids at the origin (x0 = -Nx hx/2, no offsets), cells hx x hy; the
   rms entries are not read once hx, hy are set */
__global__ void k_pic_fixed_params(double *p, double hx, double hy)
{
    if (threadIdx.x != 0) return;
    for (int q = 0; q < 8; q++) p[q] = 0.;
    p[8] = hx;
    p[9] = hy;
}

hipError_t st_pic_fixed_params(double *params, double hx, double hy, hipStream_t stream)
{
    hipLaunchKernelGGL(k_pic_fixed_params, dim3(1), dim3(64), 0, stream, params, hx, hy);
    return hipGetLastError();
}

hipError_t st_pic_params(const double *part, int nb0, int nb1, int m0, int m1, double *params, hipStream_t stream)
{
    hipLaunchKernelGGL(k_pic_params, dim3(1), dim3(256), 0, stream, part, nb0, nb1, m0, m1, params);
    return hipGetLastError();
}

hipError_t st_pic_drift(double *x, int start, int m, double S, int sgn, hipStream_t stream)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pic_drift, dim3((m + 255) / 256), dim3(256), 0, stream, x, start, m, S, sgn);
    return hipGetLastError();
}
