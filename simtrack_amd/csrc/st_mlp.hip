/*
 * st_mlp.hip -- the survival surrogate: a small MLP (n_in -> H -> H -> H -> 1,
 * ReLU, sigmoid output) trained and evaluated on the GPU with hand-written
 * BF16 MFMA kernels (FP32 accumulation), binary cross-entropy loss and Adam.
 *
 * GEMM: one kernel for the three products of a layer,
 *   forward        Y[m x n]  = X[m x k] . W^T       (W is n x k, row-major)
 *   data gradient  dX[m x k] = dY[m x n] . W
 *   weight grad.   dW[n x k] = dY^T . X            (split over m, FP32 partials)
 * expressed as C = A . B with element strides for A(i, l) and B(l, j).  A
 * workgroup of four wavefronts computes a 64 x 64 tile of C; each wavefront
 * a 32 x 32 quarter as 2 x 2 tiles of v_mfma_f32_16x16x32_bf16, staging
 * 64 x 32 slices of A and 32 x 64 of B in LDS with k contiguous, the layout
 * both MFMA operands read (lane l: row/column l % 16, k = 8 (l / 16) .. +7;
 * result row 4 (l / 16) + r, column l % 16).
 *
 * Numbers kept in BF16: the inputs, weights, hidden activations and the
 * back-propagated gradients that feed the next MFMA.  FP32: accumulators,
 * logits, bias terms, master weights, Adam moments, weight gradients.
 */
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "simtrack_amd.h"

extern "C" int st_internal_fail(int code, const char *msg);
extern "C" hipStream_t st_internal_stream(st_ctx *c, int *device);

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

namespace {

/* round to nearest even FP32 -> BF16 (NaN kept quiet) */
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

enum Epi {
    EPI_RELU_BF16 = 0,   /* C = relu(acc + bias[j]) -> bf16                     */
    EPI_BIAS_F32 = 1,    /* C = acc + bias[j] -> fp32 (logits)                  */
    EPI_MASK_BF16 = 2,   /* C = acc * (mask(i, j) > 0) -> bf16 (data gradient)  */
    EPI_F32 = 3          /* C = acc -> fp32 (weight-gradient partials)          */
};

struct GemmArgs {
    const uint16_t *A;
    int64_t sa_i, sa_k;
    const uint16_t *B;
    int64_t sb_k, sb_j;
    void *C;
    int64_t sc_i, sc_j;
    int64_t sc_split;        /* EPI_F32: offset between split partials        */
    int M, N, K;
    int k_per_split;         /* K range of one blockIdx.z                      */
    const float *bias;
    const uint16_t *mask;    /* EPI_MASK_BF16: bf16 activation, same layout as C */
    int epi;
};

constexpr int TM = 64, TN = 64, TK = 32;

__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
    /* row stride 40 halves = 80 bytes keeps every 8-element fragment 16-byte aligned */
    __shared__ __attribute__((aligned(16))) uint16_t sA[TM][TK + 8];   /* row-major, k contiguous */
    __shared__ __attribute__((aligned(16))) uint16_t sB[TN][TK + 8];   /* column-major, k contiguous */
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
    const int kbeg = blockIdx.z * g.k_per_split;
    const int kend = std::min(g.K, kbeg + g.k_per_split);
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        /* stage A (64 x 32) and B (32 x 64): 2048 elements each, 8 per thread */
        for (int e = tid; e < TM * TK; e += 256) {
            int r = e / TK, kk = e % TK;
            int gi = i0 + r, gk = k0 + kk;
            sA[r][kk] = (gi < g.M && gk < kend) ? g.A[gi * g.sa_i + gk * g.sa_k] : (uint16_t)0;
        }
        for (int e = tid; e < TN * TK; e += 256) {
            int c = e / TK, kk = e % TK;
            int gj = j0 + c, gk = k0 + kk;
            sB[c][kk] = (gj < g.N && gk < kend) ? g.B[gk * g.sb_k + gj * g.sb_j] : (uint16_t)0;
        }
        __syncthreads();
        bf16x8 af[2], bfr[2];
        for (int t = 0; t < 2; t++) {
            const u16x8 va = *(const u16x8 *)&sA[wi + t * 16 + (lane & 15)][(lane >> 4) * 8];
            const u16x8 vb = *(const u16x8 *)&sB[wj + t * 16 + (lane & 15)][(lane >> 4) * 8];
            af[t] = __builtin_bit_cast(bf16x8, va);
            bfr[t] = __builtin_bit_cast(bf16x8, vb);
        }
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    /* epilogue */
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int r = 0; r < 4; r++) {
                int i = i0 + wi + a * 16 + 4 * (lane >> 4) + r;
                int j = j0 + wj + b * 16 + (lane & 15);
                if (i >= g.M || j >= g.N) continue;
                float v = acc[a][b][r];
                int64_t o = i * g.sc_i + j * g.sc_j;
                switch (g.epi) {
                case EPI_RELU_BF16: {
                    float t = v + g.bias[j];
                    ((uint16_t *)g.C)[o] = f2bf(t > 0.f ? t : 0.f);
                    break;
                }
                case EPI_BIAS_F32:
                    ((float *)g.C)[o] = v + g.bias[j];
                    break;
                case EPI_MASK_BF16:
                    ((uint16_t *)g.C)[o] = f2bf(bf2f(g.mask[o]) > 0.f ? v : 0.f);
                    break;
                default:
                    ((float *)g.C)[o + blockIdx.z * g.sc_split] = v;
                    break;
                }
            }
}

/* FP32 -> BF16 copy */
__global__ void k_to_bf16(const float *x, uint16_t *y, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

/* logits -> probabilities */
__global__ void k_sigmoid(const float *z, float *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 1.f / (1.f + expf(-z[i]));
}

/* BCE with logits: dz = (sigmoid(z) - y) / n as bf16; per-block loss sums */
__global__ __launch_bounds__(256) void k_bce(const float *z, const float *y, int n, uint16_t *dz, float *blk_loss)
{
    __shared__ float s[256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        float zi = z[i], yi = y[i];
        float p = 1.f / (1.f + expf(-zi));
        dz[i] = f2bf((p - yi) / (float)n);
        /* softplus(z) - y z, stable form */
        l = fmaxf(zi, 0.f) - yi * zi + log1pf(expf(-fabsf(zi)));
    }
    s[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) blk_loss[blockIdx.x] = s[0];
}

/* sum of nsplit partials of length len -> out (deterministic order) */
__global__ void k_sum_splits(const float *part, int nsplit, int64_t len, float *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    float s = 0.f;
    for (int k = 0; k < nsplit; k++) s += part[k * len + i];
    out[i] = s;
}

/* bias gradient: column sums of a bf16 m x n matrix, one thread per column */
__global__ void k_colsum(const uint16_t *g, int m, int n, float *out)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float s = 0.f;
    for (int i = 0; i < m; i++) s += bf2f(g[(int64_t)i * n + j]);
    out[j] = s;
}

/* Adam on FP32 master parameters, then the BF16 copy of the weights */
__global__ void k_adam(float *w, float *m, float *v, const float *g, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, uint16_t *wbf, int64_t nbf)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (i < nbf) wbf[i] = f2bf(wi);
}

} // namespace

/* ------------------------------------------------------------- host */

struct st_mlp {
    st_ctx *ctx = nullptr;
    hipStream_t s = nullptr;
    int device = 0;
    int nin = 0, nh = 0;
    /* layer l: W_l (out x in), b_l; l = 0..3 with sizes (nh x nin), (nh x nh) x2, (1 x nh) */
    int out_[4], in_[4];
    int64_t woff[4], boff[4], nparam = 0, nweights = 0;  /* weights first, then biases */
    float *d_w = nullptr, *d_m = nullptr, *d_v = nullptr, *d_g = nullptr;  /* fp32, nparam each */
    uint16_t *d_wbf = nullptr;                                            /* bf16 weights */
    int64_t step = 0;
    /* workspaces sized for cap rows */
    int cap = 0;
    uint16_t *d_x = nullptr, *d_h[3] = {nullptr, nullptr, nullptr}, *d_dh[3] = {nullptr, nullptr, nullptr};
    uint16_t *d_dz = nullptr;
    float *d_z = nullptr, *d_loss = nullptr, *d_part = nullptr;
    size_t cap_part = 0;
    ~st_mlp()
    {
        hipSetDevice(device);
        hipFree(d_w); hipFree(d_m); hipFree(d_v); hipFree(d_g); hipFree(d_wbf);
        hipFree(d_x); hipFree(d_dz); hipFree(d_z); hipFree(d_loss); hipFree(d_part);
        for (int l = 0; l < 3; l++) { hipFree(d_h[l]); hipFree(d_dh[l]); }
    }
};

#define MLPCHK(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return st_internal_fail(ST_ERR_DEVICE, hipGetErrorString(e_));            \
    } while (0)

static int mlp_ensure(st_mlp *m, int n)
{
    if (n <= m->cap) return ST_OK;
    hipFree(m->d_x); hipFree(m->d_dz); hipFree(m->d_z); hipFree(m->d_loss);
    for (int l = 0; l < 3; l++) { hipFree(m->d_h[l]); hipFree(m->d_dh[l]); m->d_h[l] = m->d_dh[l] = nullptr; }
    m->d_x = m->d_dz = nullptr;
    m->d_z = m->d_loss = nullptr;
    m->cap = 0;
    MLPCHK(hipMalloc(&m->d_x, (size_t)n * m->nin * 2));
    for (int l = 0; l < 3; l++) {
        MLPCHK(hipMalloc(&m->d_h[l], (size_t)n * m->nh * 2));
        MLPCHK(hipMalloc(&m->d_dh[l], (size_t)n * m->nh * 2));
    }
    MLPCHK(hipMalloc(&m->d_dz, (size_t)n * 2));
    MLPCHK(hipMalloc(&m->d_z, (size_t)n * 4));
    MLPCHK(hipMalloc(&m->d_loss, (size_t)((n + 255) / 256) * 4));
    m->cap = n;
    return ST_OK;
}

static hipError_t gemm(st_mlp *m, const uint16_t *A, int64_t sa_i, int64_t sa_k, const uint16_t *B, int64_t sb_k,
                       int64_t sb_j, void *C, int64_t sc_i, int64_t sc_j, int M, int N, int K, int epi,
                       const float *bias, const uint16_t *mask, int nsplit = 1, int64_t sc_split = 0)
{
    GemmArgs g;
    g.A = A; g.sa_i = sa_i; g.sa_k = sa_k;
    g.B = B; g.sb_k = sb_k; g.sb_j = sb_j;
    g.C = C; g.sc_i = sc_i; g.sc_j = sc_j; g.sc_split = sc_split;
    g.M = M; g.N = N; g.K = K;
    int kps = (K + nsplit - 1) / nsplit;
    kps = ((kps + TK - 1) / TK) * TK;
    g.k_per_split = kps;
    g.bias = bias; g.mask = mask; g.epi = epi;
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, nsplit);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, m->s, g);
    return hipGetLastError();
}

static inline const float *bias_of(const st_mlp *m, int l) { return m->d_w + m->boff[l]; }
static inline const uint16_t *wbf_of(const st_mlp *m, int l) { return m->d_wbf + m->woff[l]; }

extern "C" int st_mlp_create(st_ctx *ctx, int n_in, int n_hidden, uint64_t seed, st_mlp **out)
{
    if (!ctx || !out || n_in < 1 || n_in > 4096 || n_hidden < 1 || n_hidden > 4096)
        return st_internal_fail(ST_ERR_ARG, "st_mlp_create: bad argument");
    std::unique_ptr<st_mlp> m(new st_mlp);
    m->ctx = ctx;
    m->s = st_internal_stream(ctx, &m->device);
    MLPCHK(hipSetDevice(m->device));
    m->nin = n_in;
    m->nh = n_hidden;
    int ins[4] = {n_in, n_hidden, n_hidden, n_hidden}, outs[4] = {n_hidden, n_hidden, n_hidden, 1};
    int64_t o = 0;
    for (int l = 0; l < 4; l++) {
        m->in_[l] = ins[l];
        m->out_[l] = outs[l];
        m->woff[l] = o;
        o += (int64_t)ins[l] * outs[l];
    }
    m->nweights = o;
    for (int l = 0; l < 4; l++) {
        m->boff[l] = o;
        o += outs[l];
    }
    m->nparam = o;
    /* He-uniform initialisation from a 64-bit LCG (reproducible), zero biases */
    std::vector<float> w((size_t)o, 0.f);
    uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
    for (int l = 0; l < 4; l++) {
        float lim = std::sqrt(6.f / (float)ins[l]);
        for (int64_t i = 0; i < (int64_t)ins[l] * outs[l]; i++) {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            double u = (double)(st >> 11) * 0x1p-53;
            w[m->woff[l] + i] = (float)((2.0 * u - 1.0) * lim);
        }
    }
    MLPCHK(hipMalloc(&m->d_w, o * 4));
    MLPCHK(hipMalloc(&m->d_m, o * 4));
    MLPCHK(hipMalloc(&m->d_v, o * 4));
    MLPCHK(hipMalloc(&m->d_g, o * 4));
    MLPCHK(hipMalloc(&m->d_wbf, m->nweights * 2));
    MLPCHK(hipMemcpy(m->d_w, w.data(), o * 4, hipMemcpyHostToDevice));
    MLPCHK(hipMemset(m->d_m, 0, o * 4));
    MLPCHK(hipMemset(m->d_v, 0, o * 4));
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((m->nweights + 255) / 256)), dim3(256), 0, m->s, m->d_w,
                       m->d_wbf, m->nweights);
    MLPCHK(hipGetLastError());
    MLPCHK(hipStreamSynchronize(m->s));
    *out = m.release();
    return ST_OK;
}

extern "C" int st_mlp_destroy(st_mlp *m)
{
    delete m;
    return ST_OK;
}

extern "C" int st_mlp_num_params(const st_mlp *m, int64_t *n)
{
    if (!m || !n) return st_internal_fail(ST_ERR_ARG, "st_mlp_num_params: bad argument");
    *n = m->nparam;
    return ST_OK;
}

extern "C" int st_mlp_get_params(st_mlp *m, float *params, int64_t n)
{
    if (!m || !params || n != m->nparam) return st_internal_fail(ST_ERR_ARG, "st_mlp_get_params: bad argument");
    MLPCHK(hipSetDevice(m->device));
    MLPCHK(hipMemcpyAsync(params, m->d_w, n * 4, hipMemcpyDeviceToHost, m->s));
    MLPCHK(hipStreamSynchronize(m->s));
    return ST_OK;
}

extern "C" int st_mlp_set_params(st_mlp *m, const float *params, int64_t n)
{
    if (!m || !params || n != m->nparam) return st_internal_fail(ST_ERR_ARG, "st_mlp_set_params: bad argument");
    MLPCHK(hipSetDevice(m->device));
    MLPCHK(hipMemcpyAsync(m->d_w, params, n * 4, hipMemcpyHostToDevice, m->s));
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((m->nweights + 255) / 256)), dim3(256), 0, m->s, m->d_w,
                       m->d_wbf, m->nweights);
    MLPCHK(hipGetLastError());
    MLPCHK(hipStreamSynchronize(m->s));
    return ST_OK;
}

/* forward through the hidden layers and the logits; d_xf: fp32 features (n x nin) */
static int mlp_forward(st_mlp *m, const float *d_xf, int n)
{
    int rc = mlp_ensure(m, n);
    if (rc) return rc;
    int64_t cnt = (int64_t)n * m->nin;
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, m->s, d_xf, m->d_x, cnt);
    MLPCHK(hipGetLastError());
    const uint16_t *in = m->d_x;
    int kin = m->nin;
    for (int l = 0; l < 3; l++) {
        /* H_l = relu(in . W_l^T + b_l): A = in (n x kin), B(k, j) = W_l[j][k] */
        MLPCHK(gemm(m, in, kin, 1, wbf_of(m, l), 1, kin, m->d_h[l], m->nh, 1, n, m->nh, kin, EPI_RELU_BF16,
                    bias_of(m, l), nullptr));
        in = m->d_h[l];
        kin = m->nh;
    }
    MLPCHK(gemm(m, in, kin, 1, wbf_of(m, 3), 1, kin, m->d_z, 1, 1, n, 1, kin, EPI_BIAS_F32, bias_of(m, 3), nullptr));
    return ST_OK;
}

extern "C" int st_mlp_forward(st_mlp *m, const float *d_x, int n, float *d_prob)
{
    if (!m || n < 0 || (n > 0 && (!d_x || !d_prob))) return st_internal_fail(ST_ERR_ARG, "st_mlp_forward: bad argument");
    if (n == 0) return ST_OK;
    MLPCHK(hipSetDevice(m->device));
    int rc = mlp_forward(m, d_x, n);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sigmoid, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->s, m->d_z, d_prob, n);
    MLPCHK(hipGetLastError());
    return ST_OK;
}

/* forward + backward into m->d_g (weights, then biases); mean loss to *loss */
static int mlp_backward(st_mlp *m, const float *d_x, const float *d_y, int n, float *loss)
{
    int rc = mlp_forward(m, d_x, n);
    if (rc) return rc;
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_bce, dim3((unsigned)nb), dim3(256), 0, m->s, m->d_z, d_y, n, m->d_dz, m->d_loss);
    MLPCHK(hipGetLastError());
    /* weight gradients are split over the batch: nsplit partial products */
    const int nsplit = std::max(1, std::min(64, n / 1024));
    size_t need = (size_t)nsplit * (size_t)std::max<int64_t>((int64_t)m->nh * m->nh, (int64_t)m->nh * m->nin);
    if (need > m->cap_part) {
        hipFree(m->d_part);
        m->d_part = nullptr;
        m->cap_part = 0;
        MLPCHK(hipMalloc(&m->d_part, need * 4));
        m->cap_part = need;
    }
    const uint16_t *dy = m->d_dz;  /* gradient at the output of layer l (n x out_l) */
    for (int l = 3; l >= 0; l--) {
        const int kin = m->in_[l], kout = m->out_[l];
        const uint16_t *xin = l > 0 ? m->d_h[l - 1] : m->d_x;
        /* dW_l (kout x kin) = dy^T . xin: A(i=o, k=r) = dy[r][o], B(k=r, j=c) = xin[r][c] */
        MLPCHK(gemm(m, dy, 1, kout, xin, kin, 1, m->d_part, kin, 1, kout, kin, n, EPI_F32, nullptr, nullptr, nsplit,
                    (int64_t)kout * kin));
        int64_t len = (int64_t)kout * kin;
        hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, m->s, m->d_part, nsplit,
                           len, m->d_g + m->woff[l]);
        MLPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_colsum, dim3((unsigned)((kout + 255) / 256)), dim3(256), 0, m->s, dy, n, kout,
                           m->d_g + m->boff[l]);
        MLPCHK(hipGetLastError());
        if (l > 0) {
            /* d h_{l-1} = (dy . W_l) * relu'(h_{l-1}): A = dy (n x kout), B(k=o, j=c) = W_l[o][c] */
            MLPCHK(gemm(m, dy, kout, 1, wbf_of(m, l), kin, 1, m->d_dh[l - 1], kin, 1, n, kin, kout, EPI_MASK_BF16,
                        nullptr, m->d_h[l - 1]));
            dy = m->d_dh[l - 1];
        }
    }
    if (loss) {
        std::vector<float> bl((size_t)nb);
        MLPCHK(hipMemcpyAsync(bl.data(), m->d_loss, (size_t)nb * 4, hipMemcpyDeviceToHost, m->s));
        MLPCHK(hipStreamSynchronize(m->s));
        double s = 0.;
        for (float v : bl) s += v;
        *loss = (float)(s / n);
    }
    return ST_OK;
}

extern "C" int st_mlp_gradients(st_mlp *m, const float *d_x, const float *d_y, int n, float *grads, float *loss)
{
    if (!m || n <= 0 || !d_x || !d_y || !grads) return st_internal_fail(ST_ERR_ARG, "st_mlp_gradients: bad argument");
    MLPCHK(hipSetDevice(m->device));
    int rc = mlp_backward(m, d_x, d_y, n, loss);
    if (rc) return rc;
    MLPCHK(hipMemcpyAsync(grads, m->d_g, m->nparam * 4, hipMemcpyDeviceToHost, m->s));
    MLPCHK(hipStreamSynchronize(m->s));
    return ST_OK;
}

extern "C" int st_mlp_train_step(st_mlp *m, const float *d_x, const float *d_y, int n, float lr, float *loss)
{
    if (!m || n <= 0 || !d_x || !d_y || !(lr > 0.f)) return st_internal_fail(ST_ERR_ARG, "st_mlp_train_step: bad argument");
    MLPCHK(hipSetDevice(m->device));
    int rc = mlp_backward(m, d_x, d_y, n, loss);
    if (rc) return rc;
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->nparam + 255) / 256)), dim3(256), 0, m->s, m->d_w, m->d_m, m->d_v,
                       m->d_g, m->nparam, lr, b1, b2, 1e-8f, c1, c2, m->d_wbf, m->nweights);
    MLPCHK(hipGetLastError());
    return ST_OK;
}
