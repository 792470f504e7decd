/*
 * st_ss_host.cpp -- strong-strong beam-beam (soft-Gaussian and
 * particle-in-cell) host drivers: the reference's slice schedule, slice
 * pairs on up to SS_LANES streams, the C ABI entry points.
 */
#include "st_capi_internal.h"

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
    /* as many lanes as the smaller beam has slices, at most the context's limit */
    const int lanes = std::max(1, std::min(std::min(SS_LANES, c->ss_max_lanes), std::min(ns1, ns2)));
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
            if (rc) {
                /* join the lanes already launched before the context stream
                   goes on, so a later call cannot reuse their buffers while
                   their kernels run */
                for (size_t k = i0 + 1; k <= i; ++k) {
                    const int lk = (int)(k - i0);
                    if (hipEventRecord(c->lane_ev[lk], c->lane[lk]) == hipSuccess)
                        (void)hipStreamWaitEvent(c->stream, c->lane_ev[lk], 0);
                }
                return rc;
            }
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
                    double Np2, double gamma1, double gamma2, double bbscale1, double bbscale2, int lane = 0,
                    const double *fixed_h = nullptr)
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
    if (fixed_h) {
        /* BeamBeam4D_PIC_NonOffset: the grid is given, no moments */
        HIPCHK(st_pic_fixed_params(k.params, fixed_h[0], fixed_h[1], st));
    } else {
        HIPCHK(st_ss_partials(a, st));
        HIPCHK(st_pic_params(k.part, a.nb[0], a.nb[1], m1, m2, k.params, st));
    }
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

/* BeamBeam4D_PIC_NonOffset, simtrack.cpp:24226-24280 */
extern "C" int st_bb_pic_4d_nonoffset(st_ctx *c, double Np1, int Mp1, double *d_x1, double Np2, int Mp2,
                                      double *d_x2, double gamma1, double gamma2, double bbscale1, double bbscale2,
                                      int Nx, int Ny, double hx, double hy)
{
    int rc = ss_check(c, Mp1, d_x1, Mp2, d_x2, "st_bb_pic_4d_nonoffset");
    if (rc) return rc;
    if (Nx != 64 || Ny != 64) return fail(ST_ERR_PARAM, "st_bb_pic_4d_nonoffset: the grid is 64 x 64 (Nx = Ny = 64)");
    if (!(hx > 0. && hx < 1e300) || !(hy > 0. && hy < 1e300))
        return fail(ST_ERR_PARAM, "st_bb_pic_4d_nonoffset: hx and hy must be positive and finite");
    HIPCHK(hipSetDevice(c->device));
    ss_work w;
    if ((rc = ss_prepare(c, Mp1, Mp2, 1, 1, false, w, std::max(std::max(Mp1, Mp2), 1)))) return rc;
    const double h[2] = {hx, hy};
    if ((rc = pic_pair(c, w, d_x1, 0, Mp1, d_x2, 0, Mp2, Np1, Np2, gamma1, gamma2, bbscale1, bbscale2, 0, h)))
        return rc;
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

extern "C" int st_bb_pic_4d_nonoffset_host(st_ctx *c, double Np1, int Mp1, double *x1, double Np2, int Mp2,
                                           double *x2, double gamma1, double gamma2, double bbscale1,
                                           double bbscale2, int Nx, int Ny, double hx, double hy)
{
    int rc = ss_check(c, Mp1, x1, Mp2, x2, "st_bb_pic_4d_nonoffset_host");
    if (rc) return rc;
    if (Nx != 64 || Ny != 64)
        return fail(ST_ERR_PARAM, "st_bb_pic_4d_nonoffset_host: the grid is 64 x 64 (Nx = Ny = 64)");
    if (!(hx > 0. && hx < 1e300) || !(hy > 0. && hy < 1e300))
        return fail(ST_ERR_PARAM, "st_bb_pic_4d_nonoffset_host: hx and hy must be positive and finite");
    HIPCHK(hipSetDevice(c->device));
    ss_work w;
    const double h[2] = {hx, hy};
    if ((rc = ss_prepare(c, Mp1, Mp2, 1, 1, true, w, std::max(std::max(Mp1, Mp2), 1))) ||
        (rc = ss_host_in(c, w, Mp1, x1, Mp2, x2)) ||
        (rc = pic_pair(c, w, w.hx[0], 0, Mp1, w.hx[1], 0, Mp2, Np1, Np2, gamma1, gamma2, bbscale1, bbscale2, 0, h)))
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
