/*
 * st_track.cpp -- the tracking drivers of the C ABI (Track, Track_spin,
 * Track_tbt, Track_Radial, Element::Pass): turn blocks of the tracking
 * kernel with survivor compaction in between.
 */
#include "st_capi_internal.h"

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
int turn_block(int n_elems, int np)
{
    (void)np;
    long work = (long)n_elems > 0 ? (long)n_elems : 1;
    long tb = 200000 / work;
    if (tb < 1) tb = 1;
    if (tb > 256) tb = 256;
    return (int)tb;
}

/* device-side core: d_x/d_state/d_lost already on the device */
/* The work buffers of one tracking call: the survivor index lists, the
   per-block counts and total of their compaction and Track_Radial's x04.
   Given (fixed != nullptr), nothing is allocated and no timing event is
   recorded; otherwise the context's grow-only buffers are used. */
struct track_scratch {
    int *idx[2];
    int *block;
    int *total;
    double *x04;
};

static int track_device_run(st_table *t, double *d_x, int np, int turn0, int nturn, int *d_state, int *d_lost,
                            bool spin, int radial, double *d_tbt, int tbt_bpm, int tbt_nturn, bool bare,
                            const track_scratch *fixed, bool perf = false);

int track_device_core(st_table *t, double *d_x, int np, int turn0, int nturn, int *d_state, int *d_lost,
                             bool spin, int radial, double *d_tbt, int tbt_bpm, int tbt_nturn, bool bare)
{
    return track_device_run(t, d_x, np, turn0, nturn, d_state, d_lost, spin, radial, d_tbt, tbt_bpm, tbt_nturn, bare,
                            nullptr);
}

static int track_device_run(st_table *t, double *d_x, int np, int turn0, int nturn, int *d_state, int *d_lost,
                            bool spin, int radial, double *d_tbt, int tbt_bpm, int tbt_nturn, bool bare,
                            const track_scratch *fixed, bool perf)
{
    auto launch = perf ? st_perf::st_launch_track : st_launch_track;
    st_ctx *c = t->ctx;
    if (!fixed && ensure_particles(c, (size_t)np, 1)) return ST_ERR_NOMEM;
    const bool timing = c->timing && !fixed;
    st_track_args a = base_args(t, spin);
    if (bare) a.ap = nullptr; /* Element::Pass turn after turn: no loss test */
    a.x = d_x;
    a.state = d_state;
    a.lost = d_lost;
    a.xtbt = d_tbt;
    a.tbt_bpm = tbt_bpm;
    a.tbt_nturn = tbt_nturn;
    a.tbt_turn0 = 0;
    if (radial == RADIAL_GIVEN && !fixed) {
        a.x04 = c->d_x04; /* filled by the caller (st_track_dz) */
    } else if (radial && fixed) {
        st_track_args ar = base_args(t, false);
        HIPCHK(st_launch_radial_x04(ar, fixed->x04, t->gp.H_expand != 0, c->stream));
        a.x04 = fixed->x04;
    } else if (radial) {
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
    int *bufs[2] = {fixed ? fixed->idx[0] : c->d_idx0, fixed ? fixed->idx[1] : c->d_idx1};
    int *d_block = fixed ? fixed->block : c->d_block;
    int *d_total = fixed ? fixed->total : c->d_total;
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
        if (timing) {
            int rc2 = timed_event(c, &ev);
            if (rc2) return rc2;
            HIPCHK(hipEventRecord(c->ev_pool[ev], c->stream));
        }
        HIPCHK(launch(a, spin, hexp, rad, c->stream));
        if (timing) {
            HIPCHK(hipEventRecord(c->ev_pool[ev + 1], c->stream));
            c->pending.push_back({ev, ev + 1});
            c->pending_work.push_back((int64_t)nact * nt);
        }
        first = false;
        t0 += nt;
        if (t0 < turn0 + nturn) {
            int *out = bufs[which];
            HIPCHK(st_launch_compact(idx, nact, d_state, d_block, out, d_total, c->stream));
            HIPCHK(hipMemcpyAsync(c->h_total, d_total, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (timing) {
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
        HIPCHK(launch(a, spin, hexp, rad, c->stream));
    }
    return ST_OK;
}

int track_host(st_table *t, double *x, int np, int nturn, int *stable, int *lost_turn, int *lost_post,
                      bool spin, int radial, double *x_tbt, int tbt_bpm, bool bare)
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

/* Track_wrt_OnMomentumPart (simtrack.cpp:20942-20966): Track with dz0[j]
   subtracted from z after element j, as Track_Radial does with its own x04 */
extern "C" int st_track_dz(st_table *t, const double *dz0, double *x, int np, int nturn, int *stable, int *lost_turn,
                           int *lost_post)
{
    if (!t || !dz0) return fail(ST_ERR_ARG, "st_track_dz: bad argument");
    st_ctx *c = t->ctx;
    HIPCHK(hipSetDevice(c->device));
    if ((size_t)t->n > c->cap_x04) {
        hipFree(c->d_x04);
        c->d_x04 = nullptr;
        c->cap_x04 = 0;
        HIPCHK(hipMalloc(&c->d_x04, (size_t)(t->n > 0 ? t->n : 1) * sizeof(double)));
        c->cap_x04 = (size_t)t->n;
    }
    HIPCHK(hipMemcpyAsync(c->d_x04, dz0, (size_t)t->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return track_host(t, x, np, nturn, stable, lost_turn, lost_post, false, RADIAL_GIVEN, nullptr, -1);
}

extern "C" int st_track_device(st_table *t, double *d_x, int np, int turn0, int nturn, int *d_state, int *d_lost,
                               int flags)
{
    if (!t || np < 0 || nturn < 0 || turn0 < 0 || (np > 0 && (!d_x || !d_state || !d_lost)))
        return fail(ST_ERR_ARG, "st_track_device: bad argument");
    if (np == 0) return ST_OK;
    HIPCHK(hipSetDevice(t->ctx->device));
    return track_device_run(t, d_x, np, turn0, nturn, d_state, d_lost, (flags & ST_TRACK_SPIN) != 0,
                            (flags & ST_TRACK_RADIAL) != 0, nullptr, -1, 0, (flags & ST_TRACK_BARE) != 0, nullptr,
                            (flags & ST_TRACK_PERF) != 0);
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

/* ------------------------------------------- allocation-free tracking */

/* Scratch of st_track_device_noalloc, one copy per GPU, held until the
   process ends: Track_Radial's x04 (ST_NOALLOC_MAX_ELEMENTS doubles), then
   two survivor index lists of ST_NOALLOC_MAX_PARTICLES ints, the per-block
   counts of their compaction (one int per 1,024 particles, plus one) and the
   survivor total.  Each launch writes what the next one reads, so nothing in
   it has to start at zero. */
#define NA_BLOCKS (ST_NOALLOC_MAX_PARTICLES / 1024 + 1)
#define NA_INTS (2 * (size_t)ST_NOALLOC_MAX_PARTICLES + NA_BLOCKS + 1)
#define NA_DOUBLES ((size_t)ST_NOALLOC_MAX_ELEMENTS + (NA_INTS + 1) / 2)
static __device__ __align__(256) double scratch[NA_DOUBLES];

/* device address of each GPU's scratch (the array's name is only a host-side
   stand-in), looked up once per device by st_ctx_create */
#define NA_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NA_MAX_DEVICES];

int noalloc_scratch_init(int device)
{
    if (device < 0 || device >= NA_MAX_DEVICES)
        return fail(ST_ERR_ARG, "device %d outside the %d the allocation-free path supports", device, NA_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return ST_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)); /* the current device's copy */
    if (e != hipSuccess || !ptr)
        return fail(ST_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = (double *)ptr;
    return ST_OK;
}

extern "C" int st_track_device_noalloc(st_table *t, double *d_x, int np, int turn0, int nturn, int *d_state,
                                       int *d_lost, int flags)
{
    if (!t || np < 0 || nturn < 0 || turn0 < 0 || (np > 0 && (!d_x || !d_state || !d_lost)))
        return fail(ST_ERR_ARG, "st_track_device_noalloc: bad argument");
    if (np > ST_NOALLOC_MAX_PARTICLES)
        return fail(ST_ERR_ARG, "st_track_device_noalloc: %d particles exceed the scratch maximum of %d", np,
                    ST_NOALLOC_MAX_PARTICLES);
    if ((flags & ST_TRACK_RADIAL) && t->n > ST_NOALLOC_MAX_ELEMENTS)
        return fail(ST_ERR_ARG, "st_track_device_noalloc: radial tracking of %d elements exceeds the scratch "
                    "maximum of %d", t->n, ST_NOALLOC_MAX_ELEMENTS);
    if (np == 0) return ST_OK;
    const int dev = t->ctx->device;
    HIPCHK(hipSetDevice(dev));
    double *base = (dev >= 0 && dev < NA_MAX_DEVICES) ? g_scratch_ptr[dev] : nullptr;
    if (!base) return fail(ST_ERR_DEVICE, "st_track_device_noalloc: no scratch address for device %d", dev);
    track_scratch s;
    s.x04 = base;
    int *ints = (int *)(base + ST_NOALLOC_MAX_ELEMENTS);
    s.idx[0] = ints;
    s.idx[1] = ints + ST_NOALLOC_MAX_PARTICLES;
    s.block = ints + 2 * (size_t)ST_NOALLOC_MAX_PARTICLES;
    s.total = s.block + NA_BLOCKS;
    return track_device_run(t, d_x, np, turn0, nturn, d_state, d_lost, (flags & ST_TRACK_SPIN) != 0,
                            (flags & ST_TRACK_RADIAL) != 0, nullptr, -1, 0, (flags & ST_TRACK_BARE) != 0, &s,
                            (flags & ST_TRACK_PERF) != 0);
}
