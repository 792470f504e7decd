/*
 * st_batch.cpp -- many lattices (error seeds) of uneven length tracked by
 * one persistent launch per turn block, and the launch-shape autotuner.
 */
#include "st_capi_internal.h"

/* --------------------------------------------- many tables, one launch */

struct st_batch {
    st_ctx *c = nullptr;
    std::vector<st_table *> tabs;
    st_seed_desc *d_seeds[2] = {nullptr, nullptr}; /* orbit, spin */
    int *d_items = nullptr;
    size_t cap_items = 0;
    int *d_counter = nullptr;
    int n_cu = 0;
    bool hexp = false, rad = false;
    st_kglobals kg;
    ~st_batch()
    {
        if (!c) return;
        hipSetDevice(c->device);
        hipFree(d_seeds[0]);
        hipFree(d_seeds[1]);
        hipFree(d_items);
        hipFree(d_counter);
    }
};

extern "C" int st_batch_create(st_ctx *c, st_table *const *tables, int ntab, st_batch **out)
{
    if (!c || !tables || ntab <= 0 || !out) return fail(ST_ERR_ARG, "st_batch_create: bad argument");
    *out = nullptr;
    for (int k = 0; k < ntab; ++k) {
        if (!tables[k] || tables[k]->ctx != c) return fail(ST_ERR_ARG, "st_batch_create: table %d not on this context", k);
        const st_table *t0 = tables[0], *t = tables[k];
        if (std::memcmp(&t->kg, &t0->kg, sizeof(st_kglobals)) != 0 || t->gp.H_expand != t0->gp.H_expand ||
            t->gp.radiate != t0->gp.radiate)
            return fail(ST_ERR_ARG, "st_batch_create: table %d has other globals than table 0 (one launch shares them)", k);
    }
    HIPCHK(hipSetDevice(c->device));
    std::unique_ptr<st_batch> b(new st_batch);
    b->c = c;
    b->tabs.assign(tables, tables + ntab);
    b->hexp = tables[0]->gp.H_expand != 0;
    b->rad = tables[0]->gp.radiate != 0;
    b->kg = tables[0]->kg;
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    b->n_cu = ncu > 0 ? ncu : 1;
    for (int s = 0; s < 2; ++s) {
        std::vector<st_seed_desc> d((size_t)ntab);
        for (int k = 0; k < ntab; ++k) {
            st_table *t = tables[k];
            dev_table &dt = s ? t->spin : t->orbit;
            d[k].ops = dt.d_ops;
            d[k].par = dt.d_par;
            d[k].ap = t->d_ap;
            d[k].nelem = t->n;
            d[k].reserved = 0;
            d[k].draws_per_turn = s ? 0 : t->draws_per_turn;
        }
        HIPCHK(hipMalloc(&b->d_seeds[s], (size_t)ntab * sizeof(st_seed_desc)));
        HIPCHK(hipMemcpy(b->d_seeds[s], d.data(), (size_t)ntab * sizeof(st_seed_desc), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&b->d_counter, sizeof(int)));
    *out = b.release();
    return ST_OK;
}

extern "C" int st_batch_destroy(st_batch *b)
{
    delete b;
    return ST_OK;
}

/* items of the alive particles, 64 per item, never mixing tables; the
   longest tables first so that the tail of the launch is short */
static void build_items(const st_batch *b, const std::vector<int> &state, const int *offsets, std::vector<int> &idx,
                        std::vector<int> &items)
{
    const int ntab = (int)b->tabs.size();
    std::vector<int> order((size_t)ntab);
    for (int k = 0; k < ntab; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return b->tabs[u]->n > b->tabs[v]->n; });
    idx.clear();
    items.clear();
    for (int k : order) {
        int start = (int)idx.size();
        for (int p = offsets[k]; p < offsets[k + 1]; ++p)
            if (state[p] != 0) idx.push_back(p);
        for (int s0 = start; s0 < (int)idx.size(); s0 += 64) {
            items.push_back(k);
            items.push_back(s0);
            items.push_back(std::min(64, (int)idx.size() - s0));
        }
    }
}

static int batch_core(st_batch *b, double *d_x, const int *offsets, int turn0, int nturn, int *d_state, int *d_lost,
                      bool spin, const st_batch_config &cfg, double *d_tbt = nullptr, bool perf = false)
{
    st_ctx *c = b->c;
    const int ntab = (int)b->tabs.size();
    const int np = offsets[ntab];
    if (np <= 0 || nturn <= 0) return ST_OK;
    if (cfg.wg_waves < 1 || cfg.wg_waves > 4 || cfg.waves_per_simd < 1 || cfg.waves_per_simd > 8 || cfg.turn_block < 1)
        return fail(ST_ERR_ARG, "st_batch: bad config (wg_waves %d, waves_per_simd %d, turn_block %d)", cfg.wg_waves,
                    cfg.waves_per_simd, cfg.turn_block);
    int rc = ensure_particles(c, (size_t)np, 1);
    if (rc) return rc;
    std::vector<int> state((size_t)np), idx, items;
    HIPCHK(hipMemcpyAsync(state.data(), d_state, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int grid = std::max(1, b->n_cu * 4 * cfg.waves_per_simd / cfg.wg_waves);
    const int block = 64 * cfg.wg_waves;
    bool first = true;
    for (int t0 = turn0; t0 < turn0 + nturn;) {
        const int nt = std::min(cfg.turn_block, turn0 + nturn - t0);
        /* the pre-check runs on every particle, alive or not, before turn 0 */
        if (first && turn0 == 0) {
            std::vector<int> all((size_t)np, 1);
            build_items(b, all, offsets, idx, items);
        } else {
            build_items(b, state, offsets, idx, items);
        }
        const int nitems = (int)items.size() / 3;
        if (nitems == 0) break;
        if ((size_t)items.size() > b->cap_items) {
            hipFree(b->d_items);
            b->d_items = nullptr;
            b->cap_items = 0;
            HIPCHK(hipMalloc(&b->d_items, items.size() * sizeof(int)));
            b->cap_items = items.size();
        }
        HIPCHK(hipMemcpyAsync(c->d_idx0, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(b->d_items, items.data(), items.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemsetAsync(b->d_counter, 0, sizeof(int), c->stream));
        st_batch_args a;
        std::memset(&a, 0, sizeof a);
        a.seeds = b->d_seeds[spin ? 1 : 0];
        a.items = b->d_items;
        a.nitems = nitems;
        a.counter = b->d_counter;
        a.idx = c->d_idx0;
        a.x = d_x;
        a.state = d_state;
        a.lost = d_lost;
        a.turn0 = t0;
        a.nturn = nt;
        a.precheck = (first && turn0 == 0) ? 1 : 0;
        a.xtbt = d_tbt;
        a.tbt_nturn = nturn;
        a.tbt_turn0 = turn0;
        a.g = b->kg;
        a.gdev = b->tabs[0]->d_kg;
        int ev = -1;
        if (c->timing) {
            rc = timed_event(c, &ev);
            if (rc) return rc;
            HIPCHK(hipEventRecord(c->ev_pool[ev], c->stream));
        }
        if (perf)
            HIPCHK(st_perf::st_launch_track_batch(a, grid, block, spin, b->hexp, b->rad, c->stream));
        else
            HIPCHK(st_launch_track_batch(a, grid, block, spin, b->hexp, b->rad, c->stream));
        if (c->timing) {
            HIPCHK(hipEventRecord(c->ev_pool[ev + 1], c->stream));
            c->pending.push_back({ev, ev + 1});
            c->pending_work.push_back((int64_t)idx.size() * nt);
        }
        HIPCHK(hipMemcpyAsync(state.data(), d_state, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->timing) {
            rc = drain_timing(c);
            if (rc) return rc;
        }
        first = false;
        t0 += nt;
    }
    return ST_OK;
}

static int batch_check(const st_batch *b, const int *offsets)
{
    if (!b || !offsets) return fail(ST_ERR_ARG, "st_batch: bad argument");
    const int ntab = (int)b->tabs.size();
    if (offsets[0] != 0) return fail(ST_ERR_ARG, "st_batch: offsets[0] must be 0");
    for (int k = 0; k < ntab; ++k)
        if (offsets[k + 1] < offsets[k]) return fail(ST_ERR_ARG, "st_batch: offsets must not decrease");
    return ST_OK;
}

extern "C" int st_batch_default_config(st_batch_config *cfg)
{
    if (!cfg) return fail(ST_ERR_ARG, "st_batch_default_config: null");
    cfg->wg_waves = 1;
    cfg->waves_per_simd = 2;
    cfg->turn_block = 256;
    return ST_OK;
}

extern "C" int st_batch_track_device(st_batch *b, double *d_x, const int *offsets, int turn0, int nturn,
                                     int *d_state, int *d_lost, int flags, const st_batch_config *cfg)
{
    int rc = batch_check(b, offsets);
    if (rc) return rc;
    if (!d_x || !d_state || !d_lost || turn0 < 0 || nturn < 0) return fail(ST_ERR_ARG, "st_batch_track_device: bad argument");
    st_batch_config dc;
    st_batch_default_config(&dc);
    HIPCHK(hipSetDevice(b->c->device));
    return batch_core(b, d_x, offsets, turn0, nturn, d_state, d_lost, (flags & ST_TRACK_SPIN) != 0, cfg ? *cfg : dc,
                      nullptr, (flags & ST_TRACK_PERF) != 0);
}

extern "C" int st_batch_track_tbt_device(st_batch *b, double *d_x, const int *offsets, int turn0, int nturn,
                                         double *d_tbt, int *d_state, int *d_lost, int flags,
                                         const st_batch_config *cfg)
{
    int rc = batch_check(b, offsets);
    if (rc) return rc;
    if (!d_x || !d_tbt || !d_state || !d_lost || turn0 < 0 || nturn < 0)
        return fail(ST_ERR_ARG, "st_batch_track_tbt_device: bad argument");
    st_batch_config dc;
    st_batch_default_config(&dc);
    HIPCHK(hipSetDevice(b->c->device));
    return batch_core(b, d_x, offsets, turn0, nturn, d_state, d_lost, (flags & ST_TRACK_SPIN) != 0, cfg ? *cfg : dc,
                      d_tbt);
}

extern "C" int st_batch_track(st_batch *b, double *x, const int *offsets, int nturn, int *stable, int *lost_turn,
                              int *lost_post, int flags, const st_batch_config *cfg)
{
    int rc = batch_check(b, offsets);
    if (rc) return rc;
    const int np = offsets[b->tabs.size()];
    if (np < 0 || nturn < 0 || (np > 0 && (!x || !stable || !lost_turn || !lost_post)))
        return fail(ST_ERR_ARG, "st_batch_track: bad argument");
    if (np == 0) return ST_OK;
    st_ctx *c = b->c;
    HIPCHK(hipSetDevice(c->device));
    const bool spin = (flags & ST_TRACK_SPIN) != 0;
    const int ncol = spin ? 9 : 6;
    rc = ensure_particles(c, (size_t)np, ncol);
    if (rc) return rc;
    std::vector<int> lost(2 * (size_t)np);
    for (int p = 0; p < np; ++p) {
        lost[2 * p] = lost_turn[p];
        lost[2 * p + 1] = lost_post[p];
    }
    HIPCHK(hipMemcpyAsync(c->d_x, x, (size_t)np * ncol * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_state, stable, (size_t)np * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_lost, lost.data(), 2 * (size_t)np * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    st_batch_config dc;
    st_batch_default_config(&dc);
    rc = batch_core(b, c->d_x, offsets, 0, nturn, c->d_state, c->d_lost, spin, cfg ? *cfg : dc);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipMemcpyAsync(x, c->d_x, (size_t)np * ncol * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(stable, c->d_state, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(lost.data(), c->d_lost, 2 * (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
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

/*
 * Autotuner: times every candidate launch shape on the same calibration run
 * (calib_turns turns of the given particles, re-uploaded for each candidate)
 * and returns the fastest.  Candidates: 1, 2 or 4 wavefronts per workgroup;
 * 1 or 2 resident wavefronts per SIMD; turn blocks of 16, 64 or 256 turns.
 */
extern "C" int st_batch_autotune(st_batch *b, const double *x, const int *offsets, int calib_turns, int flags,
                                 st_batch_config *best, double *seconds_out, int *n_candidates)
{
    int rc = batch_check(b, offsets);
    if (rc) return rc;
    const int np = offsets[b->tabs.size()];
    if (np <= 0 || !x || !best || calib_turns <= 0) return fail(ST_ERR_ARG, "st_batch_autotune: bad argument");
    st_ctx *c = b->c;
    HIPCHK(hipSetDevice(c->device));
    const bool spin = (flags & ST_TRACK_SPIN) != 0;
    const int ncol = spin ? 9 : 6;
    rc = ensure_particles(c, (size_t)np, ncol);
    if (rc) return rc;
    const int wgs[3] = {1, 2, 4}, wps[2] = {1, 2}, tbs[3] = {16, 64, 256};
    std::vector<int> ones((size_t)np, 1), zeros(2 * (size_t)np, 0);
    double tbest = 1e300;
    int k = 0;
    for (int w : wgs)
        for (int o : wps)
            for (int tb : tbs) {
                st_batch_config cfg{w, o, tb};
                HIPCHK(hipMemcpyAsync(c->d_x, x, (size_t)np * ncol * sizeof(double), hipMemcpyHostToDevice, c->stream));
                HIPCHK(hipMemcpyAsync(c->d_state, ones.data(), (size_t)np * sizeof(int), hipMemcpyHostToDevice, c->stream));
                HIPCHK(hipMemcpyAsync(c->d_lost, zeros.data(), 2 * (size_t)np * sizeof(int), hipMemcpyHostToDevice,
                                      c->stream));
                HIPCHK(hipEventRecord(c->ev0, c->stream));
                rc = batch_core(b, c->d_x, offsets, 0, calib_turns, c->d_state, c->d_lost, spin, cfg);
                if (rc) return rc;
                HIPCHK(hipEventRecord(c->ev1, c->stream));
                HIPCHK(hipEventSynchronize(c->ev1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
                if (seconds_out) seconds_out[k] = ms * 1e-3;
                if (ms * 1e-3 < tbest) {
                    tbest = ms * 1e-3;
                    *best = cfg;
                }
                ++k;
            }
    if (n_candidates) *n_candidates = k;
    return ST_OK;
}

extern "C" __attribute__((visibility("hidden"))) hipStream_t st_internal_stream(st_ctx *c, int *device)
{
    if (device) *device = c ? c->device : 0;
    return c ? c->stream : nullptr;
}

extern "C" int st_surrogate_features(st_ctx *c, const double *d_x0, int ncol, const double *d_tunes, int n,
                                     double beta_x, double alpha_x, double beta_y, double alpha_y, float *d_feat)
{
    if (!c || n < 0 || (n > 0 && (!d_x0 || !d_tunes || !d_feat)) || ncol < 4 || !(beta_x > 0.) || !(beta_y > 0.))
        return fail(ST_ERR_ARG, "st_surrogate_features: bad argument");
    if (n == 0) return ST_OK;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(st_launch_features(d_x0, ncol, d_tunes, n, beta_x, alpha_x, beta_y, alpha_y, d_feat, c->stream));
    return ST_OK;
}
