/*
 * st_capi.cpp -- host side of the C ABI (include/simtrack_amd.h): error
 * reporting, the reference's global variables (GP) and the device context
 * with its grow-only work buffers.  The other parts of the ABI live in
 * st_lower.cpp (lowering to the packed table), st_track.cpp (tracking
 * drivers), st_scan.cpp (tune finder, closed orbit, one-turn and element
 * maps, turn-by-turn scans), st_batch.cpp (many lattices per launch) and
 * st_ss_host.cpp (strong-strong beam-beam).
 */
#include "st_capi_internal.h"

/* ------------------------------------------------------------ errors */

static thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}


extern "C" const char *st_last_error(void) { return g_err.c_str(); }

/* error reporting for the library's other translation units (not exported
   in the public header) */
extern "C" __attribute__((visibility("hidden"))) int st_internal_fail(int code, const char *msg)
{
    return fail(code, "%s", msg);
}

struct st_ctx;
/* the context's stream and device, for the library's other translation units */
extern "C" __attribute__((visibility("hidden"))) hipStream_t st_internal_stream(st_ctx *c, int *device);
extern "C" int st_abi_version(void) { return ST_ABI_VERSION; }
extern "C" void st_struct_sizes(int *element_bytes, int *globals_bytes)
{
    if (element_bytes) *element_bytes = (int)sizeof(st_element);
    if (globals_bytes) *globals_bytes = (int)sizeof(st_globals);
}

/* ------------------------------------------------------ reference GP */

/* 4th-order Yoshida coefficients, simtrack.cpp:29-32 (run-time pow, as at the
   reference's static initialisation) */
double Fdrift1, Fdrift2, Fkick1, Fkick2;
void init_yoshida()
{
    static bool done = false;
    if (done) return;
    volatile double two = 2.e0, third = 1.0 / 3.0;
    Fdrift1 = 1.e0 / (2.e0 * (2.e0 - std::pow(two, third)));
    Fdrift2 = 0.5e0 - Fdrift1;
    Fkick1 = 2.e0 * Fdrift1;
    Fkick2 = 1.e0 - 2e0 * Fkick1;
    done = true;
}

extern "C" void st_globals_default(st_globals *gp)
{
    /* GlobalVariables(), simtrack.h:31-58 */
    std::memset(gp, 0, sizeof *gp);
    gp->A = 1;
    gp->Q = 1;
    gp->m0 = 1.6726216e-27;
    gp->m = 1.6726216e-27;
    gp->q = 1.6021765e-19;
    gp->gamma = 106.;
    gp->beta = 0.999955;
    gp->energy0 = 938.27201;
    gp->energy = 99456.8;
    gp->p = 1.;
    gp->cp = 1.;
    gp->brho = 331.738;
    gp->G = 1.7928474;
    gp->H_expand = 0;
    gp->radiate = 0;
    gp->quad_fringe = 0;
    gp->rng_seed = 0;
}

extern "C" void st_globals_set_energy(st_globals *gp, double gamma, double E0, double Q, double A)
{
    /* Set_RefPartEnergy, simtrack.cpp:87-108 */
    gp->Q = (int32_t)Q;
    gp->A = (int32_t)A;
    gp->m0 = E0 * 1.0e6 * 1.6021765e-19 / 299792458. / 299792458.;
    gp->m = gamma * gp->m0;
    gp->q = Q * 1.6021765e-19;
    gp->gamma = gamma;
    gp->beta = std::sqrt(1.0 - 1.0 / gamma / gamma);
    gp->energy0 = E0;
    gp->energy = gamma * gp->energy0;
    gp->cp = gp->energy * gp->beta;
    gp->p = gp->energy * gp->beta * 1.0e6 * 1.6021765e-19 / 299792458.;
    gp->brho = 3.33564095198 * gp->beta * (gp->energy / 1000) / gp->Q;
}

/* ----------------------------------------------------------- context */

extern "C" int st_ctx_create(int device, st_ctx **out)
{
    if (!out) return fail(ST_ERR_ARG, "st_ctx_create: out is null");
    *out = nullptr;
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(ST_ERR_ARG, "st_ctx_create: device %d not in [0, %d)", device, n);
    HIPCHK(hipSetDevice(device));
    st_ctx *c = new st_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipHostMalloc((void **)&c->h_total, sizeof(int)) != hipSuccess) {
        st_ctx_destroy(c);
        return fail(ST_ERR_DEVICE, "st_ctx_create: stream/event/pinned allocation failed");
    }
    init_yoshida();
    int rc = noalloc_scratch_init(device);
    if (rc) {
        st_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return ST_OK;
}

/* fold finished timed launches into the counters (stream must be idle) */
int drain_timing(st_ctx *c)
{
    for (size_t i = 0; i < c->pending.size(); ++i) {
        float ms = 0.f;
        hipError_t e = hipEventElapsedTime(&ms, c->ev_pool[c->pending[i].first], c->ev_pool[c->pending[i].second]);
        if (e != hipSuccess) return fail(ST_ERR_DEVICE, "kernel timing: %s", hipGetErrorString(e));
        c->k_seconds += ms * 1e-3;
        c->k_launches += 1;
        c->k_work += c->pending_work[i];
    }
    c->pending.clear();
    c->pending_work.clear();
    return ST_OK;
}

int timed_event(st_ctx *c, int *idx)
{
    size_t used = 2 * c->pending.size();
    if (used + 1 >= c->ev_pool.size()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c->ev_pool.push_back(e);
        HIPCHK(hipEventCreate(&e));
        c->ev_pool.push_back(e);
    }
    *idx = (int)used;
    return ST_OK;
}

extern "C" int st_ctx_set_ss_lanes(st_ctx *c, int lanes)
{
    if (!c || lanes < 1 || lanes > SS_LANES)
        return fail(ST_ERR_ARG, "st_ctx_set_ss_lanes: lanes %d not in [1, %d]", lanes, SS_LANES);
    c->ss_max_lanes = lanes;
    return ST_OK;
}

extern "C" int st_ctx_set_timing(st_ctx *c, int enable)
{
    if (!c) return fail(ST_ERR_ARG, "st_ctx_set_timing: null context");
    c->timing = enable != 0;
    return ST_OK;
}

extern "C" int st_ctx_kernel_stats(st_ctx *c, double *sec, int64_t *launches, int64_t *work)
{
    if (!c) return fail(ST_ERR_ARG, "st_ctx_kernel_stats: null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    int rc = drain_timing(c);
    if (rc) return rc;
    if (sec) *sec = c->k_seconds;
    if (launches) *launches = c->k_launches;
    if (work) *work = c->k_work;
    c->k_seconds = 0.0;
    c->k_launches = c->k_work = 0;
    return ST_OK;
}

static void free_bufs(st_ctx *c)
{
    hipFree(c->d_x);
    hipFree(c->d_state);
    hipFree(c->d_lost);
    hipFree(c->d_idx0);
    hipFree(c->d_idx1);
    hipFree(c->d_block);
    hipFree(c->d_total);
    hipFree(c->d_tbt);
    hipFree(c->d_x04);
    hipFree(c->d_ss);
    c->d_ss = nullptr;
    c->cap_ss = 0;
    c->d_x = nullptr;
    c->d_state = c->d_lost = c->d_idx0 = c->d_idx1 = c->d_block = c->d_total = nullptr;
    c->d_tbt = c->d_x04 = nullptr;
    c->cap_x = c->cap_p = c->cap_tbt = c->cap_x04 = 0;
}

extern "C" int st_ctx_destroy(st_ctx *c)
{
    if (!c) return ST_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    free_bufs(c);
    if (c->h_total) hipHostFree(c->h_total);
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    for (int l = 0; l < SS_LANES; ++l) {
        if (c->lane[l]) {
            hipStreamSynchronize(c->lane[l]);
            hipStreamDestroy(c->lane[l]);
        }
        if (c->lane_ev[l]) hipEventDestroy(c->lane_ev[l]);
    }
    if (c->fork_ev) hipEventDestroy(c->fork_ev);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return ST_OK;
}

extern "C" int st_ctx_sync(st_ctx *c)
{
    if (!c) return fail(ST_ERR_ARG, "st_ctx_sync: null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ST_OK;
}

extern "C" int st_last_gpu_seconds(st_ctx *c, double *s)
{
    if (!c || !s) return fail(ST_ERR_ARG, "st_last_gpu_seconds: null argument");
    *s = c->last_gpu_s;
    return ST_OK;
}

int ensure_particles(st_ctx *c, size_t np, int ncol)
{
    size_t need_x = np * (size_t)ncol;
    if (need_x > c->cap_x) {
        hipFree(c->d_x);
        c->d_x = nullptr;
        c->cap_x = 0;
        HIPCHK(hipMalloc(&c->d_x, need_x * sizeof(double)));
        c->cap_x = need_x;
    }
    if (np > c->cap_p) {
        hipFree(c->d_state);
        hipFree(c->d_lost);
        hipFree(c->d_idx0);
        hipFree(c->d_idx1);
        hipFree(c->d_block);
        hipFree(c->d_total);
        c->d_state = c->d_lost = c->d_idx0 = c->d_idx1 = c->d_block = c->d_total = nullptr;
        c->cap_p = 0;
        size_t nb = (np + 1023) / 1024 + 1;
        HIPCHK(hipMalloc(&c->d_state, np * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_lost, 2 * np * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_idx0, np * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_idx1, np * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_block, nb * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_total, sizeof(int)));
        c->cap_p = np;
    }
    return ST_OK;
}
