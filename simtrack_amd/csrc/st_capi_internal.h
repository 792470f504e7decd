/*
 * st_capi_internal.h -- what the host-side translation units of the C ABI
 * share: the context and table structures, error reporting and the helpers
 * that more than one of them calls.  Not installed; the public interface is
 * include/simtrack_amd.h.
 */
#ifndef ST_CAPI_INTERNAL_H
#define ST_CAPI_INTERNAL_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "simtrack_amd.h"

#define ST_FN static inline
#define ST_CONST_TABLE static const
#define ST_FMA(a, b, c) std::fma((a), (b), (c))
#define ST_RINT(x) std::nearbyint(x)
#define ST_SQRT(x) std::sqrt(x)
#include "st_libm.h"
#include "st_kernels.h"
#include "st_ssbb.h"
#include "st_tune.h"
#include "st_table_format.h"

#define ST_HIDDEN __attribute__((visibility("hidden")))

/* Sets the message st_last_error() returns and returns code. */
ST_HIDDEN int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

#define HIPCHK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return fail(ST_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

/* 4th-order Yoshida coefficients (set by init_yoshida) */
extern ST_HIDDEN double Fdrift1, Fdrift2, Fkick1, Fkick2;
ST_HIDDEN void init_yoshida();

/* ------------------------------------------------------- structures */

#define SS_LANES 8 /* slice pairs in flight at once */

struct st_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_gpu_s = 0.0;
    /* grow-only work buffers */
    double *d_x = nullptr;
    size_t cap_x = 0;
    int *d_state = nullptr, *d_lost = nullptr, *d_idx0 = nullptr, *d_idx1 = nullptr, *d_block = nullptr;
    int *d_total = nullptr, *h_total = nullptr;
    size_t cap_p = 0;
    double *d_tbt = nullptr;
    size_t cap_tbt = 0;
    double *d_x04 = nullptr;
    size_t cap_x04 = 0;
    /* per-launch kernel timing */
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> pending;  /* indices into ev_pool */
    std::vector<int64_t> pending_work;
    double k_seconds = 0.0;
    int64_t k_launches = 0, k_work = 0;
    /* strong-strong beam-beam work area (grow-only) */
    char *d_ss = nullptr;
    size_t cap_ss = 0;
    /* streams for slice pairs that touch disjoint slices (created on first
       use, default priority) and the events that fork and join them */
    hipStream_t lane[SS_LANES] = {};
    hipEvent_t lane_ev[SS_LANES] = {};
    hipEvent_t fork_ev = nullptr;
    int ss_max_lanes = SS_LANES;   /* st_ctx_set_ss_lanes */
};

struct dev_table {
    st_op *d_ops = nullptr;
    double *d_par = nullptr;
    std::vector<st_op> ops;
    std::vector<double> par;
};

struct st_table {
    st_ctx *ctx = nullptr;
    st_globals gp;
    st_kglobals kg;
    int n = 0;
    dev_table orbit, spin;
    double *d_ap = nullptr;
    st_kglobals *d_kg = nullptr; /* kg in device memory, read by the kernels */
    int64_t draws_per_turn = 0;
    int radial_ok = 1;
};


/* context helpers (st_capi.cpp) */
ST_HIDDEN int ensure_particles(st_ctx *c, size_t np, int ncol);
ST_HIDDEN int drain_timing(st_ctx *c);
ST_HIDDEN int timed_event(st_ctx *c, int *idx);

/* tracking helpers (st_track.cpp); radial: 0 none, 1 Track_Radial's own
   x04, RADIAL_GIVEN the per-element z offsets already in the context's x04
   buffer */
#define RADIAL_GIVEN 2
ST_HIDDEN int turn_block(int n_elems, int np);
ST_HIDDEN int track_device_core(st_table *t, double *d_x, int np, int turn0, int nturn, int *d_state, int *d_lost,
                                bool spin, int radial, double *d_tbt, int tbt_bpm, int tbt_nturn, bool bare = false);
ST_HIDDEN int noalloc_scratch_init(int device);
ST_HIDDEN int track_host(st_table *t, double *x, int np, int nturn, int *stable, int *lost_turn, int *lost_post,
                         bool spin, int radial, double *x_tbt, int tbt_bpm, bool bare = false);

#endif
