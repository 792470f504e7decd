/*
 * st_kernels.h -- launch interface between the C-ABI layer and the kernels.
 */
#ifndef ST_KERNELS_H
#define ST_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "st_table_format.h"

/* arguments of one tracking launch (passed by value as the kernel argument) */
struct st_track_args {
    const st_op *ops;       /* nelem ops                                      */
    const double *par;      /* payload                                        */
    const double *ap;       /* APx, APy per element (null: no loss test)      */
    const double *x04;      /* radial mode: on-momentum z per element, else 0 */
    int nelem;
    int nact;               /* threads: active particles                      */
    const int *idx;         /* active particle indices (null: 0..nact-1)      */
    double *x;              /* particle state, AoS (6 or 9 per particle)      */
    int *state;             /* stable flag per particle                       */
    int *lost;              /* (lost_turn, lost_post) per particle            */
    int64_t draws_per_turn; /* radiation: quantum-excitation draws per turn   */
    int turn0, nturn;       /* turns [turn0, turn0 + nturn)                   */
    int precheck;           /* apply the 0.1 m start-of-tracking check        */
    double *xtbt;           /* turn-by-turn store or null                     */
    int tbt_bpm;            /* -1: at turn start; else after this element     */
    int tbt_nturn;          /* rows per particle in xtbt                      */
    int tbt_turn0;          /* turn of row 0                                  */
    st_kglobals g;
    const st_kglobals *gdev; /* the same globals in device memory             */
};

/* one table (lattice seed) of a multi-table launch */
struct st_seed_desc {
    const st_op *ops;
    const double *par;
    const double *ap;
    int nelem;
    int reserved;
    int64_t draws_per_turn;
};

/* arguments of the persistent multi-table launch */
struct st_batch_args {
    const st_seed_desc *seeds;
    const int *items;       /* 3 ints per item: table, start in idx, count (<= 64) */
    int nitems;
    int *counter;           /* next item to claim (zero at launch)            */
    const int *idx;         /* particle indices                               */
    double *x;
    int *state;
    int *lost;
    int turn0, nturn;
    int precheck;
    double *xtbt;           /* turn-by-turn store at turn start, or null      */
    int tbt_nturn, tbt_turn0;
    st_kglobals g;
    const st_kglobals *gdev; /* the same globals in device memory             */
};

hipError_t st_launch_track(const st_track_args &a, bool spin, bool hexp, bool rad, hipStream_t stream);
hipError_t st_launch_track_batch(const st_batch_args &b, int grid, int block, bool spin, bool hexp, bool rad,
                                 hipStream_t stream);
hipError_t st_launch_radial_x04(const st_track_args &a, double *x04, bool hexp, hipStream_t stream);
/* the perf-mode kernels (st_kernels_perf.hip): the same maps with FMA
   contraction and division by reciprocal multiplication */
namespace st_perf {
hipError_t st_launch_track(const st_track_args &a, bool spin, bool hexp, bool rad, hipStream_t stream);
hipError_t st_launch_track_batch(const st_batch_args &b, int grid, int block, bool spin, bool hexp, bool rad,
                                 hipStream_t stream);
}
hipError_t st_launch_compact(const int *idx, int n, const int *state, int *block_buf, int *idx_out, int *d_total,
                             hipStream_t stream);

#endif
