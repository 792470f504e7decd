/*
 * st_ssbb.h -- launch interface of the strong-strong beam-beam kernels.
 */
#ifndef ST_SSBB_H
#define ST_SSBB_H

#include <hip/hip_runtime.h>
#include <stddef.h>

#define SS_CHUNK 2048 /* rows per partial-sum block */

/* one slice pair: beam 0 drifts by +S, beam 1 by -S (drift = 0: no drift) */
struct ss_pair_args {
    double *x[2];        /* beam rows (6 doubles each)                        */
    const int *idx[2];   /* row list of the beam, or null: rows in order      */
    int start[2];        /* first entry of the slice in the row list           */
    int count[2];        /* rows in the slice                                  */
    int nb[2];           /* partial blocks per beam (set by st_ss_collide)     */
    double S;
    int drift;
    double c2Nrg[2];     /* 2 N_other rp / gamma of the kicked beam            */
    double bbscale[2];
    double *part;        /* (nb[0] + nb[1]) x 4 partial sums                   */
};

struct ss_sort_buffers {
    double *dkey_in, *dkey_out;
    int *ikey_in, *ikey_out;
    int *row_in, *row_out;
    double *rows;        /* gathered rows (n x 6)                              */
    int *counts;         /* rows per slice number (nslice + 1)                  */
    void *temp;
    size_t temp_bytes;
};

hipError_t st_ss_collide(const ss_pair_args &a, hipStream_t stream);
int st_ss_partial_blocks(int count);
hipError_t st_ss_zsums(const double *x, const int *idx, const int *seg_start, const int *seg_count,
                       const int *blk_seg, const int *blk_chunk, int nblk, const int *seg_first_blk,
                       const int *seg_nblk, int nseg, double *part, double *Z, hipStream_t stream);
hipError_t st_ss_sort_by_z(double *x, int n, ss_sort_buffers &s, hipStream_t stream);
hipError_t st_ss_group_by_slice(const double *x, int n, const double *d_zb, int nslice, ss_sort_buffers &s,
                                hipStream_t stream);
hipError_t st_ss_sort_temp_bytes(int n, size_t *bytes);

#endif
