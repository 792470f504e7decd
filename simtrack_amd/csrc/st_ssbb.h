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

/* one particle-in-cell pair (st_pic.hip): beam s's charge on the grid,
   felt by beam 1 - s */
struct pic_args {
    double *x[2];         /* beam rows; the slice is rows [start, start + count) */
    int start[2], count[2];
    const double *params; /* mean, rms of both beams, hx, hy (k_pic_params)      */
    int *key_in, *key_out, *val_in, *val_out; /* 9 x count contributions          */
    double *w, *ws;       /* weights in contribution order and in cell order      */
    void *temp;
    size_t temp_bytes;
    double *src, *gr, *gi, *sr, *si, *pr, *pi, *phi, *Ex, *Ey;
    const double *tw_fwd, *tw_inv; /* FFT1D twiddles for dir = 1 and dir = -1   */
    double nsm[2];        /* charge / rows of beam s as the source               */
    double gamma[2], bbscale[2];
};

hipError_t st_ss_collide(const ss_pair_args &a, hipStream_t stream);
/* only the partial sums of st_ss_collide (no kick); sets a.nb */
hipError_t st_ss_partials(ss_pair_args &a, hipStream_t stream);
hipError_t st_sort_cells_bytes(int n, size_t *bytes);
hipError_t st_sort_cells(void *temp, size_t bytes, const int *kin, int *kout, const int *vin, int *vout, int n,
                         hipStream_t stream);
hipError_t st_pic_params(const double *part, int nb0, int nb1, int m0, int m1, double *params, hipStream_t stream);
hipError_t st_pic_solve_and_kick(const pic_args &a, int s, hipStream_t stream);
/* params of the fixed, origin-centred grid (BeamBeam4D_PIC_NonOffset) */
hipError_t st_pic_fixed_params(double *params, double hx, double hy, hipStream_t stream);
hipError_t st_pic_drift(double *x, int start, int m, double S, int sgn, hipStream_t stream);
int st_ss_partial_blocks(int count);
hipError_t st_ss_zsums(const double *x, const int *idx, const int *seg_start, const int *seg_count,
                       const int *blk_seg, const int *blk_chunk, int nblk, const int *seg_first_blk,
                       const int *seg_nblk, int nseg, double *part, double *Z, hipStream_t stream);
hipError_t st_ss_sort_by_z(double *x, int n, ss_sort_buffers &s, hipStream_t stream);
hipError_t st_ss_group_by_slice(const double *x, int n, const double *d_zb, int nslice, ss_sort_buffers &s,
                                hipStream_t stream);
hipError_t st_ss_sort_temp_bytes(int n, size_t *bytes);

#endif
