/*
 * st_table_format.h -- packed element table read by the tracking kernels.
 *
 * A Line of st_element records is lowered to one op per element: a 32-byte
 * integer header (kind, integration steps, order, flags, payload offset)
 * and a run of doubles in one payload array.  Every lane of a wavefront
 * reads the same op at the same time, so the header and payload are loaded
 * through the scalar unit into SGPRs; nothing per element lives in VGPRs.
 *
 * Payload contents are the element constants the reference recomputes on
 * every call (Fkick*KNL/Nint, Fdrift*L/Nint, L/GP.beta, cos/sin of bend
 * angles, beam-beam slice centroids, misalignment matrices, ...), evaluated
 * once on the host with the reference's own expressions, so each value is
 * bit-identical to what the reference computes per particle.
 */
#ifndef ST_TABLE_FORMAT_H
#define ST_TABLE_FORMAT_H

#include <stdint.h>

/* op kinds and payload layouts (offsets from st_op.po) */
enum st_kind {
    K_NOP = 0,
    K_DRIFT,          /* 0 L, 1 L/beta                                            */
    K_DRIFT0,         /* zero-length drift: no payload                            */
    K_QUAD,           /* thick: 0 d1,1 d1b,2 d2,3 d2b, 4 kn*F1/N,5 ks*F1/N,       */
    K_SEXT,           /*        6 kn*F2/N,7 ks*F2/N, 8 kn,9 ks, 10 Nint, 11 k1l/L */
    K_OCT,
    K_QUAD_THIN,      /* 0 kn, 1 ks, 2 Nint                                       */
    K_SEXT_THIN,
    K_OCT_THIN,
    K_MULT,           /* 0 d1,1 d1b,2 d2,3 d2b, 4 Nint, 5 mask0, 6.. kn1 ks1 kn2 ks2 KNL KNSL */
    K_MULT_THIN,      /* 0.. KNL KNSL (mask1 = non-zero orders)                   */
    K_SBEND_EXACT,    /* 0 ef1,1 ef2, 2 L,3 href,4 1/href,5 cos,6 sin,7 L/beta    */
    K_SBEND_SLICED,   /* 0 ef1,1 ef2,2 href,3 Lint,4 Angle/Nint,5 Nint,
                         exact: 6 hs,7 1/hs,8 cos,9 sin,10 Lint/beta
                         H_expand: 6 d1,7 d1b,8 d2,9 d2b,10 kl1,11 kl2           */
    K_GSBEND_EXACT,   /* 0 ef1,1 ef2, 2.. general piece (7); radiating: the piece
                         of one slice, 9 href,10 K0L/Nint,11 Lint                 */
    K_GSBEND_YOSHIDA, /* 0 ef1,1 ef2, 2 d1,3 d1b,4 d2,5 d2b,6 kl1,7 kl2,8 href,9 hreal;
                         radiating: 10 K0L/Nint,11 Lint                           */
    K_BENDMULT,       /* 0 ef1,1 ef2,2 href,3 Lint,4 Nint,5 Angle,6 Angle/Nint,7 mask0,
                         exact: 8..13 piece1 (L,h,1/h,cos,sin,L/beta), 14..19 piece2
                         H_expand: 8 d1,9 d1b,10 d2,11 d2b,12 kl1,13 kl2
                         20.. kn1 ks1 kn2 ks2 KNL KNSL                            */
    K_BENDMULT_THIN,  /* 0 Angle, 1.. KNL KNSL                                    */
    K_GSBENDMULT,     /* 0 ef1,1 ef2,2 href,3 hreal,4 Lint,5 Nint,6 K0L/Nint,
                         exact: 8..14 piece1 (general), 15..21 piece2
                         H_expand: 8 d1,9 d1b,10 d2,11 d2b,12 kl1,13 kl2
                         22.. kn1 ks1 kn2 ks2                                     */
    K_KICK,           /* 0 L/2,1 (L/2)/beta, 2 hkick, 3 vkick                     */
    K_RFCAV,          /* 0 L/2,1 Lb, 2 |VRF*Q|, 3 -2*PI*FRF, 4 PHASE0              */
    K_CRABRF,         /* 0 L/2,1 Lb, 2 VRF*Q, 3 2*PI*FRF, 4 PHASE0, 5 2*PI*FRF/c    */
    K_CCMULT,         /* 0 L/2,1 Lb, 2 2*PI*FRF, 3 PHASE0, 4 scale, 5 brho, 6..13 B2..A5 */
    K_LBT,            /* 0 sin,1 cos,2 tan,3 cos(theta)                           */
    K_ILBT,           /* 0 sin,1 cos                                              */
    K_BB4D,           /* 0 2N rp/gamma, 1 sigmax, 2 sigmay, 3 BBSCALE             */
    K_BB6D,           /* see bb6d()                                               */
    K_BB6DA,          /* see bb6d_angle()                                         */
    K_MATRIX,         /* 0..5 XCO_IN, 6..41 M66, 42..47 XCO_OUT                   */
    K_ROTAT,          /* 0 L/2, 1 Lb, 2..10 rotation matrix                       */
    K_PATCH,          /* 0 DX,1 DY,2 DS, 3 rotate (1/0), 4 cos,5 sin,6 tan of THETA */
    K_SROTAT,         /* 0 cos, 1 sin of PSI                                      */
    K_XROTAT,         /* 0 cos, 1 sin, 2 tan of PHI                               */
    K_COOLING,        /* 0 1 - ALPHA                                              */
    K_LRBB,           /* 0 SEPX, 1 SEPY, 2 2N rp/gamma, 3 sigmax, 4 sigmay, 5 BBSCALE */
    K_ELENS,          /* 0 Le/N/2, 1 (Le/N/2)/beta, 2 2(Ne/N) rp/gamma, 3 sigmax,
                         4 sigmay, 5 1+beta_e, 6 BBSCALE (nint = NSLICE)          */
    K_HELENS,         /* 0 Le/N/2, 1 (Le/N/2)/beta, 2 Theta_max, 3 RIN, 4 ROUT,
                         5 RIN*RIN, 6 ROUT*ROUT-RIN*RIN (nint = NSLICE)           */
    K_ACDIP,          /* 0 L, 1 L/beta, 2 L/2, 3 (L/2)/beta, 4 HKICKMAX, 5 VKICKMAX,
                         6 2*3.14159265*NUD, 7 PHID, 8 TURNS, 9 TURNE             */
    K_ACMULT,         /* 0 L/2, 1 (L/2)/beta, 2 KL, 3 KSL, 4 2*PI, 5 TTURNS, 6 PHI0
                         (norder = Norder)                                        */
    K_SOLEN,          /* 0 L, 1 L/Nint, 2 q*Bz/2, 3 m0, 4 p, 5 L/beta/c,
                         6 Bz*Lint/brho (nint = Nint)                             */
    K_WIGGLER,        /* 0 h = L/Nint, 1..8 the field expansion's coefficients,
                         9 kx, 10 ky, 11 kz, 12 phiz0, 13 brho, 14 Wu's Intp
                         Ax coefficient, 15 Wu's Intp Ay coefficient,
                         16 (kx/ky)*b0, 17 -(kz*b0/ky) (nint = Nint)              */
    K_COUNT
};

#define ST_BM_ARR 20   /* K_BENDMULT kick arrays */
#define ST_GBM_ARR 22  /* K_GSBENDMULT kick arrays */
#define ST_GPIECE 7    /* general_sbend_exact piece: L, rhoc, b0, b0*rhoc, cos, sin, L/beta */

/* op flags */
#define ST_F_MIS1       0x1   /* GtoL1/LtoG1: offset and/or roll             */
#define ST_F_MIS1_OFS   0x2   /*   |DX|+|DY| > 1e-10                          */
#define ST_F_MIS1_ROT   0x4   /*   |DPSI| > 1e-10                             */
#define ST_F_MIS2       0x8   /* GtoL2/LtoG2: pitched (general) frame change  */
#define ST_F_FRINGE     0x10  /* quadrupole hard-edge fringe (GP.quad_fringe) */
#define ST_F_ROUND      0x20  /* beam-beam: BB4D/BB6D kick evaluation flags   */
#define ST_F_SPIN_THIN  0x40  /* spin pass of a thin element                  */
#define ST_F_RAD        0x80  /* element radiates (GP.radiate branch taken)   */

typedef struct st_op {
    int32_t kind;
    int32_t nint;     /* integration steps                                    */
    int32_t norder;   /* multipole order, or number of beam-beam slices       */
    int32_t flags;
    int32_t po;       /* payload offset of the element body                   */
    int32_t pm;       /* payload offset of the misalignment block (-1: none)  */
    int32_t mask1;    /* multipoles: bit i set if kick-1 order i is non-zero  */
    int32_t mask2;    /*             bit i set if kick-2 order i is non-zero  */
} st_op;

/* misalignment blocks: MIS1 -> DX, DY, cosT, sinT
                        MIS2 -> RMin[36], addin[5], RMout[36], addout[5]      */
#define ST_MIS2_SIZE 82

/* kernel-wide reference-particle constants */
typedef struct st_kglobals {
    double beta, gamma, sg;       /* GP.beta, GP.gamma, sqrt(gamma^2-1)       */
    double G;                     /* GP.G                                     */
    double energy;                /* GP.energy                                */
    double brho;
    double bg2;                   /* GP.beta*GP.gamma*GP.beta*GP.gamma        */
    double rad_c2;                /* Cr*(..)^... helpers for radiation        */
    double e3;                    /* pow(GP.energy/1000., 3)                  */
    double Cr, Cf;
    uint64_t rng_seed;
    int32_t H_expand, radiate;
} st_kglobals;

#endif
