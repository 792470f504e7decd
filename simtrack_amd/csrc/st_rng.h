/*
 * st_rng.h -- counter-based normal deviates for synchrotron-radiation
 * quantum excitation.
 *
 * The reference draws gsl_ran_gaussian(gsl_r2, 1.0) from one global,
 * sequential MT19937 stream shared by all particles (simtrack.cpp:4275,
 * seeded in Init_GSL_Seeds, cpp:316-329).  A stream consumed in particle
 * order cannot be reproduced by parallel particles, so both the GPU kernels
 * and the CPU oracle use this stream instead: draw number `counter` of
 * particle `particle` under key `seed` is a pure function of the three,
 * independent of scheduling.
 *
 * Generator: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11), counter (draw low 32 bits, particle,
 * 0x53544e44, draw high 32 bits), key (seed low, seed high).  The tracking
 * drivers number the draws of a particle as turn * D + k, D the number of
 * draws per turn and k the draw within the turn, so a run split into turn
 * ranges draws the same numbers as one call.  Normal deviate: Box-Muller
 * z = sqrt(-2 log u1) cos(2 pi u2) with u1, u2 53-bit uniforms in (0,1),
 * evaluated with st_libm.h so host and device agree bit for bit.
 *
 * Requires st_libm.h (and its ST_* macros) to be included first.
 */
#ifndef ST_RNG_H
#define ST_RNG_H

#include <stdint.h>

typedef struct { uint32_t v[4]; } st_u32x4;

ST_FN void st_mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

ST_FN st_u32x4 st_philox4x32_10(st_u32x4 c, uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        st_mulhilo32(0xD2511F53u, c.v[0], &hi0, &lo0);
        st_mulhilo32(0xCD9E8D57u, c.v[2], &hi1, &lo1);
        st_u32x4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

/* (k + 1/2) 2^-53 from the top 53 bits of (a,b): uniform in (0,1) */
ST_FN double st_u01_53(uint32_t a, uint32_t b)
{
    uint64_t k = (((uint64_t)a << 32) | (uint64_t)b) >> 11;
    return ((double)k + 0.5) * 0x1p-53;
}

ST_FN double st_rng_normal(uint64_t seed, uint32_t particle, uint64_t draw)
{
    st_u32x4 c;
    c.v[0] = (uint32_t)draw;
    c.v[1] = particle;
    c.v[2] = 0x53544e44u;
    c.v[3] = (uint32_t)(draw >> 32);
    st_u32x4 r = st_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double u1 = st_u01_53(r.v[0], r.v[1]);
    double u2 = st_u01_53(r.v[2], r.v[3]);
    double rad = ST_SQRT(-2.0 * st_log(u1));
    double s, co;
    st_sincos(2.0 * ST_PIO2_1 * 2.0 * u2, &s, &co);  /* 2 pi u2, 2 pi rounded */
    return rad * co;
}

#endif /* ST_RNG_H */
