/*
 * st_rng.h -- counter-based random numbers for synchrotron radiation:
 * normal deviates for quantum excitation, photon counts and energies.
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

/* ------------------------------------------------- photon emission */

/*
 * The photon sampling of radiate2 (simtrack.cpp:4290-4366).  The reference
 * draws a Poisson number of photons (gsl_ran_poisson) and one uniform per
 * photon (gsl_rng_uniform) from its global GSL streams.  Here they come from
 * a sub-stream of the draw that radiate2's call takes, so the numbers do not
 * depend on scheduling: uniform j of draw d of a particle is half (j & 1) of
 * the Philox block with counter (d low, particle, 0x50480000 | (j >> 1),
 * d high), distinct from the normals' 0x53544e44.  The Poisson count uses
 * Knuth's product-of-uniforms method (exact), on chunks of mean <= 30 summed
 * (a sum of independent Poisson counts is Poisson); the sub-stream holds at
 * most 2^17 uniforms, after which the count stops.
 */
ST_FN double st_rng_photon_u(uint64_t seed, uint32_t particle, uint64_t draw, uint32_t j)
{
    st_u32x4 c;
    c.v[0] = (uint32_t)draw;
    c.v[1] = particle;
    c.v[2] = 0x50480000u | ((j >> 1) & 0xffffu);
    c.v[3] = (uint32_t)(draw >> 32);
    st_u32x4 r = st_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return (j & 1) ? st_u01_53(r.v[2], r.v[3]) : st_u01_53(r.v[0], r.v[1]);
}

#define ST_PHOTON_UMAX (1u << 17)

/* the photon-energy table of radiate2 (cpp:4307-4323): the fraction xt of
   photons below energy yt (in units of the critical energy) */
ST_CONST_TABLE double st_rad2_xt[43] = {
    0.00123, 0.0123, 0.0265, 0.0571, 0.1228, 0.1544, 0.194, 0.221, 0.2619, 0.2893, 0.327, 0.36914, 0.4074,
    0.4417, 0.47273, 0.5209, 0.5791, 0.62385, 0.6517, 0.6905, 0.7165, 0.7494, 0.7726, 0.8104, 0.8401,
    0.85954, 0.8835, 0.892, 0.89665, 0.8997, 0.90397, 0.9067, 0.9132, 0.93444, 0.9561, 0.9661, 0.9768,
    0.98734, 0.993, 0.99776, 0.999574, 0.9999156, 1.0};
ST_CONST_TABLE double st_rad2_yt[43] = {
    1e-09, 1e-06, 1e-05, 0.0001, 0.001, 0.002, 0.004, 0.006, 0.01, 0.0136, 0.02, 0.0292, 0.04, 0.052, 0.065,
    0.09, 0.13, 0.17, 0.2, 0.25, 0.29, 0.35, 0.4, 0.5, 0.6, 0.68, 0.8, 0.85, 0.88, 0.9, 0.93, 0.95, 1.0, 1.2,
    1.5, 1.7, 2.0, 2.5, 3.0, 4.0, 5.5, 7.0, 10.0};

/* splint, cpp:333-356: linear interpolation in the table after a bisection */
ST_FN double st_rad2_splint(double X)
{
    int klo = 0, khi = 42;
    while (khi - klo > 1) {
        int k = (khi + klo) / 2;
        if (st_rad2_xt[k] > X)
            khi = k;
        else
            klo = k;
    }
    double H = st_rad2_xt[khi] - st_rad2_xt[klo];
    double A = (st_rad2_xt[khi] - X) / H;
    double B = (X - st_rad2_xt[klo]) / H;
    return A * st_rad2_yt[klo] + B * st_rad2_yt[khi];
}

/* Poisson(nr) photon count of one radiate2 call (gsl_ran_poisson, cpp:4343);
   *j returns the next unused uniform of the sub-stream */
ST_FN int st_rad2_count(uint64_t seed, uint32_t particle, uint64_t draw, double nr, uint32_t *j)
{
    uint32_t u = 0;
    int ng = 0;
    double rem = nr;
    while (rem > 0. && u < ST_PHOTON_UMAX) {
        double lam = rem > 30. ? 30. : rem;
        rem = rem - lam;
        double lim = st_exp(-lam), p = 1.0;
        int k = 0;
        do {
            k++;
            p = p * st_rng_photon_u(seed, particle, draw, u++);
        } while (p > lim && u < ST_PHOTON_UMAX);
        ng += k - 1;
    }
    *j = u;
    return ng;
}

/* photon energy in units of the critical energy from a uniform r
   (cpp:4348-4354): (r / facg)^3 below r = 0.26, the table above */
ST_FN double st_rad2_energy(double r)
{
    const double facg = 0x1.3b496a0938c91p+0; /* 12 sqrt(3) / (2^(1/3) 5 Gamma(1/3)) */
    return r < 0.26 ? st_powi(r / facg, 3) : st_rad2_splint(r);
}

/* the summed energy, in units of the critical energy, of the photons one
   radiate2 call emits (cpp:4343-4358) */
ST_FN double st_rad2_photons(uint64_t seed, uint32_t particle, uint64_t draw, double nr)
{
    uint32_t j;
    int ng = st_rad2_count(seed, particle, draw, nr, &j);
    double etot = 0.;
    for (int i = 0; i < ng && j < ST_PHOTON_UMAX; i++)
        etot = etot + st_rad2_energy(st_rng_photon_u(seed, particle, draw, j++));
    return etot;
}

#endif /* ST_RNG_H */
