/* Generated by gen_libm_tables.py from mpmath (400 digits). Do not edit. */
#ifndef ST_LIBM_TABLES_H
#define ST_LIBM_TABLES_H

/* pi/2 = P1 + P2 + P3 + P4 + O(2^-210) */
#define ST_PIO2_1 0x1.921fb54442d18p+0 /* 1.5707963267948966 */
#define ST_PIO2_2 0x1.1a62633145c07p-54 /* 6.123233995736766e-17 */
#define ST_PIO2_3 -0x1.f1976b7ed8fbcp-110 /* -1.4973849048591698e-33 */
#define ST_PIO2_4 0x1.4cf98e804177dp-164 /* 5.562271104316826e-50 */
#define ST_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define ST_PIO4 0x1.921fb54442d18p-1
#define ST_INV_LN2 0x1.71547652b82fep+0
#define ST_LN2_HI 0x1.62e42fefa39efp-1
#define ST_LN2_LO 0x1.abc9e3b39803fp-56

#define ST_SIN_1 -0x1.5555555555555p-3 /* (-1)^1/3! */
#define ST_SIN_2 0x1.1111111111111p-7 /* (-1)^2/5! */
#define ST_SIN_3 -0x1.a01a01a01a01ap-13 /* (-1)^3/7! */
#define ST_SIN_4 0x1.71de3a556c734p-19 /* (-1)^4/9! */
#define ST_SIN_5 -0x1.ae64567f544e4p-26 /* (-1)^5/11! */
#define ST_SIN_6 0x1.6124613a86d09p-33 /* (-1)^6/13! */
#define ST_SIN_7 -0x1.ae7f3e733b81fp-41 /* (-1)^7/15! */
#define ST_SIN_8 0x1.952c77030ad4ap-49 /* (-1)^8/17! */
#define ST_COS_1 0x1.5555555555555p-5 /* (-1)^2/4! */
#define ST_COS_2 -0x1.6c16c16c16c17p-10 /* (-1)^3/6! */
#define ST_COS_3 0x1.a01a01a01a01ap-16 /* (-1)^4/8! */
#define ST_COS_4 -0x1.27e4fb7789f5cp-22 /* (-1)^5/10! */
#define ST_COS_5 0x1.1eed8eff8d898p-29 /* (-1)^6/12! */
#define ST_COS_6 -0x1.93974a8c07c9dp-37 /* (-1)^7/14! */
#define ST_COS_7 0x1.ae7f3e733b81fp-45 /* (-1)^8/16! */
#define ST_COS_8 -0x1.6827863b97d97p-53 /* (-1)^9/18! */
#define ST_EXP_2 0x1.0000000000000p-1 /* 1/2! */
#define ST_EXP_3 0x1.5555555555555p-3 /* 1/3! */
#define ST_EXP_4 0x1.5555555555555p-5 /* 1/4! */
#define ST_EXP_5 0x1.1111111111111p-7 /* 1/5! */
#define ST_EXP_6 0x1.6c16c16c16c17p-10 /* 1/6! */
#define ST_EXP_7 0x1.a01a01a01a01ap-13 /* 1/7! */
#define ST_EXP_8 0x1.a01a01a01a01ap-16 /* 1/8! */
#define ST_EXP_9 0x1.71de3a556c734p-19 /* 1/9! */
#define ST_EXP_10 0x1.27e4fb7789f5cp-22 /* 1/10! */
#define ST_EXP_11 0x1.ae64567f544e4p-26 /* 1/11! */
#define ST_EXP_12 0x1.1eed8eff8d898p-29 /* 1/12! */
#define ST_EXP_13 0x1.6124613a86d09p-33 /* 1/13! */
#define ST_EXP_14 0x1.93974a8c07c9dp-37 /* 1/14! */

#define ST_LOG_1 0x1.5555555555555p-1 /* 2/3 */
#define ST_LOG_2 0x1.999999999999ap-2 /* 2/5 */
#define ST_LOG_3 0x1.2492492492492p-2 /* 2/7 */
#define ST_LOG_4 0x1.c71c71c71c71cp-3 /* 2/9 */
#define ST_LOG_5 0x1.745d1745d1746p-3 /* 2/11 */
#define ST_LOG_6 0x1.3b13b13b13b14p-3 /* 2/13 */
#define ST_LOG_7 0x1.1111111111111p-3 /* 2/15 */
#define ST_LOG_8 0x1.e1e1e1e1e1e1ep-4 /* 2/17 */
#define ST_LOG_9 0x1.af286bca1af28p-4 /* 2/19 */
#define ST_LOG_10 0x1.8618618618618p-4 /* 2/21 */
#define ST_LOG_11 0x1.642c8590b2164p-4 /* 2/23 */
#define ST_LOG_12 0x1.47ae147ae147bp-4 /* 2/25 */
#define ST_LOG_13 0x1.2f684bda12f68p-4 /* 2/27 */
#define ST_LOG_14 0x1.1a7b9611a7b96p-4 /* 2/29 */
#define ST_SQRT_HALF 0x1.6a09e667f3bcdp-1

/* asin(x) = x + x*t*sum_{n>=1} c_n t^(n-1), t = x^2, c_n = C(2n,n)/(4^n (2n+1)) */
#define ST_ASIN_NCOEF 27
#define ST_ASIN_C0 0x1.5555555555555p-3
#define ST_ASIN_C1 0x1.3333333333333p-4
#define ST_ASIN_C2 0x1.6db6db6db6db7p-5
#define ST_ASIN_C3 0x1.f1c71c71c71c7p-6
#define ST_ASIN_C4 0x1.6e8ba2e8ba2e9p-6
#define ST_ASIN_C5 0x1.1c4ec4ec4ec4fp-6
#define ST_ASIN_C6 0x1.c99999999999ap-7
#define ST_ASIN_C7 0x1.7a87878787878p-7
#define ST_ASIN_C8 0x1.3fde50d79435ep-7
#define ST_ASIN_C9 0x1.12ef3cf3cf3cfp-7
#define ST_ASIN_C10 0x1.df3bd37a6f4dfp-8
#define ST_ASIN_C11 0x1.a6863d70a3d71p-8
#define ST_ASIN_C12 0x1.782dda12f684cp-8
#define ST_ASIN_C13 0x1.51ba308d3dcb1p-8
#define ST_ASIN_C14 0x1.31683bdef7bdfp-8
#define ST_ASIN_C15 0x1.15ee9d45d1746p-8
#define ST_ASIN_C16 0x1.fcaf8fb6db6dbp-9
#define ST_ASIN_C17 0x1.d3d2a8e0dd67dp-9
#define ST_ASIN_C18 0x1.b026f57b13b14p-9
#define ST_ASIN_C19 0x1.90cb77f60c7cep-9
#define ST_ASIN_C20 0x1.750de64d7d05fp-9
#define ST_ASIN_C21 0x1.5c5f56efaaaabp-9
#define ST_ASIN_C22 0x1.464c0950f7d47p-9
#define ST_ASIN_C23 0x1.3275586c5f2f0p-9
#define ST_ASIN_C24 0x1.208d3570ae5a6p-9
#define ST_ASIN_C25 0x1.1052bc5fa960ap-9
#define ST_ASIN_C26 0x1.018f963c229bfp-9

#define ST_TWO_OVER_PI_NWORDS 42
/* word j (j >= 2) holds bits 32(j-2)+1 .. 32(j-2)+32 of the binary fraction 2/pi */
#define ST_TWO_OVER_PI_WORDS \
  { 0x00000000u, 0x00000000u, 0xa2f9836eu, 0x4e441529u, 0xfc2757d1u, 0xf534ddc0u, \
    0xdb629599u, 0x3c439041u, 0xfe5163abu, 0xdebbc561u, 0xb7246e3au, 0x424dd2e0u, \
    0x06492eeau, 0x09d1921cu, 0xfe1deb1cu, 0xb129a73eu, 0xe88235f5u, 0x2ebb4484u, \
    0xe99c7026u, 0xb45f7e41u, 0x3991d639u, 0x835339f4u, 0x9c845f8bu, 0xbdf9283bu, \
    0x1ff897ffu, 0xde05980fu, 0xef2f118bu, 0x5a0a6d1fu, 0x6d367ecfu, 0x27cb09b7u, \
    0x4f463f66u, 0x9e5fea2du, 0x7527bac7u, 0xebe5f17bu, 0x3d0739f7u, 0x8a5292eau, \
    0x6bfb5fb1u, 0x1f8d5d08u, 0x56033046u, 0xfc7b6babu, 0xf0cfbc20u, 0x9af4361du }

#endif
