/*
 * The constant-divisor division of the multipole field sum: for b = n!
 * (n = 1..10) and y = 1/b rounded to nearest, the kernels compute
 *     q0 = a*y, r = fma(-b, q0, a), q = fma(r, y, q0)
 * and keep q when 2^-767 <= |q0| < 2^700 (else they divide).  This program
 * checks that q equals the IEEE quotient a / b bit for bit on random
 * arguments over that window, on quotients next to rounding midpoints and
 * on small integers.  With argv[1] = "nocorrect" it keeps q0 (no residual
 * step), which must give mismatches (so the check can fail).
 *
 * Build: cc -O2 -ffp-contract=off kdiv_check.c -lm; exit status 0 = all equal.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t state = 88172645463325252ull;
static uint64_t next(void)
{
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return state;
}
static double from_bits(uint64_t u)
{
    double d;
    memcpy(&d, &u, 8);
    return d;
}

int main(int argc, char **argv)
{
    const int nocorrect = argc > 1 && strcmp(argv[1], "nocorrect") == 0;
    const long per_divisor = argc > 2 ? atol(argv[2]) : 2000000;
    long tested = 0, mismatch = 0;
    double fac = 1.0;
    for (int n = 1; n <= 10; n++) {
        fac *= n;
        const double b = fac;
        const double y = 1.0 / b;
        for (long k = 0; k < per_divisor; k++) {
            double a;
            switch (k % 4) {
            case 0: /* any bit pattern */
                a = from_bits(next());
                break;
            case 1: /* any sign and magnitude inside the window */
                a = ldexp(1.0 + (double)(next() >> 11) * 0x1p-53, (int)(next() % 1490) - 767);
                if (next() & 1) a = -a;
                break;
            case 2: { /* quotients next to a rounding midpoint */
                double q = ldexp(1.0 + (double)(next() >> 11) * 0x1p-53, (int)(next() % 400) - 200);
                a = nextafter(b * q, (next() & 1) ? INFINITY : -INFINITY);
                if (next() & 1) a = -a;
                break;
            }
            default: /* small integers times powers of two */
                a = (double)((int64_t)(next() % 2000001) - 1000000) * ldexp(1.0, (int)(next() % 60) - 30);
            }
            double q0 = a * y;
            double aq = fabs(q0);
            if (!(aq >= 0x1p-767 && aq < 0x1p700))
                continue;
            double r = fma(-b, q0, a);
            double q = nocorrect ? q0 : fma(r, y, q0);
            double ref = a / b;
            tested++;
            if (memcmp(&q, &ref, sizeof q) != 0) {
                if (mismatch < 5 && !nocorrect)
                    printf("n=%d a=%a fast=%a divide=%a\n", n, a, q, ref);
                mismatch++;
            }
        }
    }
    printf("tested %ld quotients, %ld differ\n", tested, mismatch);
    return mismatch == 0 ? 0 : 1;
}
