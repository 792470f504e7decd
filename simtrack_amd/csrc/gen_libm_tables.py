#!/usr/bin/env python3
"""Generate st_libm_tables.h: the constants of the project-owned FP64 libm.

Every constant is derived here from the mathematics with mpmath at 400
decimal digits and rounded once to IEEE double (or cut into exact bit
words), so the header can be regenerated and audited:

  * pi/2 split into four doubles P1..P4, each the round-to-nearest of what
    the previous ones leave (Cody-Waite reduction with FMA),
  * 2/pi, 1/ln2, ln2 split into HI/LO,
  * Taylor coefficients 1/n! (sin, cos, exp kernels),
  * arcsine series coefficients C(2n,n) / (4^n (2n+1)),
  * the binary expansion of 2/pi as 32-bit words for Payne-Hanek reduction
    of huge arguments (two leading zero words so negative bit positions read
    as zero).

Usage: python gen_libm_tables.py > st_libm_tables.h
"""
import sys
from fractions import Fraction
from math import comb, factorial

import mpmath

mpmath.mp.dps = 400


def rn(v):
    """Round an mpf (or Fraction) to the nearest double (ties-to-even)."""
    if isinstance(v, Fraction):
        return float(v)  # Python's Fraction->float is correctly rounded
    return float(mpmath.mpf(v))


def hexd(d):
    return float.hex(d)


def split(v, nparts):
    parts = []
    rem = mpmath.mpf(v)
    for _ in range(nparts):
        p = rn(rem)
        parts.append(p)
        rem = rem - mpmath.mpf(p)
    return parts


def main(out):
    pi = mpmath.pi
    p = split(pi / 2, 4)
    ln2 = split(mpmath.log(2), 2)
    lines = []
    w = lines.append
    w("/* Generated by gen_libm_tables.py from mpmath (400 digits). Do not edit. */")
    w("#ifndef ST_LIBM_TABLES_H")
    w("#define ST_LIBM_TABLES_H")
    w("")
    w("/* pi/2 = P1 + P2 + P3 + P4 + O(2^-210) */")
    for i, v in enumerate(p):
        w(f"#define ST_PIO2_{i+1} {hexd(v)} /* {v!r} */")
    w(f"#define ST_TWO_OVER_PI {hexd(rn(2 / pi))}")
    w(f"#define ST_PIO4 {hexd(rn(pi / 4))}")
    w(f"#define ST_INV_LN2 {hexd(rn(1 / mpmath.log(2)))}")
    w(f"#define ST_LN2_HI {hexd(ln2[0])}")
    w(f"#define ST_LN2_LO {hexd(ln2[1])}")
    w("")
    # sin: x + x^3*(S1 + z*(S2 + ...)), S_k = (-1)^k / (2k+1)!, k = 1..8
    for k in range(1, 9):
        v = Fraction((-1) ** k, factorial(2 * k + 1))
        w(f"#define ST_SIN_{k} {hexd(rn(v))} /* (-1)^{k}/{2*k+1}! */")
    # cos: 1 - z/2 + z^2*(C1 + z*(C2 + ...)), C_k = (-1)^(k+1) / (2k+2)!, k = 1..8
    for k in range(1, 9):
        v = Fraction((-1) ** (k + 1), factorial(2 * k + 2))
        w(f"#define ST_COS_{k} {hexd(rn(v))} /* (-1)^{k+1}/{2*k+2}! */")
    # exp: e^t - 1 - t = t^2 * (E2 + t*(E3 + ...)), E_n = 1/n!, n = 2..14
    for n in range(2, 15):
        w(f"#define ST_EXP_{n} {hexd(rn(Fraction(1, factorial(n))))} /* 1/{n}! */")
    w("")
    # exp, table-driven (Tang): x = k ln2/128 + r, 2^(j/128) as a triple of doubles
    ln2_128 = mpmath.log(2) / 128
    l1 = rn(mpmath.nint(ln2_128 * mpmath.mpf(2) ** 42) / mpmath.mpf(2) ** 42)   # 35 significant bits
    l2 = rn(ln2_128 - l1)
    l3 = rn(ln2_128 - l1 - l2)
    w("/* exp: k = rint(x * 128/ln2); ln2/128 = L1 + L2 + L3 + O(2^-150), L1 with 35")
    w("   significant bits so that k*L1 is exact for |k| < 2^18 */")
    w(f"#define ST_EXP_INVL {hexd(rn(128 / mpmath.log(2)))}")
    w(f"#define ST_EXP_L1 {hexd(l1)}")
    w(f"#define ST_EXP_L2 {hexd(l2)}")
    w(f"#define ST_EXP_L3 {hexd(l3)}")
    # 1/n! for n = 2..6 as double-double (hi, lo)
    for n in range(2, 7):
        hi = rn(Fraction(1, factorial(n)))
        lo = rn(Fraction(1, factorial(n)) - Fraction(hi))
        w(f"#define ST_EXPC{n}_HI {hexd(hi)}")
        w(f"#define ST_EXPC{n}_LO {hexd(lo)}")
    w("/* 2^(j/128) = T[3j] + T[3j+1] + T[3j+2] + O(2^-160), j = 0..127 */")
    w("#define ST_EXP_TAB \\")
    trip = []
    for j in range(128):
        trip += split(mpmath.mpf(2) ** (mpmath.mpf(j) / 128), 3)
    rows = []
    for j in range(0, len(trip), 3):
        rows.append(", ".join(hexd(v) for v in trip[j:j + 3]))
    w("  { " + ", \\\n    ".join(rows) + " }")
    w("")
    # log: log(1+f) = f - s*(f - R), s = f/(2+f), R = sum_{n>=1} 2/(2n+1) s^(2n)
    for n in range(1, 15):
        w(f"#define ST_LOG_{n} {hexd(rn(Fraction(2, 2 * n + 1)))} /* 2/{2*n+1} */")
    w(f"#define ST_SQRT_HALF {hexd(rn(mpmath.sqrt(2) / 2))}")
    w("")
    # asin(x) = x + sum_{n>=1} c_n x^(2n+1)
    w("/* asin(x) = x + x*t*sum_{n>=1} c_n t^(n-1), t = x^2, c_n = C(2n,n)/(4^n (2n+1)) */")
    nas = 27
    w(f"#define ST_ASIN_NCOEF {nas}")
    coefs = [rn(Fraction(comb(2 * n, n), 4 ** n * (2 * n + 1))) for n in range(1, nas + 1)]
    for i, c in enumerate(coefs):
        w(f"#define ST_ASIN_C{i} {hexd(c)}")
    w("")
    # 2/pi bits, 32 per word, most significant first; words 0 and 1 are zero pads.
    nbits = 32 * 40
    frac = 2 / pi
    bits = []
    for _ in range(nbits):
        frac *= 2
        b = int(mpmath.floor(frac))
        bits.append(b)
        frac -= b
    words = [0, 0]
    for j in range(0, nbits, 32):
        v = 0
        for b in bits[j:j + 32]:
            v = (v << 1) | b
        words.append(v)
    w(f"#define ST_TWO_OVER_PI_NWORDS {len(words)}")
    w("/* word j (j >= 2) holds bits 32(j-2)+1 .. 32(j-2)+32 of the binary fraction 2/pi */")
    w("#define ST_TWO_OVER_PI_WORDS \\")
    rows = []
    for j in range(0, len(words), 6):
        rows.append(", ".join(f"0x{v:08x}u" for v in words[j:j + 6]))
    w("  { " + ", \\\n    ".join(rows) + " }")
    w("")
    w("#endif")
    out.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main(sys.stdout)
