"""Pins the project elementary functions and RNG (st_libm.h, st_rng.h) that both
the oracle and the kernels use, against mpmath, scipy, glibc and published
known-answer vectors."""
import ctypes
import math
import zlib

import mpmath
import numpy as np
import pytest
from scipy.stats import norm

import pyoracle

mpmath.mp.dps = 60


@pytest.fixture(scope="module")
def lib():
    return pyoracle.load(glibc=False)


@pytest.fixture(scope="module")
def glibc_lib():
    return pyoracle.load(glibc=True)


def ulp_err(got, exact):
    exact = mpmath.mpf(exact)
    if exact == 0:
        return 0.0 if got == 0 else math.inf
    return float(abs(mpmath.mpf(got) - exact) / mpmath.mpf(math.ulp(float(exact))))


RNG = np.random.default_rng(20261203)
CASES = {
    "orc_sin": (mpmath.sin, np.concatenate([RNG.uniform(-4, 4, 600), RNG.uniform(-3e5, 3e5, 300),
                                            10 ** RNG.uniform(6, 300, 300) * RNG.choice([-1, 1], 300),
                                            RNG.uniform(-1e-6, 1e-6, 50), [0.0, 1e-300, math.pi / 4]])),
    "orc_cos": (mpmath.cos, np.concatenate([RNG.uniform(-4, 4, 600), RNG.uniform(-3e5, 3e5, 300),
                                            10 ** RNG.uniform(6, 300, 300), [0.0, math.pi / 2]])),
    "orc_asin": (mpmath.asin, np.concatenate([RNG.uniform(-1, 1, 800), RNG.uniform(-0.01, 0.01, 200),
                                              1 - 10 ** RNG.uniform(-16, -1, 100), [0.5, -0.5, 0.125, 1.0, -1.0]])),
    "orc_exp": (mpmath.exp, np.concatenate([RNG.uniform(-700, 700, 600), RNG.uniform(-1, 1, 400),
                                            RNG.uniform(-745, -708, 60), [0.0, 1e-20, -1e-20]])),
    "orc_log": (mpmath.log, np.concatenate([10 ** RNG.uniform(-300, 300, 600), RNG.uniform(0.5, 2, 400),
                                            [5e-324, 1e-310, 1.0, 2.0]])),
    "orc_tan": (mpmath.tan, np.concatenate([RNG.uniform(-1.5, 1.5, 500), [1e-9, 0.0]])),
    "orc_atan": (mpmath.atan, np.concatenate([RNG.uniform(-3, 3, 800), 10 ** RNG.uniform(-30, 30, 300) *
                                              RNG.choice([-1, 1], 300),
                                              [0.4375, np.nextafter(0.4375, 0), 0.6875, np.nextafter(0.6875, 0),
                                               1.1875, np.nextafter(1.1875, 0), 2.4375, np.nextafter(2.4375, 0),
                                               1.0, 0.5, 1.5, 2.0 ** 66, 2.0 ** -27, 1e-300]])),
}
CASES["orc_cosh"] = (mpmath.cosh, np.concatenate([RNG.uniform(-20, 20, 600), RNG.uniform(-1, 1, 300), [0.0, 1e-300]]))
CASES["orc_sinh"] = (mpmath.sinh, np.concatenate([RNG.uniform(-20, 20, 500), RNG.uniform(-0.6, 0.6, 400),
                                                  [0.5, np.nextafter(0.5, 0), 1e-9, -1e-200]]))
BOUND = {"orc_tan": 2.0,    # tan = sin/cos: two roundings plus the quotient
         "orc_cosh": 2.0, "orc_sinh": 2.0}   # (e^x +- e^-x)/2 from the correctly rounded exp


@pytest.mark.parametrize("name", list(CASES))
def test_accuracy_vs_mpmath(lib, name):
    f = getattr(lib, name)
    ref, xs = CASES[name]
    worst = 0.0
    for x in xs:
        x = float(x)
        worst = max(worst, ulp_err(f(x), ref(mpmath.mpf(x))))
    assert worst < BOUND.get(name, 1.0), f"{name}: {worst} ulp"


def test_powi(lib):
    worst = 0.0
    for h in RNG.uniform(0.3, 30., 300):
        for n in range(-30, 31, 3):
            worst = max(worst, ulp_err(lib.orc_powi(float(h), n), mpmath.mpf(float(h)) ** n))
    assert worst <= 0.5001


def test_reduction_hard_cases(lib):
    """Arguments closest to multiples of pi/2 (Payne-Hanek path) and integers k*pi/2."""
    hard = [6381956970095103.0 * 2.0 ** 797, 5209563731060755.0 * 2.0 ** 237, 1.0e22, 2 ** 1000,
            8248.251512623145, 355.0, 103993.0]
    for x in hard:
        assert ulp_err(lib.orc_sin(x), mpmath.sin(mpmath.mpf(x))) < 1.0
        assert ulp_err(lib.orc_cos(x), mpmath.cos(mpmath.mpf(x))) < 1.0


def test_special_values(lib):
    assert math.isnan(lib.orc_sin(math.inf)) and math.isnan(lib.orc_cos(math.nan))
    assert math.isnan(lib.orc_asin(1.0000001)) and math.isnan(lib.orc_log(-1.0))
    assert lib.orc_exp(1000.) == math.inf and lib.orc_exp(-1000.) == 0.0
    assert lib.orc_log(0.0) == -math.inf and lib.orc_asin(-0.0) == 0.0


def correctly_rounded(fn, x):
    """fn(x) rounded once to the nearest double (ties to even), subnormals
    on the 2^-1074 grid."""
    with mpmath.workprec(300):
        v = fn(mpmath.mpf(x))
        if abs(v) < mpmath.mpf(2) ** -1022:
            n = v * mpmath.mpf(2) ** 1074
            fl = mpmath.floor(n)
            fr = n - fl
            if fr > 0.5 or (fr == 0.5 and int(fl) % 2 == 1):
                fl += 1
            return float(fl * mpmath.mpf(2) ** -1074)
        return float(v)


EXP_ARGS = {
    "beam-beam [-18, 0]": RNG.uniform(-18, 0, 6000),
    "full range": RNG.uniform(-745.2, 709.79, 2000),
    "near 0": RNG.uniform(-1e-3, 1e-3, 1000) * 10 ** RNG.uniform(-12, 0, 1000),
    "subnormal results": RNG.uniform(-745.2, -708.3, 1000),
    "near overflow": RNG.uniform(709.7, 709.79, 300),
}


@pytest.mark.parametrize("rng_name", list(EXP_ARGS))
def test_exp_correctly_rounded(lib, rng_name):
    """st_exp (Tang's table reduction + Ziv's rounding test) returns the
    correctly rounded exp; so does its accurate phase on its own, and the
    fast phase either agrees or leaves the rounding undecided (nan)."""
    undecided = 0
    for x in EXP_ARGS[rng_name]:
        x = float(x)
        want = correctly_rounded(mpmath.exp, x)
        assert lib.orc_exp(x) == want, f"exp({x.hex()}) = {lib.orc_exp(x)!r}, want {want!r}"
        assert lib.orc_exp_phase(x, 1) == want, f"accurate phase at {x.hex()}"
        f = lib.orc_exp_phase(x, 2)
        if f != f:
            undecided += 1
        else:
            assert f == want, f"fast phase at {x.hex()}"
    # the fast phase's error bound 2^-64 leaves about one call in 2^10 undecided
    assert undecided <= 0.01 * len(EXP_ARGS[rng_name])


def test_exp_special_values(lib):
    assert lib.orc_exp(0.0) == 1.0 and lib.orc_exp(-0.0) == 1.0
    assert lib.orc_exp(709.79) == math.inf and lib.orc_exp(math.inf) == math.inf
    assert lib.orc_exp(-745.2) == 0.0 and lib.orc_exp(-math.inf) == 0.0
    assert math.isnan(lib.orc_exp(math.nan))
    assert lib.orc_exp(709.782712893383973096) == 1.7976931348622732e308   # largest finite result
    assert lib.orc_exp(-745.1332191019411) == 5e-324                      # smallest subnormal
    for x in (1.0, -1.0, 0.5, 2.0, 10.0, -10.0):
        assert lib.orc_exp(x) == correctly_rounded(mpmath.exp, x)


# Ceilings on how often the project libm returns another double than the C
# library (glibc is not correctly rounded everywhere: on these ranges it
# misrounds exp 0.08 %, asin 0.07-0.2 %, sin/cos 0.1-0.3 % of arguments).
GLIBC_CEILING = [
    # (function, argument range, largest share of arguments that may differ)
    ("orc_exp", (-18.0, 0.0), 0.002),          # beam-beam kicks (BB4D/BB6D)
    ("orc_exp", (-700.0, 700.0), 0.002),
    ("orc_asin", (-1e-2, 1e-2), 0.002),        # exact bends: px/pz-sized arguments
    ("orc_asin", (-1.0, 1.0), 0.03),
    ("orc_sin", (-1e-2, 1e-2), 0.001),         # spin rotation angles per slice
    ("orc_cos", (-1e-2, 1e-2), 0.001),
    ("orc_sin", (-10.0, 10.0), 0.035),         # RF and crab phases
    ("orc_cos", (-10.0, 10.0), 0.035),
]


@pytest.mark.parametrize("name,rng,ceiling", GLIBC_CEILING, ids=lambda v: str(v))
def test_agrees_with_glibc(lib, glibc_lib, name, rng, ceiling):
    """Per function and argument range: the project libm and the C library
    agree to within one ulp, and return different doubles for at most the
    stated share of arguments."""
    xs = np.random.default_rng(zlib.crc32(f"{name}{rng}".encode())).uniform(rng[0], rng[1], 5000)
    a, b = getattr(lib, name), getattr(glibc_lib, name)
    differ = 0
    for x in xs:
        u, v = a(float(x)), b(float(x))
        assert abs(int(np.float64(u).view(np.int64)) - int(np.float64(v).view(np.int64))) <= 1
        differ += u != v
    assert differ <= ceiling * len(xs), f"{name} on {rng}: {differ / len(xs):.3%} differ"


def test_yoshida_constants():
    """Fdrift/Fkick (simtrack.cpp:29-32): consistency of the 4th-order
    integrator and the correctly rounded value of Fdrift1."""
    lib = pyoracle.load()
    out = (ctypes.c_double * 4)()
    lib.orc_get_yoshida(out)
    f1, f2, k1, k2 = list(out)
    assert 2 * f1 + 2 * f2 == 1.0 and abs(2 * k1 + k2 - 1.0) < 1e-16
    exact = 1 / (2 * (2 - mpmath.cbrt(2)))
    assert ulp_err(f1, exact) <= 1.0
    # 4th order: sum of cubes of kick weights vanishes
    assert abs(2 * k1 ** 3 + k2 ** 3) < 1e-15


PHILOX_KAT = [  # Random123 known-answer vectors for philox4x32_10
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_kat(lib, ctr, key, expect):
    out = (ctypes.c_uint32 * 4)()
    lib.orc_philox((ctypes.c_uint32 * 4)(*ctr), key[0], key[1], out)
    assert tuple(out) == expect


def test_normal_stream(lib):
    z = np.array([lib.orc_rng_normal(7, p, c) for p in range(300) for c in range(100)])
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    # a pure function of (seed, particle, draw)
    assert lib.orc_rng_normal(7, 3, 5) == lib.orc_rng_normal(7, 3, 5) != lib.orc_rng_normal(8, 3, 5)
    # Kolmogorov-Smirnov against N(0,1)
    from scipy.stats import kstest
    assert kstest(z, "norm").pvalue > 1e-3


def test_gaussian_quantile(lib):
    """GSL's gsl_cdf_ugaussian_Pinv is AS 241; pinned against scipy's ndtri."""
    ps = np.concatenate([np.linspace(1e-6, 1 - 1e-6, 4001), [0.2, 0.4, 0.6, 0.8], 10.0 ** -np.arange(2, 300, 13)])
    for p in ps:
        got, ref = lib.orc_ugaussian_pinv(float(p)), norm.ppf(p)
        assert abs(got - ref) <= 2e-15 * max(abs(ref), 1.0)


def _small_inputs(limit, rng):
    """Dense, adversarial arguments below `limit`: uniform, log-spaced down to
    subnormals, signed zeros, and the neighbourhood of the tiny-argument
    thresholds and of `limit` itself."""
    xs = [rng.uniform(-limit, limit, 20000),
          10.0 ** rng.uniform(-323, np.log10(limit), 20000) * rng.choice([-1.0, 1.0], 20000),
          np.array([0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308])]
    for thr in (2.0 ** -27, 2.0 ** -26, limit):
        v = np.float64(thr)
        xs.append(np.array([np.nextafter(v, 0.0), np.nextafter(np.nextafter(v, 0.0), 0.0)] if thr == limit else
                           [v, np.nextafter(v, 0.0), np.nextafter(v, 1.0)]))
    x = np.concatenate(xs)
    x = x[np.abs(x) < limit] if limit == 0.125 else x[np.abs(x) <= limit]
    return np.concatenate([x, -x])


def test_branch_free_asin_equals_asin(lib):
    """st_asin_small (the GPU kernels' first pass) equals st_asin bit for bit
    for every |x| < 0.125."""
    xs = _small_inputs(0.125, np.random.default_rng(5))
    for x in xs:
        a, b = lib.orc_asin_small(float(x)), lib.orc_asin(float(x))
        assert np.float64(a).view(np.int64) == np.float64(b).view(np.int64), x


def test_branch_free_sincos_equals_sincos(lib):
    """st_sincos_small equals st_sincos bit for bit for every |x| <= pi/4."""
    pio4 = float.fromhex("0x1.921fb54442d18p-1")
    xs = _small_inputs(pio4, np.random.default_rng(6))
    s1, c1, s2, c2 = (ctypes.c_double() for _ in range(4))
    for x in xs:
        lib.orc_sincos_small(ctypes.c_double(x), ctypes.byref(s1), ctypes.byref(c1))
        lib.orc_sincos(ctypes.c_double(x), ctypes.byref(s2), ctypes.byref(c2))
        for u, v in ((s1.value, s2.value), (c1.value, c2.value)):
            assert np.float64(u).view(np.int64) == np.float64(v).view(np.int64), x
