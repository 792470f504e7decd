"""Dynamic-aperture drivers with the reference's names, arguments and output
files, running on the batched GPU tracker (st_da_bisect / st_da_uniform).

Track_DA (simtrack.cpp:21389), Track_DA1 (cpp:21443), Track_DA_Radial
(cpp:21682): twin-particle bisection of the amplitude per angle; Track_DA2
(cpp:21497) and Track_DA3 (cpp:21535, after Cal_Orbit_Num): the same with
one particle per amplitude; each writes "angle sigma" lines with
"%f %f \\n" (angle in degrees).
Track_DA_Uniform (cpp:21574), Track_DA_Radial_Uniform (cpp:21736),
Track_DA_Radial_Uniform_Test (Example2/DA_cal.cpp:61): a uniform amplitude
scan at one angle (radians) that stops at the first loss; each writes
"angle sigma stable" lines with "%f %f %f \\n".  Where the reference calls
exit(0) after the first loss, these return.

The first argument is a Table (the built line) instead of a Line; the line's
elements are those the table was built from.
"""
import ctypes

import numpy as np

from ._capi import (ST_DA_RADIAL, ST_DA_SINGLE, ST_DA_TRACK_DA, ST_DA_TRACK_DA1, ST_DA_TRACK_DA2, ST_DA_TRACK_DA3,
                    ST_DA_TRACK_DA_RADIAL, _ptr, check)


def da_angles(lib, variant):
    angles = np.zeros(10)
    n = ctypes.c_int()
    smax = ctypes.c_double()
    check(lib.st_da_angles(variant, _ptr(angles), ctypes.byref(n), ctypes.byref(smax)), "st_da_angles")
    return angles[:n.value].copy(), smax.value


def da_bisect(tab, nturn, deltap0, sigmax0, sigmay0, angles, sigma_max, radial=False, single=False):
    """Last bisection midpoint per angle (the value the reference writes);
    single: one particle per midpoint instead of twins."""
    angles = np.ascontiguousarray(angles, dtype=np.float64)
    out = np.zeros(len(angles))
    flags = (ST_DA_RADIAL if radial else 0) | (ST_DA_SINGLE if single else 0)
    check(tab.lib.st_da_bisect(tab.h, int(nturn), float(deltap0), float(sigmax0), float(sigmay0), _ptr(angles),
                               len(angles), float(sigma_max), flags, _ptr(out)), "st_da_bisect")
    return out


def da_uniform(tab, nturn, deltap0, angle, sigmax0, sigmay0, sigma0, dsigma, n, radial=False):
    """(sigmas, stable flags) of the scan up to and including the first loss."""
    sig = np.zeros(n)
    st = np.zeros(n, np.int32)
    m = ctypes.c_int()
    check(tab.lib.st_da_uniform(tab.h, int(nturn), float(deltap0), float(angle), float(sigmax0), float(sigmay0),
                                float(sigma0), float(dsigma), int(n), 1 if radial else 0, _ptr(sig), _ptr(st),
                                ctypes.byref(m)), "st_da_uniform")
    return sig[:m.value], st[:m.value]


def _bisect_driver(tab, variant, nturn, deltap0, sigmax0, sigmay0, filename, degrees):
    angles, smax = da_angles(tab.lib, variant)
    sig = da_bisect(tab, nturn, deltap0, sigmax0, sigmay0, angles, smax,
                    radial=(variant == ST_DA_TRACK_DA_RADIAL),
                    single=(variant in (ST_DA_TRACK_DA2, ST_DA_TRACK_DA3)))
    rows = list(zip(degrees, sig))
    if filename is not None:
        with open(filename, "w") as f:
            for a, s in rows:
                f.write("%f %f \n" % (a, s))
    return rows


def Track_DA(tab, nturn, deltap0, sigmax0, sigmay0, filename=None):
    return _bisect_driver(tab, ST_DA_TRACK_DA, nturn, deltap0, sigmax0, sigmay0, filename,
                          [(90 / 11.) * (i + 1) for i in range(10)])


def Track_DA1(tab, nturn, deltap0, sigmax0, sigmay0, filename=None):
    return _bisect_driver(tab, ST_DA_TRACK_DA1, nturn, deltap0, sigmax0, sigmay0, filename,
                          [15. * (i + 1) for i in range(5)])


def Track_DA_Radial(tab, nturn, deltap0, sigmax0, sigmay0, filename=None):
    return _bisect_driver(tab, ST_DA_TRACK_DA_RADIAL, nturn, deltap0, sigmax0, sigmay0, filename,
                          [(90 / 11.) * (i + 1) for i in range(10)])


def Track_DA2(tab, nturn, deltap0, sigmax0, sigmay0, filename=None):
    return _bisect_driver(tab, ST_DA_TRACK_DA2, nturn, deltap0, sigmax0, sigmay0, filename,
                          [(90 / 11.) * (i + 1) for i in range(10)])


def Track_DA3(tab, nturn, deltap0, sigmax0, sigmay0, filename=None):
    """Track_DA3 calls Cal_Orbit_Num(linename, 0) first (cpp:21545); its
    result only fills the elements' X, which the scan does not read, so the
    scan here is the bisection alone."""
    return _bisect_driver(tab, ST_DA_TRACK_DA3, nturn, deltap0, sigmax0, sigmay0, filename,
                          [15. * (i + 1) for i in range(5)])


def _uniform_driver(tab, nturn, deltap0, angle, sigmax0, sigmay0, filename, sigma0, n, radial):
    sig, st = da_uniform(tab, nturn, deltap0, angle, sigmax0, sigmay0, sigma0, 0.2, n, radial)
    rows = [(angle, s, 1.0 * k) for s, k in zip(sig, st)]
    if filename is not None:
        with open(filename, "w") as f:
            for a, s, k in rows:
                f.write("%f %f %f \n" % (a, s, k))
    return rows


def Track_DA_Uniform(tab, nturn, deltap0, angle, sigmax0, sigmay0, filename=None):
    return _uniform_driver(tab, nturn, deltap0, angle, sigmax0, sigmay0, filename, 1., 150, False)


def Track_DA_Radial_Uniform(tab, nturn, deltap0, angle, sigmax0, sigmay0, filename=None):
    return _uniform_driver(tab, nturn, deltap0, angle, sigmax0, sigmay0, filename, 1.0, 150, True)


def Track_DA_Radial_Uniform_Test(tab, nturn, deltap0, angle, sigmax0, sigmay0, filename=None):
    return _uniform_driver(tab, nturn, deltap0, angle, sigmax0, sigmay0, filename, 3.0, 100, True)
