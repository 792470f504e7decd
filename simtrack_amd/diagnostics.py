"""Closed orbit and turn-by-turn diagnostics on the GPU, with the reference's
driver names and output files.

- cal_orbit_num: Cal_Orbit_Num (simtrack.cpp:14947-15030), the closed orbit
  by Newton iteration on the GPU-tracked one-turn map (st_closed_orbit).
- Track_tbt_FMA (cpp:21179-21241): frequency-map scan, tunes of two
  windows of turn-by-turn data for every stable grid particle
  (st_fma_scan), written to FMA-output.dat.
- Track_tbt_Lyapunov (cpp:21243-21303): twin-particle divergence scan
  (st_lyapunov_scan), written to Lmap-output.dat.

The reference writes its files through one fstream that it reopens per
line without resetting its format flags: the first line prints the grid
amplitudes in the default format, every later line in scientific notation
with 18 digits, which the writers below reproduce.
"""
import ctypes

import numpy as np

from ._capi import STError, check, load


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def cal_orbit_num(ctx, gp, records, deltap):
    """Cal_Orbit_Num: returns (x_co, x_end, iterations).  x_co is the fixed
    point of the one-turn map with every RF, crab and CCMULT voltage off;
    x_end is x_co passed once more through the line (the reference's
    Cell[Ncell-1]->X, which its drivers add to their grids)."""
    lib = load()
    rec = np.ascontiguousarray(records)
    xc, xe, it = np.zeros(6), np.zeros(6), ctypes.c_int()
    check(lib.st_closed_orbit(ctx.h, ctypes.byref(gp), _ptr(rec), len(rec), float(deltap), _ptr(xc), _ptr(xe),
                              ctypes.byref(it)), "st_closed_orbit")
    return xc, xe, it.value


def fma_scan(tab, deltap0, sigmax0, sigmay0, x_end, tune_low=0.660, tune_high=0.70, nturn=2048, sigma_step=0.1,
             sigma_max=6.0):
    """Rows (i*step, j*step, tunex1, tunex2, tuney1, tuney2, dqx, dqy) of
    Track_tbt_FMA.  The defaults are the reference's: tune_low/tune_high
    are its global search window (simtrack.cpp:27)."""
    lib = tab.lib
    xe = np.ascontiguousarray(np.asarray(x_end, dtype=np.float64)[:4])
    n = ctypes.c_int()
    check(lib.st_fma_scan(tab.h, deltap0, sigmax0, sigmay0, _ptr(xe), tune_low, tune_high, nturn, sigma_step,
                          sigma_max, None, 0, ctypes.byref(n)), "st_fma_scan")
    rows = np.zeros((max(n.value, 1), 8))
    check(lib.st_fma_scan(tab.h, deltap0, sigmax0, sigmay0, _ptr(xe), tune_low, tune_high, nturn, sigma_step,
                          sigma_max, _ptr(rows), n.value, ctypes.byref(n)), "st_fma_scan")
    return rows[:n.value]


def lyapunov_scan(tab, deltap0, sigmax0, sigmay0, nturn=2048, sigma_step=0.05, sigma_max=6.0):
    """Rows (i*step, j*step, mean Lyapunov exponent) of Track_tbt_Lyapunov."""
    lib = tab.lib
    n = ctypes.c_int()
    check(lib.st_lyapunov_scan(tab.h, deltap0, sigmax0, sigmay0, nturn, sigma_step, sigma_max, None, 0,
                               ctypes.byref(n)), "st_lyapunov_scan")
    rows = np.zeros((max(n.value, 1), 3))
    check(lib.st_lyapunov_scan(tab.h, deltap0, sigmax0, sigmay0, nturn, sigma_step, sigma_max, _ptr(rows), n.value,
                               ctypes.byref(n)), "st_lyapunov_scan")
    return rows[:n.value]


def _amp(v, first):
    return f"{v:g}" if first else f"{v:.18e}"


def _sci(v):
    return f"{v:.18e}".rjust(25)


def Track_tbt_FMA(tab, deltap0, sigmax0, sigmay0, x_end, tune_low=0.660, tune_high=0.70,
                  filename="./FMA-output.dat", **kw):
    """Track_tbt_FMA with the reference's output lines (cpp:21228-21235);
    x_end: the closed orbit from cal_orbit_num, which the reference computes
    first (cpp:21198).  Returns the rows."""
    rows = fma_scan(tab, deltap0, sigmax0, sigmay0, x_end, tune_low, tune_high, **kw)
    with open(filename, "w") as f:
        for k, r in enumerate(rows):
            f.write(f"{_amp(r[0], k == 0)} {_amp(r[1], k == 0)} {_sci(r[2])} {_sci(r[3])} {_sci(r[4])} "
                    f"{_sci(r[5])}  {_sci(r[6])} {_sci(r[7])}\n")
    return rows


def Track_tbt_Lyapunov(tab, deltap0, sigmax0, sigmay0, filename="./Lmap-output.dat", **kw):
    """Track_tbt_Lyapunov with the reference's output lines (cpp:21295-21297)."""
    rows = lyapunov_scan(tab, deltap0, sigmax0, sigmay0, **kw)
    with open(filename, "w") as f:
        for k, r in enumerate(rows):
            f.write(f"{_amp(r[0], k == 0)}  {_amp(r[1], k == 0)}{_sci(r[2])}\n")
    return rows


__all__ = ["cal_orbit_num", "fma_scan", "lyapunov_scan", "Track_tbt_FMA", "Track_tbt_Lyapunov", "STError"]
