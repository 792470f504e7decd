"""Closed orbit and turn-by-turn diagnostics on the GPU, with the reference's
driver names and output files.

- cal_orbit_num: Cal_Orbit_Num (simtrack.cpp:14947-15030), the closed orbit
  by Newton iteration on the GPU-tracked one-turn map (st_closed_orbit).
- Track_tbt_FMA (cpp:21179-21241): frequency-map scan, tunes of two
  windows of turn-by-turn data for every stable grid particle
  (st_fma_scan), written to FMA-output.dat.
- Track_tbt_Lyapunov (cpp:21243-21303): twin-particle divergence scan
  (st_lyapunov_scan), written to Lmap-output.dat.
- Track_tbt_ActionDiff (cpp:21305-21383): action diffusion of particles at
  fixed amplitude and random phases (st_action_diff), Diff-output.dat.
- twiss_at_start: closed orbit, one-turn matrix and uncoupled Twiss from
  GPU-tracked finite differences (what Cal_Twiss gives those drivers).

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


def one_turn_matrix(ctx, gp, records, x_co, h=1e-6):
    """The 6x6 one-turn matrix about x_co (st_one_turn_matrix): symmetric
    finite differences of bare GPU-tracked turns (Element::Pass, no loss
    test) with the RFCAV voltages off, as Cal_OneTurnMap (simtrack.cpp:
    15122-15137) turns them off for its TPS map.  The step h = 1e-6 balances
    truncation (O(h^2) from the sextupoles) against roundoff (O(eps/h),
    largest in the delta column): on RHIC the symplectic defect
    |M^T S M - S| is 1e-7 at h = 1e-6 and 2.7e-6 at h = 1e-8."""
    lib = load()
    rec = np.ascontiguousarray(records)
    xc = np.ascontiguousarray(np.asarray(x_co, dtype=np.float64)[:6])
    m = np.zeros((6, 6))
    check(lib.st_one_turn_matrix(ctx.h, ctypes.byref(gp), _ptr(rec), len(rec), _ptr(xc), float(h), _ptr(m)),
          "st_one_turn_matrix")
    return m


def twiss_from_matrix(m):
    """(beta_x, alpha_x, tune_x, beta_y, alpha_y, tune_y) of the uncoupled
    2x2 blocks of a one-turn matrix (tunes fractional)."""
    out = []
    for b in (m[0:2, 0:2], m[2:4, 2:4]):
        cosmu = (b[0, 0] + b[1, 1]) / 2
        if abs(cosmu) >= 1:
            raise STError("one-turn matrix: unstable plane")
        sinmu = np.sign(b[0, 1]) * np.sqrt(1 - cosmu * cosmu)
        mu = np.arctan2(sinmu, cosmu) % (2 * np.pi)
        out += [b[0, 1] / sinmu, (b[0, 0] - b[1, 1]) / (2 * sinmu), mu / (2 * np.pi)]
    return tuple(float(v) for v in out)


def twiss_at_start(ctx, gp, records, deltap=0.):
    """Closed orbit, one-turn matrix and uncoupled Twiss at the start of the
    line, the quantities Cal_Twiss (simtrack.cpp:15540) gives the reference's
    drivers: the matrix is expanded about the closed orbit at the end of the
    line with z = 0 and delta = deltap (Cal_OneTurnMap, cpp:15140-15142)."""
    x_co, x_end, _ = cal_orbit_num(ctx, gp, records, deltap)
    x0 = x_end.copy()
    x0[4], x0[5] = 0., deltap
    m = one_turn_matrix(ctx, gp, records, x0)
    return x_co, m, twiss_from_matrix(m)


def action_diff(tab, betx, alfx, bety, alfy, nsigma, phase, emitx0, emity0, nturn, nstep, npart, seed=1234.):
    """Track_tbt_ActionDiff on the GPU: (rows, seed after).  Rows: turns,
    surviving count, mean and rms of 2Jx, mean and rms of 2Jy per step."""
    rows = np.zeros((nstep, 6))
    sd = np.array([seed], dtype=np.float64)
    check(tab.lib.st_action_diff(tab.h, betx, alfx, bety, alfy, nsigma, phase, emitx0, emity0, nturn, nstep, npart,
                                 _ptr(sd), _ptr(rows)), "st_action_diff")
    return rows, float(sd[0])


def Track_tbt_ActionDiff(tab, twiss, nsigma, phase, emitx0, emity0, nturn, nstep, npart, seed=1234.,
                         filename="./Diff-output.dat"):
    """Track_tbt_ActionDiff with the reference's output lines (cpp:21377-21379);
    twiss = (beta_x, alpha_x, beta_y, alpha_y) at the start.  Returns (rows, seed)."""
    rows, sd = action_diff(tab, *twiss, nsigma, phase, emitx0, emity0, nturn, nstep, npart, seed)
    with open(filename, "w") as f:
        for r in rows:
            f.write(f"{int(r[0])}   {int(r[1])}    {r[2]:g}  {r[3]:g}     {r[4]:g}   {r[5]:g}\n")
    return rows, sd


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


__all__ = ["cal_orbit_num", "fma_scan", "lyapunov_scan", "Track_tbt_FMA", "Track_tbt_Lyapunov", "one_turn_matrix",
           "twiss_from_matrix", "twiss_at_start", "action_diff", "Track_tbt_ActionDiff", "STError"]
