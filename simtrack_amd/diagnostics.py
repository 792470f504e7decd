"""Closed orbit and turn-by-turn diagnostics on the GPU, with the reference's
driver names and output files.

- cal_orbit_num: Cal_Orbit_Num (simtrack.cpp:14947-15030), the closed orbit
  by Newton iteration on the GPU-tracked one-turn map (st_closed_orbit).
- Track_tbt_tune_footprint (cpp:21100-21176): tunes of two windows of
  turn-by-turn data on a 5-angle x 6-amplitude grid normalised by the
  Twiss at the start (footprint_scan), footprint-output.dat.
- cal_tunes_track(_x), cal_chrom_track: Cal_Tunes_Track (cpp:20971,
  20992) and Cal_Chrom_Track (cpp:21028), tunes of tracked turns.
- cal_spin_orbit, cal_spin_tune, set_initial_spin: Cal_Spin_Orbit
  (cpp:20510-20572), Cal_Spin_Tune (cpp:20574-20601) and Set_Initial_Spin
  (cpp:20603-20635) on GPU-tracked sPass turns (st_spin_orbit,
  st_spin_tune, st_set_initial_spin).
- Track_tbt_FMA (cpp:21179-21241): frequency-map scan, tunes of two
  windows of turn-by-turn data for every stable grid particle
  (st_fma_scan), written to FMA-output.dat.
- Track_tbt_Lyapunov (cpp:21243-21303): twin-particle divergence scan
  (st_lyapunov_scan), written to Lmap-output.dat.
- Track_tbt_ActionDiff (cpp:21305-21383): action diffusion of particles at
  fixed amplitude and random phases (st_action_diff), Diff-output.dat.
- twiss_at_start: closed orbit, one-turn matrix and uncoupled Twiss from
  GPU-tracked finite differences (what Cal_Twiss gives those drivers).
- cal_twiss: Cal_Twiss (cpp:15540-15603) along the whole line, with the
  reference's coupled normal-mode convention: Cal_A's normalising matrix
  of the one-turn map (cpp:15299-15477), carried through every element's
  GPU-differentiated map by Trace_A (cpp:15480-15538), and the mode
  betas, alphas, phases and the coupling r, c11..c22 read off it;
  cal_dispersion and cal_chrom (Cal_Dispersion cpp:17049, Cal_Chrom
  cpp:16964) on its maps and tunes.

The reference writes its files through one fstream that it reopens per
line without resetting its format flags: the first line prints the grid
amplitudes in the default format, every later line in scientific notation
with 18 digits, which the writers below reproduce.
"""
import ctypes
import math
from dataclasses import dataclass

import numpy as np

from ._capi import STError, check, load

PI = 3.14159265358979323    # the reference's PI, simtrack.cpp:23


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


def cal_spin_orbit(ctx, gp, records, deltap):
    """Cal_Spin_Orbit: returns (x0, n0, iterations).  x0 is the closed orbit
    at the start (Cell[Ncell-1]->X) with the periodic spin; n0[j] is the
    normalised spin after element j (the reference's Cell[j]->n0)."""
    lib = load()
    rec = np.ascontiguousarray(records)
    x0, n0, it = np.zeros(9), np.zeros((len(rec), 3)), ctypes.c_int()
    check(lib.st_spin_orbit(ctx.h, ctypes.byref(gp), _ptr(rec), len(rec), float(deltap), _ptr(x0), _ptr(n0),
                            ctypes.byref(it)), "st_spin_orbit")
    return x0, n0, it.value


def cal_spin_tune(ctx, gp, records, deltap):
    """Cal_Spin_Tune: (spin tune on [0.001, 0.499], n0 of the last element)."""
    lib = load()
    rec = np.ascontiguousarray(records)
    tune, n0 = np.zeros(1), np.zeros(3)
    check(lib.st_spin_tune(ctx.h, ctypes.byref(gp), _ptr(rec), len(rec), float(deltap), _ptr(tune), _ptr(n0)),
          "st_spin_tune")
    return float(tune[0]), n0


def set_initial_spin(ctx, gp, records, nturn, n0, x):
    """Set_Initial_Spin: x (9 coordinates) with its spin set to the
    normalised turn-averaged projection of the three unit spins on n0."""
    lib = load()
    rec = np.ascontiguousarray(records)
    x = np.array(x, dtype=np.float64).reshape(9).copy()
    n0 = np.ascontiguousarray(n0, dtype=np.float64).reshape(3)
    check(lib.st_set_initial_spin(ctx.h, ctypes.byref(gp), _ptr(rec), len(rec), int(nturn), _ptr(n0), _ptr(x)),
          "st_set_initial_spin")
    return x


def cal_tunes_track_x(tab, x, nturn, tune_low=0.660, tune_high=0.70):
    """Cal_Tunes_Track(Line, x, Nturn, tune1, tune2), simtrack.cpp:20971-20990:
    x[0] and x[2] below 1e-9 are raised by 1e-9, Track_tbt for nturn turns,
    FineTuneFinder of the x and y samples in [tune_low, tune_high] (one window
    for both, the reference's globals).  Returns (tune1, tune2, x as
    started); the tunes are NaN for a lost particle (the reference leaves
    them unset)."""
    x = np.array(x, dtype=np.float64).reshape(1, 6).copy()
    if x[0, 0] < 1.0e-9:
        x[0, 0] = x[0, 0] + 1.0e-9
    if x[0, 2] < 1.0e-9:
        x[0, 2] = x[0, 2] + 1.0e-9
    x0 = x[0].copy()
    _, tbt, stable, _, _ = tab.track_tbt(x, int(nturn))
    if stable[0] != 1:
        return math.nan, math.nan, x0
    q = tab.ctx.fine_tune(np.stack([tbt[0, :, 0], tbt[0, :, 2]]), tune_low, tune_high)
    return float(q[0]), float(q[1]), x0


def cal_tunes_track(ctx, gp, records, deltap0, tune_low=0.660, tune_high=0.70, tab=None):
    """Cal_Tunes_Track(Line, deltap0), simtrack.cpp:20992-21026: Cal_Orbit_Num,
    then the tunes of 1,024 tracked turns from the closed orbit (the line as
    given, RF on).  Returns (Tune1, Tune2)."""
    from ._capi import Table
    _, xe, _ = cal_orbit_num(ctx, gp, records, deltap0)
    x = np.array([xe[0], xe[1], xe[2], xe[3], 0., deltap0])
    own = tab is None
    tab = Table(ctx, gp, records) if own else tab
    try:
        q1, q2, _ = cal_tunes_track_x(tab, x, 1024, tune_low, tune_high)
    finally:
        if own:
            tab.close()
    return q1, q2


def cal_chrom_track(ctx, gp, records, tune_low=0.660, tune_high=0.70):
    """Cal_Chrom_Track, simtrack.cpp:21028-21051: tracked tunes at deltap
    +0.0003, -0.0003 and 0.  Returns (Chromx1, Chromy1, Chromx2, Chromy2)."""
    from ._capi import Table
    tab = Table(ctx, gp, records)
    try:
        qxp, qyp = cal_tunes_track(ctx, gp, records, 0.0003, tune_low, tune_high, tab)
        qxm, qym = cal_tunes_track(ctx, gp, records, -0.0003, tune_low, tune_high, tab)
        qx0, qy0 = cal_tunes_track(ctx, gp, records, 0.00, tune_low, tune_high, tab)
    finally:
        tab.close()
    return ((qxp - qxm) / 2. / 0.0003, (qyp - qym) / 2. / 0.0003,
            (qxp + qxm - 2 * qx0) / 2. / 0.0003 / 0.0003, (qyp + qym - 2 * qy0) / 2. / 0.0003 / 0.0003)


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


def delta_to_pt(gp, delta):
    """DeltaToPt, simtrack.cpp:125-129."""
    return (-2.0 / gp.beta + math.sqrt(4.0 / gp.beta / gp.beta - 4.0 * (1.0 - (1 + delta) * (1 + delta)))) / 2.


def element_maps(ctx, gp, records, x_co, x5, h=1e-6):
    """(T, X): every element's 6x6 map about the closed orbit entering it
    (z = 0, x[5] = x5) and the closed orbit after it (st_element_maps)."""
    lib = load()
    rec = np.ascontiguousarray(records)
    n = len(rec)
    T, X = np.zeros((n, 6, 6)), np.zeros((n, 6))
    xc = np.ascontiguousarray(np.asarray(x_co, dtype=np.float64)[:6])
    check(lib.st_element_maps(ctx.h, ctypes.byref(gp), _ptr(rec), n, _ptr(xc), float(x5), float(h), _ptr(T),
                              _ptr(X)), "st_element_maps")
    return T, X


def cal_a(m):
    """Cal_A (simtrack.cpp:15299-15477): the normalising matrix A (4x4) of
    the transverse block of a one-turn matrix, columns a1..a4 built from
    the eigenvectors.  Mode 1 is the eigenvector pair with the larger x, px
    content (cpp:15318-15345); each eigenvector is turned so that A[0][1] =
    A[2][3] = 0 (cpp:15347-15381); a2 and a4 are flipped to make A[1][1],
    A[3][3] positive (cpp:15400-15410); each pair is scaled to unit
    symplectic product (cpp:15440-15452).  The eigenvectors come from numpy
    instead of GSL; the conventions fix A, so the result agrees to rounding."""
    w, v = np.linalg.eig(np.asarray(m, dtype=np.float64)[:4, :4])
    up = [i for i in range(4) if w[i].imag > 0]
    if len(up) != 2:
        raise STError("one-turn map: unstable (eigenvalues off the unit circle)")
    content = [abs(v[0, i].real) + abs(v[0, i].imag) + abs(v[1, i].real) + abs(v[1, i].imag) for i in up]
    m1, m2 = (up[0], up[1]) if content[0] >= content[1] else (up[1], up[0])
    a = np.zeros((4, 4))
    for row, i, comp in ((0, m1, 0), (2, m2, 2)):
        vec = v[:, i]
        th = math.atan2(vec[comp].imag, vec[comp].real)
        vec = vec * complex(math.cos(-math.pi / 2 - th), math.sin(-math.pi / 2 - th))
        a[row] = -2 * vec.imag
        a[row + 1] = 2 * vec.real
    if a[1, 1] < 0:
        a[1] = -a[1]
    if a[3, 3] < 0:
        a[3] = -a[3]
    t1 = -a[0, 1] * a[1, 0] + a[0, 0] * a[1, 1] - a[0, 3] * a[1, 2] + a[0, 2] * a[1, 3]
    t2 = -a[2, 1] * a[3, 0] + a[2, 0] * a[3, 1] - a[2, 3] * a[3, 2] + a[2, 2] * a[3, 3]
    if not (t1 > 0 and t2 > 0):
        raise STError("one-turn map: normal modes not normalisable")
    a[0:2] /= math.sqrt(t1)
    a[2:4] /= math.sqrt(t2)
    return a.T.copy()


def trace_a(T, a_end):
    """Trace_A (simtrack.cpp:15480-15538): A after every element (each
    element's map applied, then each mode turned back to A[0][1] = A[2][3] =
    0) and the accumulated phase advances in turns.  Returns (A, mu1, mu2)."""
    n = len(T)
    A = np.array(a_end, dtype=np.float64)
    out = np.zeros((n, 4, 4))
    mu = np.zeros((n, 2))
    m1 = m2 = 0.
    for k in range(n):
        G = T[k][:4, :4] @ A
        d1 = math.atan2(G[0, 1], G[0, 0])
        d2 = math.atan2(G[2, 3], G[2, 2])
        s1 = math.sqrt(G[0, 0] * G[0, 0] + G[0, 1] * G[0, 1])
        s2 = math.sqrt(G[2, 2] * G[2, 2] + G[2, 3] * G[2, 3])
        A = np.empty((4, 4))
        A[:, 0] = (G[:, 0] * G[0, 0] + G[:, 1] * G[0, 1]) / s1
        A[:, 1] = (-G[:, 0] * G[0, 1] + G[:, 1] * G[0, 0]) / s1
        A[:, 2] = (G[:, 2] * G[2, 2] + G[:, 3] * G[2, 3]) / s2
        A[:, 3] = (-G[:, 2] * G[2, 3] + G[:, 3] * G[2, 2]) / s2
        out[k] = A
        m1 += d1 / 2. / math.pi
        m2 += d2 / 2. / math.pi
        mu[k] = (m1, m2)
    return out, mu[:, 0], mu[:, 1]


@dataclass
class LineTwiss:
    """Per-element optics after each element, the Element fields Cal_Twiss
    fills (Beta1, Alfa1, Beta2, Alfa2, Mu1, Mu2, r, c11, c12, c21, c22, X),
    plus A, the tunes (Line.Tune1, Line.Tune2), the one-turn matrix M (the
    last element's M) and the element maps T (each element's T)."""
    beta1: np.ndarray
    alfa1: np.ndarray
    beta2: np.ndarray
    alfa2: np.ndarray
    mu1: np.ndarray
    mu2: np.ndarray
    r: np.ndarray
    c: np.ndarray
    X: np.ndarray
    A: np.ndarray
    tune1: float
    tune2: float
    M: np.ndarray
    T: np.ndarray


def cal_twiss(ctx, gp, records, deltap=0.):
    """Cal_Twiss (simtrack.cpp:15540-15603): deltap is dp/p; without
    H_expand it enters x[5] as pt (DeltaToPt, cpp:15556).  The closed orbit,
    one-turn matrix and element maps come from the GPU (st_closed_orbit,
    st_one_turn_matrix, st_element_maps); the 4x4 analysis runs on the host,
    as in the reference."""
    x5 = deltap if gp.H_expand else delta_to_pt(gp, deltap)
    x_co, x_end, _ = cal_orbit_num(ctx, gp, records, x5)
    x0 = x_end.copy()
    x0[4], x0[5] = 0., x5
    m = one_turn_matrix(ctx, gp, records, x0)
    T, X = element_maps(ctx, gp, records, x_co, x5)
    A, mu1, mu2 = trace_a(T, cal_a(m))
    beta1 = A[:, 0, 0] / A[:, 1, 1]
    alfa1 = -A[:, 1, 0] / A[:, 1, 1]
    beta2 = A[:, 2, 2] / A[:, 3, 3]
    alfa2 = -A[:, 3, 2] / A[:, 3, 3]
    a22 = A[:, 2:4, 2:4]
    r = np.sqrt(a22[:, 0, 0] * a22[:, 1, 1] - a22[:, 0, 1] * a22[:, 1, 0])
    p = a22[:, 0, 0] * a22[:, 1, 1]
    inv = np.stack([np.stack([a22[:, 1, 1] / p, -a22[:, 0, 1] / p], -1),
                    np.stack([-a22[:, 1, 0] / p, a22[:, 0, 0] / p], -1)], -2)
    c = (A[:, 0:2, 2:4] @ inv) * r[:, None, None]
    return LineTwiss(beta1, alfa1, beta2, alfa2, mu1, mu2, r, c, X, A, float(mu1[-1]), float(mu2[-1]), m, T)


def cal_dispersion(M, T, gp):
    """Cal_Dispersion (simtrack.cpp:17049-17105): the periodic dispersion
    -(M44 - I)^-1 M[0:4, 5] at the end of the line, carried through every
    element's map (x[5] = 1), times beta without H_expand.  M, T: the
    one-turn matrix and element maps of cal_twiss.  Returns an (n, 4) array
    of (Etax, Etaxp, Etay, Etayp) after every element."""
    r44 = -np.linalg.inv(np.asarray(M)[:4, :4] - np.eye(4))
    x = r44 @ np.asarray(M)[:4, 5]
    out = np.zeros((len(T), 4))
    for k, t in enumerate(T):
        x = t[:4, :4] @ x + t[:4, 5]
        out[k] = x
    if not gp.H_expand:
        out *= gp.beta
    return out


def cal_chrom(ctx, gp, records, step_deltap=0.0001):
    """Cal_Chrom (simtrack.cpp:16964-16992): first and second order
    chromaticities from the tunes of cal_twiss at +-step and 0.  As in the
    reference, without H_expand the step is converted to pt before
    cal_twiss converts it again (cpp:16973).  Returns (Chromx1, Chromy1,
    Chromx2, Chromy2); step_deltap is GP.step_deltap (default 1e-4,
    simtrack.h:48)."""
    d = step_deltap if gp.H_expand else delta_to_pt(gp, step_deltap)
    p = cal_twiss(ctx, gp, records, d)
    m = cal_twiss(ctx, gp, records, -d)
    z = cal_twiss(ctx, gp, records, 0.)
    return ((p.tune1 - m.tune1) / 2. / step_deltap, (p.tune2 - m.tune2) / 2. / step_deltap,
            (p.tune1 + m.tune1 - 2 * z.tune1) / 2. / step_deltap / step_deltap,
            (p.tune2 + m.tune2 - 2 * z.tune2) / 2. / step_deltap / step_deltap)


def set_line_optics(line, tw, eta=None):
    """Write a cal_twiss result into the Line's elements as Cal_Twiss does
    (Beta1, Alfa1, Beta2, Alfa2, Mu1, Mu2, r, c11..c22, X) and the tunes into
    Line.Tune1, Line.Tune2; eta (cal_dispersion) fills Etax..Etayp."""
    for k, e in enumerate(line.Cell):
        e.Beta1, e.Alfa1, e.Beta2, e.Alfa2 = (float(tw.beta1[k]), float(tw.alfa1[k]), float(tw.beta2[k]),
                                              float(tw.alfa2[k]))
        e.Mu1, e.Mu2, e.r = float(tw.mu1[k]), float(tw.mu2[k]), float(tw.r[k])
        e.c11, e.c12, e.c21, e.c22 = (float(v) for v in tw.c[k].ravel())
        e.X = [float(v) for v in tw.X[k]]
        if eta is not None:
            e.Etax, e.Etaxp, e.Etay, e.Etayp = (float(v) for v in eta[k])
    line.Tune1, line.Tune2 = tw.tune1, tw.tune2


def Print_Twiss(line, filename):
    """Print_Twiss (simtrack.cpp:19846-19862): a header and one line per
    element, each field right-aligned in 15 columns in C++'s default stream
    format (%g).  Reads the fields set_line_optics wrote."""
    def f(v):
        return (f"{v:g}" if isinstance(v, float) else str(v)).rjust(15)
    line.Update()
    with open(filename, "w") as out:
        out.write("".join(h.rjust(15) for h in ("NAME", "TYPE", "S", "BETA1", "ALFA1", "MU1", "BETA2", "ALFA2", "MU2",
                                                 "ETAX", "ETAXP", "ETAY", "ETAYP")) + "\n")
        for e in line.Cell:
            out.write("".join(f(v) for v in (e.NAME, e.TYPE, float(e.S), e.Beta1, e.Alfa1, e.Mu1, e.Beta2, e.Alfa2,
                                             e.Mu2, getattr(e, "Etax", 0.), getattr(e, "Etaxp", 0.),
                                             getattr(e, "Etay", 0.), getattr(e, "Etayp", 0.))) + "\n")


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


def footprint_scan(tab, deltap0, emitx, emity, twiss, x_end, tune_low=0.660, tune_high=0.70, nturn=2048):
    """The grid of Track_tbt_tune_footprint (simtrack.cpp:21100-21176): at
    angles 15(i+1) deg and nsigma = j (0.1 for j = 0), i < 5, j < 6, the
    particle (sqrt(beta) xn, -alpha xn / sqrt(beta), ...) with
    xn = sqrt(emit) nsigma cos or sin, plus the closed orbit x_end, x or y
    exactly 0 raised by 1e-9; Track_tbt for nturn turns, FineTuneFinder of
    turns [0, 1024) and [1023, 2047).  twiss: (betx, alfx, bety, alfy) at the
    start.  Rows (angle deg, j, tunex1, tunex2, tuney1, tuney2, dqx, dqy) of
    the stable particles, in the reference's order."""
    betax0, alfax0, betay0, alfay0 = twiss
    xs, keys = [], []
    for i in range(5):
        for j in range(6):
            nsigma = float(j)
            if nsigma == 0:
                nsigma = 0.1
            nsigmax = nsigma * math.cos((i + 1) * 15 * PI / 180)
            nsigmay = nsigma * math.sin((i + 1) * 15 * PI / 180)
            xn = math.sqrt(emitx) * nsigmax
            pxn = 0.
            yn = math.sqrt(emity) * nsigmay
            pyn = 0.
            x = [math.sqrt(betax0) * xn, -alfax0 * xn / math.sqrt(betax0) + 1.0 * pxn / math.sqrt(betax0),
                 math.sqrt(betay0) * yn, -alfay0 * yn / math.sqrt(betay0) + 1.0 * pyn / math.sqrt(betay0), 0.,
                 deltap0]
            for k in range(4):
                x[k] = x[k] + x_end[k]
            if x[0] == 0.:
                x[0] = x[0] + 1.0e-09
            if x[2] == 0.:
                x[2] = x[2] + 1.0e-09
            xs.append(x)
            keys.append(((i + 1) * 15, j))
    _, tbt, stable, _, _ = tab.track_tbt(np.array(xs), int(nturn))
    rows = []
    good = [p for p in range(len(xs)) if stable[p] == 1]
    if good:
        sig = []
        for p in good:
            sig += [tbt[p, 0:1024, 0], tbt[p, 1023:2047, 0], tbt[p, 0:1024, 2], tbt[p, 1023:2047, 2]]
        q = tab.ctx.fine_tune(np.stack(sig), tune_low, tune_high).reshape(len(good), 4)
        for p, (qx1, qx2, qy1, qy2) in zip(good, q):
            a, j = keys[p]
            rows.append((a, j, qx1, qx2, qy1, qy2, qx1 - qx2, qy1 - qy2))
    return rows


def Track_tbt_tune_footprint(ctx, gp, records, deltap0, emitx, emity, tune_low=0.660, tune_high=0.70,
                             filename="./footprint-output.dat", tab=None):
    """Track_tbt_tune_footprint (simtrack.cpp:21100-21176): Cal_Twiss at
    deltap0 (the last element's Beta/Alfa), Cal_Orbit_Num at deltap0, the
    grid of footprint_scan, and the reference's footprint-output.dat lines.
    emitx, emity: unnormalised emittances.  Returns the rows."""
    from ._capi import Table
    tw = cal_twiss(ctx, gp, records, deltap0)
    twiss = (float(tw.beta1[-1]), float(tw.alfa1[-1]), float(tw.beta2[-1]), float(tw.alfa2[-1]))
    _, x_end, _ = cal_orbit_num(ctx, gp, records, deltap0)
    own = tab is None
    tab = Table(ctx, gp, records) if own else tab
    try:
        rows = footprint_scan(tab, deltap0, emitx, emity, twiss, x_end, tune_low, tune_high)
    finally:
        if own:
            tab.close()
    with open(filename, "w") as f:
        for r in rows:
            f.write(f"{r[0]} {r[1]} {_sci(r[2])} {_sci(r[3])} {_sci(r[4])} {_sci(r[5])}  {_sci(r[6])} {_sci(r[7])}\n")
    return rows


def Track_tbt_Lyapunov(tab, deltap0, sigmax0, sigmay0, filename="./Lmap-output.dat", **kw):
    """Track_tbt_Lyapunov with the reference's output lines (cpp:21295-21297)."""
    rows = lyapunov_scan(tab, deltap0, sigmax0, sigmay0, **kw)
    with open(filename, "w") as f:
        for k, r in enumerate(rows):
            f.write(f"{_amp(r[0], k == 0)}  {_amp(r[1], k == 0)}{_sci(r[2])}\n")
    return rows


__all__ = ["cal_orbit_num", "fma_scan", "lyapunov_scan", "Track_tbt_FMA", "Track_tbt_Lyapunov", "one_turn_matrix",
           "footprint_scan", "Track_tbt_tune_footprint", "cal_tunes_track", "cal_tunes_track_x", "cal_chrom_track",
           "cal_spin_orbit", "cal_spin_tune", "set_initial_spin",
           "twiss_from_matrix", "twiss_at_start", "action_diff", "Track_tbt_ActionDiff", "STError", "delta_to_pt",
           "element_maps", "cal_a", "trace_a", "cal_twiss", "LineTwiss", "cal_dispersion", "cal_chrom",
           "set_line_optics", "Print_Twiss"]
