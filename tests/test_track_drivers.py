"""Tracking drivers on the oracle (CPU): Track_wrt_OnMomentumPart
(simtrack.cpp:20942-20966), Cal_Tunes_Track (cpp:20971-21026) and
Cal_Chrom_Track (cpp:21028-21051).

Track_wrt_OnMomentumPart is Track_Radial (cpp:21641) with the caller's
per-element z offsets: with the on-momentum z after every element it is
Track_Radial bit for bit, with zeros it is Track bit for bit.  The tracked
tunes of a linear FODO ring are its thick-lens tunes, and the tracked
chromaticity is the momentum derivative of the one-turn matrix's tunes."""
import math

import numpy as np

import pyoracle
from simtrack_amd import diagnostics as D
from simtrack_amd import optics as O
from test_diagnostics import _fodo


def x04_of(orc, n):
    """the on-momentum z after every element, one pass from zero (cpp:21655-21661)"""
    out = np.zeros(n)
    x = np.zeros((1, 6))
    for j in range(n):
        x[0, 4] = 0.
        x[0, 5] = 0.
        x = orc.element_pass(j, x)
        out[j] = x[0, 4]
    return out


def _rf_ring():
    gp, ln = _fodo(2e-5)
    import simtrack_amd as st
    from simtrack_amd import workloads as W
    cells = list(ln.Cell) + [st.RFCAV("RF", 0., 3.0, 0., 0.)]
    ln2 = st.Line(cells)
    W.set_rf_frequency(ln2, gp)
    return gp, ln2


def _particles():
    r = np.random.Generator(np.random.PCG64(5))
    x = np.zeros((16, 6))
    x[:, 0] = r.normal(0, 1e-3, 16)
    x[:, 2] = r.normal(0, 1e-3, 16)
    x[:, 4] = r.normal(0, 1e-2, 16)
    x[:, 5] = r.normal(0, 3e-4, 16)
    return x


def test_track_dz_with_on_momentum_z_is_track_radial():
    gp, ln = _rf_ring()
    rec = ln.records()
    orc = pyoracle.Oracle(gp, rec, nthreads=1)
    x = _particles()
    dz0 = x04_of(orc, len(rec))
    assert np.abs(dz0).max() > 0, "the ring's on-momentum z must move for this check to bite"
    a, sa, la, pa = orc.track_dz(x, 40, dz0)
    b, sb, lb, pb = orc.track(x.copy(), 40, radial=True)
    assert np.array_equal(a.view(np.int64), b.view(np.int64))
    assert np.array_equal(sa, sb) and np.array_equal(la, lb) and np.array_equal(pa, pb)
    c, sc, _, _ = orc.track_dz(x, 40, np.zeros(len(rec)))
    d, sd, _, _ = orc.track(x.copy(), 40)
    assert np.array_equal(c.view(np.int64), d.view(np.int64)) and np.array_equal(sc, sd)
    assert not np.array_equal(a, c)


def oracle_tunes_track(orc, x, nturn, lo, hi):
    """Cal_Tunes_Track(Line, x, Nturn, ...), cpp:20971-20990, on the oracle"""
    x = np.array(x, dtype=np.float64).reshape(1, 6).copy()
    if x[0, 0] < 1.0e-9:
        x[0, 0] = x[0, 0] + 1.0e-9
    if x[0, 2] < 1.0e-9:
        x[0, 2] = x[0, 2] + 1.0e-9
    _, tbt, s, _, _ = orc.track(x, nturn, tbt=True)
    if s[0] != 1:
        return math.nan, math.nan
    return pyoracle.fine_tune(tbt[0, :, 0], lo, hi), pyoracle.fine_tune(tbt[0, :, 2], lo, hi)


def oracle_tunes_at(orc, deltap, lo, hi):
    """Cal_Tunes_Track(Line, deltap0), cpp:20992-21026, on the oracle"""
    _, xe, _ = orc.closed_orbit(deltap)
    return oracle_tunes_track(orc, [xe[0], xe[1], xe[2], xe[3], 0., deltap], 1024, lo, hi)


def tune_ring():
    """24 FODO cells with fractional tunes 0.258 / 0.333 (their aliases
    1 - q lie outside the search window [0.2, 0.4])"""
    import simtrack_amd as st
    from simtrack_amd import workloads as W
    cells = []
    ang = 2 * math.pi / 48
    for _ in range(24):
        cells += [st.QUAD("QF", 1.0, 0.03, 0.), st.DRIFT("D", 2.0), st.SBEND("B", 10.0, ang),
                  st.DRIFT("D", 2.0), st.QUAD("QD", 1.0, -0.028, 0.), st.DRIFT("D", 2.0),
                  st.SBEND("B", 10.0, ang), st.DRIFT("D", 2.0)]
    for e in cells:
        if e.TYPE == "QUAD":
            e.Nint = 8
    return W.proton_globals(100000.), st.Line(cells)


WINDOW = (0.2, 0.4)


def test_tracked_tunes_are_the_linear_tunes():
    gp, ln = tune_ring()
    orc = pyoracle.Oracle(gp, ln.records(), nthreads=1)
    opt = O.Optics(ln)
    assert WINDOW[0] < opt.tune_x < WINDOW[1] and WINDOW[0] < opt.tune_y < WINDOW[1]
    qx, qy = oracle_tunes_at(orc, 0., *WINDOW)
    assert abs(qx - opt.tune_x) < 2e-3 and abs(qy - opt.tune_y) < 2e-3


def test_tracked_chromaticity_is_the_matrix_chromaticity():
    """the tracked tunes (Cal_Tunes_Track's first form, 1,024 turns) about
    the closed orbit at +-0.0003 with a betatron offset, differenced as
    Cal_Chrom_Track does, against the tunes of the one-turn matrix there
    (the offset, 2 mm, keeps the betatron line above the leakage of the
    dispersive orbit, which FineTuneFinder does not subtract).
    (Cal_Chrom_Track's own start, the closed orbit raised by 1e-9 only where
    it is below 1e-9, leaves no betatron motion in x where the dispersion is
    positive: its x tunes are then the window's leakage peak, as in the
    reference.)"""
    gp, ln = tune_ring()
    orc = pyoracle.Oracle(gp, ln.records(), nthreads=1)

    def tracked(dp):
        _, xe, _ = orc.closed_orbit(dp)
        return oracle_tunes_track(orc, [xe[0] + 2e-3, xe[1], xe[2] + 2e-3, xe[3], 0., dp], 1024, *WINDOW)

    def matrix(dp):
        xc, _, _ = orc.closed_orbit(dp)
        _, _, qx, _, _, qy = D.twiss_from_matrix(orc.one_turn_matrix(xc, 1e-6))
        return qx, qy
    qp, qm = tracked(0.0003), tracked(-0.0003)
    cx, cy = (qp[0] - qm[0]) / 2. / 0.0003, (qp[1] - qm[1]) / 2. / 0.0003
    mp, mm = matrix(0.0003), matrix(-0.0003)
    mx, my = (mp[0] - mm[0]) / 2. / 0.0003, (mp[1] - mm[1]) / 2. / 0.0003
    assert mx < -0.3 and my < -0.3, "a FODO ring without sextupoles has a negative natural chromaticity"
    assert abs(cx / mx - 1) < 0.02 and abs(cy / my - 1) < 0.02


class OracleTable:
    """The oracle behind the Table interface footprint_scan uses
    (track_tbt, ctx.fine_tune), so the same host code runs on both."""

    def __init__(self, gp, rec):
        self.orc = pyoracle.Oracle(gp, rec, nthreads=1)
        self.ctx = self

    def track_tbt(self, x, nturn):
        xo, tbt, s, lt, lp = self.orc.track(np.asarray(x, dtype=np.float64), nturn, tbt=True)
        return xo, tbt, s, lt, lp

    def fine_tune(self, sig, lo, hi):
        return np.array([pyoracle.fine_tune(r, lo, hi) for r in np.atleast_2d(sig)])


def test_footprint_of_a_linear_ring_is_the_linear_tunes():
    """Track_tbt_tune_footprint's grid (cpp:21100-21176) on a ring without
    sextupoles: every amplitude has the linear tunes, in both windows"""
    gp, ln = tune_ring()
    rec = ln.records()
    orc = pyoracle.Oracle(gp, rec, nthreads=1)
    opt = O.Optics(ln)
    m = orc.one_turn_matrix(np.zeros(6), 1e-6)
    bx, ax, _, by, ay, _ = D.twiss_from_matrix(m)
    emit = 5e-9
    rows = D.footprint_scan(OracleTable(gp, rec), 0., emit, emit, (bx, ax, by, ay), np.zeros(6), *WINDOW)
    assert [(r[0], r[1]) for r in rows] == [(15 * (i + 1), j) for i in range(5) for j in range(6)]
    for r in rows:
        assert abs(r[2] - opt.tune_x) < 2e-3 and abs(r[4] - opt.tune_y) < 2e-3
        assert abs(r[6]) < 1e-5 and abs(r[7]) < 1e-5, "no tune drift between the windows of a linear ring"
