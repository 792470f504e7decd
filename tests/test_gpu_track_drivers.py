"""Tracking drivers on the GPU against the oracle, bit for bit:
Track_wrt_OnMomentumPart (st_track_dz, simtrack.cpp:20942-20966), the
single-particle DA bisection lives in test_gpu_da.py, and Cal_Tunes_Track /
Cal_Chrom_Track (cpp:20971-21051) through the GPU's closed orbit,
turn-by-turn tracking and FineTuneFinder."""
import numpy as np
import pytest

import pyoracle
import simtrack_amd as st
from simtrack_amd import diagnostics as D
from test_track_drivers import WINDOW, _particles, _rf_ring, oracle_tunes_at, tune_ring, x04_of

pytestmark = pytest.mark.gpu


def test_track_dz_equals_oracle(ctx):
    gp, ln = _rf_ring()
    rec = ln.records()
    orc = pyoracle.Oracle(gp, rec)
    tab = st.Table(ctx, gp, rec)
    x = _particles()
    r = np.random.Generator(np.random.PCG64(3))
    for dz0 in (x04_of(orc, len(rec)), r.normal(0, 1e-6, len(rec))):
        xg, sg, lg, pg = tab.track_dz(x.copy(), 40, dz0)
        xo, so, lo, po = orc.track_dz(x, 40, dz0)
        assert np.array_equal(xg.view(np.int64), xo.view(np.int64))
        assert np.array_equal(sg, so) and np.array_equal(lg, lo) and np.array_equal(pg, po)
    tab.close()


def test_tunes_and_chromaticity_equal_oracle(ctx):
    gp, ln = tune_ring()
    rec = ln.records()
    orc = pyoracle.Oracle(gp, rec)
    for dp in (0.0003, -0.0003, 0.):
        qg = D.cal_tunes_track(ctx, gp, rec, dp, *WINDOW)
        qo = oracle_tunes_at(orc, dp, *WINDOW)
        assert qg == qo, f"deltap {dp}: {qg} vs {qo}"
    cg = D.cal_chrom_track(ctx, gp, rec, *WINDOW)
    qp, qm, q0 = (oracle_tunes_at(orc, d, *WINDOW) for d in (0.0003, -0.0003, 0.))
    co = ((qp[0] - qm[0]) / 2. / 0.0003, (qp[1] - qm[1]) / 2. / 0.0003,
          (qp[0] + qm[0] - 2 * q0[0]) / 2. / 0.0003 / 0.0003, (qp[1] + qm[1] - 2 * q0[1]) / 2. / 0.0003 / 0.0003)
    assert cg == co


def test_footprint_equals_oracle(ctx, tmp_path):
    """Track_tbt_tune_footprint (cpp:21100-21176): the GPU's turn-by-turn
    tracking and FineTuneFinder against the oracle's on the same grid, bit
    for bit; and the full driver (Cal_Twiss and Cal_Orbit_Num on the GPU)
    writes one line per stable grid particle"""
    from test_track_drivers import OracleTable
    gp, ln = tune_ring()
    rec = ln.records()
    tab = st.Table(ctx, gp, rec)
    twiss = (33.0, 0.4, 28.0, -0.3)
    x_end = np.array([1e-4, -2e-6, 5e-5, 1e-6, 0., 0.])
    for emit in (5e-9, 2e-6):
        rg = D.footprint_scan(tab, 0., emit, emit, twiss, x_end, *WINDOW)
        ro = D.footprint_scan(OracleTable(gp, rec), 0., emit, emit, twiss, x_end, *WINDOW)
        assert rg == ro
    f = tmp_path / "footprint.dat"
    rows = D.Track_tbt_tune_footprint(ctx, gp, rec, 0., 5e-9, 5e-9, *WINDOW, filename=str(f), tab=tab)
    lines = f.read_text().splitlines()
    assert len(lines) == len(rows) == 30 and lines[0].startswith("15 0 ")
    tab.close()
