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
