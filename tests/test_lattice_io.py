"""Lattice readers (CPU): the reference's two example lattices read into a Line.

The MAD-X table of RHIC carries MAD-X's own optics in its header, which pins
the reader's element mapping: the host optics of the Line must reproduce the
fractional tunes Q1 = 28.23, Q2 = 29.22, the momentum compaction ALFA and
the LENGTH that MAD-X computed from the same elements.  The Bmad table of the
HSR is pinned by its element census, its patch rows and its total length.
The PATCH map (TRANS_Pass + YROTAT_Pass, simtrack.cpp:6248) is pinned by its
group property (a rotation by theta, then by -theta, then the opposite
translation, is the identity) and by symplecticity.
"""
import gzip
import math
import os
from collections import Counter

import numpy as np
import pytest

import pyoracle
import simtrack_amd as st
from simtrack_amd import lattice_io as LI
from simtrack_amd import optics as O
from simtrack_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
RHIC = os.path.join(HERE, "golden", "lattices", "rhic_2012_madx_twiss.tfs.gz")
HSR = os.path.join(HERE, "golden", "lattices", "hsr_bmad_table.tab.gz")


def _madx_header(path):
    out = {}
    with gzip.open(path, "rt") as f:
        for ln in f:
            if not ln.startswith("@"):
                break
            tok = ln.split()
            if tok[2] == "%le":
                out[tok[1]] = float(tok[3])
    return out


@pytest.fixture(scope="module")
def rhic():
    return LI.read_madx_twiss(RHIC)


@pytest.fixture(scope="module")
def hsr():
    return LI.read_bmad_table(HSR)


def test_rhic_census_and_length(rhic):
    hdr = _madx_header(RHIC)
    assert len(rhic.Cell) == 6858
    c = Counter(e.TYPE for e in rhic.Cell)
    # SOLENOID rows (all KS = 0), HMONITOR, INSTRUMENT, KICKER, RCOLLIMATOR become drifts
    assert c == {"DRIFT": 3648, "MULT": 1320, "MARKER": 588, "QUAD": 439, "SBEND": 213, "VBPM": 166,
                 "SEXT": 144, "HKICKER": 139, "VKICKER": 118, "BPM": 73, "RFCAV": 10}
    assert abs(rhic.circumference - hdr["LENGTH"]) < 1e-6


def test_rhic_optics_match_madx(rhic):
    """The Line's linear optics reproduce MAD-X's tunes and momentum compaction."""
    hdr = _madx_header(RHIC)
    opt = O.Optics(rhic)
    assert abs(opt.tune_x - hdr["Q1"] % 1) < 2e-7
    assert abs(opt.tune_y - hdr["Q2"] % 1) < 2e-7
    assert abs(opt.alpha_c / hdr["ALFA"] - 1) < 1e-4


def test_hsr_census_patches_and_length(hsr):
    assert len(hsr.Cell) == 2838          # 2 header lines read as rows + 2,836 elements
    assert hsr.Cell[0].TYPE == "DRIFT" and hsr.Cell[0].L == 0. and hsr.Cell[1].L == 0.
    c = Counter(e.TYPE for e in hsr.Cell)
    assert c == {"DRIFT": 1522, "MULT": 414, "QUAD": 287, "KICKER": 233, "SBEND": 193, "SEXT": 144,
                 "PATCH": 34, "BPM": 7, "SOLEN": 2, "RFCAV": 2}
    p = [e for e in hsr.Cell if e.NAME == "PE_SOL_R"][0]
    assert (p.DX, p.DY, p.DS, p.THETA, p.L) == (-0.0499947918, 0.0, -0.0006249674, -0.0, -0.0006249674)
    q = [e for e in hsr.Cell if e.NAME == "PB_SOL_R"][0]
    assert q.THETA == -0.025 and q.L == 0.
    assert all(e.KS == 0. for e in hsr.Cell if e.TYPE == "SOLEN")
    # the table's last s is 3833.824773 m; the Line's patches have L = z_offset
    assert abs(hsr.Length - 3833.824773) < 1e-3


def test_unknown_keyword_becomes_drift(tmp_path):
    f = tmp_path / "t.tfs"
    cols = " ".join(["0"] * 29)
    f.write_text('@ NAME %s "x"\n* NAME KEYWORD S L\n$ %s %s %le %le\n'
                 f'"A" "WIGGLER" 1 1.5 {" ".join(["0"] * 27)}\n'
                 f'"B" "QUADRUPOLE" 2 0.5 0 0 0 0 0 0 0.02 0 {" ".join(["0"] * 19)}\n')
    ln = LI.read_madx_twiss(str(f))
    assert [e.TYPE for e in ln.Cell] == ["DRIFT", "QUAD"]
    assert ln.Cell[0].L == 1.5 and ln.Cell[1].KN == 0.02 and ln.Length == 2.0
    del cols


def _orc_line(gp, cells):
    return pyoracle.Oracle(gp, st.Line(cells).records(), nthreads=1)


def test_patch_group_property_and_symplecticity():
    gp = W.proton_globals(275000.)
    r = np.random.Generator(np.random.PCG64(3))
    x = np.zeros((8, 6))
    x[:, :4] = 1e-3 * r.normal(size=(8, 4))
    x[:, 5] = 1e-4 * r.normal(size=8)
    th, dx, ds = 0.025, -0.05, 6.25e-4
    fwd = [st.PATCH("P", dx, 0.001, ds, th)]
    back = [st.PATCH("R", 0., 0., 0., -th), st.PATCH("T", -dx, -0.001, -ds, 0.)]
    out, s, _, _ = _orc_line(gp, fwd + back).track(x.copy(), 1)
    assert s.all()
    assert np.abs(out - x).max() < 1e-15
    # a pure rotation changes px by sin(theta) at first order; a pure translation only shifts
    out1, _, _, _ = _orc_line(gp, [st.PATCH("T", dx, 0., 0., 0.)]).track(x.copy(), 1)
    assert np.array_equal(out1[:, 1:], x[:, 1:]) and np.allclose(out1[:, 0], x[:, 0] - dx, atol=0)
    # symplectic: J^T S J = S for the 6-D map at a point
    orc = _orc_line(gp, fwd)
    p0 = x[0]
    eps = 1e-7
    pts = np.repeat(p0[None, :], 12, axis=0)
    for k in range(6):
        pts[2 * k, k] += eps
        pts[2 * k + 1, k] -= eps
    o, _, _, _ = orc.track(pts, 1)
    j = np.stack([(o[2 * k] - o[2 * k + 1]) / (2 * eps) for k in range(6)], axis=1)
    sm = np.zeros((6, 6))
    for k in range(3):
        sm[2 * k, 2 * k + 1], sm[2 * k + 1, 2 * k] = 1., -1.
    assert np.abs(j.T @ sm @ j - sm).max() < 1e-7


def test_example_lattices_track_on_the_oracle(rhic, hsr):
    for gp, ln in (W.rhic_example(RHIC), W.hsr_example(HSR)):
        opt_ok = ln.Length > 3000
        x = np.zeros((4, 6))
        x[:, 0] = [1e-4, -1e-4, 2e-4, 0.]
        x[:, 2] = [1e-4, 2e-4, -1e-4, 0.]
        out, s, lt, lp = pyoracle.Oracle(gp, ln.records(), nthreads=4).track(x, 2)
        assert opt_ok and s.all(), (lt, lp)
        assert np.isfinite(out).all()
