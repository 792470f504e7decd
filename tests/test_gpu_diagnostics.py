"""Closed orbit and turn-by-turn scans on the GPU against the oracle's
one-particle-at-a-time restatement of the reference's drivers, bit for bit."""
import math
import os

import numpy as np
import pytest

import pyoracle
import simtrack_amd as st
from simtrack_amd import diagnostics as D
from simtrack_amd import optics as O
from simtrack_amd import workloads as W

from test_diagnostics import RHIC, _fodo
from test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("deltap", [0., 1e-3])
def test_closed_orbit_rhic(ctx, deltap):
    gp, ln = W.rhic_example(RHIC)
    rec = ln.records()
    xc, xe, it = D.cal_orbit_num(ctx, gp, rec, deltap)
    oc, oe, oit = pyoracle.Oracle(gp, rec).closed_orbit(deltap)
    assert it == oit
    assert_bitwise(xc, oc, "closed orbit")
    assert_bitwise(xe, oe, "closed orbit after one more turn")


def test_closed_orbit_with_kick(ctx):
    gp, ln = _fodo(2e-6)
    rec = ln.records()
    xc, xe, _ = D.cal_orbit_num(ctx, gp, rec, 0.)
    oc, oe, _ = pyoracle.Oracle(gp, rec).closed_orbit(0.)
    assert_bitwise(xc, oc, "closed orbit")
    assert abs(xc[0]) > 1e-7


def test_fma_scan_bitwise(ctx):
    gp, ln = W.config0()[:2]
    rec = ln.records()
    opt = O.Optics(ln)
    eps, _ = W.beam_sizes(gp)
    sx, sy = math.sqrt(eps * opt.beta_x), math.sqrt(eps * opt.beta_y)
    xc, xe, _ = D.cal_orbit_num(ctx, gp, rec, 0.)
    lo, hi = min(opt.tune_x, opt.tune_y) - 0.03, max(opt.tune_x, opt.tune_y) + 0.03
    args = (3e-4, sx, sy, xe, lo, hi, 128, 1.0, 6.0)
    got = D.fma_scan(st.Table(ctx, gp, rec), *args)
    ref = pyoracle.Oracle(gp, rec).fma_scan(*args)
    assert len(got) == len(ref) > 10
    assert_bitwise(got, ref, "FMA rows")


def test_lyapunov_scan_bitwise(ctx):
    gp, ln = W.config0()[:2]
    rec = ln.records()
    eps, sig = W.beam_sizes(gp)
    args = (3e-4, 3 * sig, 3 * sig, 64, 1.5, 6.0)
    got = D.lyapunov_scan(st.Table(ctx, gp, rec), *args)
    ref = pyoracle.Oracle(gp, rec).lyapunov_scan(*args)
    assert len(got) == len(ref) > 5
    assert_bitwise(got, ref, "Lyapunov rows")


def test_twiss_at_start_rhic(ctx):
    """One-turn matrix from GPU-tracked finite differences: GPU = oracle bit
    for bit; RHIC's tunes equal MAD-X's (header Q1, Q2), Twiss agrees with
    the thick-lens optics, and the matrix is symplectic to the accuracy of
    the differences."""
    gp, ln = W.rhic_example(RHIC)
    rec = ln.records()
    x_co, m, (bx, ax, qx, by, ay, qy) = D.twiss_at_start(ctx, gp, rec)
    _, x_end, _ = D.cal_orbit_num(ctx, gp, rec, 0.)
    x0 = x_end.copy()
    x0[4:6] = 0.
    assert_bitwise(m, pyoracle.Oracle(gp, rec).one_turn_matrix(x0), "RHIC one-turn matrix")
    opt = O.Optics(ln)
    assert abs(qx - 0.23) < 2e-6 and abs(qy - 0.22) < 2e-6
    assert abs(bx / opt.beta_x - 1) < 1e-5 and abs(by / opt.beta_y - 1) < 1e-5
    s = np.zeros((6, 6))
    for k in range(3):
        s[2 * k, 2 * k + 1], s[2 * k + 1, 2 * k] = 1., -1.
    assert np.abs(m.T @ s @ m - s).max() < 1e-6


def test_cal_twiss_rhic(ctx):
    """Cal_Twiss along RHIC's 6,858 elements: the element maps and orbits
    (st_element_maps) equal the oracle's bit for bit and the last orbit is
    Cal_Orbit_Num's; the accumulated phases are MAD-X's tunes Q1 = 28.23,
    Q2 = 29.22 with their integer parts; the mode betas after every element
    equal the thick-lens uncoupled optics propagated element by element, and
    r = 1, C = 0 (RHIC has no coupling)."""
    gp, ln = W.rhic_example(RHIC)
    rec = ln.records()
    tw = D.cal_twiss(ctx, gp, rec)
    x_co, x_end, _ = D.cal_orbit_num(ctx, gp, rec, 0.)
    T, X = D.element_maps(ctx, gp, rec, x_co, 0.)
    To, Xo = pyoracle.Oracle(gp, rec).element_maps(x_co, 0.)
    assert_bitwise(T, To, "RHIC element maps")
    assert_bitwise(X, Xo, "RHIC element orbits")
    assert np.array_equal(X[-1], x_end)
    assert abs(tw.tune1 - 28.23000005) < 1e-6 and abs(tw.tune2 - 29.22000006) < 1e-6
    opt = O.Optics(ln)
    tx = np.array([[opt.beta_x, -opt.alpha_x], [-opt.alpha_x, (1 + opt.alpha_x ** 2) / opt.beta_x]])
    ty = np.array([[opt.beta_y, -opt.alpha_y], [-opt.alpha_y, (1 + opt.alpha_y ** 2) / opt.beta_y]])
    worst = 0.
    for k, e in enumerate(ln.Cell):
        a, b = O.element_matrices(e)
        tx = a[:2, :2] @ tx @ a[:2, :2].T
        ty = b @ ty @ b.T
        worst = max(worst, abs(tw.beta1[k] / tx[0, 0] - 1), abs(tw.beta2[k] / ty[0, 0] - 1))
    assert worst < 1e-5, worst
    assert np.abs(tw.r - 1).max() < 1e-9 and np.abs(tw.c).max() < 1e-6


def test_dispersion_and_chromaticity_rhic(ctx):
    """Cal_Dispersion and Cal_Chrom on RHIC against the MAD-X table's header:
    DXMAX = 1.831440463 and DXRMS = 0.9981399112 to 2e-5, no vertical
    dispersion, and the linear chromaticities DQ1 = DQ2 = 1 to 1 %."""
    gp, ln = W.rhic_example(RHIC)
    rec = ln.records()
    tw = D.cal_twiss(ctx, gp, rec)
    eta = D.cal_dispersion(tw.M, tw.T, gp)
    assert abs(eta[:, 0].max() / 1.831440463 - 1) < 2e-5
    assert abs(np.sqrt((eta[:, 0] ** 2).mean()) / 0.9981399112 - 1) < 2e-5
    assert np.abs(eta[:, 2]).max() < 1e-10
    cx1, cy1, cx2, cy2 = D.cal_chrom(ctx, gp, rec)
    assert abs(cx1 - 1) < 0.01 and abs(cy1 - 1) < 0.01
    assert math.isfinite(cx2) and math.isfinite(cy2)


def test_action_diff_bitwise(ctx):
    gp, ln = W.config0()[:2]
    rec = ln.records()
    opt = O.Optics(ln)
    eps, _ = W.beam_sizes(gp)
    args = (opt.beta_x, opt.alpha_x, opt.beta_y, opt.alpha_y, 4.0, 0.7, eps, eps, 128, 3, 40)
    got, sd = D.action_diff(st.Table(ctx, gp, rec), *args)
    ref, sdo = pyoracle.Oracle(gp, rec).action_diff(*args)
    assert sd == sdo
    assert_bitwise(got, ref, "action diffusion rows")
    assert (got[:, 1] > 30).all()


def test_cpp_action_diff_matches_python(ctx, tmp_path):
    """Cal_Orbit_Num, Cal_OneTurnMap, Track_tbt_ActionDiff, Cal_Twiss,
    Cal_Dispersion, Cal_Chrom and Print_Twiss of the C++ API
    (tests/cpp/rhic_diagnostics.cpp) on RHIC: the one-turn matrix, the Twiss,
    the seed, Diff-output.dat and the optics along the line equal the Python
    path's."""
    import gzip
    import subprocess
    exe = os.path.join(os.path.dirname(RHIC), "..", "..", "cpp", "build", "rhic_diagnostics")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/build/rhic_diagnostics is not built (run __graft_entry__.build())")
    plain = tmp_path / "rhic.tfs"
    plain.write_bytes(gzip.open(RHIC).read())
    r = subprocess.run([exe, str(plain)], cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    v = np.frombuffer(r.stdout, dtype=np.float64)
    gp, ln = W.rhic_example(RHIC)
    rec = ln.records()
    n = len(rec)
    assert v.size == 6 + 36 + 5 + 6 + 8 * n
    x_last, m_cpp, tw, seed_cpp = v[:6], v[6:42].reshape(6, 6), v[42:46], v[46]
    q, per = v[47:53], v[53:].reshape(8, n)
    _, x_end, _ = D.cal_orbit_num(ctx, gp, rec, 0.)
    assert np.abs(x_last - x_end).max() < 1e-15
    x0 = x_last.copy()
    x0[4:6] = 0.
    assert_bitwise(m_cpp, D.one_turn_matrix(ctx, gp, rec, x0), "C++ Cal_OneTurnMap")
    bx, ax, _, by, ay, _ = D.twiss_from_matrix(m_cpp)
    assert np.allclose(tw, [bx, ax, by, ay], rtol=1e-14, atol=0)
    out = tmp_path / "py-Diff-output.dat"
    rows, sd = D.Track_tbt_ActionDiff(st.Table(ctx, gp, rec), tuple(tw), 4.0, 0.7, 1e-8, 1e-8, 64, 2, 16,
                                      filename=str(out))
    assert sd == seed_cpp
    assert (tmp_path / "Diff-output.dat").read_text() == out.read_text()
    assert rows[:, 1].min() == 16
    # Cal_Twiss / Cal_Dispersion / Cal_Chrom of the C++ API (its own 4x4
    # eigenvectors) against the Python path (numpy's): equal to rounding
    pt = D.cal_twiss(ctx, gp, rec)
    eta = D.cal_dispersion(pt.M, pt.T, gp)
    ref = np.array([pt.beta1, pt.alfa1, pt.beta2, pt.alfa2, pt.mu1, pt.mu2, pt.r, eta[:, 0]])
    for f in range(8):
        assert np.abs(per[f] - ref[f]).max() < 1e-8 * max(1., np.abs(ref[f]).max()), f
    assert abs(q[0] - pt.tune1) < 1e-10 and abs(q[1] - pt.tune2) < 1e-10
    chrom = D.cal_chrom(ctx, gp, rec)
    assert np.allclose(q[2:], chrom, rtol=1e-6, atol=1e-6)
    lines = (tmp_path / "twiss0.dat").read_text().splitlines()
    assert len(lines) == n + 1 and lines[0].split()[:4] == ["NAME", "TYPE", "S", "BETA1"]
    assert float(lines[-1].split()[3]) == pytest.approx(pt.beta1[-1], rel=1e-5)
    # the Python writer gives the same table from the Python optics
    D.set_line_optics(ln, pt, eta)
    D.Print_Twiss(ln, str(tmp_path / "py-twiss0.dat"))
    py = (tmp_path / "py-twiss0.dat").read_text().splitlines()
    assert len(py) == len(lines) and py[0] == lines[0]
    # numbers compared as printed (6 digits): a value within rounding of a
    # printed boundary may differ in its last digit between the two paths
    same = sum(a.split()[-11:] == b.split()[-11:] for a, b in zip(py[1:], lines[1:]))
    assert same >= 0.99 * n, same


def test_gpu_maps_against_exact_first_order_maps(ctx):
    """st_one_turn_matrix and st_element_maps (symmetric differences of step
    1e-6 on the GPU) against the exact Jacobians of the same maps (the dual
    build of the oracle, a first-order tps as Cal_OneTurnMap/Cal_ElementMap
    take, simtrack.cpp:15122, 15205): within the finite-difference error."""
    gp, ln = W.rhic_example(RHIC)
    rec = ln.records()
    x_co, x_end, _ = D.cal_orbit_num(ctx, gp, rec, 0.)
    x0 = x_end.copy()
    x0[4], x0[5] = 0., 0.
    m = D.one_turn_matrix(ctx, gp, rec, x0)
    md = pyoracle.dual_one_turn_matrix(rec, gp, x0)
    assert np.abs(m - md).max() < 5e-7
    T, X = D.element_maps(ctx, gp, rec, x_co, 0.)
    Td = pyoracle.dual_element_maps(rec, gp, X, 0.)
    assert np.abs(T - Td).max() < 2e-7
