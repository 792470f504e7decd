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
    """One-turn matrix from GPU-tracked finite differences: RHIC's tunes equal
    MAD-X's (header Q1, Q2), Twiss agrees with the thick-lens optics."""
    gp, ln = W.rhic_example(RHIC)
    x_co, m, (bx, ax, qx, by, ay, qy) = D.twiss_at_start(ctx, gp, ln.records())
    opt = O.Optics(ln)
    assert abs(qx - 0.23) < 2e-6 and abs(qy - 0.22) < 2e-6
    assert abs(bx / opt.beta_x - 1) < 1e-3 and abs(by / opt.beta_y - 1) < 1e-3
    s = np.zeros((6, 6))
    for k in range(3):
        s[2 * k, 2 * k + 1], s[2 * k + 1, 2 * k] = 1., -1.
    assert np.abs(m.T @ s @ m - s).max() < 1e-6


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
