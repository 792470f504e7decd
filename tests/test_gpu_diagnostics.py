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
