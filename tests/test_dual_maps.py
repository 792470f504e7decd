"""The first-order maps, differentiated exactly.

The reference takes its one-turn and element maps from tps objects passed
through its maps (Cal_OneTurnMap simtrack.cpp:15122-15179, Cal_ElementMap
cpp:15205-15253).  oracle/dualmap.cpp runs the oracle's maps on dual numbers
(a value and its six first derivatives: a first-order tps), so its Jacobians
are exact to rounding.  These tests pin that build on a ring whose linear
map is known in closed form, and measure the error of the symmetric
finite differences (step 1e-6) that st_one_turn_matrix and st_element_maps
use on the GPU (tests/test_gpu_diagnostics.py compares the GPU with the
dual maps).
"""
import math

import numpy as np

import pyoracle
import simtrack_amd as st
from simtrack_amd import workloads as W

from test_diagnostics import RHIC

S6 = np.kron(np.eye(3), np.array([[0., 1.], [-1., 0.]]))


def _exact_ring():
    """Drifts, exact sector bends and thin quadrupoles: every linear map is
    known exactly."""
    ang = 2 * math.pi / 32
    cells = []
    for c in range(16):
        cells += [st.QUAD("QF", 0., 0.05, 0.), st.DRIFT("D", 2.0), st.SBEND("B", 10.0, ang), st.DRIFT("D", 2.0),
                  st.QUAD("QD", 0., -0.05, 0.), st.DRIFT("D", 2.0), st.SBEND("B", 10.0, ang), st.DRIFT("D", 2.0)]
    return W.proton_globals(100000.), st.Line(cells), ang


def _closed_form(ang):
    """4x4 transverse one-turn matrix: thin lens, drift, sector bend of
    bending radius rho (rotation by the bend angle in x, a drift of the arc
    length in y)."""
    def drift(L):
        m = np.eye(4)
        m[0, 1] = m[2, 3] = L
        return m

    def thin(k):
        m = np.eye(4)
        m[1, 0], m[3, 2] = -k, k
        return m

    rho = 10.0 / ang
    b = np.eye(4)
    b[0:2, 0:2] = [[math.cos(ang), rho * math.sin(ang)], [-math.sin(ang) / rho, math.cos(ang)]]
    b[2, 3] = 10.0
    cell = drift(2.) @ b @ drift(2.) @ thin(-0.05) @ drift(2.) @ b @ drift(2.) @ thin(0.05)
    return np.linalg.matrix_power(cell, 16)


def test_dual_one_turn_matrix_equals_the_closed_form():
    gp, ln, ang = _exact_ring()
    m = pyoracle.dual_one_turn_matrix(ln.records(), gp, np.zeros(6))
    assert np.abs(m[:4, :4] - _closed_form(ang)).max() < 1e-12
    assert np.abs(m.T @ S6 @ m - S6).max() < 1e-13


def test_finite_difference_error_on_rhic():
    """RHIC (6,858 elements, sextupoles and multipoles): the dual one-turn
    matrix is symplectic to 3e-13; the symmetric differences of step 1e-6
    differ from it by 1.2e-7 (1.7e-8 of the largest entry) and are symplectic
    to 1e-7.  Element maps: 3e-8."""
    gp, ln = W.rhic_example(RHIC)
    rec = ln.records()
    o = pyoracle.Oracle(gp, rec)
    xc, xe, _ = o.closed_orbit(0.)
    x0 = xe.copy()
    x0[4], x0[5] = 0., 0.
    md = pyoracle.dual_one_turn_matrix(rec, gp, x0)
    mf = o.one_turn_matrix(x0)
    assert np.abs(md.T @ S6 @ md - S6).max() < 1e-12
    err = np.abs(mf - md).max()
    print(f"\none-turn matrix: finite differences - exact = {err:.2e}; symplecticity "
          f"exact {np.abs(md.T @ S6 @ md - S6).max():.1e}, differences {np.abs(mf.T @ S6 @ mf - S6).max():.1e}")
    assert err < 5e-7
    T, X = o.element_maps(xc, 0.)
    Td = pyoracle.dual_element_maps(rec, gp, X, 0.)
    e_err = np.abs(T - Td).max()
    print(f"element maps: finite differences - exact = {e_err:.2e}")
    assert e_err < 2e-7
    assert max(np.abs(t.T @ S6 @ t - S6).max() for t in Td) < 1e-12
