"""The spin drivers on the oracle (CPU): Cal_Spin_Orbit
(simtrack.cpp:20510-20572), Cal_Spin_Tune (cpp:20574-20601) and
Set_Initial_Spin (cpp:20603-20635), pinned by Thomas-BMT closed forms.

On the closed orbit of a flat ring (no field on the axis but the dipoles')
the one-turn spin map is a rotation about y by 2 pi G gamma: n0 = (0, 1, 0)
and the spin tune is G gamma folded into [0, 1/2].  A partial snake (a
rotation by theta about s) after the ring composes, as quaternions, to
cos(pi nu) = cos(theta/2) cos(pi G gamma) and an axis whose components have
magnitudes (sin(theta/2) |sin(pi G gamma)|, cos(theta/2) |sin(pi G gamma)|,
sin(theta/2) |cos(pi G gamma)|) / sin(pi nu).  Set_Initial_Spin averages
the projections (R^t e_k) . n0 = e_k . n0 over turns, so on the true n0 it
returns n0."""
import math

import numpy as np
import pytest

import pyoracle
import simtrack_amd as st
from simtrack_amd import workloads as W


def flat_ring(snake_angle=None):
    """24 thick-quad FODO cells with sector bends closing 2 pi, and
    optionally a zero-length partial snake about s at the end."""
    gp = W.proton_globals(100000.)
    cells = []
    ang = 2 * math.pi / 48
    for _ in range(24):
        cells += [st.QUAD("QF", 1.0, 0.025, 0.), st.DRIFT("D", 2.0), st.SBEND("B", 10.0, ang),
                  st.DRIFT("D", 2.0), st.QUAD("QD", 1.0, -0.025, 0.), st.DRIFT("D", 2.0),
                  st.SBEND("B", 10.0, ang), st.DRIFT("D", 2.0)]
    for e in cells:
        if e.TYPE == "QUAD":
            e.Nint = 8
    if snake_angle is not None:
        cells.append(st.SNAKE("SN", 0., (0., 0., 1.), snake_angle))
    return gp, st.Line(cells)


def folded(nu):
    nu = nu % 1.0
    return min(nu, 1.0 - nu)


@pytest.fixture(scope="module")
def flat():
    gp, ln = flat_ring()
    return gp, ln.records()


@pytest.fixture(scope="module")
def snake():
    theta = 0.2          # the reference's Newton search keeps Sy = sqrt(1 - Sx^2 - Ss^2) > 0: n0 near vertical
    gp, ln = flat_ring(theta)
    return gp, ln.records(), theta


def test_flat_ring_n0_is_vertical(flat):
    gp, rec = flat
    o = pyoracle.Oracle(gp, rec)
    x0, n0, it = o.spin_orbit(0.)
    assert it <= 3
    assert np.all(np.abs(n0[:, 1] - 1.0) < 1e-12)
    assert np.all(np.abs(n0[:, [0, 2]]) < 1e-9)
    assert np.all(np.abs(x0[:4]) < 1e-12)      # on the design orbit


def test_flat_ring_spin_tune_is_g_gamma(flat):
    gp, rec = flat
    tune, n0 = pyoracle.Oracle(gp, rec).spin_tune(0.)
    assert abs(tune - folded(gp.G * gp.gamma)) < 2e-5
    assert abs(n0[1] - 1.0) < 1e-12


def test_partial_snake_n0_closed_form_and_fixed_point(snake):
    gp, rec, theta = snake
    o = pyoracle.Oracle(gp, rec)
    x0, n0, _ = o.spin_orbit(0.)
    gg = gp.G * gp.gamma
    c = math.cos(theta / 2) * math.cos(math.pi * gg)
    s_nu = math.sqrt(1 - c * c)
    want = np.array([math.sin(theta / 2) * abs(math.sin(math.pi * gg)), math.cos(theta / 2) * abs(math.sin(math.pi * gg)),
                     math.sin(theta / 2) * abs(math.cos(math.pi * gg))]) / s_nu
    assert np.allclose(np.abs(n0[-1]), want, rtol=0, atol=1e-9)
    # n0 is the fixed point of the one-turn spin map
    x = np.concatenate([x0[:6], n0[-1]])[None, :]
    xs, ss, _, _ = o.track(x.copy(), 1, spin=True)
    assert ss[0] == 1
    assert np.max(np.abs(xs[0, 6:] - n0[-1])) < 1e-10
    # and n0 after element j is n0 of the end carried through elements 0..j
    j = len(rec) // 2
    y = x.copy()
    for k in range(j + 1):
        y = o.element_pass(k, y, spin=True)
    assert np.max(np.abs(y[0, 6:] / np.linalg.norm(y[0, 6:]) - n0[j])) < 1e-10


def test_partial_snake_spin_tune(snake):
    gp, rec, theta = snake
    tune, _ = pyoracle.Oracle(gp, rec).spin_tune(0.)
    nu = math.acos(math.cos(theta / 2) * math.cos(math.pi * gp.G * gp.gamma)) / math.pi
    assert abs(tune - folded(nu)) < 2e-5
    # the tune without the snake differs: the snake is seen
    assert abs(tune - folded(gp.G * gp.gamma)) > 3e-3


def test_set_initial_spin_returns_n0(snake):
    gp, rec, _ = snake
    o = pyoracle.Oracle(gp, rec)
    x0, n0, _ = o.spin_orbit(0.)
    x = np.concatenate([x0[:6], [0.3, 0.4, 0.5]])
    out = o.set_initial_spin(50, n0[-1], x)
    assert np.array_equal(out[:6], x[:6])
    assert np.max(np.abs(out[6:] - n0[-1])) < 1e-11
    # projecting on another axis gives another vector
    other = o.set_initial_spin(50, np.array([1., 0., 0.]), x)
    assert np.max(np.abs(other[6:] - n0[-1])) > 1e-2
