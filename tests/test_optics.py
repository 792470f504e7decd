"""Host optics and the electron ring (CPU).

The linear model must be the linear part of the tracked maps: the oracle's
one-turn map, differentiated numerically, equals the product of the
optics module's matrices.  The radiation integrals of an isomagnetic ring
have closed forms.  The radiation constants are pinned against CODATA:
the classical loss constant C_gamma = 4 pi r_e / 3 (m c^2)^3, and the mean
and the variance of radiate1's energy change on one bend slice against
Sands' photon-emission rates.  eigen_emittances is checked on a beam of
known emittances and under a symplectic map.
"""
import math

import numpy as np
import pytest

import pyoracle
import simtrack_amd as st
from simtrack_amd import optics as O
from simtrack_amd import workloads as W

R_E = 2.8179403262e-15      # classical electron radius [m], CODATA 2018
HBAR = 1.054571817e-34      # J s
M_E = 9.1093837015e-31      # kg
C = 299792458.0


def _test_ring(gp):
    return W.electron_ring(gp, n_arc_cells=40, n_straight_cells=4, bend_length=4.0, chromaticity=None,
                           harmonic=400, overvoltage=2.0, mu_deg=90., straight_kq_scale=0.85)


@pytest.fixture(scope="module")
def ring():
    gp = W.electron_globals_ring()
    ln = _test_ring(gp)
    return gp, ln, O.Optics(ln)


def test_constants_against_codata():
    cgamma = 4 * math.pi * R_E / (3 * (O.ELECTRON_MEV * 1e-3) ** 3)     # m / GeV^3
    assert abs(O.CR / cgamma - 1) < 2e-6
    assert abs(O.CQ / 3.8319e-13 - 1) < 1e-4


def test_isomagnetic_radiation_integrals(ring):
    gp, ln, opt = ring
    rho = ln.info["rho"]
    assert abs(opt.I2 - 2 * math.pi / rho) < 1e-12 * opt.I2
    assert abs(opt.I3 - 2 * math.pi / rho ** 2) < 1e-12 * opt.I3
    eq = opt.equilibrium(gp)
    # U0 = C_gamma E^4 / rho for an isomagnetic ring
    assert abs(eq["U0_MeV"] - O.CR * (gp.energy / 1e3) ** 4 / rho * 1e3) < 1e-9 * eq["U0_MeV"]
    # Robinson: Jx + Jy + Jz = 4
    assert abs(eq["Jx"] + eq["Jy"] + eq["Jz"] - 4) < 1e-12
    # sigma_delta of an isomagnetic ring: gamma sqrt(Cq / (Jz rho))
    assert abs(eq["sigma_delta"] - gp.gamma * math.sqrt(O.CQ / (eq["Jz"] * rho))) < 1e-12


def _one_turn_jacobian(gp, ln, eps=1e-8):
    g = W.electron_globals(18000.0, H_expand=1)     # no radiation: a symplectic map
    orc = pyoracle.Oracle(g, ln.records(), nthreads=1)
    base = np.zeros((13, 6))
    for k in range(6):
        base[1 + 2 * k, k] = eps
        base[2 + 2 * k, k] = -eps
    out, s, _, _ = orc.track(base, 1)
    assert s.all()
    j = np.zeros((6, 6))
    for k in range(6):
        j[:, k] = (out[1 + 2 * k] - out[2 + 2 * k]) / (2 * eps)
    return j


def test_linear_model_is_the_tracked_map(ring):
    """One-turn Jacobian of the oracle (no RF: the test line's cavities are
    removed) against the optics module's matrices."""
    gp, ln, opt = ring
    cells = [e for e in ln.Cell if e.TYPE != "RFCAV"]
    bare = st.Line(cells)
    o = O.Optics(bare)
    j = _one_turn_jacobian(gp, bare)
    mx = j[np.ix_([0, 1, 5], [0, 1, 5])]
    my = j[np.ix_([2, 3], [2, 3])]
    # the maps integrate the quadrupoles with 4 Yoshida slices: agreement to ~1e-4
    assert np.abs(mx - o.mx).max() < 3e-4 * np.abs(o.mx).max(), (mx, o.mx)
    assert np.abs(my - o.my).max() < 3e-4 * np.abs(o.my).max(), (my, o.my)
    assert abs(math.acos((mx[0, 0] + mx[1, 1]) / 2) / (2 * math.pi) - min(o.tune_x % 1, 1 - o.tune_x % 1)) < 1e-4
    # uncoupled: no x-y terms
    assert np.abs(j[np.ix_([0, 1], [2, 3])]).max() < 1e-9
    # momentum compaction from the map: dz/ddelta = -alpha_c C (+ the tiny 1/gamma^2 slip)
    slip = bare.Length / gp.gamma ** 2
    assert abs(-j[4, 5] - (o.alpha_c * o.length - slip)) < 2e-3 * o.alpha_c * o.length


def test_twiss_periodic_and_symplectic(ring):
    gp, ln, opt = ring
    mx = opt.mx[:2, :2]
    assert abs(np.linalg.det(mx) - 1) < 1e-10 and abs(np.linalg.det(opt.my) - 1) < 1e-10
    t = np.array([[opt.beta_x, -opt.alpha_x], [-opt.alpha_x, (1 + opt.alpha_x ** 2) / opt.beta_x]])
    assert np.allclose(mx @ t @ mx.T, t, rtol=1e-9, atol=1e-12)
    eta = np.array([opt.D, opt.Dp, 1.])
    assert np.allclose(opt.mx @ eta, eta, atol=1e-12)


def test_radiate1_photon_statistics():
    """Energy change of on-axis electrons through one bend slice: the mean
    is the classical loss C_gamma E^3 / (2 pi) theta^2 / L and the variance
    is Sands' (55 / 24 sqrt 3) r_e hbar / m  gamma^5 h^3 L / c."""
    g = W.electron_globals(18000.0, H_expand=1, radiate=1, rng_seed=4242)
    rho, L = 50.0, 0.2
    b = st.SBEND("B", L, L / rho)
    b.Nint = 1
    rec = st.Line([b]).records()
    orc = pyoracle.Oracle(g, rec)
    n = 20000
    out = orc.element_pass(0, np.zeros((n, 6)))
    dpt = out[:, 5]
    h = 1 / rho
    mean_ref = -O.CR * (g.energy / 1e3) ** 3 / (2 * math.pi) * (L * h) ** 2 / L
    var_ref = 55 / (24 * math.sqrt(3)) * R_E * HBAR / M_E * g.gamma ** 5 * h ** 3 * L / C
    assert abs(dpt.mean() / mean_ref - 1) < 4 * math.sqrt(var_ref / n) / abs(mean_ref) + 1e-6
    assert abs(dpt.var() / var_ref - 1) < 0.04


def test_eigen_emittances_recover_known_beam(ring):
    gp, ln, opt = ring
    x = opt.matched_beam(200000, 3e-6, 1e-7, 2e-3, 0.01, seed=9)
    em = sorted(O.eigen_emittances(x))
    assert abs(em[0] / 1e-7 - 1) < 0.02
    assert abs(em[1] / 3e-6 - 1) < 0.02
    assert abs(em[2] / (0.01 * 2e-3) - 1) < 0.02
    # invariant under a symplectic map, exp(S H) with H symmetric (couples all planes)
    from scipy.linalg import expm
    s = np.zeros((6, 6))
    for k in range(3):
        s[2 * k, 2 * k + 1], s[2 * k + 1, 2 * k] = 1., -1.
    r = np.random.Generator(np.random.PCG64(5))
    h = r.normal(size=(6, 6))
    m = expm(s @ (0.3 * (h + h.T)))
    assert np.allclose(m.T @ s @ m, s, atol=1e-10)
    em2 = sorted(O.eigen_emittances(x @ m.T))
    assert np.allclose(em2, em, rtol=1e-9)


def test_config4_ring_shape():
    gp = W.electron_globals_ring()
    ln = W.electron_ring(gp)
    n = len(ln.Cell)
    assert 3300 <= n <= 3700
    opt = O.Optics(ln)
    eq = opt.equilibrium(gp)
    assert 30 < eq["U0_MeV"] < 50
    for q in (opt.tune_x, opt.tune_y):
        f = q % 1.0
        assert 0.05 < f < 0.45 or 0.55 < f < 0.95
    zs, qs = W.rf_synchronous(opt, gp, ln)
    assert 0 < qs < 0.2
    assert gp.G == W.ELECTRON_G and gp.radiate == 1 and gp.H_expand == 1
