"""The oracle's frame changes, cooling, long-range beam-beam, electron lenses,
AC elements and solenoid (TRANS, SROTAT, YROTAT, XROTAT, COOLING, LRBB,
ELENS, HELENS, HACDIP/VACDIP, HACMULT/VACMULT, SOLEN) pinned by what the
mathematics fixes, not by retyping their expressions: exact translations and scalings,
rotation matrices, the x<->y mirror between the two Forest rotations and
their inverses, symplecticity, the round-beam and Bassetti-Erskine kicks at
the separated position, the thin-lens limits of the lenses, and the AC
kicks turn by turn (the reference's GP.turn counts from 0 in Track,
simtrack.cpp:20799) against closed forms and against KICKER's own spin
pass, and the solenoid's linear map against the textbook hard-edge matrix,
its rotational symmetry about s and the (1 + G) KS spin precession."""
import math

import numpy as np
import pytest
from scipy.special import wofz

import pyoracle
import simtrack_amd as st
from simtrack_amd import workloads as W

PI = 3.14159265358979323
RP = 1.53469826769e-18
GP = W.proton_globals(100000.0)


def orc(*els):
    return pyoracle.Oracle(GP, st.Line(list(els)).records())


def particles(n=16, seed=5, scale=3e-4, spin=False):
    return W.random_particles(n, seed=seed, scale=scale, spin=spin)


def one_pass(el, x, spin=False):
    return orc(el).element_pass(0, x, spin=spin)


def test_trans_and_cooling_are_exact():
    x = particles()
    y = one_pass(st.TRANS("T", 2e-4, -1e-4, 0.05), x)
    assert np.array_equal(y[:, 0], x[:, 0] - 2e-4) and np.array_equal(y[:, 2], x[:, 2] + 1e-4)
    assert np.array_equal(y[:, 4], x[:, 4] - 0.05)
    assert np.array_equal(y[:, [1, 3, 5]], x[:, [1, 3, 5]])
    y = one_pass(st.COOLING("C", 3e-3), x)
    assert np.array_equal(y[:, :4], (1. - 3e-3) * x[:, :4]) and np.array_equal(y[:, 4:], x[:, 4:])


def test_srotat_is_a_rotation_of_both_planes():
    x = particles()
    psi = 0.37
    y = one_pass(st.SROTAT("S", psi), x)
    c, s = math.cos(psi), math.sin(psi)
    for a, b in ((0, 2), (1, 3)):
        np.testing.assert_allclose(y[:, a], c * x[:, a] + s * x[:, b], rtol=0, atol=1e-18)
        np.testing.assert_allclose(y[:, b], c * x[:, b] - s * x[:, a], rtol=0, atol=1e-18)
    back = one_pass(st.SROTAT("S", -psi), y)
    np.testing.assert_allclose(back, x, rtol=0, atol=1e-18)
    assert np.array_equal(one_pass(st.SROTAT("S0", 0.), x), x)


def _swap(x):
    return x[:, [2, 3, 0, 1, 4, 5]]


@pytest.mark.parametrize("angle", [0.011, -0.2])
def test_xrotat_mirrors_yrotat_and_both_invert(angle):
    x = particles(scale=1e-3)
    yx = one_pass(st.XROTAT("X", angle), x)
    yy = one_pass(st.YROTAT("Y", angle), _swap(x))
    # x <-> y exchanges the two rotations (Forest's frame rotations about x and y)
    np.testing.assert_allclose(_swap(yy), yx, rtol=0, atol=1e-16)
    for el, inv in ((st.XROTAT("X", angle), st.XROTAT("X", -angle)), (st.YROTAT("Y", angle), st.YROTAT("Y", -angle))):
        back = one_pass(inv, one_pass(el, x))
        np.testing.assert_allclose(back, x, rtol=0, atol=2e-16)
    # a rotation tilts px by the angle for the on-axis particle: px' = pz sin(theta)
    x0 = np.zeros((1, 6))
    y = one_pass(st.YROTAT("Y", angle), x0)
    assert y[0, 1] == pytest.approx(math.sin(angle), rel=1e-15)
    assert np.array_equal(one_pass(st.YROTAT("Y", 5e-10), x), x)


def _jacobian(el, x0, h=1e-7):
    J = np.zeros((6, 6))
    for k in range(6):
        d = np.zeros((2, 6))
        d[0, k], d[1, k] = h, -h
        y = one_pass(el, x0[None, :] + d)
        J[:, k] = (y[0] - y[1]) / (2 * h)
    return J


@pytest.mark.parametrize("el", [st.XROTAT("X", 0.05), st.YROTAT("Y", -0.08), st.SROTAT("S", 0.4),
                                st.LRBB("LR", 2e11, 1.0, 5e-4, 2e-4, 2e-4, 1e-4)])
def test_maps_are_symplectic(el):
    S = np.zeros((6, 6))
    for i in range(3):
        S[2 * i, 2 * i + 1], S[2 * i + 1, 2 * i] = 1., -1.
    J = _jacobian(el, np.array([3e-4, -2e-5, 1e-4, 4e-5, 0.01, 2e-4]))
    np.testing.assert_allclose(J.T @ S @ J, S, rtol=0, atol=2e-8)


def _be_kick(x, y, sx, sy, c2N):
    """Bassetti-Erskine field of an elliptic Gaussian (sigmax > sigmay) with
    scipy's Faddeeva function; returns (Dpx, Dpy) for x, y >= 0."""
    d = math.sqrt(2 * (sx * sx - sy * sy))
    w1 = wofz(complex(x, y) / d)
    w2 = wofz(complex(x * sy / sx, y * sx / sy) / d)
    e = math.exp(-x * x / 2 / sx / sx - y * y / 2 / sy / sy)
    f = c2N * math.sqrt(PI / 2 / (sx * sx - sy * sy)) * (w1 - e * w2)
    return f.imag, f.real


def test_lrbb_is_the_4d_kick_at_the_separated_position():
    x = particles(8, scale=2e-4)
    N, scale, sepx, sepy, sig = 1.7e11, 0.8, 6e-4, -2e-4, 1.5e-4
    y = one_pass(st.LRBB("LR", N, scale, sepx, sepy, sig, sig), x)
    assert np.array_equal(y[:, 0], (x[:, 0] + sepx) - sepx) and np.array_equal(y[:, 2], (x[:, 2] + sepy) - sepy)
    X, Y = x[:, 0] + sepx, x[:, 2] + sepy
    r2 = X * X + Y * Y
    k = 2 * N * RP / GP.gamma / r2 * (1 - np.exp(-r2 / 2 / sig ** 2)) * scale
    np.testing.assert_allclose(y[:, 1] - x[:, 1], k * X, rtol=1e-9, atol=0)
    np.testing.assert_allclose(y[:, 3] - x[:, 3], k * Y, rtol=1e-9, atol=0)
    # elliptic beam: Bassetti-Erskine with the Faddeeva function
    sx, sy = 3e-4, 1e-4
    y = one_pass(st.LRBB("LR", N, 1.0, sepx, sepy, sx, sy), x)
    for p in range(len(x)):
        X, Y = x[p, 0] + sepx, x[p, 2] + sepy
        dpx, dpy = _be_kick(abs(X), abs(Y), sx, sy, 2 * N * RP / GP.gamma)
        assert y[p, 1] - x[p, 1] == pytest.approx(math.copysign(dpx, X), rel=1e-8)
        assert y[p, 3] - x[p, 3] == pytest.approx(math.copysign(dpy, Y), rel=1e-8)


def test_elens_thin_limit_and_zero_current():
    x = particles(8, scale=2e-4)
    Ne, scale, be, sig = 3e11, 0.7, 0.4, 2.5e-4
    y = one_pass(st.ELENS("E", 0., Ne, scale, 1, be, sig, sig), x)
    r2 = x[:, 0] ** 2 + x[:, 2] ** 2
    k = -2 * Ne * RP / GP.gamma / r2 * (1 - np.exp(-r2 / 2 / sig ** 2)) * (1 + be) * scale
    np.testing.assert_allclose(y[:, 1] - x[:, 1], k * x[:, 0], rtol=1e-9)
    np.testing.assert_allclose(y[:, 3] - x[:, 3], k * x[:, 2], rtol=1e-9)
    # four thin slices at one place: four quarter kicks
    y4 = one_pass(st.ELENS("E", 0., Ne, scale, 4, be, sig, sig), x)
    np.testing.assert_allclose(y4[:, [1, 3]] - x[:, [1, 3]], y[:, [1, 3]] - x[:, [1, 3]], rtol=1e-12)
    # no current: a drift of the lens length
    np.testing.assert_array_equal(one_pass(st.ELENS("E", 1.3, 0., scale, 3, be, sig, sig), x),
                                  one_pass(st.DRIFT("D", 1.3), x))


def test_helens_radial_kick_in_its_three_regions():
    L, Ne, be, rout, rin = 1e-4, 4e11, 0.5, 2e-3, 5e-4
    ang = np.linspace(0.1, 6.2, 7)
    rad = np.array([2e-4, 8e-4, 1.5e-3, 3e-3, 6e-3, 4.5e-4, 1.99e-3])
    x = np.zeros((7, 6))
    x[:, 0], x[:, 2] = rad * np.cos(ang), rad * np.sin(ang)
    y = one_pass(st.HELENS("H", L, Ne, 1.0, 1, be, rout, rin), x)
    I_e = 1.6021765e-19 * Ne * be * 299792458. / L
    th = -2e-7 * L * I_e * (1 + be * GP.beta) / be / GP.beta / rout / GP.brho
    scale = np.where(rad <= rin, 0., np.where(rad <= rout, (rad ** 2 - rin ** 2) / (rout ** 2 - rin ** 2) * rout / rad,
                                                rout / rad))
    # the kick is radial with magnitude scale * Theta_max (the drifts of L/2 move r by < 1e-12)
    np.testing.assert_allclose(y[:, 1], scale * th * np.cos(ang), rtol=1e-9, atol=1e-22)
    np.testing.assert_allclose(y[:, 3], scale * th * np.sin(ang), rtol=1e-9, atol=1e-22)
    assert np.all(y[rad <= rin][:, [1, 3]] == 0.)


def _ac_kicks(kmax, nud, phid, ts, te, nturn):
    out = []
    for t in range(nturn):
        if t < ts:
            out.append(0.)
            continue
        s = math.sin(2.0 * 3.14159265 * nud * (t - ts) + phid)
        out.append(((t - ts) / (te - ts) if ts <= t <= te else 1.) * kmax * s)
    return np.array(out)


@pytest.mark.parametrize("kind", ["H", "V"])
def test_acdip_kick_turn_by_turn(kind):
    kmax, nud, phid, ts, te, nturn = 2e-5, 0.31, 0.2, 3, 9, 16
    el = (st.HACDIP if kind == "H" else st.VACDIP)("AC", 0., kmax, nud, phid, ts, te)
    o = orc(el)
    x0 = np.zeros((1, 6))
    _, tbt, *_ = o.track(x0, nturn + 1, tbt=True)
    p = 1 if kind == "H" else 3
    kicks = np.diff(tbt[0, :, p])
    np.testing.assert_allclose(kicks, _ac_kicks(kmax, nud, phid, ts, te, nturn), rtol=1e-12, atol=1e-24)
    assert np.all(tbt[0, :, 4 - p] == 0.), "the other plane is not kicked"
    assert kicks[ts] == 0. and kicks[ts - 1] == 0. and kicks[ts + 1] != 0.


@pytest.mark.parametrize("kind,order", [("H", 2), ("V", 3), ("H", 0), ("V", 0)])
def test_acmult_kick_turn_by_turn(kind, order):
    kl, tt, phi0, nturn = 0.04, 7, 0.3, 9
    el = (st.HACMULT if kind == "H" else st.VACMULT)("AM", 0., order, kl, tt, phi0)
    o = orc(el)
    x0 = np.array([[3e-4, 0., -2e-4, 0., 0., 0.]])
    _, tbt, *_ = o.track(x0, nturn + 1, tbt=True)
    z = complex(x0[0, 0], x0[0, 2]) ** order / math.factorial(order)
    # (By, Bx) = Re/Im of (KL + i KSL)(x + iy)^n / n! for the normal (H) or skew (V) strength
    f = (complex(kl, 0.) if kind == "H" else complex(0., kl)) * z
    By, Bx = f.real, f.imag
    s = np.array([math.sin(2.0 * PI * t / tt + phi0) for t in range(nturn)])
    np.testing.assert_allclose(np.diff(tbt[0, :, 1]), -By * s, rtol=1e-12, atol=1e-25)
    np.testing.assert_allclose(np.diff(tbt[0, :, 3]), Bx * s, rtol=1e-12, atol=1e-25)


def test_ac_spin_passes_equal_a_kicker_with_the_same_kick():
    """ACDIP_sPass and ACMULT_sPass (order 0) rotate the spin as KICK_sPass
    does for the kick they give at that turn (the same BLbrho)."""
    x = particles(4, spin=True)
    kmax, nud, phid, ts, te = 3e-5, 0.23, 0.4, 0, 4
    for t in range(1, 7):
        k = _ac_kicks(kmax, nud, phid, ts, te, t + 1)[t]
        line = st.Line([st.HACDIP("AC", 0.2, kmax, nud, phid, ts, te)])
        o = pyoracle.Oracle(GP, line.records())
        _, tbt, *_ = o.track(x, t + 2, spin=True, tbt=True)
        ref = pyoracle.Oracle(GP, st.Line([st.KICKER("K", 0.2, k, 0.)]).records()).element_pass(0, tbt[:, t], spin=True)
        np.testing.assert_allclose(tbt[:, t + 1], ref, rtol=0, atol=2e-15)
    kl, tt, phi0 = 2e-5, 5, 0.1
    for t in range(1, 5):
        s = math.sin(2.0 * PI * t / tt + phi0)
        o = pyoracle.Oracle(GP, st.Line([st.HACMULT("AM", 0.3, 0, kl, tt, phi0)]).records())
        _, tbt, *_ = o.track(x, t + 2, spin=True, tbt=True)
        ref = pyoracle.Oracle(GP, st.Line([st.KICKER("K", 0.3, -kl * s, 0.)]).records()).element_pass(0, tbt[:, t],
                                                                                                          spin=True)
        np.testing.assert_allclose(tbt[:, t + 1], ref, rtol=0, atol=2e-15)


def test_spin_stays_a_unit_vector():
    x = particles(6, spin=True)
    for el in (st.HACDIP("A", 0.2, 5e-5, 0.3, 0., 0, 2), st.VACMULT("B", 0.2, 2, 0.5, 4, 0.2),
               st.ELENS("E", 1.0, 2e11, 1.0, 2, 0.5, 3e-4, 2e-4)):
        o = pyoracle.Oracle(GP, st.Line([el]).records())
        y, *_ = o.track(x, 5, spin=True)
        np.testing.assert_allclose(np.linalg.norm(y[:, 6:9], axis=1), 1., atol=1e-14)


def _transverse_jacobian(o, spin=False, h=1e-7):
    J = np.zeros((4, 4))
    for k in range(4):
        d = np.zeros((2, 9 if spin else 6))
        d[0, k], d[1, k] = h, -h
        if spin:
            d[:, 7] = 1.
        y = o.element_pass(0, d, spin=spin)
        J[:, k] = (y[0, :4] - y[1, :4]) / (2 * h)
    return J


def _solenoid_matrix(K, L):
    """Hard-edge solenoid (fringes included), Larmor wave number K, in
    (x, px, y, py) for the reference particle: the textbook matrix."""
    C, S = math.cos(K * L), math.sin(K * L)
    return np.array([[C * C, S * C / K, S * C, S * S / K], [-K * S * C, C * C, -K * S * S, S * C],
                     [-S * C, -S * S / K, C * C, S * C / K], [K * S * S, -S * C, -K * S * C, C * C]])


@pytest.mark.parametrize("nint", [1, 4])
def test_solenoid_linear_map_is_the_textbook_matrix(nint):
    L, KS = 1.2, 0.8
    e = st.SOLEN("S", L, KS)
    e.Nint = nint
    o = orc(e)
    # SOLEN_Pass: Bz = KS Brho, so K = KS / 2
    np.testing.assert_allclose(_transverse_jacobian(o), _solenoid_matrix(KS / 2, L), rtol=0, atol=1e-11)
    # SOLEN_sPass takes Bz = KS Brho / L (simtrack.cpp:8122): K = KS / (2 L)
    np.testing.assert_allclose(_transverse_jacobian(o, spin=True), _solenoid_matrix(KS / 2 / L, L), rtol=0,
                               atol=1e-11)


def test_solenoid_commutes_with_rotations_about_s():
    x = particles(8, scale=1e-3)
    sol, rot = st.SOLEN("S", 0.9, 0.3), st.SROTAT("R", 0.7)
    a = one_pass(sol, one_pass(rot, x))
    b = one_pass(rot, one_pass(sol, x))
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-15)
    x0 = np.zeros((1, 6))
    y = one_pass(sol, x0)
    assert np.all(y[0, :4] == 0.) and abs(y[0, 4]) < 1e-14 and y[0, 5] == 0.


@pytest.mark.parametrize("nint", [1, 3])
def test_solenoid_spin_precesses_by_one_plus_g_times_ks(nint):
    """Thomas-BMT in a longitudinal field: the spin of the on-axis reference
    particle turns about s by (1 + G) B L / Brho = (1 + G) KS."""
    KS = 0.05
    e = st.SOLEN("S", 1.3, KS)
    e.Nint = nint
    x = np.zeros((1, 9))
    x[0, 6] = 1.
    y = orc(e).element_pass(0, x, spin=True)
    ang = (1 + GP.G) * KS
    assert y[0, 6] == pytest.approx(math.cos(ang), abs=1e-13)
    assert abs(y[0, 7]) == pytest.approx(math.sin(ang), abs=1e-13)
    assert y[0, 8] == 0.
    np.testing.assert_allclose(np.linalg.norm(y[0, 6:9]), 1., atol=1e-15)


def test_wiggler_forest_step_is_symplectic():
    S = np.zeros((6, 6))
    for i in range(3):
        S[2 * i, 2 * i + 1], S[2 * i + 1, 2 * i] = 1., -1.
    for kx in (2.0, 0.0):
        J = _jacobian(st.WIGGLER("W", 0.4, 8, 1.5, kx, 2 * PI / 0.1, 0.2),
                      np.array([3e-4, -2e-5, 1e-4, 4e-5, 0.01, 2e-4]))
        np.testing.assert_allclose(J.T @ S @ J, S, rtol=0, atol=2e-8)


def test_wiggler_integrators_agree():
    """Forest's implicit step (Pass) and Wu's splitting (sPass) integrate the
    same field: with fine steps they converge to one map, so a wrong field
    coefficient in either would show; with B0 = 0 and pt = 0 both are the
    same expanded drift."""
    x = particles(6, scale=5e-4)
    x[:, 5] = 0.
    for kx in (3.0, 0.0):
        el = st.WIGGLER("W", 0.3, 600, 1.2, kx, 2 * PI / 0.1, 0.3)
        a = orc(el).element_pass(0, x)
        xs = np.zeros((len(x), 9))
        xs[:, :6] = x
        xs[:, 7] = 1.
        b = orc(el).element_pass(0, xs, spin=True)
        np.testing.assert_allclose(b[:, :4], a[:, :4], rtol=0, atol=2e-10)
        assert np.array_equal(b[:, 6:], xs[:, 6:]), "sPass leaves the spin unrotated"
    el0 = st.WIGGLER("W0", 0.5, 4, 0., 3.0, 2 * PI / 0.1, 0.)
    a = orc(el0).element_pass(0, x)
    np.testing.assert_allclose(a[:, 0], x[:, 0] + 0.5 * x[:, 1], rtol=1e-14, atol=1e-20)
    np.testing.assert_allclose(a[:, 2], x[:, 2] + 0.5 * x[:, 3], rtol=1e-14, atol=1e-20)
    np.testing.assert_allclose(a[:, 4], x[:, 4] - 0.5 * (x[:, 1] ** 2 + x[:, 3] ** 2) / 2, rtol=1e-12, atol=1e-22)


def test_planar_wiggler_focuses_vertically():
    """A planar wiggler (KX = 0) focuses vertically with k = (B0/Brho)^2/2
    on average over its periods (the natural wiggler focusing)."""
    g = W.electron_globals(18000.0)
    B0, lam, L = 1.2, 0.1, 2.0
    el = st.WIGGLER("W", L, 400, B0, 0.0, 2 * PI / lam, 0.0)
    o = pyoracle.Oracle(g, st.Line([el]).records())
    h = 1e-6
    x = np.zeros((2, 6))
    x[0, 2], x[1, 2] = h, -h
    y = o.element_pass(0, x)
    dpy_dy = (y[0, 3] - y[1, 3]) / (2 * h)
    k = (B0 / g.brho) ** 2 / 2
    assert dpy_dy == pytest.approx(-k * L, rel=0.05)
    assert abs((y[0, 1] - y[1, 1]) / (2 * h)) < 1e-3 * k * L, "no x-y coupling in a planar wiggler"
