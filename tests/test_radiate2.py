"""Pins the oracle's radiate2 (simtrack.cpp:4290-4366), the photon emission
of a radiating GSBEND, with checks from the physics of synchrotron radiation
rather than its own formulas:

- the photon-energy table and its low-energy power law against the exact
  photon-number spectrum n(y) = int_y^inf K_5/3(t) dt (y = E / Ec), whose
  CDF is F(y) = (int_0^y t K_5/3 dt + y int_y^inf K_5/3 dt) / (5 pi / 3);
- the photon count against the Poisson law (mean, variance, probabilities),
  including the means above 30 that are drawn in chunks;
- the summed photon energy against the spectrum's moments <y> = 8/(15 sqrt 3)
  and <y^2> = 11/27 (a compound Poisson sum);
- the energy loss and its spread per pass against radiate1 (the Gaussian
  model of the same emission, cpp:4236) in the same bend;
- a bend with no field emits nothing.
"""
import math

import numpy as np
import pytest
from scipy import integrate, special, stats

import pyoracle
import simtrack_amd as st
from simtrack_amd import workloads as W


def _k53(t):
    return special.kv(5. / 3., t)


def _tail(y):
    pts = [y] + [p for p in (1e-3, 1e-2, 0.1, 1., 5., 20., 60.) if p > y]
    return sum(integrate.quad(_k53, a, b, limit=200, epsabs=0, epsrel=1e-12)[0] for a, b in zip(pts[:-1], pts[1:]))


def _cdf(y):
    """Fraction of synchrotron photons with energy below y Ec."""
    head = integrate.quad(lambda s: s ** 3 * _k53(s ** 3) * 3 * s * s, 0, y ** (1. / 3.), limit=200, epsabs=0,
                          epsrel=1e-12)[0]
    return (head + y * _tail(y)) / (5 * math.pi / 3)


@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
def test_photon_energy_table_is_the_synchrotron_spectrum():
    """The inverse CDF that radiate2 samples: at its nodes the table (used
    for r >= 0.26) agrees with the exact spectrum to 1e-3 in probability,
    between them (linear interpolation) to 2.5e-3, and the power law
    (r / facg)^3 below it is the spectrum's y^(1/3) limit."""
    assert abs(_cdf(60.) - 1.) < 1e-12
    nodes = [.2619, .2893, .327, .36914, .4074, .4417, .47273, .5209, .5791, .62385, .6517, .6905, .7165, .7494,
             .7726, .8104, .8401, .85954, .8835, .892, .89665, .8997, .90397, .9067, .9132, .93444, .9561, .9661,
             .9768, .98734, .9930, .99776, .999574, .9999156]
    for r in nodes:
        y = pyoracle.rad2_energy(r)
        assert abs(_cdf(y) - r) < 1e-3, (r, y, _cdf(y))
    for r in np.linspace(0.26, 0.9999, 40):
        y = pyoracle.rad2_energy(r)
        assert abs(_cdf(y) - r) < 2.5e-3, (r, y, _cdf(y))
    for r in (1e-3, 1e-2, 0.05, 0.1, 0.2, 0.2599):
        y = pyoracle.rad2_energy(r)
        assert abs(_cdf(y) / r - 1) < 0.025, (r, y, _cdf(y))
    y = pyoracle.rad2_energy(1e-3)
    assert abs(_cdf(y) / 1e-3 - 1) < 1e-5
    # monotone over the whole range, and the top of the table is 10 Ec
    r = np.linspace(1e-4, 1 - 1e-9, 2001)
    y = np.array([pyoracle.rad2_energy(v) for v in r])
    assert np.all(np.diff(y) > 0) and abs(pyoracle.rad2_energy(1.0) - 10.) < 1e-12


@pytest.mark.parametrize("lam", [0.5, 7.0, 45.0, 100.0])
def test_photon_count_is_poisson(lam):
    n = 100000
    cnt, _ = pyoracle.rad2_sample(11, n, 5, lam)
    assert abs(cnt.mean() - lam) < 5 * math.sqrt(lam / n)
    assert abs(cnt.var() / lam - 1) < 5 * math.sqrt((2 + 1 / lam) / n)
    k = np.arange(max(0, int(lam - 4 * math.sqrt(lam))), int(lam + 4 * math.sqrt(lam)) + 2)
    p = stats.poisson.pmf(k, lam)
    f = np.array([(cnt == v).mean() for v in k])
    assert np.all(np.abs(f - p) < 5 * np.sqrt(p * (1 - p) / n) + 1e-4)


def test_photon_streams_are_independent():
    """Particles and draw numbers give uncorrelated counts; the same key gives
    the same photons."""
    a, ea = pyoracle.rad2_sample(3, 50000, 8, 4.0)
    b, _ = pyoracle.rad2_sample(3, 50000, 9, 4.0)
    c, ec = pyoracle.rad2_sample(3, 50000, 8, 4.0)
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.02 and abs(np.corrcoef(a[1:], a[:-1])[0, 1]) < 0.02
    assert np.array_equal(a, c) and np.array_equal(ea, ec)


@pytest.mark.parametrize("lam", [3.0, 60.0])
def test_summed_photon_energy_moments(lam):
    """Compound Poisson sum: mean lam <y>, variance lam <y^2>, with the
    spectrum's <y> = 8/(15 sqrt 3), <y^2> = 11/27; the table's linear
    interpolation raises them by 0.7 % and 3 %."""
    n = 200000
    _, e = pyoracle.rad2_sample(21, n, 1, lam)
    assert abs(e.mean() / (lam * 8 / (15 * math.sqrt(3))) - 1) < 0.012
    assert abs(e.var() / (lam * 11 / 27) - 1) < 0.05


def _bend_pass(gp, el, n):
    rec = st.Line([el]).records()
    return pyoracle.Oracle(gp, rec).element_pass(0, np.zeros((n, 6)))


@pytest.mark.parametrize("hexp", [1, 0], ids=["yoshida", "exact"])
def test_radiate2_loss_and_spread_match_radiate1(hexp):
    """The same 18 GeV electrons through a 300 m-radius bend (6 m, 4 slices,
    about 7.7 photons each): radiate2's mean loss and its variance agree
    with radiate1's classical loss and Gaussian excitation (Sands' rates) to
    the table's interpolation error."""
    gp = W.electron_globals(18000.0, radiate=1, H_expand=hexp, rng_seed=77)
    rho, ang, n = 300., 0.02, 100000
    g = st.GSBEND("G", rho * ang, ang, ang)
    b = st.SBEND("B", rho * ang, ang)
    g.Nint = b.Nint = 4
    d2 = _bend_pass(gp, g, n)[:, 5]
    d1 = _bend_pass(gp, b, n)[:, 5]
    assert abs(d2.mean() / d1.mean() - 1) < 0.015
    assert abs(d2.var() / d1.var() - 1) < 0.05
    assert (d2 == 0).mean() < 2e-3 and np.all(d2 <= 0)


def test_bend_without_field_emits_nothing():
    """K0L = 0 in the expanded-Hamiltonian branch: no transverse field, no
    photons, and the pass equals the non-radiating one bit for bit.  (The
    exact branch divides by K0L / L, cpp:3980-3999, so it has no such case.)"""
    el = st.GSBEND("G", 2.0, 0.01, 0.)
    el.Nint = 3
    r = np.random.default_rng(4)
    x = np.zeros((64, 6))
    x[:, :4] = r.normal(0, 1e-4, (64, 4))
    x[:, 5] = r.normal(0, 1e-4, 64)
    rec = st.Line([el]).records()
    on = pyoracle.Oracle(W.electron_globals(18000.0, radiate=1, H_expand=1), rec).element_pass(0, x.copy())
    off = pyoracle.Oracle(W.electron_globals(18000.0, radiate=0, H_expand=1), rec).element_pass(0, x.copy())
    assert np.isfinite(on).all() and np.array_equal(on, off)
