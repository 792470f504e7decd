"""GPU dynamic-aperture drivers against the reference's sequential loops
restated on the CPU oracle: the same particles, one Track call at a time,
must give the same bisection results and the same uniform-scan records."""
import math

import numpy as np
import pytest

import pyoracle
import simtrack_amd as st
from simtrack_amd import da
from simtrack_amd import workloads as W

from test_gpu_parity import _fodo_ring

pytestmark = pytest.mark.gpu

PI = 3.14159265358979323
NTURN = 60


def particle(sigma, angle, sx, sy, dp):
    return np.array([[sigma * math.cos(angle) * sx, 0., sigma * math.sin(angle) * sy, 0., 0., dp]])


def oracle_stable(orc, x, radial):
    _, s, _, _ = orc.track(x.copy(), NTURN, radial=radial)
    return int(s[0])


def oracle_bisect(orc, angle, sigma_max, sx, sy, dp, radial, single=False):
    """do-while loop of Track_DA, simtrack.cpp:21403-21435 (twins), or of
    Track_DA2, cpp:21511-21526 (one particle)"""
    s, e = 0., sigma_max
    while True:
        sigma = (s + e) / 2.0
        st1 = oracle_stable(orc, particle(sigma, angle, sx, sy, dp), radial)
        st2 = 1 if single else oracle_stable(orc, particle(sigma + 0.05, angle, sx, sy, dp), radial)
        stable = 1 if st1 * st2 == 1 else 0
        if stable == 0:
            e = sigma
        if stable == 1:
            s = sigma
        if not abs(e - s) > 0.1:
            return sigma


@pytest.fixture(scope="module")
def ring(ctx):
    ln = _fodo_ring(ncell=12)
    g = W.proton_globals(100000.0)
    rec = ln.records()
    return st.Table(ctx, g, rec), pyoracle.Oracle(g, rec)


@pytest.mark.parametrize("variant", [st.ST_DA_TRACK_DA, st.ST_DA_TRACK_DA1, st.ST_DA_TRACK_DA_RADIAL])
def test_bisection_matches_sequential_loop(ring, variant):
    tab, orc = ring
    sx, sy, dp = 6e-4, 4e-4, 2e-4
    angles, smax = da.da_angles(tab.lib, variant)
    if variant == st.ST_DA_TRACK_DA1:
        assert np.array_equal(angles, [15 * (i + 1) * PI / 180 for i in range(5)]) and smax == 50.0
    else:
        assert np.array_equal(angles, [(90 / 11.) * (i + 1) * PI / 180 for i in range(10)])
        assert smax == (15.0 if variant == st.ST_DA_TRACK_DA_RADIAL else 50.0)
    radial = variant == st.ST_DA_TRACK_DA_RADIAL
    got = da.da_bisect(tab, NTURN, dp, sx, sy, angles, smax, radial=radial)
    ref = [oracle_bisect(orc, a, smax, sx, sy, dp, radial) for a in angles]
    assert list(got) == ref
    assert 0 < min(got) and max(got) < smax, "the scan should find a finite aperture"


@pytest.mark.parametrize("variant", [st.ST_DA_TRACK_DA2, st.ST_DA_TRACK_DA3])
def test_single_particle_bisection_matches_sequential_loop(ring, variant):
    """Track_DA2 (cpp:21497) and Track_DA3 (cpp:21535): one particle per
    midpoint, angles (90/11)(i+1) deg on [0, 50] and 15(i+1) deg on [0, 30]"""
    tab, orc = ring
    sx, sy, dp = 6e-4, 4e-4, 2e-4
    angles, smax = da.da_angles(tab.lib, variant)
    if variant == st.ST_DA_TRACK_DA3:
        assert np.array_equal(angles, [15 * (i + 1) * PI / 180 for i in range(5)]) and smax == 30.0
    else:
        assert np.array_equal(angles, [(90 / 11.) * (i + 1) * PI / 180 for i in range(10)]) and smax == 50.0
    got = da.da_bisect(tab, NTURN, dp, sx, sy, angles, smax, single=True)
    ref = [oracle_bisect(orc, a, smax, sx, sy, dp, False, single=True) for a in angles]
    assert list(got) == ref
    assert 0 < min(got) and max(got) < smax, "the scan should find a finite aperture"
    twins = da.da_bisect(tab, NTURN, dp, sx, sy, angles, smax)
    assert np.all(twins <= got + 0.1), "the twin criterion is the stricter one"


@pytest.mark.parametrize("radial,sigma0,n", [(False, 1.0, 150), (True, 3.0, 100)])
def test_uniform_scan_matches_sequential_loop(ring, radial, sigma0, n):
    tab, orc = ring
    sx, sy, dp, angle = 1.2e-3, 8e-4, 0.0, 0.6
    sig, stable = da.da_uniform(tab, NTURN, dp, angle, sx, sy, sigma0, 0.2, n, radial=radial)
    ref_sig, ref_st = [], []
    for i in range(n):
        sigma = sigma0 + i * 0.2
        s = oracle_stable(orc, particle(sigma, angle, sx, sy, dp), radial)
        ref_sig.append(sigma)
        ref_st.append(s)
        if s == 0:
            break
    assert list(sig) == ref_sig and list(stable) == ref_st
    assert ref_st[-1] == 0, "the scan should reach the aperture"


def test_driver_files(ring, tmp_path):
    tab, _ = ring
    f = tmp_path / "da.txt"
    rows = da.Track_DA1(tab, NTURN, 2e-4, 6e-4, 4e-4, str(f))
    lines = f.read_text().splitlines()
    assert len(lines) == 5 and lines[0].startswith("15.000000 ")
    assert float(lines[2].split()[1]) == pytest.approx(rows[2][1], abs=1e-6)
