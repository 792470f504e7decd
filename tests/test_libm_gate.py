"""The parity gate against a glibc-linked restatement, measured on the CPU.

The reference links the C library's libm; the oracle and the GPU kernels
share the project libm (st_libm.h), which is what lets them agree bit for
bit.  These tests run the oracle twice on the benchmark workloads, once
with the project libm (liboracle.so, bit-equal to the GPU) and once with the
C library's (liboracle_glibc.so, what the reference calls), for 1,000 turns:

  * survival, loss turn and loss element are identical, so the DA boundary
    on config 0's 5 x 205 grid is identical at every angle;
  * the per-turn share of surviving particles whose coordinates agree within
    1e-11 relative is measured and reported (scale: max(|x|, the column's
    rms at that turn); spins: 1).  Where it falls below 1 the cause is a
    C-library result that is not correctly rounded (exp, asin, sin, cos)
    amplified by the ring: the same oracle with one initial coordinate moved
    by one ulp diverges as far (test_one_ulp_sensitivity).

The project-libm runs must also reproduce the committed fixtures that the
GPU tests compare against (tests/golden/make_oracle_fixtures.py).
"""
import numpy as np
import pytest

import pyoracle
from simtrack_amd import workloads as W

NTURN = 1000
_cache = {}


def _golden(name):
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_oracle_fixtures as mk
    return mk.load(name)


def _workload(name):
    if name == "config0":
        gp, ln, x = W.config0()
        return gp, ln.records(), x, False
    gp, ln, x = W.config1(n_particles=256)
    return gp, ln.records(), np.ascontiguousarray(x[::16]), True


def _run(name, variant):
    key = (name, variant)
    if key not in _cache:
        gp, rec, x, spin = _workload(name)
        _cache[key] = pyoracle.Oracle(gp, rec, glibc=variant).track(x.copy(), NTURN, spin=spin, tbt=True)
    return _cache[key]


def agreement(a, b):
    """Per-turn share of particles alive in run b whose coordinates in run a
    agree within 1e-11 relative, and the largest relative difference."""
    _, ta, _, _, _ = a
    _, tb, sb, ltb, _ = b
    alive = (sb[:, None] == 1) | (ltb[:, None] >= np.arange(ta.shape[1])[None, :])
    masked = np.where(alive[..., None], tb, np.nan)
    rms = np.sqrt(np.nanmean(masked ** 2, axis=0))
    scale = np.maximum(np.maximum(np.abs(tb), rms[None]), 1e-300)
    if tb.shape[2] == 9:
        scale[..., 6:9] = 1.0
    rel = np.where(alive[..., None], np.abs(ta - tb) / scale, 0.0).max(axis=2)
    frac = ((rel <= 1e-11) & alive).sum(axis=0) / np.maximum(alive.sum(axis=0), 1)
    return frac, rel


def da_boundary(stable, n_amp=205):
    """Per angle: the number of grid amplitudes below the first lost one
    (Track_DA_Uniform's boundary, cpp:21574)."""
    s = stable.reshape(-1, n_amp)
    return np.array([int(np.argmin(row)) if not row.all() else n_amp for row in s])


@pytest.mark.slow
@pytest.mark.parametrize("name,fixture", [("config0", "oracle_config0_grid_1000.npz"),
                                          ("config1", "oracle_config1_256_1000.npz")])
def test_project_libm_reproduces_fixture(name, fixture):
    g = _golden(fixture)
    xf, _, s, lt, lp = _run(name, False)
    sel = slice(None) if name == "config0" else slice(None, None, 16)
    assert np.array_equal(s, g["stable"][sel]) and np.array_equal(lt, g["lost_turn"][sel])
    assert np.array_equal(lp, g["lost_post"][sel])
    a, b = xf, g["x"][sel]
    same = (a.view(np.int64) == b.view(np.int64)) | (np.isnan(a) & np.isnan(b))
    assert same.all()


@pytest.mark.slow
@pytest.mark.parametrize("name", ["config0", "config1"])
def test_glibc_records_and_da_identical(name):
    a, b = _run(name, True), _run(name, False)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])
    frac, rel = agreement(a, b)
    print(f"\n{name}: share within 1e-11 at turns 1, 10, 100, 1000: "
          f"{frac[0]:.3f} {frac[9]:.3f} {frac[99]:.3f} {frac[-1]:.3f}; largest relative difference {rel.max():.2e}")
    assert frac[0] == 1.0
    if name == "config0":
        assert np.array_equal(da_boundary(a[2]), da_boundary(b[2]))
        # only sin/cos differ from the C library on this ring (no beam-beam)
        assert frac[99] == 1.0 and frac[-1] >= 0.95
    else:
        assert frac[9] == 1.0 and frac[-1] >= 0.25


@pytest.mark.slow
def test_exp_drives_the_beam_beam_gap():
    """Config 1: the glibc build that differs only in exp diverges as far as
    the all-glibc build; the one differing only in sin/cos/tan stays within
    1e-11 for all 1,000 turns."""
    base = _run("config1", False)
    f_exp, _ = agreement(_run("config1", "exp"), base)
    f_trig, r_trig = agreement(_run("config1", "trig"), base)
    f_all, _ = agreement(_run("config1", True), base)
    assert f_trig[-1] == 1.0 and r_trig.max() < 1e-11
    assert f_exp[-1] < 1.0 and abs(f_exp[-1] - f_all[-1]) <= 0.1


@pytest.mark.slow
def test_one_ulp_sensitivity():
    """The same oracle, one initial x moved by one ulp: on config 1 the share
    within 1e-11 after 1,000 turns drops as far as the C-library build's, so
    the 1e-11 gate over 1,000 turns does not survive any one-ulp difference
    on this beam-beam workload."""
    gp, rec, x, spin = _workload("config1")
    xp = x.copy()
    xp[:, 0] = np.nextafter(xp[:, 0], np.inf)
    pert = pyoracle.Oracle(gp, rec).track(xp, NTURN, spin=spin, tbt=True)
    f_pert, r_pert = agreement(pert, _run("config1", False))
    print(f"\none-ulp perturbation: share within 1e-11 after {NTURN} turns {f_pert[-1]:.3f}, "
          f"largest relative difference {r_pert.max():.2e}")
    assert f_pert[-1] < 0.9 and r_pert.max() > 1e-8
