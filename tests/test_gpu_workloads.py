"""GPU parity at the benchmark workloads' full grids, 1,000 turns, against
the oracle's results committed under tests/golden/ (written by
make_oracle_fixtures.py; tests/test_libm_gate.py checks on the CPU that the
oracle still reproduces them).

  config 0: the 5 x 205 DA grid on the 998-element ring (Track,
            simtrack.cpp:20791), coordinates and loss records bit for bit,
            hence the same DA boundary at every angle (Track_DA_Uniform's
            first loss, cpp:21574);
  config 1: 256 Gaussian particles with 6-D beam-beam and spin (Track_spin,
            cpp:20816), coordinates, spins and loss records bit for bit.
"""
import os
import sys

import numpy as np
import pytest

import simtrack_amd as st
from simtrack_amd import workloads as W
from test_gpu_parity import assert_bitwise

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_oracle_fixtures as mk  # noqa: E402

pytestmark = pytest.mark.gpu


def _check(ctx, make, fixture, spin):
    gp, rec, x, _ = make()
    g = mk.load(fixture)
    assert np.array_equal(g["x0"].view(np.int64), x.view(np.int64)), "workload generator changed"
    tab = st.Table(ctx, gp, rec)
    xg, sg, ltg, lpg = tab.track(x.copy(), mk.NTURN, spin=spin)
    tab.close()
    assert np.array_equal(sg, g["stable"]) and np.array_equal(ltg, g["lost_turn"])
    assert np.array_equal(lpg, g["lost_post"])
    assert_bitwise(xg, g["x"], fixture)
    return sg


def test_config0_grid_1000_turns(ctx):
    sg = _check(ctx, mk.config0_grid, "oracle_config0_grid_1000.npz", False)
    assert 0 < sg.sum() < len(sg)
    # DA boundary per angle: amplitudes below the first lost one
    s = sg.reshape(5, 205)
    da = [int(np.argmin(r)) if not r.all() else 205 for r in s]
    print("\nconfig 0 DA boundary (grid steps of 20/205 sigma) per angle 15..75 deg:", da)
    assert all(0 < d < 205 for d in da)


def test_config1_256_particles_1000_turns(ctx):
    sg = _check(ctx, mk.config1_sample, "oracle_config1_256_1000.npz", True)
    assert sg.sum() >= 250
