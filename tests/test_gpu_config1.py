"""GPU parity on the benchmark workload itself: the 998-element DA ring with
6-D synchro-beam beam-beam (BEAMBEAM treatment 6, 5 slices) at both IPs and
spin tracking (Track_spin, simtrack.cpp:20816-20839; BB6D cpp:6579-6678;
BEAMBEAM::sPass is orbit-only, cpp:11213-11217).

The GPU runs the composed line through the C ABI and is compared bit for bit
with the CPU oracle: coordinates, spins and the stable/lost_turn/lost_post
records.  The oracle is run live on a subset sized to finish in seconds on
the host, and the turn-by-turn check compares every turn's coordinates.
"""
import numpy as np
import pytest

import pyoracle
import simtrack_amd as st
from simtrack_amd import workloads as W
from test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu

NTURN = 1000


@pytest.fixture(scope="module")
def c1():
    gp, ln, x = W.config1(n_particles=256)
    return gp, ln.records(), x


def test_config1_track_spin_1000_turns(ctx, c1):
    """64 of config 1's Gaussian particles (every 4th of the first 256), spin
    on, 1,000 turns: GPU = oracle bit for bit, loss records identical (the
    far tails of the Gaussian reach the dynamic aperture)."""
    gp, rec, x = c1
    xs = np.ascontiguousarray(x[::4])
    tab = st.Table(ctx, gp, rec)
    stable = np.ones(len(xs), np.int32)
    stable[5] = 0                       # stable == 0 on input: returned at once, records zeroed (cpp:20795)
    lt = np.full(len(xs), 3, np.int32)
    lp = np.full(len(xs), 4, np.int32)
    xg, sg, ltg, lpg = tab.track(xs.copy(), NTURN, spin=True, stable=stable.copy(), lost_turn=lt.copy(),
                                 lost_post=lp.copy())
    tab.close()
    xo, so, lto, lpo = pyoracle.Oracle(gp, rec).track(xs.copy(), NTURN, spin=True, stable=stable.copy(),
                                                       lost_turn=lt.copy(), lost_post=lp.copy())
    assert np.array_equal(sg, so) and np.array_equal(ltg, lto) and np.array_equal(lpg, lpo)
    assert sg[5] == 0 and lto[5] == 0 and lpo[5] == 0 and np.array_equal(xg[5], xs[5])
    assert sg.sum() >= len(xs) - 5      # a few 6-sigma tail particles may be lost
    assert_bitwise(xg, xo, "config 1, 1000 turns with spin")
    # the survivors' spins stayed unit vectors
    assert np.allclose(np.linalg.norm(xg[sg == 1, 6:9], axis=1), 1.0, atol=1e-12)


def test_config1_turn_by_turn(ctx, c1):
    """Track_tbt with spin on 16 particles for 200 turns: every turn's
    coordinates and spins equal the oracle's bit for bit."""
    gp, rec, x = c1
    xs = np.ascontiguousarray(x[1:256:16])
    tab = st.Table(ctx, gp, rec)
    _, tg, sg, _, _ = tab.track_tbt(xs.copy(), 200, spin=True)
    tab.close()
    _, to, so, _, _ = pyoracle.Oracle(gp, rec).track(xs.copy(), 200, spin=True, tbt=True)
    assert np.array_equal(sg, so) and sg.all()
    assert_bitwise(tg, to, "config 1 turn-by-turn")


def test_config1_lossy_edge(ctx, c1):
    """The same ring with particles placed at 3-5 sigma in x (0.8 of that in
    y), across the dynamic aperture, so that some are lost after tens to
    hundreds of turns: identical loss turns and elements."""
    gp, rec, x = c1
    xs = np.ascontiguousarray(x[:32]).copy()
    emit, sig = W.beam_sizes(gp)
    a = np.linspace(3., 5., 32) * sig
    xs[:, 0] = a
    xs[:, 2] = a * 0.8
    tab = st.Table(ctx, gp, rec)
    xg, sg, ltg, lpg = tab.track(xs.copy(), 300, spin=True)
    tab.close()
    xo, so, lto, lpo = pyoracle.Oracle(gp, rec).track(xs.copy(), 300, spin=True)
    assert np.array_equal(sg, so) and np.array_equal(ltg, lto) and np.array_equal(lpg, lpo)
    assert 0 < sg.sum() < len(sg) and (lto[so == 0] > 10).any()
    assert_bitwise(xg, xo, "config 1 edge")
