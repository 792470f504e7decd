"""Host logic of the surrogate scan: which particles the boundary-aware
fallback tracks, and the DA boundary of a line."""
import numpy as np

from simtrack_amd import surrogate as S


def test_boundary_fallback_walks_each_line_to_the_third_predicted_loss():
    # one lattice, two angles of 8 amplitudes each
    alive = np.array([1, 1, 1, 1, 1, 0, 1, 1,
                      1, 1, 1, 1, 1, 1, 1, 1])
    prob = np.array([.95, .7, .3, .95, .2, .5, .1, .9,
                     .99, .4, .45, .95, .1, .6, .3, .05])
    sel = S.boundary_fallback(alive, prob, [0, 16], 2)
    # line 1: stops at the early loss (index 5); safe particles (p > 0.9) skipped
    assert list(np.nonzero(sel[:8])[0]) == [1, 2, 4]
    # line 2: up to and including the third p < 0.5 (indices 9, 10, 12)
    assert list(np.nonzero(sel[8:])[0]) == [1, 2, 4]


def test_boundary_fallback_with_several_lattices():
    alive = np.ones(24, dtype=int)
    prob = np.full(24, 0.95)
    prob[[3, 13, 20]] = 0.2
    sel = S.boundary_fallback(alive, prob, [0, 8, 24], 1)
    assert list(np.nonzero(sel)[0]) == [3, 13, 20]


def test_da_boundary():
    amps = np.arange(1, 5) * 1.0
    s = np.array([1, 1, 0, 1, 1, 1, 1, 1])
    assert list(S.da_boundary(s, amps, 2)) == [3.0, 5.0]
