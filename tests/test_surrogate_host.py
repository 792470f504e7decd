"""Host logic of the surrogate scan: which particles the boundary-aware
fallback tracks, and the DA boundary of a line."""
import numpy as np

from simtrack_amd import surrogate as S


def test_boundary_fallback_tracks_what_can_move_the_boundary():
    # one lattice, two angles of 8 amplitudes each
    alive = np.array([1, 1, 1, 1, 1, 0, 1, 1,
                      1, 1, 1, 1, 1, 1, 1, 1])
    prob = np.array([.95, .7, .3, .9995, .2, .5, .1, .9,
                     .99, .4, .45, .95, .1, .6, .3, .9999])
    sel = S.boundary_fallback(alive, prob, [0, 16], 2)
    # line 1: below the early loss (index 5) every particle not called safe
    # (p > 0.999); nothing beyond it, whatever its probability
    assert list(np.nonzero(sel[:8])[0]) == [0, 1, 2, 4]
    # line 2: no early loss, so the whole line except the safe last one
    assert list(np.nonzero(sel[8:])[0]) == [0, 1, 2, 3, 4, 5, 6]
    # a lower trust threshold tracks more
    sel9 = S.boundary_fallback(alive, prob, [0, 16], 2, p_safe=0.9)
    assert list(np.nonzero(sel9[:8])[0]) == [1, 2, 4]
    assert list(np.nonzero(sel9[8:])[0]) == [1, 2, 4, 5, 6]


def test_boundary_fallback_with_several_lattices():
    alive = np.ones(24, dtype=int)
    alive[[5, 14, 22]] = 0
    prob = np.full(24, 0.9999)
    prob[[3, 6, 13, 15, 20, 23]] = 0.2
    sel = S.boundary_fallback(alive, prob, [0, 8, 24], 1)
    # lattice 0: early loss at 5 hides 6; lattice 1 (16 rows): 13 < 14, 15 > 14
    assert list(np.nonzero(sel)[0]) == [3, 13]
    sel2 = S.boundary_fallback(alive, prob, [0, 8, 24], 2)
    # two angles of 8 in lattice 1: rows 8-15 (loss at 14) and 16-23 (loss at 22)
    assert list(np.nonzero(sel2)[0]) == [3, 13, 20]


def test_boundary_fallback_catches_what_the_first_loss_needs():
    """Against a brute-force definition on random lines: the boundary of
    every line is exact when every particle outside the mask is either safe
    (and really survives) or beyond the first early loss."""
    rng = np.random.default_rng(5)
    n_ang, m, nlat = 4, 32, 6
    alive = rng.random(nlat * n_ang * m) > 0.1
    prob = rng.random(nlat * n_ang * m)
    offs = np.arange(nlat + 1) * n_ang * m
    truth = alive & (prob > 0.3) & (rng.random(len(alive)) > 0.05)
    truth[prob > 0.999] = alive[prob > 0.999]     # the network is right where it is sure
    sel = S.boundary_fallback(alive, prob, offs, n_ang)
    surv = np.where(sel, truth, alive & (prob >= 0.5))
    amps = np.arange(1, m + 1) * 1.0
    for k in range(nlat):
        sl = slice(offs[k], offs[k + 1])
        assert np.array_equal(S.da_boundary(surv[sl], amps, n_ang), S.da_boundary(truth[sl], amps, n_ang))


def test_da_boundary():
    amps = np.arange(1, 5) * 1.0
    s = np.array([1, 1, 0, 1, 1, 1, 1, 1])
    assert list(S.da_boundary(s, amps, 2)) == [3.0, 5.0]


def test_state_round_trip(tmp_path):
    """save_state keeps x only for particles still tracked; everything else
    comes back as it was."""
    import torch
    n = 50
    g = torch.Generator().manual_seed(3)
    st = {"d_x": torch.randn(n, 6, generator=g, dtype=torch.float64),
          "d_s": (torch.rand(n, generator=g) > 0.3).to(torch.int32),
          "d_l": torch.randint(0, 99, (2 * n,), generator=g, dtype=torch.int32),
          "fallback": torch.rand(n, generator=g) > 0.5,
          "p": torch.rand(n, generator=g, dtype=torch.float32),
          "turn": 1024, "elapsed": 12.5, "n_fallback": 7}
    path = tmp_path / "state.npz"
    S.save_state(st, path)
    back = S.load_state(path, device="cpu")
    live = (st["d_s"] != 0) & st["fallback"]
    assert torch.equal(back["d_x"][live], st["d_x"][live])
    assert torch.all(back["d_x"][~live] == 0)
    for k in ("d_s", "d_l", "fallback", "p"):
        assert torch.equal(back[k], st[k])
    assert back["turn"] == 1024 and back["elapsed"] == 12.5 and back["n_fallback"] == 7
