"""Many lattices in one persistent launch (st_batch_*): results equal
tracking every lattice separately, bit for bit, for every launch shape the
autotuner considers; the autotuner returns one of its candidates."""
import numpy as np
import pytest

import pyoracle
import simtrack_amd as st
from simtrack_amd import workloads as W

from test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu
NTURN = 60


@pytest.fixture(scope="module")
def seeds(ctx):
    gp, lines, x, offs = W.config2(n_seeds=5, n_per_seed=256, n_angles=8, min_elements=1000, max_elements=1800)
    tabs = [st.Table(ctx, gp, ln.records()) for ln in lines]
    return gp, lines, tabs, x, offs


def separate(tabs, x, offs, nturn):
    out = []
    for k, t in enumerate(tabs):
        out.append(t.track(x[offs[k]:offs[k + 1]].copy(), nturn))
    return (np.concatenate([o[0] for o in out]), np.concatenate([o[1] for o in out]),
            np.concatenate([o[2] for o in out]), np.concatenate([o[3] for o in out]))


@pytest.mark.parametrize("wg,wps,tb", [(1, 2, 256), (2, 1, 16), (4, 2, 7), (1, 1, 60)])
def test_batch_equals_separate_tracking(ctx, seeds, wg, wps, tb):
    gp, lines, tabs, x, offs = seeds
    ref = separate(tabs, x, offs, NTURN)
    b = st.Batch(ctx, tabs)
    cfg = st.BatchConfig(wg, wps, tb)
    got = b.track(x.copy(), offs, NTURN, cfg=cfg)
    for g, r, name in zip(got[1:], ref[1:], ("stable", "lost_turn", "lost_post")):
        assert np.array_equal(g, r), name
    assert_bitwise(got[0], ref[0], "batch vs separate")
    surv = ref[1].reshape(5, -1).sum(axis=1)
    assert (surv > 0).all() and (surv < 256).all(), "each seed's grid should straddle its aperture"


def test_batch_against_oracle(ctx, seeds):
    gp, lines, tabs, x, offs = seeds
    b = st.Batch(ctx, tabs)
    got = b.track(x.copy(), offs, 20)
    k = 3
    orc = pyoracle.Oracle(gp, lines[k].records())
    xo, so, lto, lpo = orc.track(x[offs[k]:offs[k + 1]].copy(), 20)
    sl = slice(offs[k], offs[k + 1])
    assert np.array_equal(got[1][sl], so) and np.array_equal(got[2][sl], lto) and np.array_equal(got[3][sl], lpo)
    assert_bitwise(got[0][sl], xo, "batch seed 3 vs oracle")


def test_autotuner(ctx, seeds):
    gp, lines, tabs, x, offs = seeds
    b = st.Batch(ctx, tabs)
    best, table = b.autotune(x, offs, 32)
    assert len(table) == 18 and all(t > 0 for _, t in table)
    assert (best.wg_waves, best.waves_per_simd, best.turn_block) in [c for c, _ in table]
    assert min(t for _, t in table) == dict(table)[(best.wg_waves, best.waves_per_simd, best.turn_block)]
