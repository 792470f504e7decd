"""Spin drivers on the GPU against the oracle, bit for bit: Cal_Spin_Orbit
(st_spin_orbit, simtrack.cpp:20510-20572), Cal_Spin_Tune (st_spin_tune,
cpp:20574-20601) and Set_Initial_Spin (st_set_initial_spin,
cpp:20603-20635), on the flat FODO ring with a partial snake (whose closed
forms tests/test_spin_drivers.py pins on the oracle) and on the config-1
ring (998 elements, beam-beam at two IPs, RF on)."""
import numpy as np
import pytest

import pyoracle
from simtrack_amd import diagnostics as D
from simtrack_amd import workloads as W
from test_spin_drivers import flat_ring

pytestmark = pytest.mark.gpu


def _rings():
    gp, ln = flat_ring(0.2)
    yield "snake", gp, ln.records()
    gp1, ln1, _ = W.config1(n_particles=1)[:3]
    yield "config1", gp1, ln1.records()


@pytest.mark.parametrize("which", ["snake", "config1"])
def test_spin_drivers_equal_oracle(ctx, which):
    name, gp, rec = next(r for r in _rings() if r[0] == which)
    o = pyoracle.Oracle(gp, rec)
    x0g, n0g, itg = D.cal_spin_orbit(ctx, gp, rec, 0.)
    x0o, n0o, ito = o.spin_orbit(0.)
    assert itg == ito
    assert np.array_equal(x0g.view(np.int64), x0o.view(np.int64)), f"{name}: start orbit and spin differ"
    assert np.array_equal(n0g.view(np.int64), n0o.view(np.int64)), f"{name}: n0 along the line differs"
    tg, ng = D.cal_spin_tune(ctx, gp, rec, 0.)
    to, no = o.spin_tune(0.)
    assert tg == to and np.array_equal(ng, no), f"{name}: spin tune {tg!r} vs {to!r}"
    x = np.concatenate([x0o[:6], [0.3, 0.4, 0.5]])
    sg = D.set_initial_spin(ctx, gp, rec, 20, n0o[-1], x)
    so = o.set_initial_spin(20, n0o[-1], x)
    assert np.array_equal(sg.view(np.int64), so.view(np.int64)), f"{name}: Set_Initial_Spin differs"
