"""The reference's example lattices on the GPU: RHIC (a MAD-X table, 6,858
elements, Au ions, as the first example driver sets it up) and the EIC
hadron ring (a Bmad table with patches, 2,838 elements, 275 GeV protons).
GPU = oracle bit for bit, orbit and spin tracking."""
import os

import numpy as np
import pytest

import pyoracle
import simtrack_amd as st
from simtrack_amd import workloads as W

from test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
RHIC = os.path.join(HERE, "golden", "lattices", "rhic_2012_madx_twiss.tfs.gz")
HSR = os.path.join(HERE, "golden", "lattices", "hsr_bmad_table.tab.gz")


def _beam(n, spin, seed):
    r = np.random.Generator(np.random.PCG64(seed))
    x = np.zeros((n, 9 if spin else 6))
    x[:, 0:4] = r.normal(size=(n, 4)) * [1e-3, 2e-5, 1e-3, 2e-5]
    x[:, 4] = 0.1 * r.normal(size=n)
    x[:, 5] = 3e-4 * r.normal(size=n)
    if spin:
        x[:, 7] = 1.0
    return x


@pytest.mark.parametrize("which", ["rhic", "hsr"])
@pytest.mark.parametrize("spin", [False, True], ids=["orbit", "spin"])
def test_example_lattice_bitwise(ctx, which, spin):
    gp, ln = W.rhic_example(RHIC) if which == "rhic" else W.hsr_example(HSR)
    rec = ln.records()
    x = _beam(64, spin, 5)
    nturn = 2 if spin else 3
    xg, sg, ltg, lpg = st.Table(ctx, gp, rec).track(x.copy(), nturn, spin=spin)
    xo, so, lto, lpo = pyoracle.Oracle(gp, rec).track(x.copy(), nturn, spin=spin)
    assert np.array_equal(sg, so) and np.array_equal(ltg, lto) and np.array_equal(lpg, lpo)
    assert sg.sum() > 48, "most of the beam should survive a few turns"
    assert_bitwise(xg, xo, f"{which} {'spin' if spin else 'orbit'}")


@pytest.mark.parametrize("spin", [False, True])
def test_detector_solenoid_bitwise(ctx, spin):
    """A detector-like solenoid (the HSR energy, KS = 0.1 /m over 2 m):
    SOLEN_Pass / SOLEN_sPass on the GPU equal the oracle bit for bit."""
    gp = W.proton_globals(275000.)
    rec = st.Line([st.DRIFT("D", 1.0), st.SOLEN("S", 2.0, 0.1), st.DRIFT("D", 1.0)]).records()
    x = W.random_particles(256, seed=7, spin=spin, scale=1e-3)
    xg, sg, ltg, lpg = st.Table(ctx, gp, rec).track(x.copy(), 5, spin=spin)
    xo, so, lto, lpo = pyoracle.Oracle(gp, rec).track(x.copy(), 5, spin=spin)
    assert np.array_equal(sg, so)
    assert_bitwise(xg, xo, f"solenoid {'spin' if spin else 'orbit'}")


def test_unsupported_type_is_rejected(ctx):
    gp = W.proton_globals(275000.)
    rec = st.Line([st.DRIFT("D", 1.0)]).records()
    rec["type"] = 99
    with pytest.raises(st.STError, match="type"):
        st.Table(ctx, gp, rec)
