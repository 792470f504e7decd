"""The radiating 18 GeV electron ring (config 4) on the GPU.

1. Per-particle determinism: on the ~3,500-element ring the GPU equals the
   oracle bit for bit, for orbit tracking with quantum excitation (radiate1
   draws from the counter-based stream keyed by seed, particle and draw) and
   for spin tracking (classical radiation, as SBEND_sPass does,
   simtrack.cpp:7519-7529).  A run split into turn ranges equals one run.
2. Equilibrium: on a strongly damped ring (80 bends of rho = 51 m, U0/E = 1 %
   per turn), a mismatched beam of 32,768 electrons relaxes to the emittance
   and energy spread that the radiation integrals give (Cq gamma^2 I5 / (Jx I2),
   Cq gamma^2 I3 / (Jz I2)), and the vertical emittance damps at the rate
   2 / tau_y.  The arcs are dispersion suppressors (90 degree cells) and the
   RF sits in the dispersion-free straights: RF at dispersive points couples
   synchrotron and betatron motion, which the uncoupled integrals omit (it
   raises the tracked emittance by ~20 % when every arc cell has a cavity).
3. Photon emission: the same ring with radiating GSBENDs (radiate2, Poisson
   photons from the counter-based stream) equals the oracle bit for bit.
"""
import math

import numpy as np
import pytest
import torch

import pyoracle
import simtrack_amd as st
from simtrack_amd import optics as O
from simtrack_amd import workloads as W

from test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ring4():
    gp, ln, x = W.config4(n_particles=64, spin=True)
    gp.rng_seed = 20260
    return gp, ln, x


def test_config4_orbit_with_quantum_excitation_bitwise(ctx, ring4):
    gp, ln, x9 = ring4
    x = np.ascontiguousarray(x9[:, :6])
    rec = ln.records()
    tab = st.Table(ctx, gp, rec)
    xg, sg, ltg, lpg = tab.track(x.copy(), 2)
    xo, so, lto, lpo = pyoracle.Oracle(gp, rec).track(x.copy(), 2)
    assert sg.all() and np.array_equal(sg, so)
    assert_bitwise(xg, xo, "config 4 orbit + radiate1")
    # the quantum kicks are real: a second stream seed changes the result
    g2 = W.electron_globals_ring()
    g2.rng_seed = gp.rng_seed + 1
    x2, _, _, _ = st.Table(ctx, g2, rec).track(x.copy(), 2)
    assert not np.array_equal(x2, xg)


def test_config4_spin_with_radiation_bitwise(ctx, ring4):
    gp, ln, x = ring4
    rec = ln.records()
    tab = st.Table(ctx, gp, rec)
    xg, sg, _, _ = tab.track(x.copy(), 2, spin=True)
    xo, so, _, _ = pyoracle.Oracle(gp, rec).track(x.copy(), 2, spin=True)
    assert sg.all() and np.array_equal(sg, so)
    assert_bitwise(xg, xo, "config 4 spin + radiation")
    assert np.allclose(np.linalg.norm(xg[:, 6:9], axis=1), 1.0, atol=1e-13)


def test_turn_ranges_continue_the_stream(ctx, ring4):
    gp, ln, x9 = ring4
    x = np.ascontiguousarray(x9[:, :6])
    tab = st.Table(ctx, gp, ln.records())
    one, _, _, _ = tab.track(x.copy(), 3)
    dev = torch.device("cuda:0")
    d_x = torch.tensor(x, dtype=torch.float64, device=dev)
    d_s = torch.ones(len(x), dtype=torch.int32, device=dev)
    d_l = torch.zeros(2 * len(x), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    for t0, nt in ((0, 1), (1, 2)):
        tab.track_device(d_x.data_ptr(), len(x), t0, nt, d_s.data_ptr(), d_l.data_ptr(), 0)
    ctx.sync()
    assert_bitwise(d_x.cpu().numpy(), one, "turn ranges 0-1, 1-3 vs one run of 3 turns")


def _gsbend_line(ln):
    """The ring with every SBEND a GSBEND whose field is 0.1 % above the
    frame's bend (K0L != ANGLE), so it radiates through radiate2."""
    cells = []
    for e in ln.Cell:
        if e.TYPE == "SBEND":
            g = st.GSBEND(e.NAME, e.L, e.ANGLE, e.ANGLE * 1.001, e.E1, e.E2)
            g.Nint = e.Nint
            e = g
        cells.append(e)
    return st.Line(cells)


@pytest.mark.parametrize("hexp", [1, 0], ids=["yoshida", "exact"])
def test_gsbend_photon_emission_bitwise(ctx, ring4, hexp):
    """radiate2 on the ring (400 radiating GSBENDs, Poisson photons):
    GPU = oracle bit for bit over 2 turns, and a run split into turn ranges
    draws the same photons."""
    gp0, ln, x9 = ring4
    gp = W.electron_globals_ring()
    gp.rng_seed = 4242
    gp.H_expand = hexp
    x = np.ascontiguousarray(x9[:, :6])
    rec = _gsbend_line(ln).records()
    tab = st.Table(ctx, gp, rec)
    xg, sg, _, _ = tab.track(x.copy(), 2)
    xo, so, _, _ = pyoracle.Oracle(gp, rec).track(x.copy(), 2)
    assert sg.all() and np.array_equal(sg, so)
    assert_bitwise(xg, xo, "GSBEND ring + radiate2")
    dev = torch.device("cuda:0")
    d_x = torch.tensor(x, dtype=torch.float64, device=dev)
    d_s = torch.ones(len(x), dtype=torch.int32, device=dev)
    d_l = torch.zeros(2 * len(x), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    for t0, nt in ((0, 1), (1, 1)):
        tab.track_device(d_x.data_ptr(), len(x), t0, nt, d_s.data_ptr(), d_l.data_ptr(), 0)
    ctx.sync()
    assert_bitwise(d_x.cpu().numpy(), xg, "turn ranges 0-1, 1-2 vs one run of 2 turns")


def _damped_ring():
    gp = W.electron_globals_ring()
    gp.rng_seed = 777
    ln = W.electron_ring(gp, n_arc_cells=40, n_straight_cells=4, bend_length=4.0, chromaticity=None,
                         harmonic=400, overvoltage=2.0, mu_deg=90., straight_kq_scale=0.85)
    return gp, ln


def test_radiation_equilibrium(ctx):
    gp, ln = _damped_ring()
    opt = O.Optics(ln)
    eq = opt.equilibrium(gp)
    zs, _ = W.rf_synchronous(opt, gp, ln)
    sz = W._bunch_length(opt, gp, ln, eq["sigma_delta"])
    n = 32768
    # mismatched start: twice the horizontal emittance, half the energy spread
    x = opt.matched_beam(n, 2 * eq["eps_x"], eq["eps_x"], 0.5 * eq["sigma_delta"], 0.5 * sz, seed=31)
    x[:, 4] += zs
    tab = st.Table(ctx, gp, ln.records())
    dev = torch.device("cuda:0")
    d_x = torch.tensor(x, dtype=torch.float64, device=dev)
    d_s = torch.ones(n, dtype=torch.int32, device=dev)
    d_l = torch.zeros(2 * n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    step, nturn, t_eq = 25, 1200, 700
    ex, sd, ey = [], [], []
    for t0 in range(0, nturn, step):
        tab.track_device(d_x.data_ptr(), n, t0, step, d_s.data_ptr(), d_l.data_ptr(), 0)
        ctx.sync()
        snap = d_x.cpu().numpy()
        em = sorted(O.eigen_emittances(snap))
        ey.append((t0 + step, em[0]))
        if t0 + step >= t_eq:
            ex.append(em[1])
            sd.append(snap[:, 5].std())
    assert int(d_s.sum()) == n, "no electron may be lost"
    ex_m, sd_m = float(np.mean(ex)), float(np.mean(sd))
    print(f"eps_x {ex_m:.4e} (theory {eq['eps_x']:.4e}), sigma_delta {sd_m:.4e} (theory {eq['sigma_delta']:.4e})")
    assert abs(ex_m / eq["eps_x"] - 1) < 0.03
    assert abs(sd_m / eq["sigma_delta"] - 1) < 0.02
    # vertical: no excitation in a flat ring, the emittance damps as exp(-2 t / tau_y)
    t = np.array([p[0] for p in ey if p[0] <= 600])
    e = np.array([p[1] for p in ey if p[0] <= 600])
    rate = -np.polyfit(t, np.log(e), 1)[0]
    assert abs(rate / (2 / eq["tau_y_turns"]) - 1) < 0.05
