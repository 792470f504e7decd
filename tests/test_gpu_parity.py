"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

The bar is bit-exact equality of every coordinate and of the loss records:
both sides perform the reference's operations in the same order with the
same elementary functions (st_libm.h) and no FMA contraction.
"""
import math

import numpy as np
import pytest

import pyoracle
import simtrack_amd as st
from simtrack_amd import workloads as W

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def bitwise_mismatch(gpu, ref):
    """Positions whose IEEE bit patterns differ.  NaN sign and payload are
    unspecified by IEEE 754 for invalid operations (x86 returns a negative
    default NaN, gfx950 a positive one) and the reference only ever tests
    isnan(), so two NaNs count as equal."""
    gpu = np.ascontiguousarray(gpu, dtype=np.float64)
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    diff = bits(gpu) != bits(ref)
    return diff & ~(np.isnan(gpu) & np.isnan(ref))


def assert_bitwise(gpu, ref, what):
    m = bitwise_mismatch(gpu, ref)
    if m.any():
        bad = np.argwhere(m)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} mismatching values; first at {i}: gpu {gpu[i]!r} "
                             f"oracle {ref[i]!r}")


FLAG_SETS = [
    dict(),
    dict(H_expand=1),
    dict(quad_fringe=1),
    dict(radiate=1, rng_seed=12345),
    dict(H_expand=1, radiate=1, rng_seed=99),
]


def _line_and_globals(flags, electron=False):
    g = W.electron_globals(18000.0, **flags) if electron else W.proton_globals(100000.0, **flags)
    ln = W.all_types_line(seed=11)
    return g, ln


@pytest.mark.parametrize("flags", FLAG_SETS, ids=lambda f: ",".join(f"{k}={v}" for k, v in f.items()) or "default")
@pytest.mark.parametrize("spin", [False, True], ids=["pass", "spass"])
def test_element_pass_bitwise(ctx, flags, spin):
    """Element::Pass / sPass of every element type, 256 random particles each."""
    electron = bool(flags.get("radiate"))
    g, ln = _line_and_globals(flags, electron)
    rec = ln.records()
    tab = st.Table(ctx, g, rec)
    orc = pyoracle.Oracle(g, rec)
    x0 = W.random_particles(256, seed=3, spin=spin)
    for j in range(len(rec)):
        xg = tab.element_pass(j, x0.copy(), spin=spin)
        xo = orc.element_pass(j, x0.copy(), spin=spin)
        assert_bitwise(xg, xo, f"element {j} {ln.Cell[j].TYPE} ({ln.Cell[j].NAME})")


def _fodo_ring(ncell=24, errors=True, seed=5):
    r = np.random.Generator(np.random.PCG64(seed))
    cells = []
    ang = 2 * np.pi / (2 * ncell)
    for c in range(ncell):
        kn = [0.] * 11
        ks = [0.] * 11
        kn[1] = 0.0372
        if errors:
            for j in range(2, 10):
                kn[j] = 0.0372 * 0.04 * 1e-4 * r.normal() * math.factorial(j) / 0.04 ** j
                ks[j] = 0.0372 * 0.04 * 1e-4 * r.normal() * math.factorial(j) / 0.04 ** j
        qf = st.MULT("QF", 1.5, kn, ks)
        qf.Nint = 4
        kd = list(kn)
        kd[1] = -0.0372
        qd = st.MULT("QD", 1.5, kd, ks)
        qd.Nint = 4
        sf = st.SEXT("SF", 0.5, 0.12, 0.)
        sd = st.SEXT("SD", 0.5, -0.22, 0.)
        sf.Nint = sd.Nint = 4
        cells += [qf, st.DRIFT("D", 0.75), sf, st.DRIFT("D", 0.75), st.SBEND("B", 32., ang),
                  st.DRIFT("D", 0.75), st.KICKER("K", 0.5, 1e-6, -1e-6), st.DRIFT("D", 0.75), st.BPM("BPM", 0.),
                  qd, st.DRIFT("D", 0.75), sd, st.DRIFT("D", 0.75), st.SBEND("B", 32., ang),
                  st.DRIFT("D", 0.75), st.BPM("BPM", 0.), st.DRIFT("D", 1.25)]
    cells.insert(5, st.RFCAV("RF", 0., 3.0, 360 * 299792458. * 0.99995 / (76. * ncell), 0.))
    for e in cells:
        e.APx = e.APy = 0.05
    return st.Line(cells)


def _grid(n, amax=8e-3, delta=3e-4, spin=False):
    a = np.linspace(amax / n, amax, n)
    ang = np.deg2rad(np.array([15., 30., 45., 60., 75.]))
    A, T = np.meshgrid(a, ang, indexing="ij")
    m = A.size
    x = np.zeros((m, 9 if spin else 6))
    x[:, 0] = (A * np.cos(T)).ravel()
    x[:, 2] = (A * np.sin(T)).ravel() * 0.6
    x[:, 5] = delta
    if spin:
        x[:, 7] = 1.0
    return x


@pytest.mark.parametrize("mode", ["plain", "radial", "spin"])
def test_ring_tracking_bitwise(ctx, mode):
    """Multi-turn Track / Track_Radial / Track_spin on a FODO ring with b2..b9
    errors: coordinates and loss records identical, including particles lost
    on the way (grid spans the dynamic aperture)."""
    ln = _fodo_ring()
    g = W.proton_globals(100000.0)
    rec = ln.records()
    spin = mode == "spin"
    x = _grid(40, amax=2.4e-2, spin=spin)
    tab = st.Table(ctx, g, rec)
    orc = pyoracle.Oracle(g, rec)
    nturn = 300
    xg, sg, ltg, lpg = tab.track(x.copy(), nturn, spin=spin, radial=(mode == "radial"))
    xo, so, lto, lpo = orc.track(x.copy(), nturn, spin=spin, radial=(mode == "radial"))
    assert np.array_equal(sg, so) and np.array_equal(ltg, lto) and np.array_equal(lpg, lpo)
    assert 0 < sg.sum() < len(sg), "grid should straddle the aperture"
    assert_bitwise(xg, xo, f"ring {mode}")


def test_tbt_bitwise(ctx):
    """Track_tbt at turn start and the BPM overload after one element."""
    ln = _fodo_ring(ncell=8)
    g = W.proton_globals(100000.0)
    rec = ln.records()
    tab = st.Table(ctx, g, rec)
    orc = pyoracle.Oracle(g, rec)
    x = _grid(6, amax=1e-2)
    for bpm in (-1, 8):
        _, tg, sg, ltg, lpg = tab.track_tbt(x.copy(), 50, bpm_index=bpm)
        _, to, so, lto, lpo = orc.track(x.copy(), 50, tbt=True, bpm_index=bpm)
        assert np.array_equal(sg, so)
        assert_bitwise(tg, to, f"tbt bpm={bpm}")


def test_loss_semantics(ctx):
    """Pre-check at 0.1 m, stable==0 on input, NaN and aperture losses."""
    ln = _fodo_ring(ncell=4)
    g = W.proton_globals(100000.0)
    rec = ln.records()
    tab = st.Table(ctx, g, rec)
    orc = pyoracle.Oracle(g, rec)
    x = np.zeros((6, 6))
    x[0, 0] = 0.11                  # pre-check
    x[1, 2] = -0.1000001            # pre-check on y
    x[2, 0] = 1e-3                  # stable == 0 on input
    x[3, 0] = 0.049                 # lost on aperture 0.05 early
    x[4, 1] = 2.0                   # |px| > 1+delta: sqrt of negative -> NaN in the first drift
    x[5, 0] = 1e-4
    stable = np.array([1, 1, 0, 1, 1, 1], np.int32)
    lt0 = np.full(6, 7, np.int32)
    lp0 = np.full(6, 9, np.int32)
    rg = tab.track(x.copy(), 20, stable=stable.copy(), lost_turn=lt0.copy(), lost_post=lp0.copy())
    ro = orc.track(x.copy(), 20, stable=stable.copy(), lost_turn=lt0.copy(), lost_post=lp0.copy())
    for a, b in zip(rg[1:], ro[1:]):
        assert np.array_equal(a, b)
    assert_bitwise(rg[0], ro[0], "loss cases")
    assert list(rg[1]) == [0, 0, 0, 0, 0, 1]
    assert rg[2][0] == 0 and rg[3][0] == 0 and rg[2][2] == 0 and rg[3][2] == 0
    assert rg[2][5] == 7 and rg[3][5] == 9, "survivors keep their input lost_turn/lost_post"
