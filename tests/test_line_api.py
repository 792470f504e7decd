"""Element / Line construction API against the reference's constructors and
SetP/GetP (defaults cited by file:line), and the synthetic workloads."""
import math

import numpy as np
import pytest

import simtrack_amd as st
from simtrack_amd import workloads as W


@pytest.mark.parametrize("L", [0.0, 0.05, 0.3, 1.0, 2.0, 7.77])
def test_quad_family_slices(L):
    """QUAD/SEXT/OCT/MULT: Nint = int(L/QLslice)+1 (simtrack.cpp:8635, 8782,
    8862, 8943), QLslice = 0.05."""
    for el in (st.QUAD("Q", L, 0.1, 0.), st.SEXT("S", L, 0.1, 0.), st.OCT("O", L, 0.1, 0.),
               st.MULT("M", L, [0., 0.1], [0.])):
        assert el.Nint == int(L / 0.05) + 1


@pytest.mark.parametrize("L", [0.1, 0.25, 1.0, 32.0])
def test_bend_slices(L):
    """SBEND/GSBEND/SBENDMULT: Nint = int(L/BLslice)+1, BLslice = 0.25
    (simtrack.cpp:8451, 8540, 9057)."""
    for el in (st.SBEND("B", L, 0.01), st.GSBEND("G", L, 0.01, 0.01),
               st.SBENDMULT("BM", L, 0.01, 0., 0., [0.] * 11, [0.] * 11)):
        assert el.Nint == int(L / 0.25) + 1


def test_mult_norder_is_highest_nonzero():
    """simtrack.cpp:8944-8950: Norder = highest index with KNL or KNSL != 0,
    1 when all of 1..10 vanish."""
    assert st.MULT("M", 0., [0.] * 11, [0.] * 11).Norder == 1
    kn = [0.] * 11
    kn[4] = 1e-3
    ks = [0.] * 11
    ks[7] = -2.
    assert st.MULT("M", 0., kn, [0.] * 11).Norder == 4
    assert st.MULT("M", 0., kn, ks).Norder == 7
    kn[10] = 1.
    assert st.MULT("M", 0., kn, ks).Norder == 10
    assert st.MULT("M", 0., [5.] + [0.] * 10, [0.] * 11).Norder == 1   # KNL[0] not scanned


def test_setp_getp_and_unknown_names():
    q = st.QUAD("Q", 1.0, 0.1, 0.)
    q.SetP("K1L", 0.2)
    q.SetP("Nint", 7.9)
    assert q.GetP("K1L") == 0.2 and q.GetP("Nint") == 7 and isinstance(q.Nint, int)
    m = st.MULT("M", 1.0, [0.] * 11, [0.] * 11)
    m.SetP("K3L", 1.5)
    m.SetP("K2SL", -0.5)
    m.SetP("Norder", 3)
    assert m.KNL[3] == 1.5 and m.KNSL[2] == -0.5 and m.GetP("K3L") == 1.5
    r = st.Line([m]).records()[0]
    assert r["knl"][3] == 1.5 and r["knsl"][2] == -0.5 and r["norder"] == 3
    with pytest.raises(KeyError):
        q.SetP("K2L", 1.)
    with pytest.raises(KeyError):
        st.DRIFT("D", 1.).GetP("L2")


def test_lengths_are_checked():
    with pytest.raises(ValueError):
        st.DRIFT("D", -1.)
    with pytest.raises(ValueError):
        st.SBEND("B", 0., 0.1)


def test_line_update_positions():
    """Line::Update (simtrack.cpp:12362-12372): S is the end position of each
    cell, Length the sum of L; Append/Insert/Delete/Replace keep them."""
    ln = st.Line([st.DRIFT("A", 1.), st.QUAD("Q", 0.5, 0.1, 0.), st.MARKER("M", 0.)])
    assert [c.S for c in ln.Cell] == [1., 1.5, 1.5] and ln.Length == 1.5 and ln.Ncell == 3
    ln.Append(st.DRIFT("B", 2.))
    ln.Insert(0, st.DRIFT("C", 0.25))
    assert ln.Length == 3.75 and ln.Cell[1].S == 1.25 and len(ln) == 5
    ln.Delete(1)
    assert ln.Length == 2.75 and ln.Cell[0].NAME == "C"
    ln.Replace(0, st.DRIFT("E", 1.0))
    assert ln.Length == 3.5
    ln.Delete(99)                        # out of range: ignored like the reference
    assert len(ln) == 4


def test_all_types_line_covers_every_type():
    ln = W.all_types_line(seed=1)
    types = {c.TYPE for c in ln.Cell}
    assert types == set(st.TYPES), set(st.TYPES) - types


# -------------------------------------------------------------- workloads

def test_ir_matching_reproduces_periodic_twiss():
    """match_ir: the IR quadruplet transports the arc's periodic Twiss to
    beta* at the IP with alpha* = 0, and the ring tunes are fractional and
    away from low-order resonances."""
    k = W.match_ir()
    assert np.allclose(k, W.IR_K1L, rtol=1e-6)
    (bx, ax, qx), (by, ay, qy) = W.ring_linear_tunes(k)
    # the ring's own periodic solution at the IP is the matched beta*
    assert abs(bx - W.BETA_STAR) < 1e-9 and abs(by - W.BETA_STAR) < 1e-9
    assert abs(ax) < 1e-9 and abs(ay) < 1e-9
    for q in (qx % 1, qy % 1):
        assert 0.05 < q < 0.45
        for n in (3, 4, 5):
            assert abs(n * q - round(n * q)) > 0.01


def test_da_ring_structure():
    ln = W.da_ring(seed=3)
    rec = ln.records()
    assert len(rec) == 998
    n_by = {t: int((rec["type"] == v).sum()) for t, v in st.TYPES.items()}
    assert n_by["SBEND"] == 96 and n_by["RFCAV"] == 1 and n_by["MULT"] == 256
    assert np.all(rec["apx"] == 0.5) and np.all(rec["apy"] == 0.5)
    # circumference and a deterministic lattice per seed
    assert math.isclose(ln.Length, W.da_ring(seed=3).Length)
    assert np.array_equal(W.da_ring(seed=3).records(), rec)
    assert not np.array_equal(W.da_ring(seed=4).records()["knl"], rec["knl"])


def test_da_grid_shape():
    gp = W.proton_globals(100000.0)
    x, amps, angles = W.da_grid(gp, n_amp=205)
    assert x.shape == (205 * 5, 6) and len(amps) == 205 and len(angles) == 5
    assert amps[0] == 20. / 205 and amps[-1] == 20.
    r = np.hypot(x[:, 0], x[:, 2])
    assert np.all(r > 0) and np.all(x[:, 5] == 3e-4)


def test_random_particles_deterministic():
    a = W.random_particles(10, seed=5, spin=True)
    b = W.random_particles(10, seed=5, spin=True)
    assert np.array_equal(a, b) and a.shape == (10, 9)
    assert np.allclose(np.linalg.norm(a[:, 6:9], axis=1), 1.0)
