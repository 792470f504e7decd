"""The C++ API (include/simtrack_amd.hpp) on the GPU: per-particle Track()
calls, a batched Track(), Track_spin() and Track_DA1() from a compiled C++
program, against the CPU oracle on the same ring built with the Python API."""
import math
import os
import subprocess

import numpy as np
import pytest

import pyoracle
import simtrack_amd as st
from simtrack_amd import workloads as W

from test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "build", "track_ring")
NP, NTURN = 24, 200


def ring():
    """tests/cpp/track_ring.cpp's build_ring, with the Python API"""
    ncell = 8
    ang = 2 * 3.14159265358979323 / (2 * ncell)
    els = []
    for c in range(ncell):
        kn = [0.] * 11
        ks = [0.] * 11
        kn[1] = 0.0372
        kn[2] = 0.01 * (c % 3 - 1)
        ks[3] = 0.5 * (c % 2)
        qf = st.MULT("QF", 1.5, kn, ks)
        qf.SetP("Nint", 4)
        kd = list(kn)
        kd[1] = -0.0372
        qd = st.MULT("QD", 1.5, kd, ks)
        qd.SetP("Nint", 4)
        sf = st.SEXT("SF", 0.5, 0.12, 0.)
        sf.SetP("Nint", 4)
        els += [qf, st.DRIFT("D", 0.75), sf, st.DRIFT("D", 0.75), st.SBEND("B", 32., ang, 0., 0.),
                st.DRIFT("D", 0.75), st.KICKER("K", 0.5, 1e-6, -1e-6), st.BPM("BPM", 0.), qd, st.DRIFT("D", 2.0),
                st.SBEND("B", 32., ang, 0., 0.), st.DRIFT("D", 1.25)]
    ln = st.Line(els)
    ln.Cell[3].DX = 2e-4
    ln.Cell[8].DPSI = 1e-3
    for e in ln.Cell:
        e.APx = e.APy = 0.05
    return ln


def test_cpp_api_tracks_like_the_oracle(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/build/track_ring is not built (run __graft_entry__.build())")
    da_file = tmp_path / "da1.txt"
    r = subprocess.run([EXE, str(da_file)], capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    buf = r.stdout
    o = 0

    def take(dtype, n):
        nonlocal o
        a = np.frombuffer(buf, dtype=dtype, count=n, offset=o)
        o += a.nbytes
        return a

    x1 = take(np.float64, NP * 6).reshape(NP, 6)
    s1, t1, p1 = take(np.int32, NP), take(np.int32, NP), take(np.int32, NP)
    x2 = take(np.float64, NP * 6).reshape(NP, 6)
    s2 = take(np.int32, NP)
    xs = take(np.float64, NP * 9).reshape(NP, 9)
    s3 = take(np.int32, NP)
    assert o == len(buf)

    g = W.proton_globals(100000.0)
    rec = ring().records()
    orc = pyoracle.Oracle(g, rec)
    x0 = np.zeros((NP, 6))
    for p in range(NP):
        a, t = 2e-3 * (p + 1), 0.2 + 0.05 * p
        x0[p] = [a * math.cos(t), 0., 0.6 * a * math.sin(t), 0., 0., 2e-4]
    xo, so, lto, lpo = orc.track(x0.copy(), NTURN)
    assert np.array_equal(s1, so) and np.array_equal(t1, lto) and np.array_equal(p1, lpo)
    assert 0 < so.sum() < NP
    assert_bitwise(x1, xo, "C++ Track per particle")
    assert_bitwise(x2, xo, "C++ Track batched")
    assert np.array_equal(s2, so)
    x9 = np.zeros((NP, 9))
    x9[:, :6] = x0
    x9[:, 7] = 1.0
    xso, sso, _, _ = orc.track(x9, NTURN, spin=True)
    assert np.array_equal(s3, sso)
    assert_bitwise(xs, xso, "C++ Track_spin")
    lines = da_file.read_text().splitlines()
    assert len(lines) == 5 and all(len(l.split()) == 2 for l in lines)
