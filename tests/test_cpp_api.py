"""The C++ API (include/simtrack_amd.hpp) lowers a line to the same C-ABI
records as the Python API, element for element and byte for byte, through
its own classes and through the duck-typed path a reference-side driver
uses.  Compiled with g++ and run on the CPU (no GPU call)."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import simtrack_amd as st

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def python_line():
    knl = [0., 0.031, 0.2, -1.5, 0., 12., 0., 0., 0., 0., 0.]
    knsl = [0., 0.001, -0.05, 0., 0., 0., 0., 3.0, 0., 0., 0.]
    m66 = np.array([1.0 if i % 7 == 0 else 1e-3 * i for i in range(36)])
    els = [
        st.DRIFT("D1", 1.3), st.MARKER("M0", 0.), st.BPM("B0", 0.), st.HBPM("HB", 0.2), st.VBPM("VB", 0.),
        st.SBEND("BE", 3.0, 0.02, 0.003, -0.002), st.GSBEND("GB", 2.2, 0.016, 0.0161, 0.001, 0.0),
        st.QUAD("QF", 1.1, 0.05, 0.002), st.SKEWQ("SQ", 0.4, 0.01), st.SEXT("SF", 0.5, 0.3, 0.02),
        st.OCT("OF", 0.3, 2.0, 0.5), st.MULT("MU", 0.9, knl, knsl), st.GMULT("GM", 1.5, -0.012, 0., 0., knl, knsl),
        st.SBENDMULT("SBM", 2.5, 0.018, 0.001, 0.001, knl, knsl), st.SMULT("SM", 1.2, 0.01, 0., 0., knl, knsl),
        st.GSBENDMULT("GBM", 1.9, 0.014, 0.0139, 0., 0., knl, knsl), st.KICKER("K", 0.4, 2e-5, -1e-5),
        st.HKICKER("HK", 0., 3e-5), st.VKICKER("VK", 0.2, -2e-5), st.RFCAV("RF", 0.6, 3.0, 28.15e6, 0.05),
        st.CRABRF("CR", 0.4, 2.0, 197e6, 0.1),
        st.CCMULT("CCM", 0., 25.8, 197e6, 0.0, 2.8e-3, -9.8e-7, 2.2e-3, -3.9e-5, 0.59, 2.3e-3, 0.1, 0.01),
        st.LBT("LBT", 0.0125), st.ILBT("ILBT", 0.0125),
        st.BEAMBEAM("BB6", 6, 1.72e11, -1.0, 0.02, 5, 20e-9, 0.45, 0., 1.3e-9, 0.056, 0.),
        st.CRBB("CRB", 1.5e11, 0.01, 1.0, 0.05, 5, 5e-9, 0.4, 0.02, 2.5e-9, 0.2, 0.),
        st.MATRIX("MAT", 0., np.zeros(6), [1e-4, 0., -2e-4, 0., 0., 0.], m66),
        st.ROTAT("ROT", 0.5, [0.3, 0.9, 0.1], 0.2), st.SNAKE("SNK", 0.1, [1.0, 0.0, 0.2], 1.5707963267948966),
        st.TRANS("TR", 2e-4, -1e-4, 0.05), st.SROTAT("SR", 0.013), st.YROTAT("YR", -0.007), st.XROTAT("XR", 0.009),
        st.COOLING("CO", 1e-4), st.LRBB("LR", 1.5e11, 1.0, 8e-4, -3e-4, 2e-4, 1e-4),
        st.ELENS("EL", 2.0, 1e11, 1.0, 3, 0.5, 3e-4, 2e-4), st.HELENS("HE", 2.0, 5e11, 1.0, 2, 0.5, 1e-2, 1e-6),
        st.HACDIP("HAC", 0.3, 2e-5, 0.31, 0.2, -3, 5), st.VACDIP("VAC", 0.0, 1e-5, 0.27, 0.0, -10, -2),
        st.HACMULT("HAM", 0.2, 2, 0.05, 100, 0.3), st.VACMULT("VAM", 0.0, 0, 1e-5, 50, 0.1),
        st.WIGGLER("WIG", 0.8, 8, 0.5, 2.0, 31.4, 0.1),
    ]
    ln = st.Line(els)
    ln.Cell[7].SetP("K1L", 0.06)
    ln.Cell[7].SetP("Nint", 7.9)
    ln.Cell[11].SetP("K4L", 0.5)
    ln.Cell[24].SetP("NSLICE", 3)
    ln.Cell[5].DX = 1e-4
    ln.Cell[5].DPSI = 2e-3
    ln.Cell[9].APx = 0.03
    ln.Cell[34].SetP("SEPX", 9e-4)
    ln.Cell[37].SetP("TURNE", 6)
    ln.Cell[39].SetP("PHI0", 0.25)
    ln.Cell[36].DX = 3e-5
    return ln


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_cpp_lowering_matches_python(tmp_path):
    lib = st.build.library_path()
    if not os.path.exists(lib):
        pytest.skip("library not built")
    exe = tmp_path / "lower_line"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "lower_line.cpp"), "-o", str(exe),
                    "-L", os.path.dirname(lib), "-lsimtrack_amd", "-Wl,-rpath," + os.path.dirname(lib)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, check=True)
    got = np.frombuffer(r.stdout, dtype=st.ELEMENT_DTYPE)
    ln = python_line()
    ref = ln.records()
    assert len(got) == len(ref) == 42
    for i in range(len(ref)):
        assert got[i].tobytes() == ref[i].tobytes(), ln.Cell[i].TYPE
    length = float(r.stderr.split()[1])
    assert length == ln.Length
    assert ref[11]["norder"] == 7 and ref[7]["nint"] == 7
    assert ref[15]["k0l"] == 0.0139 and ref[15]["knl"][0] == 0.0


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("fmt,fixture", [("madx", "rhic_2012_madx_twiss.tfs.gz"), ("bmad", "hsr_bmad_table.tab.gz")])
def test_cpp_lattice_readers_match_python(tmp_path, fmt, fixture):
    """Read_MADX_Lattice / Read_BMAD_Lattice of the C++ API give the same
    records, byte for byte, as simtrack_amd.lattice_io, and the same
    GP.circumference."""
    import gzip
    from simtrack_amd import lattice_io as LI
    lib = st.build.library_path()
    if not os.path.exists(lib):
        pytest.skip("library not built")
    src = os.path.join(ROOT, "tests", "golden", "lattices", fixture)
    plain = tmp_path / fixture[:-3]
    plain.write_bytes(gzip.open(src).read())
    exe = tmp_path / "read_lattice"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "read_lattice.cpp"), "-o", str(exe),
                    "-L", os.path.dirname(lib), "-lsimtrack_amd", "-Wl,-rpath," + os.path.dirname(lib)], check=True)
    r = subprocess.run([str(exe), fmt, str(plain)], capture_output=True, check=True)
    n = (len(r.stdout) - 8) // st.ELEMENT_DTYPE.itemsize
    got = np.frombuffer(r.stdout[:n * st.ELEMENT_DTYPE.itemsize], dtype=st.ELEMENT_DTYPE)
    circ = np.frombuffer(r.stdout[-8:], dtype=np.float64)[0]
    ln = LI.read_madx_twiss(src) if fmt == "madx" else LI.read_bmad_table(src)
    ref = ln.records()
    assert len(got) == len(ref)
    for i in range(len(ref)):
        assert got[i].tobytes() == ref[i].tobytes(), (i, ln.Cell[i].NAME, ln.Cell[i].TYPE)
    assert circ == ln.circumference


def test_cpp_normal_modes_match_python(tmp_path):
    """The C++ API's Cal_A (its own 4x4 eigenvectors) against the Python
    path's (numpy's) on coupled one-turn matrices R(th) diag(M1, M2) R^-1
    and products of them: the same A to rounding, in particular the same
    mode order, rotation and signs."""
    from test_diagnostics import _blk, _m2, _rot, _six

    from simtrack_amd import diagnostics as D
    lib = st.build.library_path()
    if not os.path.exists(lib):
        pytest.skip("library not built")
    exe = tmp_path / "normal_modes"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "normal_modes.cpp"), "-o", str(exe),
                    "-L", os.path.dirname(lib), "-lsimtrack_amd", "-Wl,-rpath," + os.path.dirname(lib)], check=True)
    rng = np.random.default_rng(8)
    mats = []
    for _ in range(20):
        th = rng.uniform(-0.7, 0.7)
        R = _rot(th)
        m = R @ _blk(_m2(rng.uniform(1, 50), rng.uniform(-2, 2), 2 * math.pi * rng.uniform(0.05, 0.45)),
                     _m2(rng.uniform(1, 50), rng.uniform(-2, 2), 2 * math.pi * rng.uniform(0.55, 0.95))) @ R.T
        mats.append(_six(m))
    r = subprocess.run([str(exe)], input=np.ascontiguousarray(mats).tobytes(), capture_output=True, check=True)
    got = np.frombuffer(r.stdout, dtype=np.float64).reshape(-1, 4, 4)
    assert len(got) == len(mats)
    for g, m in zip(got, mats):
        assert np.abs(g - D.cal_a(m)).max() < 1e-11
