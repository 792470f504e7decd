"""The C ABI (include/simtrack_amd.h) without a GPU: the library loads,
exports every declared function, its struct layouts agree with the header
as a C compiler sees it, the reference-particle setters reproduce
Set_RefPartEnergy, and bad arguments come back as error codes."""
import ctypes
import math
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import simtrack_amd as st
from simtrack_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "simtrack_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(st_[a-z0-9_]+)\s*\([^;{]*\)\s*;", text)))


@pytest.fixture(scope="module")
def lib():
    return st.load()


def test_every_declared_function_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/simtrack_amd.h but not exported"
    assert sorted(_capi.EXPORTS) == names, "the binding's export list drifted from the header"


def test_abi_version_and_sizes(lib):
    assert lib.st_abi_version() >= 1
    es, gs = ctypes.c_int(), ctypes.c_int()
    lib.st_struct_sizes(ctypes.byref(es), ctypes.byref(gs))
    assert es.value == st.ELEMENT_DTYPE.itemsize
    assert gs.value == ctypes.sizeof(st.Globals)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_struct_layout_matches_c_compiler(tmp_path):
    """offsetof() of every field as gcc sees the header == the numpy/ctypes
    mirrors the Python binding uses."""
    fields = [n for n in st.ELEMENT_DTYPE.names]
    gfields = [f[0] for f in st.Globals._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "simtrack_amd.h"', 'int main(void){']
    src += [f'printf("e {n} %zu\\n", offsetof(st_element, {n}));' for n in fields]
    src += [f'printf("g {n} %zu\\n", offsetof(st_globals, {n}));' for n in gfields]
    src += ['printf("E %zu\\nG %zu\\n", sizeof(st_element), sizeof(st_globals)); return 0;}']
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        p = line.split()
        if len(p) == 3:
            got[(p[0], p[1])] = int(p[2])
        elif len(p) == 2:
            got[p[0]] = int(p[1])
    for n in fields:
        assert got[("e", n)] == st.ELEMENT_DTYPE.fields[n][1], n
    for n in gfields:
        assert got[("g", n)] == getattr(st.Globals, n).offset, n
    assert got["E"] == st.ELEMENT_DTYPE.itemsize and got["G"] == ctypes.sizeof(st.Globals)


def test_globals_default_is_reference_gp(lib):
    """GlobalVariables() defaults, simtrack.h:31-58."""
    g = st.Globals()
    lib.st_globals_default(ctypes.byref(g))
    assert (g.A, g.Q) == (1, 1)
    assert g.m0 == g.m == 1.6726216e-27 and g.q == 1.6021765e-19
    assert (g.gamma, g.beta, g.energy0, g.energy) == (106., 0.999955, 938.27201, 99456.8)
    assert (g.p, g.cp, g.brho, g.G) == (1., 1., 331.738, 1.7928474)
    assert (g.H_expand, g.radiate, g.quad_fringe) == (0, 0, 0)


@pytest.mark.parametrize("gamma,E0,Q,A", [(106.57803, 938.27201, 1, 1), (35225.0, 0.51099895, -1, 1),
                                          (107.4, 183473.2, 79, 197)])
def test_set_energy_matches_python_mirror_and_physics(lib, gamma, E0, Q, A):
    """Set_RefPartEnergy (simtrack.cpp:87-108) in the library == the Python
    mirror bit for bit, and obeys E = gamma m c^2, cp = beta E, p = cp / c."""
    g = st.Globals()
    lib.st_globals_default(ctypes.byref(g))
    lib.st_globals_set_energy(ctypes.byref(g), gamma, E0, Q, A)
    h = st.set_ref_part_energy(st.globals_default(), gamma, E0, Q, A)
    for f, _ in st.Globals._fields_:
        assert getattr(g, f) == getattr(h, f), f
    assert math.isclose(g.energy, gamma * E0, rel_tol=1e-15)
    assert math.isclose(g.beta, math.sqrt(1 - 1 / gamma ** 2), rel_tol=1e-15)
    assert math.isclose(g.p * 299792458. / 1.6021765e-13, g.cp, rel_tol=1e-14)
    assert math.isclose(g.m * 299792458. ** 2 / 1.6021765e-13, g.energy, rel_tol=1e-14)
    assert g.Q == Q and g.A == A


def test_null_arguments_are_errors(lib):
    P = ctypes.c_void_p
    assert lib.st_ctx_create(0, None) != 0
    assert b"null" in lib.st_last_error()
    assert lib.st_ctx_sync(None) != 0
    assert lib.st_ctx_set_timing(None, 1) != 0
    assert lib.st_table_size(None, None, None) != 0
    g = st.Globals()
    out = P()
    assert lib.st_table_build(None, ctypes.byref(g), None, 0, ctypes.byref(out)) != 0
    for f in ("st_track", "st_track_spin", "st_track_radial"):
        assert getattr(lib, f)(None, None, 1, 1, None, None, None) != 0
    assert lib.st_element_pass(None, 0, None, 1, 0) != 0
    assert lib.st_track_device(None, None, 1, 0, 1, None, None, 0) != 0
    assert lib.st_last_error() != b""


def _gpu_visible():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-device error path")
def test_context_without_device_fails_cleanly(lib):
    h = ctypes.c_void_p()
    rc = lib.st_ctx_create(0, ctypes.byref(h))
    assert rc != 0 and not h.value
    assert lib.st_last_error() != b""
    with pytest.raises(st.STError):
        st.Context(0)


def test_library_missing_fails_loudly(tmp_path):
    with pytest.raises(st.STError, match="not built"):
        _capi.load(str(tmp_path / "nope.so"))


def test_element_record_layout_roundtrip():
    """Line.records() fills the C struct fields the table builder reads."""
    q = st.QUAD("Q", 1.2, 0.05, -0.01)
    q.DX, q.DPSI, q.APx = 1e-4, 2e-3, 0.03
    r = st.Line([q]).records()[0]
    assert r["type"] == st.TYPES["QUAD"] and r["l"] == 1.2 and r["nint"] == int(1.2 / 0.05) + 1
    assert r["norder"] == 1 and r["knl"][1] == 0.05 and r["knsl"][1] == -0.01
    assert (r["dx"], r["dpsi"], r["apx"], r["apy"]) == (1e-4, 2e-3, 0.03, 1.0)
    raw = np.frombuffer(st.Line([q]).records().tobytes(), np.uint8)
    assert raw.size == st.ELEMENT_DTYPE.itemsize
