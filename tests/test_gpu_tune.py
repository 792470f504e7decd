"""FineTuneFinder on the GPU against the oracle's restatement of
simtrack.cpp:2288-2352: the same peak tune bit for bit, both forms, on
synthetic turn-by-turn signals and on tracked particles."""
import ctypes

import numpy as np
import pytest

import pyoracle
import simtrack_amd as st
from simtrack_amd import workloads as W

pytestmark = pytest.mark.gpu


def signals(nsig, n, seed):
    r = np.random.default_rng(seed)
    i = np.arange(n)
    q = r.uniform(0.1, 0.4, nsig)
    ph = r.uniform(0, 2 * np.pi, nsig)
    x = np.cos(2 * np.pi * q[:, None] * i + ph[:, None]) * r.uniform(1e-4, 1e-3, (nsig, 1))
    x += 0.1 * x.std(axis=1, keepdims=True) * np.cos(2 * np.pi * r.uniform(0.1, 0.4, (nsig, 1)) * i)
    x += 1e-6 * r.normal(size=(nsig, n))
    px = -np.sin(2 * np.pi * q[:, None] * i + ph[:, None]) * np.abs(x).max(axis=1, keepdims=True)
    return x, px, q


@pytest.mark.parametrize("n,lo,hi", [(512, 0.05, 0.45), (1024, 0.0, 0.5), (256, 0.2, 0.3)])
def test_fine_tune_matches_oracle(ctx, n, lo, hi):
    lib = pyoracle.load()
    x, px, q = signals(40, n, seed=n)
    got = ctx.fine_tune(x, lo, hi)
    got2 = ctx.fine_tune(x, lo, hi, px=px)
    out = ctypes.c_double()
    for j in range(len(x)):
        xr = np.ascontiguousarray(x[j])
        lib.orc_fine_tune(n, xr.ctypes.data_as(ctypes.c_void_p), lo, hi, ctypes.byref(out))
        assert got[j] == out.value, j
        pr = np.ascontiguousarray(px[j])
        lib.orc_fine_tune_xpx(n, xr.ctypes.data_as(ctypes.c_void_p), pr.ctypes.data_as(ctypes.c_void_p), lo, hi,
                              ctypes.byref(out))
        assert got2[j] == out.value, j
    in_range = (q > lo + 0.01) & (q < hi - 0.01)
    assert np.all(np.abs(np.minimum(np.abs(got - q), np.abs(got + q - 1))[in_range]) < 2e-3)


def test_fine_tune_edge_cases(ctx):
    lib = pyoracle.load()
    z = np.zeros((2, 128))
    z[1, 5] = 1.0             # flat spectrum: every candidate has the same power
    got = ctx.fine_tune(z, 0.1, 0.2)
    out = ctypes.c_double()
    for j in range(2):
        r = np.ascontiguousarray(z[j])
        lib.orc_fine_tune(128, r.ctypes.data_as(ctypes.c_void_p), 0.1, 0.2, ctypes.byref(out))
        assert got[j] == out.value
    with pytest.raises(st.STError):
        ctx.fine_tune(z, 0.0, 5.0)   # more than 4096 steps


def test_tunes_of_tracked_particles(ctx):
    """Turn-by-turn x of tracked particles (Track_tbt) -> horizontal tune,
    as Track_tbt_FMA does (cpp:21218-21227)."""
    from test_gpu_parity import _fodo_ring
    ln = _fodo_ring(ncell=8)
    g = W.proton_globals(100000.0)
    tab = st.Table(ctx, g, ln.records())
    x = np.zeros((16, 6))
    x[:, 0] = np.linspace(1e-4, 2e-3, 16)
    x[:, 2] = 5e-5
    _, tbt, s, _, _ = tab.track_tbt(x.copy(), 512)
    assert s.all()
    xs = np.ascontiguousarray(tbt[:, :, 0])
    got = ctx.fine_tune(xs, 0.001, 0.499)
    lib = pyoracle.load()
    out = ctypes.c_double()
    for j in range(16):
        r = np.ascontiguousarray(xs[j])
        lib.orc_fine_tune(512, r.ctypes.data_as(ctypes.c_void_p), 0.001, 0.499, ctypes.byref(out))
        assert got[j] == out.value
    assert np.ptp(got) < 0.05 and 0.001 < got.min()


def test_windowed_fine_tune_matches_oracle_on_windowed_samples(ctx):
    """st_fine_tune_window_device on x equals the oracle's FineTuneFinder on
    w * x (the product formed in numpy, the same IEEE multiply)."""
    import torch
    from simtrack_amd.surrogate import hann
    lib = pyoracle.load()
    n, nsig, lo, hi = 512, 24, 0.05, 0.45
    x, px, q = signals(nsig, n, 11)
    w = hann(n)
    dev = torch.device("cuda:0")
    d_x = torch.tensor(x, dtype=torch.float64, device=dev)
    d_w = torch.tensor(w, dtype=torch.float64, device=dev)
    d_t = torch.zeros(nsig, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx.fine_tune_device(d_x.data_ptr(), n, 1, n, nsig, lo, hi, d_t.data_ptr(), d_window=d_w.data_ptr())
    ctx.sync()
    got = d_t.cpu().numpy()
    for j in range(nsig):
        xw = np.ascontiguousarray(w * x[j])
        ref = ctypes.c_double()
        lib.orc_fine_tune(n, xw.ctypes.data_as(ctypes.c_void_p), lo, hi, ctypes.byref(ref))
        assert got[j] == ref.value, (j, got[j], ref.value)
