"""st_track_device_noalloc (the allocation-free variant of st_track_device)
against st_track_device itself, through the C ABI, bit for bit: two calls in
a row on the same GPU scratch, the second with fewer particles, each with
losses so that the survivor compaction between turn blocks runs; and the
size limits rejected on the host."""
import numpy as np
import pytest

import simtrack_amd as st
from simtrack_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run(tab, x, nturn, flags, noalloc):
    import torch
    dev = torch.device("cuda:0")
    d_x = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    d_state = torch.ones(len(x), dtype=torch.int32, device=dev)
    d_lost = torch.zeros((len(x), 2), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    f = tab.track_device_noalloc if noalloc else tab.track_device
    f(d_x.data_ptr(), len(x), 0, nturn, d_state.data_ptr(), d_lost.data_ptr(), flags)
    tab.ctx.sync()
    return d_x.cpu().numpy(), d_state.cpu().numpy(), d_lost.cpu().numpy()


def _same(a, b):
    for u, v in zip(a, b):
        if u.dtype == np.float64:
            eq = (u.view(np.int64) == v.view(np.int64)) | (np.isnan(u) & np.isnan(v))
            assert eq.all()
        else:
            assert np.array_equal(u, v)


@pytest.mark.parametrize("flags", [0, st.TRACK_RADIAL, st.TRACK_SPIN], ids=["track", "radial", "spin"])
def test_noalloc_equals_track_device(ctx, flags):
    gp, ln, x = W.config0()
    if flags & st.TRACK_SPIN:
        xs = np.zeros((len(x), 9))
        xs[:, :6] = x
        xs[:, 7] = 1.0
        x = xs
    tab = st.Table(ctx, gp, ln.records())
    nturn = 450          # three turn blocks of this 998-element ring: two compactions
    a = _run(tab, x, nturn, flags, noalloc=False)
    b = _run(tab, x, nturn, flags, noalloc=True)
    _same(a, b)
    assert 0 < a[1].sum() < len(x)
    # a second call, fewer particles (another slice of the grid), the same scratch
    x2 = np.ascontiguousarray(x[300:700])
    a2 = _run(tab, x2, nturn, flags, noalloc=False)
    b2 = _run(tab, x2, nturn, flags, noalloc=True)
    _same(a2, b2)
    assert 0 < a2[1].sum() < len(x2)
    tab.close()


def test_noalloc_rejects_oversize(ctx):
    gp, ln, _ = W.config0()
    tab = st.Table(ctx, gp, ln.records())
    with pytest.raises(st.STError, match="scratch maximum"):
        tab.track_device_noalloc(16, st.NOALLOC_MAX_PARTICLES + 1, 0, 1, 16, 16, 0)
    tab.close()
