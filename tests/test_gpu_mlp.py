"""The surrogate network's hand-written BF16-MFMA kernels against the CPU
reference (tests/mlp_reference.py): forward probabilities, loss and every
weight and bias gradient; and training reduces the loss on a separable set."""
import numpy as np
import pytest
import torch

import simtrack_amd as st
from mlp_reference import forward, loss_and_grads

pytestmark = pytest.mark.gpu


def data(n, nin, seed):
    r = np.random.default_rng(seed)
    x = r.normal(0, 1, (n, nin)).astype(np.float32)
    w = r.normal(0, 1, nin)
    y = (x @ w + 0.3 * r.normal(size=n) > 0).astype(np.float32)
    return x, y


@pytest.mark.parametrize("n,nin,nh", [(1000, 12, 256), (4096, 12, 256), (77, 5, 40)])
def test_forward_and_gradients_match_cpu(ctx, n, nin, nh):
    m = st.Mlp(ctx, nin, nh, seed=3)
    p = m.get_params()
    layers = m.unpack(p)
    x, y = data(n, nin, seed=n)
    dx = torch.tensor(x, device="cuda")
    dy = torch.tensor(y, device="cuda")
    prob = torch.zeros(n, device="cuda")
    m.forward(dx.data_ptr(), n, prob.data_ptr())
    ctx.sync()
    z, _ = forward(layers, x)
    pref = 1 / (1 + np.exp(-z))
    assert np.max(np.abs(prob.cpu().numpy() - pref)) < 2e-3
    g, loss = m.gradients(dx.data_ptr(), dy.data_ptr(), n)
    lref, gref = loss_and_grads(layers, x, y.astype(np.float64))
    assert abs(loss - lref) < 1e-3 * max(1.0, abs(lref))
    got = m.unpack(g)
    for li, ((gw, gb), (rw, rb)) in enumerate(zip(got, gref)):
        ew = np.linalg.norm(gw - rw) / max(np.linalg.norm(rw), 1e-30)
        eb = np.linalg.norm(gb - rb) / max(np.linalg.norm(rb), 1e-30)
        assert ew < 2e-2 and eb < 2e-2, (li, ew, eb)


def test_training_reduces_loss(ctx):
    m = st.Mlp(ctx, 12, 256, seed=5)
    x, y = data(8192, 12, seed=9)
    dx = torch.tensor(x, device="cuda")
    dy = torch.tensor(y, device="cuda")
    losses = []
    for it in range(200):
        s = (it * 1024) % 8192
        losses.append(m.train_step(dx[s:s + 1024].data_ptr(), dy[s:s + 1024].data_ptr(), 1024, 1e-3))
    prob = torch.zeros(8192, device="cuda")
    m.forward(dx.data_ptr(), 8192, prob.data_ptr())
    ctx.sync()                    # the network runs on the context's stream, not torch's
    acc = ((prob.cpu().numpy() > 0.5) == (y > 0.5)).mean()
    assert np.mean(losses[-20:]) < 0.5 * np.mean(losses[:5]) and acc > 0.9
