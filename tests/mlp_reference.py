"""CPU reference of the surrogate network for the GPU gradient checks:
float64 arithmetic on the values the GPU rounds to BF16 (inputs, weights,
hidden activations, back-propagated gradients), so that the only
differences left are the GPU's FP32 accumulation order."""
import numpy as np


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(layers, x):
    """layers: [(W, b)] FP32; returns (logits, [inputs of each layer])."""
    h = bf16(x)
    ins = []
    for li, (W, b) in enumerate(layers):
        ins.append(h)
        z = h @ bf16(W).T + b.astype(np.float64)
        if li < len(layers) - 1:
            h = bf16(np.maximum(z, 0.0).astype(np.float32))
        else:
            return z[:, 0], ins


def loss_and_grads(layers, x, y):
    z, ins = forward(layers, x)
    n = len(y)
    p = 1.0 / (1.0 + np.exp(-z))
    loss = np.mean(np.maximum(z, 0) - y * z + np.log1p(np.exp(-np.abs(z))))
    dy = bf16(((p - y) / n).astype(np.float32))[:, None]
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        xin = ins[li]
        gW = dy.T @ xin
        gb = dy.sum(axis=0)
        grads.append((gW, gb))
        if li > 0:
            dh = dy @ bf16(W)
            dy = bf16(np.where(xin > 0, dh, 0.0).astype(np.float32))
    return loss, grads[::-1]
