"""Surrogate-assisted dynamic-aperture scan (config 3).

The tracker itself produces the training data and the features:
  1. every grid particle of every lattice is tracked for 2 x 512 turns with
     turn-by-turn capture (st_batch_track_tbt_device);
  2. FineTuneFinder on Hann-windowed samples (st_fine_tune_window_device)
     gives the horizontal and vertical tune of each 512-turn window; with
     the initial actions and angles they form 12 FP32 features
     (st_surrogate_features);
  3. the MLP (st_mlp_*, BF16 MFMA) predicts survival to the long horizon;
  4. the particles that can decide the DA boundary fall back to full FP64
     tracking from where step 1 left them, to the long horizon.
A particle lost within the first 1,024 turns is lost.  The DA boundary of
one angle of one lattice is the amplitude of the first particle (by
increasing amplitude) that does not survive, as Track_DA_Uniform's first
loss (simtrack.cpp:21574), so only the particles below the first early loss
can move it, and of those only the ones the network does not call safe.
The boundary-aware fallback (the default) tracks, on each line of the grid,
every particle below the first early loss whose predicted survival
probability is at most p_safe (0.999 by default), in one batched pass;
untracked particles take the network's verdict.  A lost particle that the
network calls safe is the only way the boundary can come out wrong, so
p_safe trades fallback particles for boundary errors.  The band fallback
tracks every particle with probability between 0.1 and 0.9.  Full tracking
of every particle to the long horizon is the reference, and judges the
surrogate's errors.
"""
import time

import numpy as np

from ._capi import Batch, Mlp

WINDOW = 512
TUNE_RANGE = (0.01, 0.49)


def hann(n):
    """Hann weights 1 - cos(2 pi i / n): the tune search on windowed samples
    (NAFF's choice), whose leakage error falls as 1/n^3 instead of 1/n."""
    return 1.0 - np.cos(2 * np.pi * np.arange(n) / n)
P_LOW, P_HIGH = 0.1, 0.9
P_SAFE = 0.999


def _torch():
    import torch
    return torch


def features(ctx, batch, x, offsets, beta_x, alpha_x, beta_y, alpha_y, cfg=None, window="hann"):
    """Track 2*WINDOW turns with turn-by-turn capture and build the features
    (tunes of Hann-windowed samples unless window is None).
    Returns device tensors (x after 2*WINDOW turns, state, lost, features)."""
    torch = _torch()
    dev = torch.device("cuda:0")
    n = len(x)
    d_x0 = torch.tensor(x, dtype=torch.float64, device=dev).contiguous()
    d_x = d_x0.clone()
    d_s = torch.ones(n, dtype=torch.int32, device=dev)
    d_l = torch.zeros(2 * n, dtype=torch.int32, device=dev)
    nt = 2 * WINDOW
    d_tbt = torch.zeros((n, nt, 6), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    batch.track_tbt_device(d_x.data_ptr(), offsets, 0, nt, d_tbt.data_ptr(), d_s.data_ptr(), d_l.data_ptr(), 0, cfg)
    d_q = torch.zeros((4, n), dtype=torch.float64, device=dev)
    base = d_tbt.data_ptr()
    d_w = torch.tensor(hann(WINDOW), dtype=torch.float64, device=dev) if window == "hann" else None
    torch.cuda.synchronize()
    for k, (col, win) in enumerate(((0, 0), (2, 0), (0, 1), (2, 1))):
        ctx.fine_tune_device(base + 8 * (win * WINDOW * 6 + col), nt * 6, 6, WINDOW, n, TUNE_RANGE[0], TUNE_RANGE[1],
                             d_q[k].data_ptr(), d_window=d_w.data_ptr() if d_w is not None else None)
    ctx.sync()
    d_tunes = d_q.t().contiguous()
    d_f = torch.zeros((n, 12), dtype=torch.float32, device=dev)
    ctx.surrogate_features(d_x0.data_ptr(), 6, d_tunes.data_ptr(), n, beta_x, alpha_x, beta_y, alpha_y, d_f.data_ptr())
    ctx.sync()
    del d_tbt
    return d_x, d_s, d_l, d_f


CHUNK = 2000


def _chunked(run, turn0, turn1, progress, what):
    """run(t, nt) over [turn0, turn1) in chunks (turn ranges compose exactly),
    reporting progress after each."""
    t = turn0
    while t < turn1:
        nt = min(CHUNK, turn1 - t)
        run(t, nt)
        t += nt
        if progress:
            progress(f"{what}: turn {t}/{turn1}")


def track_to(ctx, tabs, d_x, d_s, d_l, offsets, sel, turn0, turn1, cfg=None, progress=None):
    """Continue the particles of boolean device mask `sel` from turn0 to turn1
    (batched over lattices); results written back into d_x, d_s, d_l."""
    torch = _torch()
    idx = torch.nonzero(sel).flatten()
    if len(idx) == 0 or turn1 <= turn0:
        return 0
    seed_of = torch.bucketize(idx, torch.as_tensor(offsets[1:], device=idx.device), right=True)
    counts = torch.bincount(seed_of, minlength=len(tabs)).cpu().numpy()
    keep = [k for k in range(len(tabs)) if counts[k] > 0]
    sub_offs = np.concatenate([[0], np.cumsum(counts[keep])]).astype(np.int32)
    b = Batch(ctx, [tabs[k] for k in keep])
    xs = d_x[idx].contiguous()
    ss = d_s[idx].contiguous()
    ls = d_l.view(-1, 2)[idx].contiguous()
    torch.cuda.synchronize()
    _chunked(lambda t, nt: (b.track_device(xs.data_ptr(), sub_offs, t, nt, ss.data_ptr(), ls.data_ptr(), 0, cfg),
                            ctx.sync()), turn0, turn1, progress, f"tracking {len(idx)} particles")
    d_x[idx] = xs
    d_s[idx] = ss
    d_l.view(-1, 2)[idx] = ls
    torch.cuda.synchronize()
    return len(idx)


class Standardizer:
    def __init__(self, f=None, mean=None, std=None):
        """From features f (per-column mean and standard deviation), or from
        a saved mean and std."""
        if f is not None:
            mean, std = f.mean(dim=0), f.std(dim=0).clamp_min(1e-6)
        self.mean, self.std = mean, std

    def __call__(self, f):
        return ((f - self.mean) / self.std).contiguous()


def train(ctx, feats, labels, epochs=30, batch=2048, lr=1e-3, hidden=256, seed=1, log=None):
    """Train the MLP on (features, labels) device tensors; returns (mlp, standardizer)."""
    torch = _torch()
    std = Standardizer(feats)
    fz = std(feats)
    y = labels.float().contiguous()
    m = Mlp(ctx, 12, hidden, seed=seed)
    g = torch.Generator(device="cpu").manual_seed(seed)
    n = len(fz)
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(fz.device)
        fsh, ysh = fz[perm].contiguous(), y[perm].contiguous()
        torch.cuda.synchronize()
        tot = 0.
        nb = 0
        for s in range(0, n - batch + 1, batch):
            tot += m.train_step(fsh[s:s + batch].data_ptr(), ysh[s:s + batch].data_ptr(), batch,
                                lr * (0.5 ** (ep // 10)))
            nb += 1
        if log is not None:
            log.append(tot / max(nb, 1))
    ctx.sync()
    return m, std


def predict(ctx, mlp, std, feats):
    torch = _torch()
    fz = std(feats)
    p = torch.zeros(len(fz), dtype=torch.float32, device=fz.device)
    torch.cuda.synchronize()
    mlp.forward(fz.data_ptr(), len(fz), p.data_ptr())
    ctx.sync()
    return p


def da_boundary(survive, amps, n_angles):
    """Per-angle DA boundary: amplitude of the first non-surviving particle
    (grid rows are angle-major, amplitudes increasing); amps[-1] + step if
    all survive."""
    s = np.asarray(survive).reshape(n_angles, -1)
    out = np.empty(n_angles)
    for a in range(n_angles):
        lost = np.nonzero(s[a] == 0)[0]
        out[a] = amps[lost[0]] if len(lost) else amps[-1] + (amps[1] - amps[0])
    return out


def boundary_fallback(alive, prob, offsets, n_angles, p_safe=P_SAFE):
    """The particles that can decide the DA boundary: on each line (lattice,
    angle; rows angle-major, amplitude increasing) those below the first
    particle lost within the feature turns, except the ones the network
    calls safe (probability above p_safe).  alive, prob: per particle (host
    arrays).  Returns a boolean mask."""
    alive = np.asarray(alive).astype(bool)
    prob = np.asarray(prob)
    sel = np.zeros(len(alive), dtype=bool)
    for k in range(len(offsets) - 1):
        a = alive[offsets[k]:offsets[k + 1]].reshape(n_angles, -1)
        p = prob[offsets[k]:offsets[k + 1]].reshape(n_angles, -1)
        m = a.shape[1]
        first_lost = np.where((~a).any(axis=1), (~a).argmax(axis=1), m)
        below = np.arange(m)[None, :] < first_lost[:, None]
        sel[offsets[k]:offsets[k + 1]] = (below & (p <= p_safe)).reshape(-1)
    return sel


def scan_start(ctx, tabs, x, offsets, mlp, std, twiss, cfg=None, progress=None, mode="boundary", n_angles=8,
               p_safe=P_SAFE):
    """First part of the scan: the feature turns of every particle, the
    network's predictions and the choice of fallback particles (mode
    "boundary" or "band", see the module's description; p_safe: the
    boundary mode's trust threshold).  Returns the scan state, which
    scan_track continues and scan_finish reads."""
    torch = _torch()
    t0 = time.perf_counter()
    batch = Batch(ctx, tabs)
    d_x, d_s, d_l, d_f = features(ctx, batch, x, offsets, *twiss, cfg=cfg)
    t_feat = time.perf_counter() - t0
    p = predict(ctx, mlp, std, d_f)
    alive = d_s != 0
    if mode == "boundary":
        sel = boundary_fallback(alive.cpu().numpy(), p.cpu().numpy(), offsets, n_angles, p_safe)
        fallback = torch.as_tensor(sel, device=alive.device)
        decided_ok = alive & ~fallback & (p >= 0.5)
    elif mode == "band":
        fallback = alive & (p >= P_LOW) & (p <= P_HIGH)
        decided_ok = alive & (p > P_HIGH)
    else:
        raise ValueError(f"unknown fallback mode {mode!r}")
    torch.cuda.synchronize()
    if progress:
        progress(f"surrogate: features and predictions in {time.perf_counter() - t0:.1f} s")
    return dict(d_x=d_x, d_s=d_s, d_l=d_l, p=p, alive=alive, fallback=fallback, decided_ok=decided_ok,
                turn=2 * WINDOW, elapsed=time.perf_counter() - t0, t_features=t_feat,
                n_fallback=int(fallback.sum().item()))


def scan_track(ctx, tabs, st, offsets, turn1, cfg=None, progress=None):
    """Continue the scan's fallback particles from st["turn"] to turn1; the
    time is added to st["elapsed"]."""
    t0 = time.perf_counter()
    track_to(ctx, tabs, st["d_x"], st["d_s"], st["d_l"], offsets, st["fallback"], st["turn"], turn1, cfg, progress)
    st["turn"] = max(st["turn"], turn1)
    st["elapsed"] += time.perf_counter() - t0


def scan_finish(st):
    """(survive, info) of a scan whose fallback particles reached the horizon."""
    torch = _torch()
    surv = torch.where(st["fallback"], st["d_s"] != 0, st["decided_ok"])
    surv = surv.cpu().numpy().astype(np.int32)
    return surv, dict(wall=st["elapsed"], t_features=st["t_features"], n_fallback=st["n_fallback"],
                      n_alive_1k=int(st["alive"].sum().item()),
                      plain=st["alive"].cpu().numpy().astype(np.int32), prob=st["p"].cpu().numpy(),
                      fallback=st["fallback"].cpu().numpy().astype(bool))


def scan(ctx, tabs, x, offsets, n_long, mlp, std, twiss, cfg=None, progress=None, mode="boundary", n_angles=8,
         p_safe=P_SAFE):
    """The surrogate-plus-fallback scan to n_long turns in one go (scan_start,
    scan_track, scan_finish).  Returns (survive, info); info["wall"] is the
    scan's wall time."""
    st = scan_start(ctx, tabs, x, offsets, mlp, std, twiss, cfg, progress, mode, n_angles, p_safe)
    scan_track(ctx, tabs, st, offsets, n_long, cfg, progress)
    return scan_finish(st)


def save_state(st, path):
    """Write a scan or full-tracking state (a dict of device tensors and
    numbers) to an .npz, keeping x only for particles still being tracked,
    so that a long horizon can be split over several processes."""
    torch = _torch()
    live = st["d_s"] != 0
    if "fallback" in st:
        live = live & st["fallback"]
    idx = torch.nonzero(live).flatten()
    out = {"_idx": idx.cpu().numpy(), "_x": st["d_x"][idx].cpu().numpy(), "_n": len(st["d_s"])}
    for k, v in st.items():
        if k == "d_x":
            continue
        out[k] = v.cpu().numpy() if torch.is_tensor(v) else np.asarray(v)
    np.savez_compressed(path, **out)


def load_state(path, device="cuda:0"):
    """The state save_state wrote, back on the device."""
    torch = _torch()
    z = np.load(path)
    st = {}
    for k in z.files:
        if k.startswith("_"):
            continue
        v = z[k]
        st[k] = torch.as_tensor(v, device=device) if v.ndim else v.item()
    d_x = torch.zeros((int(z["_n"]), 6), dtype=torch.float64, device=device)
    d_x[torch.as_tensor(z["_idx"], device=device)] = torch.as_tensor(z["_x"], device=device)
    st["d_x"] = d_x
    return st


def full_start(x, device="cuda:0"):
    """Full-tracking state of the particles x at turn 0."""
    torch = _torch()
    d_x = torch.tensor(x, dtype=torch.float64, device=device).contiguous()
    return dict(d_x=d_x, d_s=torch.ones(len(x), dtype=torch.int32, device=device),
                d_l=torch.zeros(2 * len(x), dtype=torch.int32, device=device), turn=0, elapsed=0.)


def full_track(ctx, tabs, st, offsets, turn1, cfg=None, progress=None):
    """Continue full tracking of every particle from st["turn"] to turn1; the
    time is added to st["elapsed"]."""
    torch = _torch()
    t0 = time.perf_counter()
    batch = Batch(ctx, tabs)
    d_x, d_s, d_l = st["d_x"], st["d_s"], st["d_l"]
    torch.cuda.synchronize()
    _chunked(lambda t, nt: (batch.track_device(d_x.data_ptr(), offsets, t, nt, d_s.data_ptr(), d_l.data_ptr(), 0, cfg),
                            ctx.sync()), st["turn"], turn1, progress, "full tracking")
    st["turn"] = max(st["turn"], turn1)
    st["elapsed"] += time.perf_counter() - t0


def full(ctx, tabs, x, offsets, n_long, cfg=None, progress=None):
    """Reference: every particle tracked to n_long.  Returns (survive, wall,
    d_x, d_lost)."""
    t0 = time.perf_counter()
    st = full_start(x)
    full_track(ctx, tabs, st, offsets, n_long, cfg, progress)
    s = st["d_s"].cpu().numpy().astype(np.int32)
    return s, time.perf_counter() - t0, st["d_x"], st["d_l"]
