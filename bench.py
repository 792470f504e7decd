#!/usr/bin/env python3
"""Benchmark: FP64 particle-turns/s of the tracking hot path on one MI355X.

Workload (config 1): the 998-element DA ring (48 FODO cells, 2 low-beta IRs,
thick quadrupoles/sextupoles with b3..b10 errors, one 28 MHz RF cavity) with
6-D weak-strong beam-beam at both IPs (round strong beam, sigma = 0.15 mm,
sigma_z = 0.3 m, 5 slices, xi = 0.012 per IP), Thomas-BMT spin tracking on,
65,536 protons at 100 GeV Gaussian-sampled to 6 sigma, synthetic seeded data.

A step tracks every surviving particle through --turns-per-step turns with
the state resident in HBM (st_track_device).  value = particle-turns
actually tracked in the timed steps / wall time.  The CPU baseline is the
oracle (the reference restatement) on a bounded sample of the same
workload, timed on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]


def cpu_baseline(gp, records, x, spin, sample_particles, sample_turns, threads):
    import pyoracle
    orc = pyoracle.Oracle(gp, records, nthreads=threads)
    xs = x[:sample_particles].copy()
    t0 = time.perf_counter()
    _, stable, lt, _ = orc.track(xs, sample_turns, spin=spin)
    dt = time.perf_counter() - t0
    pturns = _tracked_turns(stable, lt, 0, sample_turns, prelost=None)
    return {"value": pturns / dt, "unit": "particle-turns/s", "cores": threads, "kind": "port",
            "sample": f"{sample_particles} particles x {sample_turns} turns of the same workload "
                      f"(oracle restatement, OpenMP over particles), {dt:.1f} s"}


def _tracked_turns(stable, lost_turn, t_begin, t_end, prelost):
    """Complete particle-turns tracked in [t_begin, t_end) from the loss records."""
    import numpy as np
    stable = np.asarray(stable)
    lt = np.asarray(lost_turn, dtype=np.int64)
    turns = np.where(stable != 0, t_end - t_begin, np.clip(lt - t_begin, 0, t_end - t_begin))
    if prelost is not None:
        turns = np.where(prelost, 0, turns)
    return int(turns.sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--turns-per-step", type=int, default=20)
    ap.add_argument("--particles", type=int, default=65536)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=str, default="1024x64", help="particles x turns for the CPU leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the tracking "
              f"path shards by particles or seeds with no collective, see the sharding notes", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import simtrack_amd as st
    from simtrack_amd import flops
    from simtrack_amd import workloads as W

    gp, ln, x = W.config1(n_particles=args.particles)
    rec = ln.records()
    spin = True
    ncol = 9
    f_turn = flops.flops_per_turn(rec, gp, spin=spin)

    ctx = st.Context(0)
    tab = st.Table(ctx, gp, rec)
    dev = torch.device("cuda:0")
    d_x = torch.tensor(x, dtype=torch.float64, device=dev).contiguous()
    d_state = torch.ones(args.particles, dtype=torch.int32, device=dev)
    d_lost = torch.zeros(2 * args.particles, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    T = args.turns_per_step
    turn = 0
    for _ in range(args.warmup):
        tab.track_device(d_x.data_ptr(), args.particles, turn, T, d_state.data_ptr(), d_lost.data_ptr(),
                         st.TRACK_SPIN)
        turn += T
    ctx.sync()
    torch.cuda.synchronize()
    state0 = d_state.cpu().numpy().copy()
    t_begin = turn

    ctx.set_timing(True)
    ctx.kernel_stats()                      # clear counters
    if torch.distributed.is_available() and torch.distributed.is_initialized():
        torch.distributed.barrier()
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tab.track_device(d_x.data_ptr(), args.particles, turn, T, d_state.data_ptr(), d_lost.data_ptr(),
                         st.TRACK_SPIN)
        turn += T
    ctx.sync()
    torch.cuda.synchronize()
    if torch.distributed.is_available() and torch.distributed.is_initialized():
        torch.distributed.barrier()
    wall = time.perf_counter() - t0
    k_sec, k_launches, k_work = ctx.kernel_stats()
    ctx.set_timing(False)

    state = d_state.cpu().numpy()
    lost = d_lost.cpu().numpy().reshape(-1, 2)
    pturns = _tracked_turns(state, lost[:, 0], t_begin, turn, prelost=(state0 == 0))
    value = pturns / wall
    kernel_rate = pturns / k_sec if k_sec > 0 else float("nan")
    achieved = f_turn * kernel_rate
    peak = flops.PEAK_FP64_VECTOR

    baseline = None
    if not args.no_cpu_baseline:
        sp, stn = (int(v) for v in args.cpu_sample.split("x"))
        threads = min(16, os.cpu_count() or 1)
        baseline = cpu_baseline(gp, rec, x, spin, sp, stn, threads)

    out = {
        "metric": "FP64 particle-turns/s",
        "value": value,
        "unit": "particle-turns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lattice errors and Gaussian particles)",
        "config": {"workload": "config1: 998-element DA ring + 6-D beam-beam at 2 IPs + spin, 100 GeV p",
                   "particles": args.particles, "turns_per_step": T, "elements": int(len(rec)),
                   "spin": True, "beam_beam_slices": 5},
        # compute-bound: the roof is MI355X's FP64 peak (78.6 TFLOP/s, the same figure for
        # the vector and matrix pipes); the kernel runs on the FP64 vector pipe
        "roofline": {"bound": "mfma", "pipe": "fp64 VALU", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "f_turn_flops": f_turn, "kernel": "k_track<spin>",
                     "avg_launch_ms": 1e3 * k_sec / max(k_launches, 1), "launches": k_launches},
        "cpu_baseline": baseline,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
