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
                      f"(oracle restatement, OpenMP over particles), {dt:.1f} s",
            "host": host_info()}


def host_threads():
    """Threads for the CPU baseline: the cores this process may run on, at
    most 16 (a one-GPU box's CPU share, whatever nproc shows there)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_info():
    """nproc, the affinity mask's size and the CPU model of this host."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = None
    return {"nproc": os.cpu_count(), "affinity": aff, "model": model}


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
    ap.add_argument("--perf-mode", action="store_true",
                    help="config1/config4: the perf-mode kernels (FMA contraction, division by reciprocal "
                         "multiplication; results differ from the reference's operations by rounding)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device (a kernel of known traffic for "
                         "calibrating the HBM byte counters)")
    ap.add_argument("--cpu-sample", type=str, default=None,
                    help="particles x turns for the CPU leg (default 1024x64; config4: 2048x24)")
    ap.add_argument("--workload", choices=("config1", "config2", "config3", "config4", "rhic", "strong_strong"),
                    default="config1",
                    help="config1: the headline workload; config2: 64 uneven lattices in one persistent launch; "
                         "config3: surrogate DA scan; config4: radiating electron ring, 1,048,576 particles; "
                         "rhic: the reference's RHIC example lattice; strong_strong: soft-Gaussian 6-D "
                         "collisions of two 1,048,576-macroparticle beams")
    ap.add_argument("--slices", type=int, default=8, help="strong_strong: z slices per beam")
    ap.add_argument("--pic", action="store_true", help="strong_strong: the particle-in-cell solver "
                    "(BeamBeam6D_PIC, z-sorted slices) instead of the soft-Gaussian one")
    ap.add_argument("--seeds", type=int, default=64, help="config2: lattices")
    ap.add_argument("--per-seed", type=int, default=4096, help="config2/3: particles per lattice")
    ap.add_argument("--train-seeds", type=int, default=16, help="config3: lattices for the training set")
    ap.add_argument("--n-long", type=int, default=30000, help="config3: survival horizon in turns")
    ap.add_argument("--epochs", type=int, default=60, help="config3: training epochs")
    ap.add_argument("--fallback", choices=("boundary", "band"), default="boundary",
                    help="config3: which particles fall back to full tracking")
    ap.add_argument("--p-safe", default="0.999",
                    help="config3, boundary fallback: trust threshold(s) of the network's 'safe' verdict; a comma "
                         "list runs one scan per value (the first is the reported one)")
    ap.add_argument("--stage", choices=("all", "train", "full", "scan"), default="all",
                    help="config3: run one part of the comparison and save its state, so that a long horizon "
                         "can be split over several processes (train -> --state-out; full and scan: from "
                         "--state-in, if given, to --turns, then --state-out; scan needs --mlp, and --ref "
                         "once it reaches --n-long)")
    ap.add_argument("--turns", type=int, default=None, help="config3 --stage full/scan: stop at this turn")
    ap.add_argument("--state-in", default=None, help="config3 --stage full/scan: state to continue")
    ap.add_argument("--state-out", default=None, help="config3 --stage: where to write the state")
    ap.add_argument("--mlp", default=None, help="config3 --stage scan: the trained network (--stage train's output)")
    ap.add_argument("--ref", default=None, help="config3 --stage scan: full tracking's final state")
    ap.add_argument("--dump", default=None,
                    help="config3: write per-particle outcomes (full tracking, network, plain method) to this .npz")
    ap.add_argument("--orbit", action="store_true", help="config4: orbit tracking with quantum excitation "
                    "instead of spin tracking")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the tracking path shards by "
              f"particles or lattice seeds with no collective", file=sys.stderr)
        sys.exit(2)

    if args.cpu_sample is None:
        args.cpu_sample = {"config4": "2048x24", "config2": "2048x512"}.get(args.workload, "1024x64")
    if args.workload == "config2":
        return bench_config2(args)
    if args.workload == "config3":
        return bench_config3(args)

    if args.workload == "config4":
        return bench_config4(args)
    if args.workload == "rhic":
        return bench_rhic(args)
    if args.workload == "strong_strong":
        return bench_strong_strong(args)

    from simtrack_amd import workloads as W
    gp, ln, x = W.config1(n_particles=args.particles)
    # HBM bytes per launch of this shape, measured with rocprofv3 FETCH_SIZE and
    # WRITE_SIZE against a copy of known size (profiles/k_track_spin_config1_traffic.txt);
    # the state is read and written once per launch, independent of the turn count
    traffic = 13.87e6 if args.particles == 65536 else None
    config = {"workload": "config1: 998-element DA ring + 6-D beam-beam at 2 IPs + spin, 100 GeV p",
              "particles": args.particles, "turns_per_step": args.turns_per_step, "elements": int(len(ln.Cell)),
              "spin": True, "beam_beam_slices": 5}
    bench_tracking(args, gp, ln.records(), x, True, config, traffic,
                   "synthetic (seeded lattice errors and Gaussian particles)")


def bench_tracking(args, gp, rec, x, spin, config, traffic, data):
    """Time args.steps steps of args.turns_per_step turns of x (resident in HBM) and print the JSON line."""
    import numpy as np
    import torch

    import simtrack_amd as st
    from simtrack_amd import flops

    npart = len(x)
    flags = st.TRACK_SPIN if spin else 0
    if getattr(args, "perf_mode", False):
        flags |= st.TRACK_PERF
        config = dict(config, arithmetic="perf mode: FMA contraction and reciprocal division (not bit-equal)")
    else:
        config = dict(config, arithmetic="bit-exact: the reference's operations, no contraction")
    f_turn = flops.flops_per_turn(rec, gp, spin=spin)

    ctx = st.Context(0)
    tab = st.Table(ctx, gp, rec)
    dev = torch.device("cuda:0")
    d_x = torch.tensor(x, dtype=torch.float64, device=dev).contiguous()
    d_state = torch.ones(npart, dtype=torch.int32, device=dev)
    d_lost = torch.zeros(2 * npart, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    T = args.turns_per_step
    turn = 0
    for _ in range(args.warmup):
        tab.track_device(d_x.data_ptr(), npart, turn, T, d_state.data_ptr(), d_lost.data_ptr(), flags)
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
        tab.track_device(d_x.data_ptr(), npart, turn, T, d_state.data_ptr(), d_lost.data_ptr(), flags)
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

    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    baseline = None
    if not args.no_cpu_baseline:
        sp, stn = (int(v) for v in args.cpu_sample.split("x"))
        threads = host_threads()
        baseline = cpu_baseline(gp, rec, x, spin, sp, stn, threads)

    ncol = 9 if spin else 6
    kname = ("k_track<spin>" if spin else "k_track") + (" (perf mode)" if flags & st.TRACK_PERF else "")
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
        "data": data,
        "config": config,
        # compute-bound on the FP64 vector ALU: the roof is MI355X's FP64 vector peak, 78.6 TFLOP/s
        "roofline": {"bound": "fp64_valu", "pipe": "fp64 VALU", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                     "algorithmic_bytes": 2 * ncol * 8 * npart,
                     "f_turn_flops": f_turn, "kernel": kname,
                     "avg_launch_ms": 1e3 * k_sec / max(k_launches, 1), "launches": k_launches},
        "cpu_baseline": baseline,
    }
    print(json.dumps(out))


def _ss_beams(n, seed):
    """Two 100 GeV proton beams at an IP with beta* = 0.5 m: sigma_x = 0.1 mm,
    sigma_y = 0.06 mm (elliptic, so BB4D takes the complex-error-function
    branch), sigma_z = 5 cm, sigma_delta = 1e-4, opposite 5 um offsets."""
    import numpy as np
    r = np.random.default_rng(seed)
    beams = []
    for k, off in enumerate((5e-6, -5e-6)):
        x = np.empty((n, 6))
        x[:, 0] = r.normal(off, 1e-4, n)
        x[:, 1] = r.normal(0, 1e-4 / 0.5, n)
        x[:, 2] = r.normal(0, 6e-5, n)
        x[:, 3] = r.normal(0, 6e-5 / 0.5, n)
        x[:, 4] = r.normal(0, 0.05, n)
        x[:, 5] = r.normal(0, 1e-4, n)
        beams.append(x)
    return beams


def bench_strong_strong(args):
    """Strong-strong beam-beam (BeamBeam6D_SoftGaussian with Gaussian slice
    boundaries, simtrack.cpp:23134): a step is one collision of the two
    beams, every slice pair in order.  value = macroparticle kicks per
    second (rows of both slices, summed over slice pairs)."""
    import numpy as np
    import torch

    import simtrack_amd as st
    from simtrack_amd import flops
    from simtrack_amd import strong_strong as SS
    from simtrack_amd import workloads as W

    n = args.particles if args.particles != 65536 else 1048576
    ns = args.slices
    gp = W.proton_globals(100000.0)
    b1, b2 = _ss_beams(n, 2031)
    zb = SS.Zboundary_slices_Gaussian(ns, 0.05)
    Np = 1.5e11
    ctx = st.Context(0)
    dev = torch.device("cuda:0")
    d1 = torch.tensor(b1, device=dev)
    d2 = torch.tensor(b2, device=dev)
    torch.cuda.synchronize()

    def collide():
        if args.pic:
            SS.BeamBeam6D_PIC(ctx, Np, d1, Np, d2, gp.gamma, gp.gamma, 1., 1., ns, ns)
        else:
            SS.BeamBeam6D_SoftGaussian(ctx, Np, d1, Np, d2, gp.gamma, gp.gamma, 1., 1., ns, ns, zb, zb)
    for _ in range(args.warmup):
        collide()
    ctx.set_timing(True)
    ctx.kernel_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        collide()
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    k_sec, k_pairs, k_rows = ctx.kernel_stats()
    ctx.set_timing(False)

    baseline = None
    if not args.no_cpu_baseline:
        import pyoracle
        m = min(n, 131072 if not args.pic else 16384)
        c1, c2 = _ss_beams(m, 2031)
        t = time.perf_counter()
        if args.pic:
            # the reference's z sort is an exchange sort (O(m^2)); sort first, untimed
            c1, c2 = c1[np.argsort(-c1[:, 4], kind="stable")], c2[np.argsort(-c2[:, 4], kind="stable")]
            t = time.perf_counter()
            pyoracle.bb_pic_6d(Np, c1, Np, c2, gp.gamma, gp.gamma, 1., 1., ns, ns)
            rows = ns * 2 * (m // ns) * ns
        else:
            pyoracle.bb_soft_gaussian_6d_zb(Np, c1, Np, c2, gp.gamma, gp.gamma, 1., 1., zb, zb)
            inside = lambda b: int(((b[:, 4] < zb[0]) & (b[:, 4] >= zb[-1])).sum())
            rows = ns * (inside(c1) + inside(c2))
        dt = time.perf_counter() - t
        baseline = {"value": rows / dt, "unit": "macroparticle kicks/s", "cores": 1,
                    "kind": "port", "sample": f"one collision of two {m}-macroparticle beams, {ns} x {ns} slices "
                                              f"(oracle restatement, one thread{', beams pre-sorted' if args.pic else ''}), "
                                              f"{dt:.1f} s"}

    bytes_ = (flops.PIC_BYTES if args.pic else flops.SS_BYTES) * k_rows
    fl = (flops.SS_MOMENTS + flops.SS_KICK_ELLIPTIC) * k_rows
    out = {
        "metric": "strong-strong macroparticle kicks/s",
        "value": k_rows / wall,
        "unit": "macroparticle kicks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian beams)",
        "config": {"workload": ("strong_strong: two 100 GeV proton beams, particle-in-cell 6-D collision with "
                                f"{ns} x {ns} z-sorted slices (BeamBeam6D_PIC, 64 x 64 grid)") if args.pic else
                               ("strong_strong: two 100 GeV proton beams, soft-Gaussian 6-D collision with "
                                f"{ns} x {ns} Gaussian z slices (BeamBeam6D_SoftGaussian, slice boundaries)"),
                   "particles_per_beam": n, "slices": ns, "charge_per_beam": Np},
        "roofline": {"bound": "hbm", "achieved": bytes_ / k_sec / 1e9, "peak": 8000.0, "unit": "GB/s",
                     "frac": bytes_ / k_sec / 8e12, "traffic": None,
                     "algorithmic_bytes_per_row": flops.PIC_BYTES if args.pic else flops.SS_BYTES,
                     "fp64_frac": None if args.pic else fl / k_sec / flops.PEAK_FP64_VECTOR,
                     "kernel": "st_pic.hip pair (two field solves)" if args.pic else "k_ss_partial + k_ss_kick",
                     "avg_pair_ms": 1e3 * k_sec / max(k_pairs, 1), "pairs": k_pairs,
                     "kernel_share_of_wall": k_sec / wall},
        "cpu_baseline": baseline,
    }
    print(json.dumps(out))


def bench_rhic(args):
    """The reference's first example: RHIC from its MAD-X table (6,858
    elements), Au ions at gamma = 100 with the 4 MV cavity, orbit tracking of
    Gaussian particles at the beam size."""
    import numpy as np

    from simtrack_amd import optics as O
    from simtrack_amd import workloads as W
    path = os.path.join(ROOT, "tests", "golden", "lattices", "rhic_2012_madx_twiss.tfs.gz")
    gp, ln = W.rhic_example(path)
    opt = O.Optics(ln)
    eps = 2.5e-6 * 0.65 / gp.gamma          # the example's sigma^2 / beta (Ex1: sqrt(2.5e-6*0.65/gamma))
    r = np.random.default_rng(2030)
    n = args.particles
    x = np.zeros((n, 6))
    u = r.standard_normal((n, 4))
    x[:, 0] = np.sqrt(eps * opt.beta_x) * u[:, 0]
    x[:, 1] = np.sqrt(eps / opt.beta_x) * (u[:, 1] - opt.alpha_x * u[:, 0])
    x[:, 2] = np.sqrt(eps * opt.beta_y) * u[:, 2]
    x[:, 3] = np.sqrt(eps / opt.beta_y) * (u[:, 3] - opt.alpha_y * u[:, 2])
    x[:, 5] = 0.0015                         # the example's deltap
    config = {"workload": "rhic: the reference's Example1 set-up (MAD-X table, 6,858 elements, Au at gamma 100, "
                          "4 MV RF), Gaussian particles at the beam size, deltap 1.5e-3",
              "particles": n, "turns_per_step": args.turns_per_step, "elements": int(len(ln.Cell)), "spin": False}
    bench_tracking(args, gp, ln.records(), x, False, config, None,
                   "the reference's RHIC lattice table; synthetic Gaussian particles")


def bench_config4(args):
    """Config 4: the ~3,500-element 18 GeV electron ring with radiation and
    spin, 1,048,576 electrons matched to the radiation equilibrium."""
    from simtrack_amd import workloads as W
    npart = args.particles if args.particles != 65536 else 1048576
    spin = not args.orbit
    gp, ln, x = W.config4(n_particles=npart, spin=spin)
    gp.rng_seed = 2029
    what = ("spin (Track_spin: classical radiation, as SBEND_sPass)" if spin
            else "orbit only (Track: radiation damping + quantum excitation)")
    config = {"workload": "config4: 18 GeV electron ring, synthetic FODO arcs + 2 straights, H_expand, "
                          "synchrotron radiation, " + what, "particles": npart,
              "turns_per_step": args.turns_per_step, "elements": int(len(ln.Cell)), "spin": spin, "radiate": True}
    bench_tracking(args, gp, ln.records(), x, spin, config, None,
                   "synthetic (Gaussian electrons matched to the radiation equilibrium)")


def bench_config2(args):
    """64 lattices of 1,000-1,800 elements (crab crossing + 6-D beam-beam),
    4,096 DA-grid particles each, tracked by the persistent multi-lattice
    launch with the autotuned shape; one launch per lattice (the plain
    per-table path) timed beside it on the same turns."""
    import numpy as np
    import torch

    import simtrack_amd as st
    from simtrack_amd import flops
    from simtrack_amd import workloads as W

    gp, lines, x, offs = W.config2(n_seeds=args.seeds, n_per_seed=args.per_seed)
    ctx = st.Context(0)
    tabs = [st.Table(ctx, gp, ln.records()) for ln in lines]
    f_seed = np.array([flops.flops_per_turn(ln.records(), gp) for ln in lines], dtype=float)
    batch = st.Batch(ctx, tabs)
    best, table = batch.autotune(x, offs, 32)
    bflags = st.TRACK_PERF if args.perf_mode else 0
    dev = torch.device("cuda:0")
    T = args.turns_per_step

    def fresh():
        d_x = torch.tensor(x, dtype=torch.float64, device=dev).contiguous()
        d_s = torch.ones(len(x), dtype=torch.int32, device=dev)
        d_l = torch.zeros(2 * len(x), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        return d_x, d_s, d_l

    def run(batched, d_x, d_s, d_l, turn0, nsteps):
        t = turn0
        for _ in range(nsteps):
            if batched:
                batch.track_device(d_x.data_ptr(), offs, t, T, d_s.data_ptr(), d_l.data_ptr(), bflags, best)
            else:
                for k, tb in enumerate(tabs):
                    o = int(offs[k])
                    tb.track_device(d_x.data_ptr() + 8 * 6 * o, int(offs[k + 1] - o), t, T, d_s.data_ptr() + 4 * o,
                                    d_l.data_ptr() + 8 * o, 0)
            t += T
        ctx.sync()
        torch.cuda.synchronize()
        return t

    results = {}
    for mode in ("batched", "per_lattice"):
        d_x, d_s, d_l = fresh()
        t = run(mode == "batched", d_x, d_s, d_l, 0, args.warmup)
        s0 = d_s.cpu().numpy().copy()
        ctx.set_timing(True)
        ctx.kernel_stats()
        t0 = time.perf_counter()
        t_end = run(mode == "batched", d_x, d_s, d_l, t, args.steps)
        wall = time.perf_counter() - t0
        k_sec, k_launch, _ = ctx.kernel_stats()
        ctx.set_timing(False)
        s1 = d_s.cpu().numpy()
        lost = d_l.cpu().numpy().reshape(-1, 2)
        per = np.where(s1 != 0, t_end - t, np.clip(lost[:, 0] - t, 0, t_end - t))
        per = np.where(s0 == 0, 0, per)
        flops_done = sum(float(per[offs[k]:offs[k + 1]].sum()) * f_seed[k] for k in range(len(tabs)))
        results[mode] = dict(pturns=int(per.sum()), wall=wall, k_sec=k_sec, launches=k_launch, flops=flops_done)
    baseline = None
    if not args.no_cpu_baseline:
        sp, stn = (int(v) for v in args.cpu_sample.split("x"))
        o0, o1 = int(offs[0]), int(offs[1])
        baseline = cpu_baseline(gp, lines[0].records(), x[o0:o1], False, min(sp, o1 - o0), stn, host_threads())
        baseline["sample"] = "lattice 0: " + baseline["sample"]
    r = results["batched"]
    value = r["pturns"] / r["wall"]
    achieved = r["flops"] / r["k_sec"] if r["k_sec"] > 0 else float("nan")
    peak = flops.PEAK_FP64_VECTOR
    out = {
        "metric": "FP64 particle-turns/s", "value": value, "unit": "particle-turns/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * r["wall"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded lattice errors, extra elements and DA grids)",
        "config": {"workload": f"config2: {len(tabs)} lattices of {min(len(l) for l in lines)}-"
                               f"{max(len(l) for l in lines)} elements, crab crossing + 6-D beam-beam, "
                               f"{args.per_seed} DA-grid particles each", "turns_per_step": T,
                   "autotuned": {"wg_waves": best.wg_waves, "waves_per_simd": best.waves_per_simd,
                                 "turn_block": best.turn_block}},
        "roofline": {"bound": "fp64_valu", "pipe": "fp64 VALU", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                     "kernel": "k_track_batch" + (" (perf mode)" if args.perf_mode else ""),
                     "launches": r["launches"]},
        "per_lattice_launches": {"value": results["per_lattice"]["pturns"] / results["per_lattice"]["wall"],
                                 "launches": results["per_lattice"]["launches"],
                                 "speedup_of_batched": (value / (results["per_lattice"]["pturns"] /
                                                                 results["per_lattice"]["wall"]))},
        "autotune_seconds": {f"{c[0]}x{c[1]}x{c[2]}": round(t_, 5) for c, t_ in table},
        "cpu_baseline": baseline,
    }
    print(json.dumps(out))


def bench_config3(args):
    """Surrogate + fallback DA scan vs full tracking: the network is trained
    on lattices the tracker generates, then both scans run on other seeds;
    DA boundaries compared per angle and lattice."""
    import numpy as np
    import torch

    import simtrack_amd as st
    from simtrack_amd import surrogate as S
    from simtrack_amd import workloads as W

    ctx = st.Context(0)
    gp = W.proton_globals(100000.0)
    twiss = (W.BETA_STAR, 0., W.BETA_STAR, 0.)

    def progress(msg):
        print(f"[config3] {msg}", file=sys.stderr, flush=True)
    n_ang = 8
    _, _, x1, _ = W.config2(n_seeds=1, n_per_seed=args.per_seed, n_angles=n_ang)
    rng = np.random.default_rng(7)

    def lattices(seeds, build=True):
        sizes = [int(rng.integers(1000, 1801)) for _ in seeds]     # drawn in the same order in every stage
        if not build:
            return None
        lines = [W.crab_ring(s, n, gp) for s, n in zip(seeds, sizes)]
        tabs = [st.Table(ctx, gp, ln.records()) for ln in lines]
        x = np.concatenate([x1] * len(seeds))
        offs = np.arange(len(seeds) + 1, dtype=np.int32) * len(x1)
        return tabs, x, offs

    if args.stage in ("full", "scan"):
        lattices(range(3000, 3000 + args.train_seeds), build=False)
        return bench_config3_stage(args, ctx, lattices, x1, n_ang, twiss, progress)
    t0 = time.perf_counter()
    tabs, x, offs = lattices(range(3000, 3000 + args.train_seeds))
    batch = st.Batch(ctx, tabs)
    d_x, d_s, d_l, d_f = S.features(ctx, batch, x, offs, *twiss)
    progress("training set: features done")
    alive = d_s != 0
    S.track_to(ctx, tabs, d_x, d_s, d_l, offs, alive, 2 * S.WINDOW, args.n_long, progress=progress)
    log = []
    mlp, std = S.train(ctx, d_f[alive], d_s[alive] != 0, epochs=args.epochs, batch=2048, log=log)
    progress(f"trained: loss {log[0]:.4f} -> {log[-1]:.4f}")
    t_train = time.perf_counter() - t0
    del d_x, d_s, d_l, d_f, batch
    if args.stage == "train":
        np.savez(args.state_out, params=mlp.get_params(), mean=std.mean.cpu().numpy(), std=std.std.cpu().numpy(),
                 log=np.array(log), t_train=t_train, n_long=args.n_long, train_seeds=args.train_seeds)
        print(json.dumps({"stage": "train", "n_long": args.n_long, "train_s": t_train,
                          "train_loss_first_last": [log[0], log[-1]], "state": args.state_out}))
        return

    etabs, ex, eoffs = lattices(range(2026, 2026 + args.seeds))
    ref, t_full, _, _ = S.full(ctx, etabs, ex, eoffs, args.n_long, progress=progress)
    amps = (np.arange(len(x1) // n_ang) + 1) * 20. / (len(x1) // n_ang)
    d_ref = [S.da_boundary(ref[eoffs[k]:eoffs[k + 1]], amps, n_ang) for k in range(len(etabs))]

    def da_errors(surv):
        return np.array([np.abs(S.da_boundary(surv[eoffs[k]:eoffs[k + 1]], amps, n_ang) - d_ref[k])
                         for k in range(len(etabs))])

    def summary(err):
        return {"max": float(err.max()), "mean": float(err.mean()),
                "lattices_within_0.25": int((err.max(axis=1) <= 0.25).sum())}

    runs = []
    for p_safe in [float(v) for v in args.p_safe.split(",")]:
        surv, info = S.scan(ctx, etabs, ex, eoffs, args.n_long, mlp, std, twiss, progress=progress,
                            mode=args.fallback, n_angles=n_ang, p_safe=p_safe)
        runs.append((p_safe, surv, info, da_errors(surv)))
        progress(f"p_safe {p_safe}: {t_full / info['wall']:.2f}x, DA error {summary(runs[-1][3])}")
    p_safe, surv, info, err_net = runs[0]
    err_plain = da_errors(info["plain"])
    if args.dump:
        np.savez_compressed(args.dump, ref=ref.astype(np.int8), surv=surv.astype(np.int8),
                            plain=info["plain"].astype(np.int8), prob=info["prob"].astype(np.float32),
                            fallback=info["fallback"], offsets=eoffs, amps=amps, n_angles=n_ang,
                            n_long=args.n_long, t_full=t_full, t_scan=info["wall"], t_features=info["t_features"])
    speedup = t_full / info["wall"]
    out = {
        "metric": "DA-scan wall time, surrogate+fallback vs full track", "value": speedup, "unit": "x (speedup)",
        "n_gpus": 1, "higher_is_better": True, "dtype": "f64 tracking, bf16/fp32 network", "data": "synthetic",
        "config": {"workload": f"config3: {args.seeds} evaluation lattices x {len(x1)} DA-grid particles, "
                               f"survival to {args.n_long} turns; network trained on {args.train_seeds} other "
                               f"lattices from their first 1,024 turns",
                   "mlp": "12-256-256-256-1 BF16 MFMA"},
        "full_track_s": t_full, "surrogate_scan_s": info["wall"], "features_s": info["t_features"],
        "fallback": args.fallback, "p_safe": p_safe if args.fallback == "boundary" else None,
        "fallback_particles": info["n_fallback"],
        "fallback_fraction": info["n_fallback"] / max(info["n_alive_1k"], 1),
        "da_error_sigma": summary(err_net),
        "plain_method_da_error_sigma": summary(err_plain),
        "survival_accuracy": {"surrogate": float((surv == ref).mean()), "plain": float((info["plain"] == ref).mean())},
        "tradeoff": [{"p_safe": ps, "speedup": t_full / inf["wall"], "surrogate_scan_s": inf["wall"],
                      "fallback_particles": inf["n_fallback"], "da_error_sigma": summary(er)}
                     for ps, _, inf, er in runs[1:]],
        "train_s": t_train, "train_loss_first_last": [log[0], log[-1]], "train_lattices": args.train_seeds,
        "features": "tunes of Hann-windowed 512-turn windows (FineTuneFinder on w*x)",
    }
    print(json.dumps(out))


def bench_config3_stage(args, ctx, lattices, x1, n_ang, twiss, progress):
    """One part of the config-3 comparison on the evaluation lattices, from a
    saved state to args.turns: full tracking, or the surrogate scan.  Each
    part's wall time accumulates in its state; the scan's last part prints
    the same line as the one-process run."""
    import numpy as np
    import torch

    import simtrack_amd as st
    from simtrack_amd import surrogate as S

    etabs, ex, eoffs = lattices(range(2026, 2026 + args.seeds))
    turn1 = min(args.turns or args.n_long, args.n_long)
    if args.stage == "full":
        state = S.load_state(args.state_in) if args.state_in else S.full_start(ex)
        S.full_track(ctx, etabs, state, eoffs, turn1, progress=progress)
        S.save_state(state, args.state_out)
        print(json.dumps({"stage": "full", "turn": state["turn"], "n_long": args.n_long, "elapsed_s": state["elapsed"],
                          "alive": int((state["d_s"] != 0).sum().item()), "state": args.state_out}))
        return
    z = np.load(args.mlp)
    if int(z["n_long"]) != args.n_long:
        raise SystemExit(f"--mlp was trained for survival to {int(z['n_long'])} turns, not {args.n_long}")
    mlp = st.Mlp(ctx, 12, 256)
    mlp.set_params(z["params"])
    dev = torch.device("cuda:0")
    std = S.Standardizer(mean=torch.as_tensor(z["mean"], device=dev), std=torch.as_tensor(z["std"], device=dev))
    p_safe = float(args.p_safe.split(",")[0])
    if args.state_in:
        state = S.load_state(args.state_in)
    else:
        state = S.scan_start(ctx, etabs, ex, eoffs, mlp, std, twiss, progress=progress, mode=args.fallback,
                             n_angles=n_ang, p_safe=p_safe)
    S.scan_track(ctx, etabs, state, eoffs, turn1, progress=progress)
    if state["turn"] < args.n_long:
        S.save_state(state, args.state_out)
        print(json.dumps({"stage": "scan", "turn": state["turn"], "n_long": args.n_long,
                          "elapsed_s": state["elapsed"], "state": args.state_out}))
        return
    surv, info = S.scan_finish(state)
    r = np.load(args.ref)
    if int(r["turn"]) != args.n_long:
        raise SystemExit(f"--ref reached turn {int(r['turn'])}, not {args.n_long}")
    ref, t_full = (r["d_s"] != 0).astype(np.int32), float(r["elapsed"])
    amps = (np.arange(len(x1) // n_ang) + 1) * 20. / (len(x1) // n_ang)
    err_net, err_plain = [], []
    for k in range(len(etabs)):
        sl = slice(eoffs[k], eoffs[k + 1])
        d_ref = S.da_boundary(ref[sl], amps, n_ang)
        err_net.append(np.abs(S.da_boundary(surv[sl], amps, n_ang) - d_ref))
        err_plain.append(np.abs(S.da_boundary(info["plain"][sl], amps, n_ang) - d_ref))
    err_net, err_plain = np.array(err_net), np.array(err_plain)

    def summary(err):
        return {"max": float(err.max()), "mean": float(err.mean()),
                "lattices_within_0.25": int((err.max(axis=1) <= 0.25).sum())}
    print(json.dumps({
        "metric": "DA-scan wall time, surrogate+fallback vs full track", "value": t_full / info["wall"],
        "unit": "x (speedup)", "n_gpus": 1, "higher_is_better": True, "dtype": "f64 tracking, bf16/fp32 network",
        "data": "synthetic",
        "config": {"workload": f"config3: {args.seeds} evaluation lattices x {len(x1)} DA-grid particles, "
                               f"survival to {args.n_long} turns, run in parts (times summed)",
                   "mlp": "12-256-256-256-1 BF16 MFMA"},
        "full_track_s": t_full, "surrogate_scan_s": info["wall"], "features_s": info["t_features"],
        "fallback": args.fallback, "p_safe": p_safe if args.fallback == "boundary" else None,
        "fallback_particles": info["n_fallback"],
        "fallback_fraction": info["n_fallback"] / max(info["n_alive_1k"], 1),
        "da_error_sigma": summary(err_net), "plain_method_da_error_sigma": summary(err_plain),
        "survival_accuracy": {"surrogate": float((surv == ref).mean()), "plain": float((info["plain"] == ref).mean())},
        "train_s": float(z["t_train"]), "train_n_long": int(z["n_long"]), "train_lattices": int(z["train_seeds"]),
    }))


if __name__ == "__main__":
    main()
