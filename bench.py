#!/usr/bin/env python
"""Benchmark: RK4 steps/s of the droplet-laden forced DNS on one MI355X.

Default workload (the metric's configuration): 256^3 FP32 fields + 10^6
droplets (FP64 positions), nu = 2.5e-3, forced shells |k| <= 2,
E(k) ~ k^4 exp(-2 (k/3)^2) random-phase initial field (seed 0), droplet St
lognormal (median 0.3, sigma 0.6), dt from CFL 0.5.  Inputs are generated on
the host, copied to HBM before timing; a step is one full RK4 step (four RHS
evaluations, droplet advance, clip_zero, reality round trip, projection,
blow-up check) advanced on the device-resident state.  Defaults: 10 warm-up
steps, then 100 timed steps (HIP events on the stream).

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

METRIC = "RK steps/s @256³+1e6 droplets, FP32 vs FP64 plain; % HBM roofline"


def algorithmic_bytes(N: int, s: int, npart: int, sp: int) -> int:
    """Algorithmic HBM bytes of one RK4 step:
    74 (R + C) + 4 [min(7R, 448 s Np) + R] + 4 Np 5 S_p: one read and one write
    per 3-D transform (74 per step), the 7-component gather and the deposit per
    stage, and 3 reads + 2 writes of the droplet state per stage."""
    R = N ** 3 * s
    C = N * N * (N // 2 + 1) * 2 * s
    return 74 * (R + C) + 4 * (min(7 * R, 448 * s * npart) + R) + 4 * npart * 5 * sp


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_inputs(cfg):
    import numpy as np
    from lucky_droplets_amd import ic
    rng = np.random.default_rng(cfg.seed)
    uk, n, u = ic.initial_flow(cfg, rng)
    dt = ic.cfl_dt(cfg, u)
    coord, pid = ic.initial_droplets(cfg, rng)
    return uk, n, coord, pid, dt


def time_gpu(cfg, uk, n, coord, dt, steps, warmup, profile_kernel=None, graph=False, autotune=False,
             extra_profile=None):
    """ms per step over `steps` timed steps (after `warmup`), plus the summed
    HIP-event duration and launch count of `profile_kernel` inside the timed region
    (profiling runs the steps eagerly; graph=True replays a captured HIP graph;
    autotune=True lets the context time and pick its launch geometries at creation)."""
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    dev = torch.device("cuda:0")
    d = ForcedDNS(cfg, graph=graph, autotune=autotune)
    tune = d.tuning_all() if autotune else None
    ukg = torch.from_numpy(uk).to(dev).to(d.cdtype)
    ng = torch.from_numpy(n).to(dev).to(d.rdtype)
    cg = torch.from_numpy(coord).to(dev)
    d.set_state(ukg, ng, cg)
    del ukg, ng
    if warmup:
        d.advance(dt, warmup)
    torch.cuda.synchronize()
    if profile_kernel:
        d.profile(profile_kernel, True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    d.advance(dt, steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    kern = d.profile_read() if profile_kernel else (0.0, 0)
    if profile_kernel:
        d.profile(profile_kernel, False)
    # further kernel classes, each timed over as many steps after the timed region
    time_gpu.extra = {}
    for name in extra_profile or ():
        d.profile(name, True)
        d.advance(dt, steps)
        time_gpu.extra[name] = d.profile_read()
        d.profile(name, False)
    lastmax, blown = d.status()
    d.close()
    time_gpu.last_tuning = tune
    return ms, kern, lastmax, blown


def large_context(cfg, uk_small, dt_small):
    """The 512^3 FP32 + 10^7-droplet workload as a resident context.  Its state
    is the smaller workload's initial field zero-padded to this grid on the
    device (ld_spectral_resample: the same band-limited velocity field, so the
    dealiasing fast path applies) with a uniform density and the workload's own
    seeded droplet population; dt scales with the grid spacing.  Returns
    (context, dt)."""
    import numpy as np
    import torch
    from lucky_droplets_amd import _lib, ic
    from lucky_droplets_amd.dns import ForcedDNS, _ptr
    dev = torch.device("cuda:0")
    d = ForcedDNS(cfg, autotune=True)
    ns = uk_small.shape[1]
    src = torch.from_numpy(uk_small).to(dev).to(d.cdtype).contiguous()
    uk = d.zeros_spec(3)
    d._stream()
    _lib.check(d.lib.ld_spectral_resample(d.ctx, _ptr(src), ns, _ptr(uk), 3))
    del src
    n = torch.ones(d.rshape(), dtype=d.rdtype, device=dev)
    coord, _ = ic.initial_droplets(cfg, np.random.default_rng(cfg.seed))
    cg = torch.from_numpy(coord).to(dev)
    del coord
    d.set_state(uk, n, cg)
    del uk, n, cg
    return d, dt_small * ns / cfg.N


def time_large(cfg, uk_small, dt_small, steps, warmup):
    """ms per step of the 512^3 FP32 + 10^7-droplet workload (large_context)."""
    import torch
    d, dt = large_context(cfg, uk_small, dt_small)
    d.advance(dt, warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    d.advance(dt, steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    _, blown = d.status()
    d.close()
    return ms, dt, blown


def time_ensemble(cfg, nreal, steps, warmup, graph=True, autotune=True):
    """ms per ensemble step of `nreal` realizations of `cfg` stepped together
    (one launch per kernel for all realizations), each with its own seeded
    initial field and droplets; with autotune the context picks its launch
    geometries for this ensemble at creation (identical results)."""
    import numpy as np
    import torch
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.dns import ForcedDNS
    dev = torch.device("cuda:0")
    d = ForcedDNS(cfg, batch=nreal, graph=graph, autotune=autotune)
    time_ensemble.last_tuning = d.tuning_all() if autotune else None
    uke = torch.empty((nreal,) + d.cshape(3), dtype=d.cdtype, device=dev)
    ne = torch.empty((nreal,) + d.rshape(), dtype=d.rdtype, device=dev)
    ce = torch.empty((nreal, cfg.n_particles, 7), dtype=torch.float64, device=dev)
    dt = None
    for b in range(nreal):
        rng = np.random.default_rng(cfg.seed + b)
        uk, n, u = ic.initial_flow(cfg, rng)
        dtb = ic.cfl_dt(cfg, u)
        dt = dtb if dt is None else min(dt, dtb)
        coord, _ = ic.initial_droplets(cfg, rng)
        uke[b].copy_(torch.from_numpy(uk))
        ne[b].copy_(torch.from_numpy(n))
        ce[b].copy_(torch.from_numpy(coord))
    d.set_state(uke, ne, ce)
    del uke, ne, ce
    d.advance(dt, warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    d.advance(dt, steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    _, blown = d.status_all()
    d.close()
    return ms, dt, bool(blown.any())


def time_cnn(N, channels, steps):
    """ms per forward pass and per training step (forward, backward, Adam) of
    the learned subgrid CNN on an N^3 grid, and the MFMA FLOPs they execute
    (padded channel counts)."""
    import numpy as np
    import torch
    from lucky_droplets_amd.cnn import SubgridCNN
    dev = torch.device("cuda:0")
    net = SubgridCNN(N, channels)
    net.init_params(seed=0)
    rng = np.random.default_rng(0)
    x = torch.from_numpy(rng.normal(size=(N ** 3, channels[0])).astype(np.float32)).to(dev)
    t = torch.from_numpy(rng.normal(size=(N ** 3, channels[-1])).astype(np.float32)).to(dev)
    for _ in range(2):
        net.train_step(x, t)
        net.forward(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        net.forward(x)
    e1.record()
    torch.cuda.synchronize()
    ms_fwd = e0.elapsed_time(e1) / steps
    e0.record()
    for _ in range(steps):
        net.train_step(x, t)
    e1.record()
    torch.cuda.synchronize()
    ms_train = e0.elapsed_time(e1) / steps
    pad = [16 if c <= 16 else 32 for c in channels]
    f_layer = [2 * 27 * ci * co * N ** 3 for ci, co in zip(pad[:-1], pad[1:])]
    flops_fwd = sum(f_layer)
    flops_train = 2 * flops_fwd + sum(f_layer[1:])        # forward + weight gradients + data gradients
    net.close()
    return ms_fwd, ms_train, flops_fwd, flops_train


def time_plain(cfg, uk, n, coord, dt, steps):
    """ms per step of the unfused reference-dataflow path (ld_advance_plain)."""
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    dev = torch.device("cuda:0")
    d = ForcedDNS(cfg)
    d.set_state(torch.from_numpy(uk).to(dev).to(d.cdtype), torch.from_numpy(n).to(dev).to(d.rdtype),
                torch.from_numpy(coord).to(dev))
    d.advance_plain(dt, 1)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    d.advance_plain(dt, steps)
    e1.record()
    torch.cuda.synchronize()
    d.close()
    return e0.elapsed_time(e1) / steps


def time_collisions(cfg, coord, reps=5):
    """ms per cell-list collision-pair search and per pair-statistics pass over
    the workload's droplets (radius r = sqrt(9 nu tau_eta St / (2 rho_p)))."""
    import torch
    from lucky_droplets_amd.dns import ForcedDNS
    from lucky_droplets_amd.particles import MPI_particles
    dev = torch.device("cuda:0")
    d = ForcedDNS(cfg)
    stb = MPI_particles(None, cfg.L, cfg.N, cfg.n_particles, cfg.sts, 0.3, cfg.g, cfg.nu_eff, cfg.tf, cfg.rhop, cfg.M0,
                        dns=d, coord=torch.from_numpy(coord).to(dev))
    rad = stb.radius(cfg.nu_eff, cfg.rhop)
    npairs = stb.collision_pairs(rad).shape[0]
    stb.pair_statistics(0.1, 64)
    torch.cuda.synchronize()
    t = time.time()
    for _ in range(reps):
        stb.collision_pairs(rad)
    torch.cuda.synchronize()
    ms_pairs = (time.time() - t) / reps * 1e3
    t = time.time()
    for _ in range(reps):
        stb.pair_statistics(0.1, 64)
    torch.cuda.synchronize()
    ms_stats = (time.time() - t) / reps * 1e3
    d.close()
    return {"droplets": cfg.n_particles, "pairs": int(npairs), "ms_collision_pairs": ms_pairs,
            "ms_pair_statistics_rmax_0.1": ms_stats,
            "note": "host wall clock per call (each call synchronises to read its count)"}


def droplet_launch_bytes(N, s, npart, sp):
    """Algorithmic bytes of one launch of the droplet stage kernel: the
    per-stage droplet terms of the step without the deposit's write (done by
    the deposit kernels), min(7R, 448 s Np) + 5 Np S_p."""
    R = N ** 3 * s
    return min(7 * R, 448 * s * npart) + 5 * npart * sp


def cpu_baseline(uk, n, coord, cfg, dt, c1_steps=10):
    """The CPU FP64 oracle (numpy + scipy.fft) timed on this host, bounded:
    one full step (RK4 + epilogue) of the benchmark workload on the same
    inputs (the reported value, steps/s), and c1_steps full steps of the 64^3
    workload with one scipy.fft worker (the reference's per-rank FFT, which
    passes no workers) and with min(16, cpu_count) workers."""
    import numpy as np
    import oracle
    import oracle.flow as oflow
    from lucky_droplets_amd import ic
    from lucky_droplets_amd.config import preset
    ncpu = os.cpu_count() or 1
    workers = min(16, ncpu)

    def flow_drops(c, crd):
        fl = oracle.Flow(c.N, c.nu_eff, c.shell_no, c.f0, c.tps, c.g, c.isforcing, 1.0 if c.explicit else 0.0,
                         c.lp, c.L)
        dr = oracle.Droplets(c.L, c.N, crd.copy(), c.sts, c.g, c.nu_eff, c.tf, c.rhop, c.M0)
        return fl, dr

    c1 = preset("c1")
    rng = np.random.default_rng(c1.seed)
    uk1, n1, u1 = ic.initial_flow(c1, rng)
    dt1 = ic.cfl_dt(c1, u1)
    coord1, _ = ic.initial_droplets(c1, rng)
    c1_times = {}
    for w in (1, workers):
        oflow.WORKERS = w
        fl, dr = flow_drops(c1, coord1)
        a, b = uk1.copy(), n1.copy()
        t = time.time()
        for i in range(c1_steps):
            a, b = fl.step(i * dt1, dt1, dr, a, b)
        c1_times[w] = (time.time() - t) / c1_steps
    oflow.WORKERS = workers
    fl, dr = flow_drops(cfg, coord)
    t = time.time()
    fl.step(0.0, dt, dr, uk.copy(), n.copy())
    step_s = time.time() - t
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": 1.0 / step_s, "unit": "steps/s", "cores": workers, "kind": "port",
            "sample": f"one full step (RK4: 4 full_RHS + droplets, epilogue) of the FP64 oracle at {cfg.N}^3 + "
                      f"{cfg.n_particles} droplets on the same inputs: {step_s:.1f} s, scipy.fft workers={workers}; "
                      f"host: {model}, os.cpu_count()={ncpu}",
            "c1_64cubed_1e4_droplets": {"s_per_step_workers_1": c1_times[1], f"s_per_step_workers_{workers}":
                                        c1_times[workers], "steps_timed": c1_steps,
                                        "est_200_steps_s_workers_1": 200 * c1_times[1],
                                        f"s_for_{c1_steps}_steps_workers_1": c1_steps * c1_times[1],
                                        f"s_for_{c1_steps}_steps_workers_{workers}": c1_steps * c1_times[workers]},
            "cpu_count": ncpu}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph timing")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64 x 64^3 ensemble timing")
    ap.add_argument("--no-large", action="store_true", help="skip the 512^3 + 10^7-droplet timing")
    ap.add_argument("--track-steps", type=int, default=0,
                    help="also run the 512^3 workload this many steps with mass tracking (off by default)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP32-against-FP64 accuracy run")
    ap.add_argument("--no-collisions", action="store_true", help="skip the collision-pair timing")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-subgrid-form comparison")
    ap.add_argument("--no-cnn", action="store_true", help="skip the learned-subgrid CNN timing")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default y-pass block shape")
    ap.add_argument("--cpu-c1-steps", type=int, default=10,
                    help="full oracle steps of the 64^3 workload in the CPU baseline (200 = the whole run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding is a "
              "design only.", file=sys.stderr)
        sys.exit(2)

    from lucky_droplets_amd.config import preset
    cfg = preset(args.config)
    log(f"[bench] generating inputs for {args.config}: N={cfg.N} droplets={cfg.n_particles}")
    t = time.time()
    uk, n, coord, pid, dt = make_inputs(cfg)
    log(f"[bench] inputs ready in {time.time() - t:.1f}s, dt={dt:.4e}")

    log(f"[bench] FP32 timing: warmup {args.warmup}, steps {args.steps}")
    tuned = not args.no_autotune
    ms32, (kms, klaunch), lastmax, blown = time_gpu(cfg, uk, n, coord, dt, args.steps, args.warmup,
                                                    profile_kernel="droplets" if cfg.n_particles else None,
                                                    autotune=tuned,
                                                    extra_profile=("deposit",) if cfg.n_particles else None)
    dep_ms, dep_launch = time_gpu.extra.get("deposit", (0.0, 0))
    tuning = time_gpu.last_tuning
    log(f"[bench] FP32 {ms32:.3f} ms/step; droplet kernel {kms / max(klaunch, 1):.3f} ms x {klaunch}"
        + (f"; launch choices {tuning[0]}" if tuning else ""))
    if tuning:
        log("[bench] autotuner median us per candidate: "
            + ", ".join(f"{k} {[round(x, 1) for x in v]}" for k, v in tuning[1].items()))
    ms_eager = ms32
    msg = None
    if not args.no_graph:
        msg, _, _, blown_g = time_gpu(cfg, uk, n, coord, dt, args.steps, args.warmup, graph=True, autotune=tuned)
        log(f"[bench] FP32 with HIP-graph replay {msg:.3f} ms/step")
        blown = blown or blown_g
        ms32 = min(ms32, msg)

    out = {
        "metric": METRIC,
        "value": 1000.0 / ms32,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "eager_ms_per_step": ms_eager,
        "graph_ms_per_step": msg,
        "launch_tuning": ({"choices": tuning[0], "median_us_per_candidate": tuning[1]} if tuning else None),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-phase field, seeded droplets)",
        "config": {"workload": f"forced DNS {cfg.N}^3 + {cfg.n_particles} droplets, RK4 step",
                   "N": cfg.N, "droplets": cfg.n_particles, "nu": cfg.nu_eff, "forced_shells": list(cfg.shell_no),
                   "dt": dt, "precision": cfg.precision},
    }
    sp = 40 if cfg.precision == "f32" else 56
    s = 4 if cfg.precision == "f32" else 8
    B = algorithmic_bytes(cfg.N, s, cfg.n_particles, sp)
    out["step_roofline"] = {"bytes_per_step": B, "achieved_GBs": B / (ms32 * 1e-3) / 1e9,
                            "peak_GBs": HBM_PEAK_GBS, "frac": B / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS}
    if klaunch:
        kb = droplet_launch_bytes(cfg.N, s, cfg.n_particles, sp)
        avg = kms / klaunch
        ach = kb / (avg * 1e-3) / 1e9
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "traffic_droplets_256.json")
        if os.path.exists(tpath) and args.config == "c2":
            with open(tpath) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        out["roofline"] = {"bound": "hbm", "kernel": "k_particles_stage (droplet RK stage)",
                           "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                           "traffic": traffic, "algorithmic_bytes_per_launch": kb, "avg_launch_ms": avg,
                           "launches": klaunch}
        if dep_launch:
            # the deposit (brick pass + fallback) writes the R-byte source field of
            # each stage; the droplet stage as a whole is both kernels
            R = cfg.N ** 3 * s
            davg = dep_ms / dep_launch
            out["roofline"]["deposit"] = {"kernels": "k_brick_deposit + k_fallback_accum/fold",
                                          "algorithmic_bytes_per_launch": R, "avg_launch_ms": davg,
                                          "achieved": R / (davg * 1e-3) / 1e9,
                                          "frac": R / (davg * 1e-3) / 1e9 / HBM_PEAK_GBS}
            tot = (kb + R) / ((avg + davg) * 1e-3) / 1e9
            out["roofline"]["droplet_stage"] = {"algorithmic_bytes_per_stage": kb + R, "ms_per_stage": avg + davg,
                                                "achieved": tot, "frac": tot / HBM_PEAK_GBS}
    if not args.no_fp64:
        c64 = cfg.replace(precision="f64")
        nsteps64 = max(2, args.steps // 4)
        log(f"[bench] FP64 timing: steps {nsteps64}")
        ms64, _, _, _ = time_gpu(c64, uk, n, coord, dt, nsteps64, 1)
        log(f"[bench] FP64 {ms64:.3f} ms/step")
        out["fp64_fused"] = {"ms_per_step": ms64, "steps_per_s": 1000.0 / ms64, "fp32_speedup": ms64 / ms32}
        nplain = max(2, args.steps // 5)
        log(f"[bench] FP64 plain timing: steps {nplain}")
        msp = time_plain(c64, uk, n, coord, dt, nplain)
        log(f"[bench] FP64 plain {msp:.3f} ms/step")
        out["fp64_plain"] = {"ms_per_step": msp, "steps_per_s": 1000.0 / msp, "fp32_speedup": msp / ms32,
                             "form": "reference dataflow on the GPU: 74 separate 3-D transforms per step and one "
                                     "kernel per array statement (ld_advance_plain)"}
        # the reference's own CPU-runnable case on the GPU, beside the CPU baseline's run of it
        c1 = preset("c1")
        uk1, n1, coord1, _, dt1 = make_inputs(c1)
        ms_c1, _, _, _ = time_gpu(c1, uk1, n1, coord1, dt1, 200, 5)
        log(f"[bench] C1 64^3 FP64 + {c1.n_particles} droplets: {ms_c1:.3f} ms/step, 200 steps in "
            f"{0.2 * ms_c1:.2f} s")
        out["c1_fp64"] = {"workload": f"forced DNS 64^3 FP64 + {c1.n_particles} droplets, 200 steps",
                          "ms_per_step": ms_c1, "s_for_200_steps": 0.2 * ms_c1}
    if not args.no_fp64 and not args.no_accuracy:
        from lucky_droplets_amd.accuracy import compare_precisions
        nacc = cfg.steps
        log(f"[bench] FP32 against FP64 over {nacc} steps")
        rep = compare_precisions(cfg, uk, n, coord, dt, nacc)
        rep.pop("ek_f32")
        rep.pop("ek_f64")
        out["fp32_accuracy"] = rep
        log(f"[bench] spectrum error max {rep['spectrum_err_max']:.2e} (shell {rep['spectrum_err_max_shell']}), "
            f"dissipation error {rep['dissipation_err']:.2e}")
    if cfg.n_particles and not args.no_collisions:
        log("[bench] collision pairs and pair statistics")
        out["collisions"] = time_collisions(cfg, coord)
        log(f"[bench] collisions {out['collisions']['ms_collision_pairs']:.2f} ms, "
            f"{out['collisions']['pairs']} pairs")
    if not args.no_learned:
        from lucky_droplets_amd.config import Config as _Config
        from lucky_droplets_amd.learned import collision_comparison
        log("[bench] learned subgrid form: 256^3 fine DNS vs 128^3 coarse (plain, learned), 10^6 droplets St 0.5")
        cf = _Config(N=256, precision="f32", nu=2.5e-3, nshells=2, profile="k4")
        lr = collision_comparison(cf, npart=1_000_000, st=0.5, spinup=500, train_states=12, train_every=10,
                                  train_iters=1000, relax=300, samples=5, sample_every=20, rmax=0.05, nbins=16)
        forms = lr.pop("forms")
        lr["K_sum"] = {k: float(forms[k]["K"].sum()) for k in forms}
        lr["workload"] = ("fine DNS 256^3 FP32 (nu 2.5e-3) against a 128^3 coarse DNS from the truncated state, "
                          "droplets one-way coupled in the frozen field of each step; pair statistics up to r = 0.05 "
                          "over 5 snapshots after 300 steps")
        out["learned"] = lr
        log(f"[bench] learned: kernel error plain {lr['kernel_err_plain']:.3f}, learned {lr['kernel_err_learned']:.4f}; "
            f"ms/step fine {lr['ms_per_step']['fine']:.2f}, learned {lr['ms_per_step']['learned']:.2f}")
    if not args.no_large and args.config == "c2":
        from lucky_droplets_amd.config import preset as _preset
        c4 = _preset("c4")
        log(f"[bench] large: {c4.N}^3 FP32 + {c4.n_particles} droplets")
        ms4, dt4, blown4 = time_large(c4, uk, dt, max(5, args.steps // 10), 2)
        b4 = algorithmic_bytes(c4.N, 4, c4.n_particles, 40)
        out["large"] = {"workload": f"forced DNS {c4.N}^3 FP32 + {c4.n_particles} droplets (bidisperse St 0.2 / 0.6, "
                                    f"nu {c4.nu}); the {cfg.N}^3 initial field zero-padded, uniform density",
                        "ms_per_step": ms4, "steps_per_s": 1000.0 / ms4, "dt": dt4, "bytes_per_step": b4,
                        "achieved_GBs": b4 / (ms4 * 1e-3) / 1e9, "frac": b4 / (ms4 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "blowup": blown4}
        blown = blown or blown4
        log(f"[bench] large {ms4:.2f} ms/step")
        if args.track_steps > 0:
            from lucky_droplets_amd import tracking
            d4, dt4 = large_context(c4, uk, dt)
            hist = tracking.mass_history(d4, dt4, args.track_steps, max(1, args.track_steps // 20), log=log)
            d4.close()
            out["large"]["mass_tracking"] = tracking.summary(hist)
            out["large"]["mass_tracking"]["history"] = hist.tolist()
            log(f"[bench] large: {args.track_steps} steps tracked, droplet mass "
                f"{hist[0, 1]:.6e} -> {hist[-1, 1]:.6e}, max drift of the conserved sum "
                f"{out['large']['mass_tracking']['max_total_drift']:.3e}")
    if not args.no_ensemble:
        from lucky_droplets_amd.config import preset as _preset
        c3 = _preset("c3")
        nreal = 64
        log(f"[bench] ensemble: {nreal} x {c3.N}^3 FP32 + {c3.n_particles} droplets each")
        mse, dte, blown_e = time_ensemble(c3, nreal, max(5, args.steps // 2), 2, autotune=tuned)
        etune = time_ensemble.last_tuning
        be = nreal * algorithmic_bytes(c3.N, 4, c3.n_particles, 40)
        out["ensemble"] = {"workload": f"{nreal} realizations of forced DNS {c3.N}^3 FP32 + {c3.n_particles} "
                                       "droplets, stepped together (one launch per kernel), HIP-graph replay",
                           "ms_per_ensemble_step": mse, "ensemble_steps_per_s": 1000.0 / mse,
                           "realization_steps_per_s": nreal * 1000.0 / mse, "dt": dte,
                           "bytes_per_ensemble_step": be, "achieved_GBs": be / (mse * 1e-3) / 1e9,
                           "frac": be / (mse * 1e-3) / 1e9 / HBM_PEAK_GBS, "blowup": blown_e,
                           "launch_tuning": etune[0] if etune else None}
        log(f"[bench] ensemble {mse:.3f} ms per ensemble step"
            + (f"; launch choices {etune[0]}" if etune else ""))
    if not args.no_cnn:
        ch = (3, 32, 32, 32, 3)
        log("[bench] learned subgrid CNN: 128^3, channels " + str(ch))
        msf, mst, ff, ft = time_cnn(128, ch, max(5, args.steps // 2))
        out["cnn"] = {"workload": "4-layer periodic 3x3x3 CNN, channels 3-32-32-32-3 (padded to 16/32), 128^3 nodes, "
                                  "bf16 MFMA with FP32 accumulation",
                      "ms_forward": msf, "ms_train_step": mst,
                      "forward_TFLOPs": ff / (msf * 1e-3) / 1e12, "train_TFLOPs": ft / (mst * 1e-3) / 1e12,
                      "peak_TFLOPs": 2500.0, "frac_forward": ff / (msf * 1e-3) / 1e12 / 2500.0}
        log(f"[bench] CNN forward {msf:.3f} ms, train step {mst:.3f} ms")
    out["blowup"] = bool(blown)
    if not args.no_cpu:
        log(f"[bench] CPU baseline (oracle: one full step at the benchmark workload, {args.cpu_c1_steps} steps at 64^3)")
        out["cpu_baseline"] = cpu_baseline(uk, n, coord, cfg, dt, c1_steps=args.cpu_c1_steps)
        log(f"[bench] CPU {out['cpu_baseline']['value']:.4f} steps/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
