#!/usr/bin/env python3
"""Benchmark of the HOLD hot path on one MI355X.

A step is one full direct-sum kick of the N = 131072 Plummer sphere with
Salpeter masses (eps = 1e-4, seed 3): every particle is a sink and a source,
N^2 = 1.718e10 interactions, acc written and vel += acc * dt.  `value` is
Ginteractions/s over the timed steps (inputs resident in HBM).  The roofline
object prices the kick at 20 flops per interaction against the FP32 vector
peak.  One HOLD root step (dt = 1/64) of the same system is timed beside it,
and the FP64 CPU oracle is timed on a bounded sample of sinks on the host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallel-hhs_amd"))

METRIC = "Ginteractions/s at N=128k; ensemble systems/s + p99 latency at |dE/E|≤1e-6"
N = 131072
EPS = 1e-4
SEED = 3
FLOPS_PER_INTERACTION = 20
FP32_PEAK_TFLOPS = 157.3


def load_traffic():
    """HBM bytes per kick launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", "kick_n131072_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def cpu_baseline(pos, m, acc_gpu, eps2, seconds=12.0):
    """FP64 oracle direct sum on a sample of sinks, 1 thread (the reference's
    CPU force is serial, tests/test_nbforce.cu:9-23)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.set_threads(1)
    rng = np.random.default_rng(0)
    probe = rng.choice(N, 64, replace=False)
    t0 = time.perf_counter()
    oracle.accel_subset(pos, probe, pos, m, eps2)
    dt_probe = time.perf_counter() - t0
    k = int(min(N, max(64, 64 * seconds / max(dt_probe, 1e-6))))
    idx = rng.choice(N, k, replace=False)
    t0 = time.perf_counter()
    ref = oracle.accel_subset(pos, idx, pos, m, eps2)
    dt = time.perf_counter() - t0
    # precision of the benchmarked kick on the same sample
    scale = np.empty(k)
    for q, i in enumerate(idx[:256]):
        d = pos - pos[i]
        r2 = (d * d).sum(1) + eps2
        scale[q] = (m / r2).sum()
    err = np.linalg.norm(acc_gpu[idx[:256]] - ref[:256], axis=1) / scale[:256]
    return {"value": k * N / dt / 1e9, "unit": "Ginteractions/s", "cores": 1, "kind": "port",
            "sample": f"{k} random sinks x {N} sources, FP64 oracle direct sum (oracle/hhs_oracle.c)",
            "seconds": round(dt, 3)}, float(err.max())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-hold", action="store_true", help="skip the HOLD root-step timing")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; the kick shards by sinks "
              "(see DESIGN.md) but no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)

    import torch
    import hhs

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    eps2 = EPS * EPS
    t0 = time.perf_counter()
    pos, vel, m = hhs.ic_plummer(N, SEED, 1, True, eps2)
    t_ic = time.perf_counter() - t0
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    stream = torch.cuda.current_stream()
    dt = 1.0 / 64

    for _ in range(args.warmup):
        hhs.kick_self(d, dt=dt, fac=1.0, eps2=eps2)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        hhs.kick_self(d, dt=dt, fac=1.0, eps2=eps2)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = ev0.elapsed_time(ev1) / args.steps
    pairs = float(N) * N
    gint = pairs / (ms * 1e-3) / 1e9
    achieved = FLOPS_PER_INTERACTION * pairs / (ms * 1e-3) / 1e12
    s_cfg, w_cfg = hhs.kick_get_config(N, N)

    out = {
        "metric": METRIC,
        "value": round(gint, 2),
        "unit": "Ginteractions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Plummer sphere, Salpeter masses)",
        "config": {"workload": "direct-sum kick, Plummer N=131072 Salpeter masses, eps=1e-4, "
                               "seed 3 (FP64 positions, FP64-derived FP32 separations, "
                               "two-sum compensated FP32 sums)",
                   "n": N, "eps": EPS, "sinks_per_lane": s_cfg, "waves_per_group": w_cfg},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "traffic": load_traffic(),
                     "flops_per_interaction": FLOPS_PER_INTERACTION,
                     "kernel": "kick_kernel<double,S,W,true>"},
        "wall_s_timed": round(wall, 4),
        "ic_seconds": round(t_ic, 2),
    }

    acc_gpu = d.acc.cpu().numpy()
    if not args.no_hold:
        # One HOLD root step of the same system (timesteps, splits, kicks,
        # drifts, scatter-back at every level).
        d2 = hhs.DeviceParticles.from_host(pos, vel, m)
        g = hhs.Globals(N, 0.01, EPS)
        ke0, pe0 = hhs.energy(d2.pos, d2.vel, d2.m, eps2)
        torch.cuda.synchronize()
        h0 = time.perf_counter()
        g.hold_step(d2, dt)
        torch.cuda.synchronize()
        hs = time.perf_counter() - h0
        ke1, pe1 = hhs.energy(d2.pos, d2.vel, d2.m, eps2)
        st = g.stats()
        out["hold_root_step"] = {
            "dt": dt, "seconds": round(hs, 4), "calls": st["calls"],
            "levels": st["max_level"] + 1, "interactions": st["interactions"],
            "ts_pairs": st["ts_pairs"],
            "kick_ginteractions_per_s_effective": round(st["interactions"] / hs / 1e9, 2),
            "dE_over_E": abs((ke1 + pe1 - ke0 - pe0) / (ke0 + pe0)),
        }
        g.close()

    if not args.no_cpu:
        cb, ferr = cpu_baseline(pos, m, acc_gpu, eps2)
        out["cpu_baseline"] = cb
        out["force_err_max_sampled"] = ferr
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
