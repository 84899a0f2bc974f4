#!/usr/bin/env python3
"""Benchmark of the HOLD hot path on one MI355X.

A step is one full direct-sum kick of the N = 131072 Plummer sphere with
Salpeter masses (eps = 1e-4, seed 3): every particle is a sink and a source,
N^2 = 1.718e10 interactions, acc written and vel += acc * dt.  `value` is
Ginteractions/s over the timed steps (inputs resident in HBM).  The roofline
object prices the kick at 20 flops per interaction against the FP32 vector
peak.  Beside the headline line the JSON carries:
  * timesteps -- the timestep kernel over the same system, three modes;
  * timestep_autotune -- the one-sided FP32 timestep kernel tuned for the
    size classes 2^8 .. 2^15 (the HOLD legs then run with it);
  * stream_kernels -- split, scan, drift and update against HBM bandwidth
    at N = 131072 and 2^20..2^24;
  * hold_root_step -- one HOLD root step (dt = 1/64) of the same system,
    beside the FP64-only HOLD of the same step (the energy reference);
  * hold_n16384 -- the plain HOLD of an N = 16384 Plummer sphere to t =
    0.125, beside its FP64-only run;
  * ensemble -- 4096 Plummer systems (N log-uniform in [8, 2048], t = 1) in
    one persistent launch, against the plain one-system-at-a-time loop;
  * served -- a Poisson request stream (N log-uniform in [64, 65536], t_end
    uniform in [0.1, 1]) through the continuous-batching server, p50/p99
    time-to-solution per arrival rate;
  * split_sweep -- the reference's split timing design: GPU split beside
    cpu_split at N = 2^10 .. 2^22;
  * cpu_baseline -- the FP64 oracle direct sum on the host, one thread, on
    a bounded sample of sinks, and the kick's force error on that sample;
  * cpu_reference -- all-threads CPU kick, cpu_split at 2^24, and the FP64
    CPU HOLD at N = 1024 beside the GPU.
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
ETA = 0.01
SEED = 3
FLOPS_PER_INTERACTION = 20
TS_FLOPS_PER_PAIR = 49
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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
    # precision of the benchmarked kick on the same sample: error relative to
    # sum_j m_j / (r_ij^2 + eps^2) (normalised) and to |a_i| (plain)
    q = min(k, 1024)
    scale = np.empty(q)
    for j, i in enumerate(idx[:q]):
        d = pos - pos[i]
        r2 = (d * d).sum(1) + eps2
        scale[j] = (m / r2).sum()
    diff = np.linalg.norm(acc_gpu[idx[:q]] - ref[:q], axis=1)
    err = diff / scale
    rel = diff / np.linalg.norm(ref[:q], axis=1)
    ferr = {"sinks": q, "normalised_max": float(err.max()),
            "normalised_rms": float(np.sqrt((err ** 2).mean())),
            "relative_max": float(rel.max()), "relative_rms": float(np.sqrt((rel ** 2).mean()))}
    return {"value": round(k * N / dt / 1e9, 4), "unit": "Ginteractions/s", "cores": 1,
            "kind": "port",
            "sample": f"{k} random sinks x {N} sources, FP64 oracle direct sum "
                      "(oracle/hhs_oracle.c), one host thread",
            "seconds": round(dt, 3)}, ferr


def box_info():
    """Host facts of the machine the CPU legs ran on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "cpu_model": model}


def cpu_reference(hhs, torch, pos, m, eps2, kick_seconds=8.0):
    """Further CPU timings (steady clock): the oracle direct sum with every
    host thread; cpu_split (src/cpu_scan.h:20-50, the oracle's restatement),
    median of 5 on 2^24 floats; and the FP64 CPU
    HOLD at N = 1024 (eps = 0, eta = 0.01, t = 1 in root steps of 1/16)
    beside the GPU FP64-only and default HOLD on the same system."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    out = {"box": box_info()}
    threads = out["box"]["affinity_cpus"]
    oracle.set_threads(threads)
    rng = np.random.default_rng(1)
    probe = rng.choice(N, 256, replace=False)
    t0 = time.perf_counter()
    oracle.accel_subset(pos, probe, pos, m, eps2)
    dtp = time.perf_counter() - t0
    k = int(min(N, max(256, 256 * kick_seconds / max(dtp, 1e-6))))
    idx = rng.choice(N, k, replace=False)
    t0 = time.perf_counter()
    oracle.accel_subset(pos, idx, pos, m, eps2)
    dt = time.perf_counter() - t0
    out["kick_all_threads"] = {"value": round(k * N / dt / 1e9, 4), "unit": "Ginteractions/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{k} random sinks x {N} sources"}
    vals = (10.0 * np.random.default_rng(2).random(1 << 24)).astype(np.float32)
    times = []
    for _ in range(5):
        arr = vals.copy()
        t0 = time.perf_counter()
        oracle.cpu_split_inplace_f32(arr, 5.0)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    out["cpu_split"] = {"n": 1 << 24, "median_ms": round(med * 1e3, 2),
                        "Melements_per_s": round((1 << 24) / med / 1e6, 1), "cores": 1,
                        "kind": "port",
                        "note": "float32 values U[0,10), pivot 5; the oracle's restatement of "
                                "src/cpu_scan.h:20-50, pinned element for element to the "
                                "reference's own header in the CPU tests"}
    oracle.set_threads(1)
    p1, v1, m1 = hhs.ic_plummer(1024, 1, 0, True, 0.0)
    e0 = oracle.energy(p1, v1, m1, 0.0)
    host = oracle.HostParticles(1024, p1, v1, m1)
    t0 = time.perf_counter()
    rc, ost = oracle.hold(host, 1.0, 1.0 / 16, ETA, 0.0)
    t_cpu = time.perf_counter() - t0
    gpu = {}
    for label, kw in (("fp64", {"kick_precision": hhs.KICK_F64_EXACT}), ("default", {}),
                      ("fp32_timesteps", {"fast_timesteps": hhs.TIMESTEP_FP32})):
        d1 = hhs.DeviceParticles.from_host(p1, v1, m1)
        g1 = hhs.Globals(1024, ETA, 0.0, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g1.hold(d1, 1.0, 1.0 / 16)
        torch.cuda.synchronize()
        gpu[label] = {"seconds": round(time.perf_counter() - t0, 3), "calls": g1.stats()["calls"],
                      "dE_over_E": abs(oracle.energy(d1.pos.cpu().numpy(), d1.vel.cpu().numpy(),
                                                     m1, 0.0) / e0 - 1.0)}
        if label == "fp64":
            gpu[label]["bit_identical_to_cpu"] = bool(
                d1.pos.cpu().numpy().tobytes() == host.pos.tobytes()
                and d1.vel.cpu().numpy().tobytes() == host.vel.tobytes())
        g1.close()
    out["hold_n1024"] = {"cpu_seconds": round(t_cpu, 3), "cpu_threads": 1, "rc": rc,
                         "calls": ost.calls, "levels": ost.max_level + 1,
                         "interactions": ost.interactions,
                         "dE_over_E": abs(oracle.energy(host.pos, host.vel, host.m, 0.0) / e0
                                          - 1.0),
                         "gpu": gpu,
                         "gpu_modes": {"fp64": "exact kick and timesteps (FP64-only)",
                                       "default": "mixed kick, EXACT timesteps (the library's "
                                                  "default globals)",
                                       "fp32_timesteps": "mixed kick, FP32 timesteps (the mode "
                                                         "of the other HOLD legs)"}}
    return out


def hold_bench(hhs, torch, pos, vel, m, dt, label, eps=EPS, t_end=None, **modes):
    """One HOLD root step of dt (or root steps of dt to t_end) of the given
    system; returns the summary and the final positions (for the comparison
    with the FP64-only run)."""
    d2 = hhs.DeviceParticles.from_host(pos, vel, m)
    g = hhs.Globals(len(m), ETA, eps, **modes)
    ke0, pe0 = hhs.energy(d2.pos, d2.vel, d2.m, eps * eps)
    torch.cuda.synchronize()
    h0 = time.perf_counter()
    if t_end is None:
        g.hold_step(d2, dt)
    else:
        g.hold(d2, t_end, dt)
    torch.cuda.synchronize()
    hs = time.perf_counter() - h0
    ke1, pe1 = hhs.energy(d2.pos, d2.vel, d2.m, eps * eps)
    st = g.stats()
    g.close()
    return {"dt": dt, "mode": label, "seconds": round(hs, 4), "calls": st["calls"],
            "levels": st["max_level"] + 1, "interactions": st["interactions"],
            "ts_pairs": st["ts_pairs"],
            "dE_over_E": abs((ke1 + pe1 - ke0 - pe0) / (ke0 + pe0))}, d2.pos.cpu().numpy()


def hold16k_bench(hhs, torch, fp64=True):
    """Plain HOLD of the N = 16384 equal-mass Plummer sphere (seed 2, eps =
    1e-4, eta = 0.01) to t = 0.125 in one root step (about ten power-of-two
    levels): the default mode (MIXED32 kicks, FP32 timesteps) beside the
    FP64-only mode, which is bit-identical with the FP64 CPU HOLD and gives
    the reference energy error."""
    n, eps, t_end = 16384, 1e-4, 0.125
    pos, vel, m = hhs.ic_plummer(n, 2, 0, True, eps * eps)
    out, p_def = hold_bench(hhs, torch, pos, vel, m, t_end, "mixed kick, FP32 timesteps",
                            eps=eps, fast_timesteps=hhs.TIMESTEP_FP32)
    out.update({"n": n, "eps": eps, "t_end": t_end})
    if not fp64:
        return out
    ref, p64 = hold_bench(hhs, torch, pos, vel, m, t_end, "FP64-only (exact kick and timesteps)",
                          eps=eps, kick_precision=hhs.KICK_F64_EXACT)
    ref["max_pos_diff_over_rms"] = float(np.abs(p_def - p64).max() / np.sqrt((p64 ** 2).mean()))
    out.update({"fp64_reference": ref,
                "dE_ratio_to_fp64": out["dE_over_E"] / ref["dE_over_E"]})
    return out


def tune_timesteps(hhs, torch, sizes=tuple(1 << k for k in range(8, 16)), reps=5):
    """Autotunes the one-sided FP32 timestep kernel (waves per group, source
    chunks) for the size classes 2^8 .. 2^15 that HOLD's levels below the
    symmetric kernel use; reports the heuristic's time beside the tuned one.
    Both are timed the same way, launches from a Python loop between two
    synchronisations (below ~4096 particles that loop, not the kernel, sets
    the time); the autotuner's own figure, from device events around
    back-to-back launches from C, is `tuned_device_ms`."""
    rows = []

    def loop_ms(d):
        for _ in range(2):
            hhs.timesteps(d, eta=ETA, eps2=EPS * EPS, fast=hhs.TIMESTEP_FP32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            hhs.timesteps(d, eta=ETA, eps2=EPS * EPS, fast=hhs.TIMESTEP_FP32)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    for n in sizes:
        p, v, m = hhs.ic_plummer(n, 777, 0, True, EPS * EPS)
        d = hhs.DeviceParticles.from_host(p, v, m)
        hw, hk = hhs.timesteps_get_config(n, hhs.TIMESTEP_FP32)
        h_ms = loop_ms(d)
        w, k, ms = hhs.autotune_timesteps(n, hhs.TIMESTEP_FP32, reps)
        rows.append({"n": n, "heuristic": [hw, hk], "heuristic_ms": round(h_ms, 4),
                     "tuned": [w, k], "tuned_ms": round(loop_ms(d), 4),
                     "tuned_device_ms": round(ms, 4)})
        del d
    return {"mode": "fp32", "config": "[waves per group, source chunks]",
            "timing": "heuristic_ms and tuned_ms: Python launch loop; tuned_device_ms: device "
                      "events around launches from C",
            "classes": rows}


def timesteps_bench(hhs, torch, d, stream, reps=3):
    """The timestep kernels over all N = 131072 particles: the one-sided
    kernel in its three modes and the symmetric kernel (each unordered pair
    once) in EXACT and FP32, priced at 49 flops per evaluated pair (the
    reference formula counting each +, -, *, /, sqrt and min as one)."""
    out = []
    pairs = float(N) * (N - 1)
    import ctypes as C
    nbytes = int(hhs.raw().hhs_timesteps_sym_workspace_bytes(N))
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    s = d.struct()

    def sym(mode):
        hhs.check(hhs.raw().hhs_timesteps_sym_f64(C.byref(s), N, ETA, EPS * EPS, mode,
                                                  C.c_void_p(ws.data_ptr()), nbytes,
                                                  C.c_void_p(stream.cuda_stream)))

    for name, mode, peak, symmetric in (
            ("exact", 0, FP64_PEAK_TFLOPS, False), ("fast", 1, FP64_PEAK_TFLOPS, False),
            ("fp32", 2, FP32_PEAK_TFLOPS, False), ("exact_symmetric", 0, FP64_PEAK_TFLOPS, True),
            ("fp32_symmetric", 2, FP32_PEAK_TFLOPS, True)):
        def go():
            if symmetric:
                sym(mode)
            else:
                hhs.timesteps(d, eta=ETA, eps2=EPS * EPS, fast=mode)
        go()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(reps):
            go()
        ev1.record(stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        # flops at the evaluated pairs: ordered pairs for the one-sided kernel,
        # unordered pairs (half) for the symmetric one
        evaluated = pairs / 2 if symmetric else pairs
        tf = TS_FLOPS_PER_PAIR * evaluated / (ms * 1e-3) / 1e12
        out.append({"mode": name, "ms": round(ms, 3),
                    "Gpairs_per_s": round(pairs / (ms * 1e-3) / 1e9, 1),
                    "evaluated_pairs": "unordered" if symmetric else "ordered",
                    "TFLOP_per_s": round(tf, 2), "peak": peak, "frac": round(tf / peak, 4)})
    return {"n": N, "pairs": "Gpairs_per_s counts ordered pairs (i != j) for every mode",
            "flops_per_pair": TS_FLOPS_PER_PAIR, "modes": out}


def stream_bench(hhs, torch, stream, sizes=(N, 1 << 20, 1 << 22, 1 << 24), reps=10):
    """HBM-bound kernels at the HOLD size and a large-N sweep, priced at
    their algorithmic bytes: fused split 148 B/particle (8 B timestep for
    the predicate + 68 B record read + 72 B record and parent_id written),
    stand-alone scan 8 B/element, drift 72 B/particle, update 132 B/particle
    (68 B read + 64 B scattered write; timed both as the scatter-back of a
    split's halves, the HOLD pattern, and under a random permutation)."""
    rows = []

    def timed(fn):
        fn()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(reps):
            fn()
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / reps

    for n in sizes:
        gen = torch.Generator(device="cuda").manual_seed(n)

        def rand_set():
            p = hhs.DeviceParticles.empty(n)
            for f in ("pos", "vel", "m", "timestep"):
                t = getattr(p, f)
                t.copy_(torch.rand(t.shape, generator=gen, dtype=torch.float64, device="cuda"))
            p.id.copy_(torch.arange(n, dtype=torch.int32, device="cuda"))
            return p

        tot, cur = rand_set(), rand_set()
        cur.parent_id.copy_(torch.randperm(n, generator=gen, device="cuda").to(torch.int32))
        slow, fast = hhs.DeviceParticles.empty(n), hhs.DeviceParticles.empty(n)
        prev = hhs.DeviceParticles.empty(n)
        ws = hhs.ScanWorkspace(n)
        part = torch.zeros(2, dtype=torch.int32, device="cuda")
        pred = torch.randint(0, 2, (n,), generator=gen, dtype=torch.int32, device="cuda")
        scanned = torch.empty_like(pred)
        row = {"n": n}
        hhs.split(tot, slow, fast, 0.5, part, ws)
        torch.cuda.synchronize()
        ns, nf = (int(x) for x in part.cpu())

        def update_split():  # the HOLD pattern: scatter back both halves of a split
            hhs.update(prev, slow, n=ns)
            hhs.update(prev, fast, n=nf)

        for name, fn, bpp in (
                ("split", lambda: hhs.split(tot, slow, fast, 0.5, part, ws), 148),
                ("scan", lambda: hhs.scan_exclusive(pred, scanned, ws), 8),
                ("drift", lambda: hhs.drift(tot.pos, tot.vel, dt=1e-9, fac=0.5), 72),
                ("update", update_split, 132),
                ("update_random_perm", lambda: hhs.update(prev, cur), 132)):
            ms = timed(fn)
            gbs = bpp * n / (ms * 1e-3) / 1e9
            row[name] = {"us": round(ms * 1e3, 2), "GB_per_s": round(gbs, 1),
                         "frac": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"stream kernels n={n}: {row}")
        rows.append(row)
        del tot, cur, slow, fast, prev, ws, pred, scanned
    return {"bytes_per_particle": {"split": 148, "scan": 8, "drift": 72, "update": 132},
            "peak_GB_per_s": HBM_PEAK_GBS, "sizes": rows}


def split_sweep(hhs, torch, stream, log2_sizes=range(10, 23), events=10):
    """The reference's split timing design (perf/perf_scan_multi_block.cu:186-307):
    float values U[0,10), pivot 5, N = 2^10 .. 2^22, the GPU split of the value
    array (hhs_split_values_f32, split.cuh naming: values <= pivot are fast)
    beside cpu_split (src/cpu_scan.h:20-50; the oracle's restatement, one host
    thread, in place) on the same values.  GPU: mean of `events` timed calls
    (HIP events on the launch stream) after one warm-up; CPU: median of
    `events` calls, each on a fresh copy.  Also the HOLD record split (148 B
    per particle) at the same sizes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.set_threads(1)
    rows = []
    for k in log2_sizes:
        n = 1 << k
        vals = (10.0 * np.random.default_rng(k).random(n)).astype(np.float32)
        dv = torch.as_tensor(vals).cuda()
        slow, fast = torch.empty_like(dv), torch.empty_like(dv)
        part = torch.zeros(2, dtype=torch.int32, device="cuda")
        ws = hhs.ScanWorkspace(n)
        hhs.split_values(dv, slow, fast, 5.0, part, ws)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(events):
            hhs.split_values(dv, slow, fast, 5.0, part, ws)
        ev1.record(stream)
        torch.cuda.synchronize()
        g_us = ev0.elapsed_time(ev1) / events * 1e3
        ns_gpu = int(part[0].item())
        # the HOLD record split on a random set of the same size
        tot = hhs.DeviceParticles.empty(n)
        gen = torch.Generator(device="cuda").manual_seed(k)
        for f in ("pos", "vel", "m", "timestep"):
            t = getattr(tot, f)
            t.copy_(torch.rand(t.shape, generator=gen, dtype=torch.float64, device="cuda"))
        rs, rf = hhs.DeviceParticles.empty(n), hhs.DeviceParticles.empty(n)
        hhs.split(tot, rs, rf, 0.5, part, ws)
        ev0.record(stream)
        for _ in range(events):
            hhs.split(tot, rs, rf, 0.5, part, ws)
        ev1.record(stream)
        torch.cuda.synchronize()
        r_us = ev0.elapsed_time(ev1) / events * 1e3
        times = []
        for _ in range(events):
            arr = vals.copy()
            t0 = time.perf_counter()
            ns_cpu, nf_cpu, _ = oracle.cpu_split_inplace_f32(arr, 5.0)
            times.append(time.perf_counter() - t0)
        c_us = float(np.median(times)) * 1e6
        assert ns_cpu == ns_gpu, (n, ns_cpu, ns_gpu)
        rows.append({"n": n, "gpu_split_values_us": round(g_us, 2),
                     "gpu_split_records_us": round(r_us, 2), "cpu_split_us": round(c_us, 2),
                     "cpu_over_gpu": round(c_us / g_us, 2),
                     "gpu_records_GB_per_s": round(148 * n / (r_us * 1e-6) / 1e9, 1)})
        del dv, slow, fast, ws, tot, rs, rf
    return {"values": "float32 U[0,10), pivot 5", "events": events,
            "cpu": "cpu_split restated (oracle/hhs_oracle.c), 1 thread, in place",
            "sizes": rows}


def ensemble_bench(hhs, torch, n_systems=4096, loop_sample=12, fp64_sample=6):
    """Ensemble workload: 4096 Plummer systems, N log-uniform in [8, 2048],
    equal masses, eta = 0.01, t = 1, eps = 1e-4, root steps of 1/16, FP32
    timesteps, one persistent launch."""
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(8), np.log(2048), n_systems)).astype(int)
    systems = [hhs.ic_plummer(int(n), 4000 + k, 0, True, EPS * EPS)
               for k, n in enumerate(sizes)]
    t_end, dtr = 1.0, 1.0 / 16
    ens = hhs.Ensemble(systems, t_end, dtr, ETA, EPS, fast_timesteps=hhs.TIMESTEP_FP32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    r = ens.results()
    de = np.abs((r["e1"] - r["e0"]) / r["e0"])
    ok = (r["status"] == 0)
    # the plain loop: the same systems one at a time through hhs_hold_f64
    pick = (np.argsort(sizes)[np.linspace(0, n_systems - 1, loop_sample).astype(int)]
            if loop_sample > 0 else [])
    loop_t = []
    for k in pick:
        p, v, mm = systems[k]
        dd = hhs.DeviceParticles.from_host(p, v, mm)
        g = hhs.Globals(len(mm), ETA, EPS, fast_timesteps=hhs.TIMESTEP_FP32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.hold(dd, t_end, dtr)
        torch.cuda.synchronize()
        loop_t.append(time.perf_counter() - t0)
        g.close()
    loop_per_system = float(np.mean(loop_t)) if loop_t else float("nan")
    sps = n_systems / wall
    # Systems that miss |dE/E| <= 1e-6, integrated again one at a time by the
    # FP64-only HOLD (exact kick and timesteps): the same error there means it
    # is the integrator's at this eta, not the mixed precision's.
    fail = np.nonzero(ok & (de > 1e-6) & (sizes <= 400))[0]
    fail = fail[np.argsort(sizes[fail])]
    fp64_check = []
    for k in (fail[np.linspace(0, len(fail) - 1, fp64_sample).astype(int)]
              if len(fail) and fp64_sample > 0 else []):
        p, v, mm = systems[k]
        dd = hhs.DeviceParticles.from_host(p, v, mm)
        g = hhs.Globals(len(mm), ETA, EPS, kick_precision=hhs.KICK_F64_EXACT)
        e0 = sum(hhs.energy(dd.pos, dd.vel, dd.m, EPS * EPS))
        g.hold(dd, t_end, dtr)
        e1 = sum(hhs.energy(dd.pos, dd.vel, dd.m, EPS * EPS))
        g.close()
        fp64_check.append({"n": int(sizes[k]), "ensemble_dE": float(de[k]),
                           "fp64_dE": abs((e1 - e0) / e0)})
    return {"systems": n_systems, "n_range": [8, 2048], "t_end": t_end, "dt_root": dtr,
            "eta": ETA, "eps": EPS, "timesteps": "fp32", "seconds": round(wall, 3),
            "systems_per_s": round(sps, 2),
            "systems_per_s_at_dE_le_1e-6": round(float(np.sum(ok & (de <= 1e-6))) / wall, 2),
            "frac_dE_le_1e-6": round(float(np.mean(de <= 1e-6)), 4),
            "dE_over_E": {"median": float(np.median(de)), "p99": float(np.quantile(de, 0.99)),
                          "max": float(de.max())},
            "status_ok": int(ok.sum()),
            "interactions": int(r["interactions"].sum()), "ts_pairs": int(r["ts_pairs"].sum()),
            "calls": int(r["calls"].sum()),
            "decisions": "exact (FP32 timesteps within 1e-5 of a pivot recomputed in FP64)",
            "plain_loop": {"systems_sampled": loop_sample,
                           "sample": "evenly spaced in sorted N (stratified)",
                           "seconds_per_system": round(loop_per_system, 4),
                           "systems_per_s": round(1.0 / loop_per_system, 3) if loop_t else None},
            "speedup_vs_plain_loop": round(sps * loop_per_system, 1) if loop_t else None,
            "fp64_check_of_misses": {
                "sample": "systems with N <= 400 that miss 1e-6, evenly spaced in N, "
                          "rerun by the FP64-only HOLD (exact kick and timesteps)",
                "systems": fp64_check}}


def serve_bench(hhs, rates, n_req, n_lo, n_hi, seed=5, batch_streams=0, big_workers=0,
                small_max_n=0, progress=False):
    """Served workload: Poisson arrivals of Plummer integrations with N
    log-uniform in [n_lo, n_hi], end time uniform in [0.1, 1], eta = 0.01,
    through the continuous-batching server; time-to-solution = completion -
    arrival (server clock)."""
    out = []
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n_req):
        n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        p, v, m = hhs.ic_plummer(n, 9000 + k, 0, True, EPS * EPS)
        reqs.append((p, v, m, float(rng.uniform(0.1, 1.0))))
    for rate in rates:
        gaps = np.random.default_rng(seed + int(rate * 1000)).exponential(1.0 / rate, n_req)
        with hhs.Server(ETA, EPS, hbm_budget_bytes=64 << 30, small_max_n=small_max_n,
                        max_n=n_hi, batch_streams=batch_streams, big_workers=big_workers,
                        fast_timesteps=hhs.TIMESTEP_FP32) as srv:
            tickets = []
            t0 = time.perf_counter()
            t_next = t0
            for (p, v, m, te), gap in zip(reqs, gaps):
                t_next += gap
                delay = t_next - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                tickets.append(srv.submit(p, v, m, te, 1.0 / 16))
            infos = []
            for k, t in enumerate(tickets):
                infos.append(srv.wait(t))
                if progress:
                    log(f"served rate {rate}: request {k} (N={len(reqs[k][2])}) done after "
                        f"{infos[-1]['t_done'] - infos[-1]['t_submit']:.2f} s")
            wall = time.perf_counter() - t0
            peak = srv.stats()["peak_bytes"]
        tts = np.array([i["t_done"] - i["t_submit"] for i in infos])
        wait = np.array([i["t_start"] - i["t_submit"] for i in infos])
        run = np.array([i["t_done"] - i["t_start"] for i in infos])
        de = np.array([abs((i["result"]["e1"] - i["result"]["e0"]) / i["result"]["e0"])
                       for i in infos])
        out.append({"rate_per_s": rate, "requests": n_req,
                    "ok": int(sum(i["state"] == 3 for i in infos)),
                    "p50_s": round(float(np.median(tts)), 4),
                    "p99_s": round(float(np.quantile(tts, 0.99)), 4),
                    "queue_wait_mean_s": round(float(wait.mean()), 4),
                    "run_mean_s": round(float(run.mean()), 4),
                    "throughput_req_per_s": round(n_req / wall, 3),
                    "wall_s": round(wall, 3), "peak_admitted_bytes": int(peak),
                    "frac_dE_le_1e-6": round(float(np.mean(de <= 1e-6)), 4)})
    return {"n_range": [n_lo, n_hi], "t_end_range": [0.1, 1.0], "dt_root": 1.0 / 16,
            "eta": ETA, "eps": EPS, "timesteps": "fp32", "small_max_n": small_max_n or 512,
            "batch_streams": batch_streams or 2, "big_workers": big_workers or 4, "sweep": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-hold", action="store_true", help="skip both HOLD legs")
    ap.add_argument("--no-root-step", action="store_true",
                    help="skip the N=131072 HOLD root step")
    ap.add_argument("--no-hold16k", action="store_true", help="skip the N=16384 plain HOLD")
    ap.add_argument("--no-timesteps", action="store_true",
                    help="skip the N=131072 timestep kernels")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--no-tune", action="store_true", help="use the built-in kick heuristic")
    ap.add_argument("--no-stream", action="store_true", help="skip the split/scan/drift/update sweep")
    ap.add_argument("--no-fp64-hold", action="store_true",
                    help="skip the FP64-only HOLD root step (the energy reference)")
    ap.add_argument("--no-cpu-extra", action="store_true",
                    help="skip the all-threads, cpu_split and N=1024 CPU HOLD timings")
    ap.add_argument("--loop-sample", type=int, default=12,
                    help="systems timed through the plain per-system loop (0 skips it)")
    ap.add_argument("--fp64-sample", type=int, default=6,
                    help="ensemble systems missing |dE/E| <= 1e-6 rerun by the FP64-only HOLD")
    ap.add_argument("--serve-requests", type=int, default=16,
                    help="requests per arrival rate in the served leg (N in [64, 65536])")
    ap.add_argument("--serve-rates", default="1,4",
                    help="comma-separated Poisson arrival rates (per s) of the served leg")
    ap.add_argument("--no-split-sweep", action="store_true",
                    help="skip the GPU split vs cpu_split sweep over N = 2^10 .. 2^22")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="first run one device-to-device copy of MB MiB (a kernel of known "
                         "HBM traffic for scaling profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only; the kick would shard by sinks "
              "across GPUs but no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)

    import torch
    import hhs

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.calibrate_copy > 0:
        src = torch.ones(args.calibrate_copy << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy} MiB read + {args.calibrate_copy} MiB written")
        del src, dst
    eps2 = EPS * EPS
    t0 = time.perf_counter()
    pos, vel, m = hhs.ic_plummer(N, SEED, 1, True, eps2)
    t_ic = time.perf_counter() - t0
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    stream = torch.cuda.current_stream()
    dt = 1.0 / 64

    tune = None
    if not args.no_tune:
        t0 = time.perf_counter()
        s_t, w_t, ms_t = hhs.autotune_kick(N, 10)
        tune = {"sinks_per_lane": s_t, "waves_per_group": w_t, "ms": round(ms_t, 4),
                "sweep_seconds": round(time.perf_counter() - t0, 2)}
        log(f"autotune: {tune}")
    for _ in range(args.warmup):
        hhs.kick_self(d, dt=dt, fac=1.0, eps2=eps2)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
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
    log(f"kick: {ms:.3f} ms, {gint:.1f} Ginteractions/s")

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
                   "n": N, "eps": EPS, "sinks_per_lane": s_cfg, "waves_per_group": w_cfg,
                   "autotune": tune},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "traffic": load_traffic(),
                     "flops_per_interaction": FLOPS_PER_INTERACTION,
                     "kernel": "kick_kernel<double,S,W,0,true>"},
        "wall_s_timed": round(wall, 4),
        "ic_seconds": round(t_ic, 2),
    }
    acc_gpu = d.acc.cpu().numpy()
    if not args.no_timesteps:
        out["timesteps"] = timesteps_bench(hhs, torch, d, stream)
        log(f"timesteps: {out['timesteps']}")
    del d
    if not args.no_stream:
        out["stream_kernels"] = stream_bench(hhs, torch, stream)
    if not args.no_split_sweep and not args.no_cpu:
        out["split_sweep"] = split_sweep(hhs, torch, stream)
        log(f"split sweep: {out['split_sweep']}")
    if not args.no_tune:
        out["timestep_autotune"] = tune_timesteps(hhs, torch)
        log(f"timestep autotune: {out['timestep_autotune']}")
    if not args.no_hold and not args.no_root_step:
        out["hold_root_step"], p_mixed = hold_bench(
            hhs, torch, pos, vel, m, dt, "mixed kick, FP32 timesteps",
            fast_timesteps=hhs.TIMESTEP_FP32)
        log(f"hold: {out['hold_root_step']}")
        if not args.no_fp64_hold:
            ref, p64 = hold_bench(hhs, torch, pos, vel, m, dt, "FP64-only (exact kick and "
                                  "timesteps)", kick_precision=hhs.KICK_F64_EXACT)
            ref["max_pos_diff_over_rms"] = float(np.abs(p_mixed - p64).max()
                                                 / np.sqrt((p64 ** 2).mean()))
            out["hold_root_step"]["fp64_reference"] = ref
            log(f"hold fp64: {ref}")
    if not args.no_hold and not args.no_hold16k:
        out["hold_n16384"] = hold16k_bench(hhs, torch, fp64=not args.no_fp64_hold)
        log(f"hold n=16384: {out['hold_n16384']}")
    if not args.no_ensemble:
        out["ensemble"] = ensemble_bench(hhs, torch, loop_sample=args.loop_sample,
                                         fp64_sample=args.fp64_sample)
        log(f"ensemble: {out['ensemble']}")
    if not args.no_serve:
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        out["served"] = serve_bench(hhs, rates, args.serve_requests, 64, 65536, progress=True)
        log(f"served: {out['served']}")
    if not args.no_cpu:
        cb, ferr = cpu_baseline(pos, m, acc_gpu, eps2)
        out["cpu_baseline"] = cb
        out["force_error"] = ferr
        log(f"cpu baseline: {cb}")
        if not args.no_cpu_extra:
            out["cpu_reference"] = cpu_reference(hhs, torch, pos, m, eps2)
            log(f"cpu reference: {out['cpu_reference']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
