#!/usr/bin/env python3
"""Writes the golden fixtures of the HOLD hot path into tests/golden/hold_golden.npz.

The reference ships no golden vectors or known-answer tests (its scan tests
check properties, tests/test_scan_unrolled.cu:191-215; tests/test_nbforce.cu
asserts nothing), so these vectors come from the FP64 CPU oracle (oracle/),
whose functions are pinned by closed forms in tests/test_oracle.py.  They
freeze, on seeded inputs of at most 1024 particles:

  * Plummer initial conditions (hhs_ic_plummer): N = 8, 64, 1024, equal
    masses m = 1/N, seed 1 (the N = 1024 workload of the plain CPU HOLD),
    and N = 64 with Salpeter masses, seed 3;
  * accelerations (src/kick.cuh:15-32 with the attractive sign of
    tests/test_nbforce.cu:14-18) at eps^2 = 0 and 1e-8;
  * timesteps (src/timestep.cuh:17-60), eta = 0.01;
  * the stable split of the N = 1024 set by its own timesteps at pivots
    2^-3 ... 2^-14 -- all fast, mixed, two fast, none fast
    (src/split.cuh:304-370): counts and parent ids;
  * one HOLD root step of dt = 1/32 (src/hold.cuh:51-181), eta = 0.01,
    eps = 0, of the N = 64 and N = 1024 sets: positions, velocities,
    timesteps and calls per level;
  * exclusive scans of seeded 0/1 predicates at the reference's scan-test
    sizes (512 ... 588021): totals and SHA-256 digests of the outputs.

Everything is stored as plain arrays (np.savez_compressed, no pickles).
Run from the repository root after building: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("parallel-hhs_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

OUT = os.path.join(HERE, "hold_golden.npz")
ETA = 0.01
IC_CASES = (("n8", 8, 1, 0), ("n64", 64, 1, 0), ("n1024", 1024, 1, 0), ("n64s", 64, 3, 1))
EPS2S = (0.0, 1e-8)
PIVOT_EXPS = (3, 6, 8, 10, 12, 13, 14)
HOLD_CASES = ("n64", "n1024")
HOLD_DT = 1.0 / 32
SCAN_SIZES = (512, 1023, 2048, 2151, 4090, 588021)


def scan_input(n):
    """Seeded 0/1 predicates (values U[0, 10) against pivot 5, as the
    reference's scan tests draw them)."""
    rng = np.random.default_rng(1000 + n)
    return (rng.random(n) * 10.0 <= 5.0).astype(np.int32)


def digest(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(),
                         dtype=np.uint8).copy()


def generate(hhs, oracle):
    """All fixtures as a dict of arrays (also used by the CPU test to check
    that the oracle still reproduces them)."""
    out = {}
    for name, n, seed, mass_model in IC_CASES:
        pos, vel, m = hhs.ic_plummer(n, seed, mass_model, True, 0.0)
        out[f"{name}_pos"], out[f"{name}_vel"], out[f"{name}_m"] = pos, vel, m
        for k, eps2 in enumerate(EPS2S):
            out[f"{name}_acc_e{k}"] = oracle.accel(pos, pos, m, eps2)
            out[f"{name}_ts_e{k}"] = oracle.timesteps(pos, vel, m, ETA, eps2)
    # split of the N = 1024 set by its own timesteps (eps = 0)
    pos, vel, m = out["n1024_pos"], out["n1024_vel"], out["n1024_m"]
    ts = out["n1024_ts_e0"]
    for p in PIVOT_EXPS:
        host = oracle.HostParticles(len(m), pos, vel, m, ts)
        slow, fast, ns, nf, _, _ = oracle.split(host, 2.0 ** -p)
        out[f"split_p{p}_counts"] = np.array([ns, nf], dtype=np.int32)
        out[f"split_p{p}_fast_parent"] = fast.parent_id[:nf].copy()
        out[f"split_p{p}_slow_parent"] = slow.parent_id[:ns].copy()
    for name in HOLD_CASES:
        host = oracle.HostParticles(len(out[f"{name}_m"]), out[f"{name}_pos"],
                                    out[f"{name}_vel"], out[f"{name}_m"])
        rc, st = oracle.hold_step(host, HOLD_DT, True, ETA, 0.0)
        assert rc == 0
        out[f"hold_{name}_pos"] = host.pos.copy()
        out[f"hold_{name}_vel"] = host.vel.copy()
        out[f"hold_{name}_ts"] = host.timestep.copy()
        out[f"hold_{name}_calls"] = np.array(st.calls_per_level[:st.max_level + 1],
                                             dtype=np.int64)
        out[f"hold_{name}_simtime"] = np.array([st.simtime])
    for n in SCAN_SIZES:
        s = oracle.exclusive_scan(scan_input(n))
        out[f"scan_{n}_total"] = np.array([int(s[-1]) + int(scan_input(n)[-1])], dtype=np.int64)
        out[f"scan_{n}_sha256"] = digest(s)
    return out


def main():
    import hhs
    import oracle
    oracle.set_threads(1)
    data = generate(hhs, oracle)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT}: {len(data)} arrays, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
