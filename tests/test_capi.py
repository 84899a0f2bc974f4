"""C ABI checks that need no GPU: the library loads, exports every function
declared in include/hhs.h, rejects bad arguments with a status code and a
message, and its host-only entry points (initial conditions) behave."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hhs.h")
LIB = os.path.join(ROOT, "parallel-hhs_amd", "lib", "libhhs.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(hhs_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("hhs_kick_self_f64", "hhs_kick_cross_f64", "hhs_timesteps_f64",
                 "hhs_split_f64", "hhs_scan_exclusive_i32", "hhs_drift_f64",
                 "hhs_update_f64", "hhs_hold_step_f64", "hhs_hold_f64", "hhs_energy_f64",
                 "hhs_globals_create", "hhs_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (hhs_[a-z0-9_]+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_every_declared_symbol(hhs):
    missing = [n for n in declared_functions() if n not in hhs.EXPORTED]
    assert not missing, f"hhs.py lacks: {missing}"
    lib = hhs.raw()
    for n in declared_functions():
        assert getattr(lib, n) is not None


def test_version_and_empty_error(hhs):
    assert b"gfx950" in hhs.raw().hhs_version()


def test_invalid_arguments_return_status(hhs):
    import ctypes as C
    lib = hhs.raw()
    assert lib.hhs_kick_set_config(3, 1) == hhs.HHS_ERR_INVALID
    assert "sinks_per_lane" in hhs.last_error()
    assert lib.hhs_kick_set_config(2, 16) == hhs.HHS_ERR_INVALID
    h = C.c_void_p()
    assert lib.hhs_globals_create(C.byref(h), 0, 0.01, 0.0, 0) == hhs.HHS_ERR_INVALID
    assert lib.hhs_globals_create(C.byref(h), 10, -1.0, 0.0, 0) == hhs.HHS_ERR_INVALID
    assert lib.hhs_globals_create(C.byref(h), 10, 0.01, 0.0, 64) == hhs.HHS_ERR_INVALID
    assert lib.hhs_kick_self_f64(None, 10, 1.0, 1.0, 0.0, None) == hhs.HHS_ERR_INVALID
    assert lib.hhs_split_f64(None, None, None, None, None, 1.0, 5, None, None, 0,
                             None) == hhs.HHS_ERR_INVALID
    assert lib.hhs_hold_step_f64(None, None, 1, 0, 0.0, 1.0, 1.0, 1, None) == hhs.HHS_ERR_INVALID
    assert lib.hhs_ic_plummer(0, 1, 0, 0, 0.0, None, None, None) == hhs.HHS_ERR_INVALID
    assert "ic_plummer" in hhs.last_error()
    assert lib.hhs_globals_set_kick_precision(None, 3) == hhs.HHS_ERR_INVALID
    assert lib.hhs_kick_set_precision(4) == hhs.HHS_ERR_INVALID
    assert lib.hhs_kick_set_precision(-1) == hhs.HHS_ERR_INVALID
    assert lib.hhs_kick_set_precision(hhs.KICK_F64_EXACT) == hhs.HHS_OK
    assert lib.hhs_kick_get_precision() == hhs.KICK_F64_EXACT
    assert lib.hhs_kick_set_precision(hhs.KICK_MIXED32) == hhs.HHS_OK
    assert lib.hhs_scan_workspace_bytes(0) >= 64
    assert lib.hhs_scan_workspace_bytes(1 << 20) == 64 + 8 * (1 << 12)


def test_kick_config_heuristic(hhs):
    s, w = hhs.kick_get_config(131072, 131072)
    assert (s, w) == (1, 8)
    s, w = hhs.kick_get_config(100, 100)
    assert s == 1 and 1 <= w <= 8
    hhs.kick_set_config(4, 2)
    assert hhs.kick_get_config(131072, 131072) == (4, 2)
    hhs.kick_set_config(0, 0)
    assert hhs.kick_get_config(131072, 131072) == (1, 8)


# ------------------------------------------------------ initial conditions
def test_plummer_deterministic_and_centred(hhs):
    a = hhs.ic_plummer(2000, 42, 0, True, 0.0)
    b = hhs.ic_plummer(2000, 42, 0, True, 0.0)
    c = hhs.ic_plummer(2000, 43, 0, True, 0.0)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    pos, vel, m = a
    assert np.allclose(m, 1 / 2000)
    assert np.abs((m[:, None] * pos).sum(0)).max() < 1e-14
    assert np.abs((m[:, None] * vel).sum(0)).max() < 1e-14


def test_plummer_henon_units_and_virial(hhs, oracle):
    pos, vel, m = hhs.ic_plummer(4096, 7, 0, True, 0.0)
    ke, pe = oracle.kinetic(vel, m), oracle.potential(pos, m)
    assert ke / -pe == pytest.approx(0.5, rel=1e-12)
    # Henon units: E = -1/4 for the model; a 4096-body sample is within a few %
    assert ke + pe == pytest.approx(-0.25, rel=0.06)
    r = np.sort(np.linalg.norm(pos, axis=1))
    # Plummer half-mass radius 1.3048 a, a = 3 pi / 16 in Henon units
    assert r[2047] == pytest.approx(1.3048 * 3 * np.pi / 16, rel=0.05)
    # radii beyond 10 scale lengths are redrawn (the recentring shift is tiny)
    assert r[-1] <= 10 * 3 * np.pi / 16 * 1.001


def test_salpeter_masses(hhs):
    _, _, m = hhs.ic_plummer(20000, 3, 1, False, 0.0)
    assert m.sum() == pytest.approx(1.0, rel=1e-12)
    ratio = m.max() / m.min()
    assert 50 < ratio <= 1000.0 + 1e-9  # [0.1, 100] before normalisation
    # slope of dN/dlog m: -1.35 for a Salpeter IMF
    hist, edges = np.histogram(np.log10(m / m.min() * 0.1), bins=np.linspace(-1, 1, 9))
    slope = np.polyfit(0.5 * (edges[1:] + edges[:-1]), np.log10(hist), 1)[0]
    assert slope == pytest.approx(-1.35, abs=0.1)


def test_autotune_cache_round_trip(hhs, tmp_path):
    """The tuning cache is plain text; loading it installs per-size-class
    kick configurations (host logic only, no GPU)."""
    path = tmp_path / "tune.txt"
    path.write_text("kick 17 4 2 1.0\nkick 12 1 8 0.5\nkick 17 2 4 0.9\ngarbage line\n")
    try:
        hhs.autotune_load(str(path))
        assert hhs.kick_get_config(131072, 131072) == (2, 4)   # class 17, later line wins
        assert hhs.kick_get_config(4096, 4096) == (1, 8)       # class 12
        hhs.kick_set_config(4, 1)                              # explicit setting overrides
        assert hhs.kick_get_config(131072, 131072) == (4, 1)
        hhs.kick_set_config(0, 0)
    finally:
        hhs.autotune_clear()
    assert hhs.kick_get_config(131072, 131072) == (1, 8)
    bad = tmp_path / "bad.txt"
    bad.write_text("nothing here\n")
    with pytest.raises(hhs.HHSError):
        hhs.autotune_load(str(bad))


def test_timestep_autotune_cache_round_trip(hhs, tmp_path):
    """"ts <mode> <class> <W> <K> <ms>" lines install one-sided timestep
    configurations per mode and size class; invalid entries are ignored and
    clearing returns to the heuristic (host logic only, no GPU)."""
    base_fp32 = hhs.timesteps_get_config(3000, hhs.TIMESTEP_FP32)
    base_exact = hhs.timesteps_get_config(3000, hhs.TIMESTEP_EXACT)
    path = tmp_path / "tune.txt"
    path.write_text("ts 2 11 2 16 0.01\nts 2 11 4 8 0.009\nts 0 11 8 2 0.03\n"
                    "ts 2 11 3 8 0.001\nts 7 11 2 2 0.001\nkick 12 1 8 0.5\n")
    try:
        hhs.autotune_load(str(path))
        assert hhs.timesteps_get_config(3000, hhs.TIMESTEP_FP32) == (4, 8)   # later line wins
        assert hhs.timesteps_get_config(2048, hhs.TIMESTEP_EXACT) == (8, 2)  # same class 11
        assert hhs.timesteps_get_config(4096, hhs.TIMESTEP_FP32) == \
            hhs.timesteps_get_config(4096, hhs.TIMESTEP_FAST)                # class 12 untouched
        assert hhs.kick_get_config(4096, 4096) == (1, 8)
    finally:
        hhs.autotune_clear()
    assert hhs.timesteps_get_config(3000, hhs.TIMESTEP_FP32) == base_fp32
    assert hhs.timesteps_get_config(3000, hhs.TIMESTEP_EXACT) == base_exact


def test_one_hip_runtime_after_importing_torch(hhs):
    """hhs.py maps PyTorch's bundled HIP runtime before libhhs, so the
    process holds one runtime whichever of the two is imported first."""
    code = ("import sys; sys.path.insert(0, %r); import hhs, torch; hhs.runtime_check();"
            "maps = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l};"
            "assert len(maps) == 1, maps; print('ok')") % os.path.join(ROOT, "parallel-hhs_amd")
    r = subprocess.run(["python3", "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr


def test_runtime_check_reports_two_runtimes():
    """Loading libhhs on its own (RUNPATH: /opt/rocm's runtime) and then
    torch maps two HIP runtimes: hhs_runtime_check says so, naming both,
    instead of the first device call crashing.  No device call is made."""
    code = ("import ctypes; lib = ctypes.CDLL(%r); import torch;"
            "lib.hhs_last_error.restype = ctypes.c_char_p;"
            "rc = lib.hhs_runtime_check(); print(rc, lib.hhs_last_error().decode())") % LIB
    r = subprocess.run(["python3", "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rc, msg = r.stdout.strip().split(" ", 1)
    assert int(rc) == 3 and "two HIP runtimes" in msg and msg.count("libamdhip64") == 2
