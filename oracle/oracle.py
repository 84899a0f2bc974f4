"""ctypes wrapper of the FP64 CPU oracle (liboracle.so) and of the
reference's own CPU partition compiled into _ref/ (when present).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_cpu_scan.so")


def build(quiet: bool = True) -> None:
    """Compile the oracle (and _ref when the reference tree is mounted)."""
    out = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


if not os.path.exists(LIB):
    build()

_lib = C.CDLL(LIB)


class Particles(C.Structure):
    _fields_ = [("pos", C.c_void_p), ("vel", C.c_void_p), ("acc", C.c_void_p),
                ("timestep", C.c_void_p), ("m", C.c_void_p), ("id", C.c_void_p),
                ("parent_id", C.c_void_p), ("n", C.c_int)]


class HoldStats(C.Structure):
    _fields_ = [("simtime", C.c_double), ("calls", C.c_longlong),
                ("interactions", C.c_longlong), ("ts_pairs", C.c_longlong),
                ("max_level", C.c_int), ("calls_per_level", C.c_longlong * 64)]


_vp, _i, _d = C.c_void_p, C.c_int, C.c_double
_PI = C.POINTER(C.c_int)
_sigs = {
    "oracle_num_threads": (_i, []),
    "oracle_set_num_threads": (None, [_i]),
    "oracle_accel": (None, [_vp, _i, _vp, _vp, _i, _d, _vp]),
    "oracle_accel_subset": (None, [_vp, _vp, _i, _vp, _vp, _i, _d, _vp]),
    "oracle_kick": (None, [_vp, _vp, _vp, _i, _vp, _vp, _i, _d, _d, _d]),
    "oracle_timesteps": (None, [_vp, _vp, _vp, _i, _d, _d, _vp]),
    "oracle_exclusive_scan_i32": (None, [_vp, _vp, _i]),
    "oracle_split": (None, [C.POINTER(Particles), C.POINTER(Particles), C.POINTER(Particles),
                            _vp, _vp, _d, _i, _PI, _PI]),
    "oracle_split_values_f64": (None, [_vp, _vp, _vp, _d, _i, _PI, _PI]),
    "oracle_cpu_split_f64": (_i, [_vp, _d, _i, _PI, _PI]),
    "oracle_cpu_split_f32": (_i, [_vp, C.c_float, _i, _PI, _PI]),
    "oracle_drift": (None, [_vp, _vp, _i, _d, _d]),
    "oracle_update": (None, [C.POINTER(Particles), C.POINTER(Particles), _i]),
    "oracle_hold_step": (_i, [C.POINTER(Particles), _d, _i, _d, _d, _i, C.POINTER(HoldStats)]),
    "oracle_hold": (_i, [C.POINTER(Particles), _d, _d, _d, _d, _i, C.POINTER(HoldStats)]),
    "oracle_energy": (_d, [_vp, _vp, _vp, _i, _d]),
    "oracle_kinetic": (_d, [_vp, _vp, _i]),
    "oracle_potential": (_d, [_vp, _vp, _i, _d]),
}
for _n, (_r, _a) in _sigs.items():
    _f = getattr(_lib, _n)
    _f.restype = _r
    _f.argtypes = _a


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    _lib.oracle_set_num_threads(int(n))


def num_threads() -> int:
    return _lib.oracle_num_threads()


def accel(si_pos, so_pos, so_m, eps2=0.0):
    si_pos, so_pos, so_m = _f64(si_pos), _f64(so_pos), _f64(so_m)
    out = np.zeros((len(si_pos), 3))
    _lib.oracle_accel(_p(si_pos), len(si_pos), _p(so_pos), _p(so_m), len(so_m), eps2, _p(out))
    return out


def accel_subset(si_pos, idx, so_pos, so_m, eps2=0.0):
    si_pos, so_pos, so_m = _f64(si_pos), _f64(so_pos), _f64(so_m)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    out = np.zeros((len(idx), 3))
    _lib.oracle_accel_subset(_p(si_pos), _p(idx), len(idx), _p(so_pos), _p(so_m), len(so_m),
                             eps2, _p(out))
    return out


def kick(si_pos, si_vel, so_pos, so_m, dt, fac, eps2=0.0):
    """Returns (vel_new, acc)."""
    si_pos, so_pos, so_m = _f64(si_pos), _f64(so_pos), _f64(so_m)
    vel = _f64(si_vel).copy()
    acc = np.zeros_like(vel)
    _lib.oracle_kick(_p(si_pos), _p(vel), _p(acc), len(si_pos), _p(so_pos), _p(so_m),
                     len(so_m), dt, fac, eps2)
    return vel, acc


def timesteps(pos, vel, m, eta, eps2=0.0):
    pos, vel, m = _f64(pos), _f64(vel), _f64(m)
    ts = np.zeros(len(m))
    _lib.oracle_timesteps(_p(pos), _p(vel), _p(m), len(m), eta, eps2, _p(ts))
    return ts


def exclusive_scan(x):
    x = np.ascontiguousarray(x, dtype=np.int32)
    out = np.zeros_like(x)
    _lib.oracle_exclusive_scan_i32(_p(x), _p(out), len(x))
    return out


class HostParticles:
    """Owns numpy arrays for one oracle_particles set."""

    def __init__(self, n, pos=None, vel=None, m=None, timestep=None, ids=None):
        self.pos = np.zeros((n, 3)) if pos is None else _f64(pos).reshape(n, 3).copy()
        self.vel = np.zeros((n, 3)) if vel is None else _f64(vel).reshape(n, 3).copy()
        self.acc = np.zeros((n, 3))
        self.timestep = np.zeros(n) if timestep is None else _f64(timestep).copy()
        self.m = np.zeros(n) if m is None else _f64(m).copy()
        self.id = (np.arange(n, dtype=np.int32) if ids is None
                   else np.ascontiguousarray(ids, dtype=np.int32).copy())
        self.parent_id = np.zeros(n, dtype=np.int32)
        self.n = n

    def struct(self):
        return Particles(self.pos.ctypes.data, self.vel.ctypes.data, self.acc.ctypes.data,
                         self.timestep.ctypes.data, self.m.ctypes.data, self.id.ctypes.data,
                         self.parent_id.ctypes.data, self.n)


def split(total: HostParticles, pivot: float):
    """Returns (slow, fast, n_slow, n_fast, predicate, scanned)."""
    n = total.n
    slow, fast = HostParticles(n), HostParticles(n)
    pred = np.zeros(n, dtype=np.int32)
    scan = np.zeros(n, dtype=np.int32)
    ns, nf = C.c_int(), C.c_int()
    t, s, f = total.struct(), slow.struct(), fast.struct()
    _lib.oracle_split(C.byref(t), C.byref(s), C.byref(f), _p(pred), _p(scan), pivot, n,
                      C.byref(ns), C.byref(nf))
    return slow, fast, ns.value, nf.value, pred, scan


def split_values(values, pivot):
    values = _f64(values)
    n = len(values)
    slow, fast = np.zeros(n), np.zeros(n)
    ns, nf = C.c_int(), C.c_int()
    _lib.oracle_split_values_f64(_p(values), _p(slow), _p(fast), pivot, n, C.byref(ns),
                                 C.byref(nf))
    return slow[:ns.value], fast[:nf.value]


def cpu_split(values, pivot):
    """Restated reference cpu_split (in place on a copy): (array, n_slow, n_fast, rc)."""
    arr = _f64(values).copy()
    ns, nf = C.c_int(), C.c_int()
    rc = _lib.oracle_cpu_split_f64(_p(arr), pivot, len(arr), C.byref(ns), C.byref(nf))
    return arr, ns.value, nf.value, rc


def cpu_split_f32(values, pivot):
    arr = np.ascontiguousarray(values, dtype=np.float32).copy()
    ns, nf = C.c_int(), C.c_int()
    rc = _lib.oracle_cpu_split_f32(_p(arr), C.c_float(pivot), len(arr), C.byref(ns),
                                   C.byref(nf))
    return arr, ns.value, nf.value, rc


def cpu_split_inplace_f32(arr, pivot):
    """Restated reference cpu_split on a float32 array, in place (no copy, for
    timing): (n_slow, n_fast, rc)."""
    assert arr.dtype == np.float32 and arr.flags.c_contiguous
    ns, nf = C.c_int(), C.c_int()
    rc = _lib.oracle_cpu_split_f32(_p(arr), C.c_float(pivot), len(arr), C.byref(ns), C.byref(nf))
    return ns.value, nf.value, rc


def drift(pos, vel, dt, fac):
    pos = _f64(pos).copy()
    vel = _f64(vel)
    _lib.oracle_drift(_p(pos), _p(vel), len(pos) if pos.ndim == 2 else len(pos) // 3, dt, fac)
    return pos


def update(prev: HostParticles, cur: HostParticles, n=None):
    n = cur.n if n is None else n
    a, b = prev.struct(), cur.struct()
    _lib.oracle_update(C.byref(a), C.byref(b), n)


def hold_step(sys: HostParticles, dt, calc, eta, eps2=0.0, max_level=48):
    st = HoldStats()
    s = sys.struct()
    rc = _lib.oracle_hold_step(C.byref(s), dt, 1 if calc else 0, eta, eps2, max_level,
                               C.byref(st))
    return rc, st


def hold(sys: HostParticles, t_end, dt_root, eta, eps2=0.0, max_level=48):
    st = HoldStats()
    s = sys.struct()
    rc = _lib.oracle_hold(C.byref(s), t_end, dt_root, eta, eps2, max_level, C.byref(st))
    return rc, st


def energy(pos, vel, m, eps2=0.0):
    pos, vel, m = _f64(pos), _f64(vel), _f64(m)
    return _lib.oracle_energy(_p(pos), _p(vel), _p(m), len(m), eps2)


def kinetic(vel, m):
    vel, m = _f64(vel), _f64(m)
    return _lib.oracle_kinetic(_p(vel), _p(m), len(m))


def potential(pos, m, eps2=0.0):
    pos, m = _f64(pos), _f64(m)
    return _lib.oracle_potential(_p(pos), _p(m), len(m), eps2)


# ------------------------------------------------ compiled reference (_ref)
def ref_available() -> bool:
    return os.path.exists(REF_LIB)


_ref = None


def _ref_lib():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_LIB)
        for name, ty in (("ref_cpu_split_f64", C.c_double), ("ref_cpu_split_f32", C.c_float)):
            f = getattr(_ref, name)
            f.restype = C.c_int
            f.argtypes = [_vp, ty, _i, _PI, _PI]
    return _ref


def ref_cpu_split(values, pivot, f32=False):
    """The reference's own cpu_split (src/cpu_scan.h:20-50), compiled from
    the mounted reference tree: (array, n_slow, n_fast)."""
    lib = _ref_lib()
    arr = (np.ascontiguousarray(values, dtype=np.float32) if f32 else _f64(values)).copy()
    ns, nf = C.c_int(), C.c_int()
    fn = lib.ref_cpu_split_f32 if f32 else lib.ref_cpu_split_f64
    fn(_p(arr), pivot, len(arr), C.byref(ns), C.byref(nf))
    return arr, ns.value, nf.value
