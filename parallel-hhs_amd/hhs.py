"""ctypes binding of libhhs.so (the C ABI declared in include/hhs.h).

Device arrays are passed as torch tensors (PyTorch only supplies device
memory and streams); every call goes straight to the C ABI.  There is no
CPU fallback: if the library is not built, importing this module raises.

Reference API mirrored (dipto4/parallel-hhs): particle_system<T>
(src/particle_system.cuh:20-54), split (src/split.cuh:573-610), kick_slow /
kick_sf (src/kick.cuh:158-181), timesteps (src/timestep.cuh:103-118), drift
(src/drift.cuh:30-43), update_system (src/update_system.cuh:29-43) and
hold_step / hold (src/hold.cuh:17-23,51-181).
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import os
from dataclasses import dataclass

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhhs.so")

HHS_OK = 0
HHS_ERR_INVALID = 1
HHS_ERR_ALLOC = 2
HHS_ERR_HIP = 3
HHS_ERR_NO_DEVICE = 4
HHS_ERR_DEPTH = 5
HHS_ERR_BUDGET = 6
HHS_ERR_NONFINITE = 7
HHS_ERR_WORKSPACE = 8
HHS_ERR_IO = 9
HHS_ERR_FORMAT = 10


class HHSError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hhs error {code}: {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `make -C parallel-hhs_amd` "
        "(or __graft_entry__.build()) first")

# PyTorch-ROCm bundles its own HIP runtime with the same soname
# (libamdhip64.so.7) as /opt/rocm's.  Two runtimes in one process crash on
# the first device call, so when PyTorch is installed its runtime file is
# mapped first (RTLD_GLOBAL, without importing torch): libhhs's
# libamdhip64.so.7 then resolves to it, and a later `import torch` finds the
# same file already loaded.  Without PyTorch, libhhs's RUNPATH picks
# /opt/rocm's.
def _preload_torch_hip():
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        C.CDLL(path, mode=C.RTLD_GLOBAL)


_preload_torch_hip()
_lib = C.CDLL(LIB_PATH)
try:
    import torch  # noqa: F401
except ImportError:  # C++/ctypes-only use without PyTorch
    torch = None


class ParticlesF64(C.Structure):
    _fields_ = [("pos", C.c_void_p), ("vel", C.c_void_p), ("acc", C.c_void_p),
                ("timestep", C.c_void_p), ("m", C.c_void_p), ("id", C.c_void_p),
                ("parent_id", C.c_void_p), ("n", C.c_int)]


ParticlesF32 = type("ParticlesF32", (C.Structure,), {"_fields_": ParticlesF64._fields_})


class Partition(C.Structure):
    _fields_ = [("N_slow", C.c_int), ("N_fast", C.c_int)]


class SystemResult(C.Structure):
    _fields_ = [("e0", C.c_double), ("e1", C.c_double), ("t", C.c_double),
                ("calls", C.c_longlong), ("interactions", C.c_longlong),
                ("ts_pairs", C.c_longlong), ("max_level", C.c_int), ("status", C.c_int)]


class EnsembleDesc(C.Structure):
    _fields_ = [("n_systems", C.c_int), ("offsets", C.c_void_p), ("pos", C.c_void_p),
                ("vel", C.c_void_p), ("m", C.c_void_p), ("t_end", C.c_void_p),
                ("dt_root", C.c_void_p), ("eta", C.c_double), ("eps", C.c_double),
                ("max_level", C.c_int), ("max_calls", C.c_longlong),
                ("wave_team_max_n", C.c_int), ("results", C.c_void_p),
                ("timestep_mode", C.c_int), ("done", C.c_void_p),
                ("approximate_decisions", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("eta", C.c_double), ("eps", C.c_double), ("hbm_budget_bytes", C.c_size_t),
                ("small_max_n", C.c_int), ("max_n", C.c_int), ("max_batch", C.c_int),
                ("batch_streams", C.c_int), ("big_workers", C.c_int), ("max_level", C.c_int),
                ("timestep_mode", C.c_int)]


class RequestInfo(C.Structure):
    _fields_ = [("t_submit", C.c_double), ("t_admit", C.c_double), ("t_start", C.c_double),
                ("t_done", C.c_double), ("result", SystemResult), ("path", C.c_int),
                ("state", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("now", C.c_double), ("used_bytes", C.c_size_t), ("peak_bytes", C.c_size_t),
                ("pending", C.c_int), ("running_batches", C.c_int), ("running_big", C.c_int)]


class SnapshotHeader(C.Structure):
    """hhs_snapshot_header: the 128-byte header of a snapshot file."""
    _fields_ = [("magic", C.c_char * 8), ("version", C.c_uint32), ("header_bytes", C.c_uint32),
                ("n", C.c_int64), ("root_steps", C.c_int64), ("simtime", C.c_double),
                ("dt_root", C.c_double), ("eta", C.c_double), ("eps", C.c_double),
                ("timestep_mode", C.c_int32), ("kick_precision", C.c_int32),
                ("payload_fnv1a", C.c_uint64), ("reserved", C.c_uint8 * 48)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f not in ("magic", "reserved")}


class HoldStats(C.Structure):
    _fields_ = [("calls", C.c_longlong), ("interactions", C.c_longlong),
                ("ts_pairs", C.c_longlong), ("max_level", C.c_int),
                ("calls_per_level", C.c_longlong * 64)]


_vp, _i, _d, _f, _sz = C.c_void_p, C.c_int, C.c_double, C.c_float, C.c_size_t
_P64, _P32 = C.POINTER(ParticlesF64), C.POINTER(ParticlesF32)

_SIGS = {
    "hhs_last_error": (C.c_char_p, []),
    "hhs_version": (C.c_char_p, []),
    "hhs_runtime_check": (C.c_int, []),
    "hhs_device_count": (_i, [C.POINTER(_i)]),
    "hhs_malloc": (_i, [C.POINTER(_vp), _sz]),
    "hhs_free": (_i, [_vp]),
    "hhs_memcpy_h2d": (_i, [_vp, _vp, _sz]),
    "hhs_memcpy_d2h": (_i, [_vp, _vp, _sz]),
    "hhs_memcpy_d2d": (_i, [_vp, _vp, _sz]),
    "hhs_memset": (_i, [_vp, _i, _sz]),
    "hhs_device_synchronize": (_i, []),
    "hhs_stream_synchronize": (_i, [_vp]),
    "hhs_host_alloc_mapped": (_i, [C.POINTER(_vp), _sz]),
    "hhs_host_free_mapped": (_i, [_vp]),
    "hhs_kick_self_f64": (_i, [_P64, _i, _d, _d, _d, _vp]),
    "hhs_kick_self_f32": (_i, [_P32, _i, _f, _f, _f, _vp]),
    "hhs_kick_cross_f64": (_i, [_P64, _P64, _i, _i, _d, _d, _d, _vp]),
    "hhs_kick_cross_f32": (_i, [_P32, _P32, _i, _i, _f, _f, _f, _vp]),
    "hhs_kick_set_config": (_i, [_i, _i]),
    "hhs_kick_set_precision": (_i, [_i]),
    "hhs_kick_get_precision": (_i, []),
    "hhs_autotune_kick": (_i, [_i, _i, C.c_char_p, C.POINTER(_i), C.POINTER(_i),
                               C.POINTER(_d)]),
    "hhs_autotune_timesteps": (_i, [_i, _i, _i, C.c_char_p, C.POINTER(_i), C.POINTER(_i),
                                    C.POINTER(C.c_double)]),
    "hhs_timesteps_get_config": (_i, [_i, _i, C.POINTER(_i), C.POINTER(_i)]),
    "hhs_autotune_load": (_i, [C.c_char_p]),
    "hhs_autotune_clear": (_i, []),
    "hhs_kick_get_config": (_i, [_i, _i, C.POINTER(_i), C.POINTER(_i)]),
    "hhs_timesteps_f64": (_i, [_P64, _i, _d, _d, _vp]),
    "hhs_timesteps_f32": (_i, [_P32, _i, _f, _f, _vp]),
    "hhs_timesteps_fast_f64": (_i, [_P64, _i, _d, _d, _vp]),
    "hhs_timesteps_mode_f64": (_i, [_P64, _i, _d, _d, _i, _vp]),
    "hhs_globals_set_timestep_mode": (_i, [_vp, _i]),
    "hhs_globals_set_kick_precision": (_i, [_vp, _i]),
    "hhs_timesteps_sym_workspace_bytes": (_sz, [_i]),
    "hhs_timesteps_sym_f64": (_i, [_P64, _i, _d, _d, _i, _vp, _sz, _vp]),
    "hhs_timesteps_sym_set_sinks_per_lane": (_i, [_i]),
    "hhs_timesteps_sym_set_unit": (_i, [_i]),
    "hhs_globals_set_subtree_max": (_i, [_vp, _i]),
    "hhs_globals_subtree_max": (_i, [_vp]),
    "hhs_globals_set_speculation": (_i, [_vp, _i]),
    "hhs_globals_set_fused_reduce": (_i, [_vp, _i]),
    "hhs_globals_set_chain_scatter": (_i, [_vp, _i]),
    "hhs_globals_timestep_mode": (_i, [_vp]),
    "hhs_globals_kick_precision": (_i, [_vp]),
    "hhs_globals_root_steps": (C.c_longlong, [_vp]),
    "hhs_globals_set_root_steps": (None, [_vp, C.c_longlong]),
    "hhs_snapshot_write_f64": (_i, [C.c_char_p, _vp, _P64, _i, _d, _vp]),
    "hhs_snapshot_read_header": (_i, [C.c_char_p, C.POINTER(SnapshotHeader)]),
    "hhs_snapshot_read_f64": (_i, [C.c_char_p, _P64, _vp, C.POINTER(SnapshotHeader), _vp]),
    "hhs_evolve_f64": (_i, [_vp, _P64, _d, _d, C.c_char_p, _i, _vp]),
    "hhs_scan_workspace_bytes": (_sz, [_i]),
    "hhs_scan_workspace_init": (_i, [_vp, _sz, _vp]),
    "hhs_scan_workspace_thread": (_i, [_i, C.POINTER(_vp), C.POINTER(_sz)]),
    "hhs_scan_exclusive_i32": (_i, [_vp, _vp, _i, _vp, _sz, _vp]),
    "hhs_split_f64": (_i, [_P64, _P64, _P64, _vp, _vp, _d, _i, _vp, _vp, _sz, _vp]),
    "hhs_split_f32": (_i, [_P32, _P32, _P32, _vp, _vp, _f, _i, _vp, _vp, _sz, _vp]),
    "hhs_split_f64_noalloc": (_i, [_P64, _P64, _P64, _vp, _vp, _d, _i, _vp, _vp]),
    "hhs_split_values_f32": (_i, [_vp, _vp, _vp, _f, _i, _vp, _vp, _sz, _vp]),
    "hhs_split_values_f64": (_i, [_vp, _vp, _vp, _d, _i, _vp, _vp, _sz, _vp]),
    "hhs_drift_f64": (_i, [_vp, _vp, _i, _d, _d, _vp]),
    "hhs_drift_f32": (_i, [_vp, _vp, _i, _f, _f, _vp]),
    "hhs_update_f64": (_i, [_P64, _P64, _i, _vp]),
    "hhs_update_f32": (_i, [_P32, _P32, _i, _vp]),
    "hhs_energy_f64": (_i, [_vp, _vp, _vp, _i, _d, C.POINTER(_d), C.POINTER(_d), _vp]),
    "hhs_globals_create": (_i, [C.POINTER(_vp), _i, _d, _d, _i]),
    "hhs_globals_destroy": (_i, [_vp]),
    "hhs_globals_predicate": (_vp, [_vp]),
    "hhs_globals_scanned_predicate": (_vp, [_vp]),
    "hhs_globals_simtime": (_d, [_vp]),
    "hhs_globals_set_simtime": (None, [_vp, _d]),
    "hhs_globals_eta": (_d, [_vp]),
    "hhs_globals_eps": (_d, [_vp]),
    "hhs_globals_stats": (_i, [_vp, C.POINTER(HoldStats)]),
    "hhs_globals_reset_stats": (_i, [_vp]),
    "hhs_hold_step_f64": (_i, [_vp, _P64, _i, _i, _d, _d, _d, _i, _vp]),
    "hhs_hold_f64": (_i, [_vp, _P64, _d, _d, _vp]),
    "hhs_hold_step_f32": (_i, [_vp, _P32, _i, _i, _f, _f, _f, _i, _vp]),
    "hhs_hold_f32": (_i, [_vp, _P32, _f, _f, _vp]),
    "hhs_ic_plummer": (_i, [_i, C.c_uint64, _i, _i, _d, _vp, _vp, _vp]),
    "hhs_ensemble_workspace_bytes": (_sz, [_i, C.c_longlong]),
    "hhs_ensemble_run_f64": (_i, [C.POINTER(EnsembleDesc), _vp, _vp, _sz, _vp]),
    "hhs_server_create": (_i, [C.POINTER(_vp), C.POINTER(ServerConfig)]),
    "hhs_server_destroy": (_i, [_vp]),
    "hhs_server_submit": (_i, [_vp, _i, _vp, _vp, _vp, _d, _d, C.POINTER(C.c_longlong)]),
    "hhs_server_poll": (_i, [_vp, C.c_longlong, C.POINTER(RequestInfo)]),
    "hhs_server_wait": (_i, [_vp, C.c_longlong, C.POINTER(RequestInfo)]),
    "hhs_server_fetch": (_i, [_vp, C.c_longlong, _vp, _vp]),
    "hhs_server_release": (_i, [_vp, C.c_longlong]),
    "hhs_server_stats": (_i, [_vp, C.POINTER(ServerStats)]),
}

for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED = tuple(_SIGS)


def last_error() -> str:
    return _lib.hhs_last_error().decode()


def runtime_check():
    """Raises HHSError when two HIP runtimes are mapped into the process."""
    check(_lib.hhs_runtime_check())


def check(rc: int) -> None:
    if rc != HHS_OK:
        raise HHSError(rc, last_error())


def raw():
    """The loaded ctypes library (for tests that call entry points directly)."""
    return _lib


# ---------------------------------------------------------------- helpers
def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)


@dataclass
class DeviceParticles:
    """Device particle set backed by torch tensors (one field per
    particle_system<T> member, AoS vec3 for pos/vel/acc)."""
    pos: object
    vel: object
    acc: object
    timestep: object
    m: object
    id: object
    parent_id: object

    @property
    def n(self) -> int:
        return int(self.m.shape[0])

    @classmethod
    def empty(cls, n: int, dtype=None, device="cuda"):
        import torch
        dtype = dtype or torch.float64
        return cls(pos=torch.zeros(n, 3, dtype=dtype, device=device),
                   vel=torch.zeros(n, 3, dtype=dtype, device=device),
                   acc=torch.zeros(n, 3, dtype=dtype, device=device),
                   timestep=torch.zeros(n, dtype=dtype, device=device),
                   m=torch.zeros(n, dtype=dtype, device=device),
                   id=torch.zeros(n, dtype=torch.int32, device=device),
                   parent_id=torch.zeros(n, dtype=torch.int32, device=device))

    @classmethod
    def from_host(cls, pos, vel, m, timestep=None, ids=None, dtype=None, device="cuda"):
        import numpy as np
        import torch
        dtype = dtype or torch.float64
        n = len(m)
        p = cls.empty(n, dtype, device)
        p.pos.copy_(torch.as_tensor(np.asarray(pos).reshape(n, 3)))
        p.vel.copy_(torch.as_tensor(np.asarray(vel).reshape(n, 3)))
        p.m.copy_(torch.as_tensor(np.asarray(m)))
        if timestep is not None:
            p.timestep.copy_(torch.as_tensor(np.asarray(timestep)))
        p.id.copy_(torch.as_tensor(np.arange(n, dtype=np.int32) if ids is None
                                   else np.asarray(ids, dtype=np.int32)))
        return p

    def struct(self):
        import torch
        cls = ParticlesF64 if self.m.dtype == torch.float64 else ParticlesF32
        return cls(self.pos.data_ptr(), self.vel.data_ptr(), self.acc.data_ptr(),
                   self.timestep.data_ptr(), self.m.data_ptr(), self.id.data_ptr(),
                   self.parent_id.data_ptr(), self.n)

    def is_f64(self) -> bool:
        import torch
        return self.m.dtype == torch.float64


def _sfx(p: DeviceParticles) -> str:
    return "f64" if p.is_f64() else "f32"


def kick_self(sys: DeviceParticles, n=None, dt=1.0, fac=1.0, eps2=0.0, stream=None):
    n = sys.n if n is None else n
    s = sys.struct()
    check(getattr(_lib, "hhs_kick_self_" + _sfx(sys))(C.byref(s), n, dt, fac, eps2, _stream(stream)))


def kick_cross(so: DeviceParticles, si: DeviceParticles, n_sink=None, n_src=None, dt=1.0,
               fac=1.0, eps2=0.0, stream=None):
    n_sink = si.n if n_sink is None else n_sink
    n_src = so.n if n_src is None else n_src
    a, b = so.struct(), si.struct()
    check(getattr(_lib, "hhs_kick_cross_" + _sfx(si))(C.byref(a), C.byref(b), n_sink, n_src, dt,
                                                      fac, eps2, _stream(stream)))


def kick_set_config(sinks_per_lane: int, waves: int):
    check(_lib.hhs_kick_set_config(sinks_per_lane, waves))


def autotune_kick(n: int, reps: int = 3, cache_path: str | None = None):
    """Sweep the kick configurations at size n; returns (S, W, ms)."""
    s, w, ms = C.c_int(), C.c_int(), C.c_double()
    path = cache_path.encode() if cache_path else None
    check(_lib.hhs_autotune_kick(n, reps, path, C.byref(s), C.byref(w), C.byref(ms)))
    return s.value, w.value, ms.value


def autotune_timesteps(n: int, mode: int, reps: int = 3, cache_path: str | None = None):
    """Sweeps (waves per group, source chunks) of the one-sided timestep
    kernel for `mode` at n particles; returns (W, K, ms) and installs it."""
    w, k, ms = C.c_int(), C.c_int(), C.c_double()
    path = None if cache_path is None else cache_path.encode()
    check(_lib.hhs_autotune_timesteps(n, _ts_mode(mode), reps, path, C.byref(w), C.byref(k),
                                      C.byref(ms)))
    return w.value, k.value, ms.value


def timesteps_get_config(n: int, mode: int):
    """(waves per group, source chunks) of the one-sided timestep kernel."""
    w, k = C.c_int(), C.c_int()
    check(_lib.hhs_timesteps_get_config(n, _ts_mode(mode), C.byref(w), C.byref(k)))
    return w.value, k.value


def autotune_load(cache_path: str):
    check(_lib.hhs_autotune_load(cache_path.encode()))


def autotune_clear():
    check(_lib.hhs_autotune_clear())


def kick_get_config(n_sink: int, n_src: int):
    s, w = C.c_int(), C.c_int()
    check(_lib.hhs_kick_get_config(n_sink, n_src, C.byref(s), C.byref(w)))
    return s.value, w.value


TIMESTEP_EXACT = 0
TIMESTEP_FAST = 1
TIMESTEP_FP32 = 2


def _ts_mode(fast):
    """fast: False/True or a mode constant (TIMESTEP_EXACT/_FAST/_FP32)."""
    if fast is True:
        return TIMESTEP_FAST
    if fast is False or fast is None:
        return TIMESTEP_EXACT
    return int(fast)


def timesteps_sym(sys: DeviceParticles, n=None, eta=0.01, eps2=0.0, mode=0, stream=None):
    """hhs_timesteps_sym_f64: each unordered pair evaluated once (EXACT stays
    bit-exact; FP32 within 1e-6).  Allocates the partial-minima workspace."""
    import torch
    n = sys.n if n is None else n
    nbytes = int(_lib.hhs_timesteps_sym_workspace_bytes(n))
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device="cuda")
    s = sys.struct()
    check(_lib.hhs_timesteps_sym_f64(C.byref(s), n, eta, eps2, _ts_mode(mode), _ptr(ws), nbytes,
                                     _stream(stream)))
    return ws


def timesteps(sys: DeviceParticles, n=None, eta=0.01, eps2=0.0, stream=None, fast=False):
    n = sys.n if n is None else n
    s = sys.struct()
    mode = _ts_mode(fast)
    if mode != TIMESTEP_EXACT:
        check(_lib.hhs_timesteps_mode_f64(C.byref(s), n, eta, eps2, mode, _stream(stream)))
        return
    check(getattr(_lib, "hhs_timesteps_" + _sfx(sys))(C.byref(s), n, eta, eps2, _stream(stream)))


class ScanWorkspace:
    """Device workspace of the single-pass scan (allocated through torch)."""

    def __init__(self, n_max: int, stream=None):
        import torch
        self.nbytes = int(_lib.hhs_scan_workspace_bytes(int(n_max)))
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device="cuda")
        check(_lib.hhs_scan_workspace_init(_ptr(self.buf), self.nbytes, _stream(stream)))

    @property
    def ptr(self):
        return _ptr(self.buf)


def scan_exclusive(inp, out, ws: ScanWorkspace, n=None, stream=None):
    n = int(inp.shape[0]) if n is None else n
    check(_lib.hhs_scan_exclusive_i32(_ptr(inp), _ptr(out), n, ws.ptr, ws.nbytes, _stream(stream)))


def split(total: DeviceParticles, slow: DeviceParticles, fast: DeviceParticles, pivot: float,
          part_dev, ws: ScanWorkspace, n=None, predicate=None, scanned=None, stream=None):
    n = total.n if n is None else n
    t, s, f = total.struct(), slow.struct(), fast.struct()
    check(getattr(_lib, "hhs_split_" + _sfx(total))(
        C.byref(t), C.byref(s), C.byref(f), _ptr(predicate), _ptr(scanned), pivot, n,
        _ptr(part_dev), ws.ptr, ws.nbytes, _stream(stream)))


def split_values(values, slow, fast, pivot: float, part_dev, ws: ScanWorkspace, n=None,
                 stream=None):
    import torch
    n = int(values.shape[0]) if n is None else n
    fn = _lib.hhs_split_values_f64 if values.dtype == torch.float64 else _lib.hhs_split_values_f32
    check(fn(_ptr(values), _ptr(slow), _ptr(fast), pivot, n, _ptr(part_dev), ws.ptr, ws.nbytes,
             _stream(stream)))


def drift(pos, vel, n=None, dt=1.0, fac=1.0, stream=None):
    import torch
    n = int(pos.shape[0]) if n is None else n
    fn = _lib.hhs_drift_f64 if pos.dtype == torch.float64 else _lib.hhs_drift_f32
    check(fn(_ptr(pos), _ptr(vel), n, dt, fac, _stream(stream)))


def update(prev: DeviceParticles, cur: DeviceParticles, n=None, stream=None):
    n = cur.n if n is None else n
    a, b = prev.struct(), cur.struct()
    check(getattr(_lib, "hhs_update_" + _sfx(cur))(C.byref(a), C.byref(b), n, _stream(stream)))


def energy(pos, vel, m, eps2=0.0, n=None, stream=None):
    n = int(m.shape[0]) if n is None else n
    ke, pe = C.c_double(), C.c_double()
    check(_lib.hhs_energy_f64(_ptr(pos), _ptr(vel), _ptr(m), n, eps2, C.byref(ke), C.byref(pe),
                              _stream(stream)))
    return ke.value, pe.value


class Globals:
    """nbodysystem_globals<double> + per-level buffers (hhs_globals)."""

    def __init__(self, n_total: int, eta: float, eps: float, max_level: int = 48,
                 fast_timesteps: bool = False, kick_precision: int = -1,
                 subtree_max: int | None = None, speculation: bool = True,
                 fused_reduce: bool = True, chain_scatter: bool = True):
        """`kick_precision` is a KICK_* mode for this HOLD (-1: the process-wide
        setting).  kick_precision=KICK_F64_EXACT with exact timesteps is the
        FP64-only HOLD mode, bit-exact with the FP64 CPU HOLD.  `speculation`
        queues each call's first fast recursion before its counts are read
        back (same results, fewer idle gaps).  `fused_reduce` lets a call's
        device subtree and scatter-back add its kick's split-K partials
        (same results, one launch fewer per such call).  `chain_scatter`
        runs the scatter-backs of a chain of second fast recursions as one
        launch (same results)."""
        h = C.c_void_p()
        check(_lib.hhs_globals_create(C.byref(h), n_total, eta, eps, max_level))
        self.h = h
        if _ts_mode(fast_timesteps) != TIMESTEP_EXACT:
            check(_lib.hhs_globals_set_timestep_mode(h, _ts_mode(fast_timesteps)))
        if kick_precision != -1:
            check(_lib.hhs_globals_set_kick_precision(h, kick_precision))
        if subtree_max is not None:
            check(_lib.hhs_globals_set_subtree_max(h, subtree_max))
        if not speculation:
            check(_lib.hhs_globals_set_speculation(h, 0))
        if not fused_reduce:
            check(_lib.hhs_globals_set_fused_reduce(h, 0))
        if not chain_scatter:
            check(_lib.hhs_globals_set_chain_scatter(h, 0))

    @property
    def subtree_max(self) -> int:
        return _lib.hhs_globals_subtree_max(self.h)

    def close(self):
        if self.h:
            _lib.hhs_globals_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def simtime(self) -> float:
        return _lib.hhs_globals_simtime(self.h)

    @simtime.setter
    def simtime(self, t: float):
        _lib.hhs_globals_set_simtime(self.h, t)

    def stats(self) -> dict:
        st = HoldStats()
        check(_lib.hhs_globals_stats(self.h, C.byref(st)))
        lv = [int(x) for x in st.calls_per_level[:st.max_level + 1]]
        return {"calls": st.calls, "interactions": st.interactions, "ts_pairs": st.ts_pairs,
                "max_level": st.max_level, "calls_per_level": lv}

    def reset_stats(self):
        check(_lib.hhs_globals_reset_stats(self.h))

    def hold_step(self, sys: DeviceParticles, dt: float, calc_timestep=True, clevel=0, n=None,
                  stream=None):
        """hhs_hold_step_f64 / _f32 by the particle set's precision."""
        n = sys.n if n is None else n
        s = sys.struct()
        check(getattr(_lib, "hhs_hold_step_" + _sfx(sys))(self.h, C.byref(s), n, clevel, 0.0, dt,
                                                        dt, 1 if calc_timestep else 0,
                                                        _stream(stream)))

    def hold(self, sys: DeviceParticles, t_end: float, dt_root: float, stream=None):
        s = sys.struct()
        check(getattr(_lib, "hhs_hold_" + _sfx(sys))(self.h, C.byref(s), t_end, dt_root,
                                                   _stream(stream)))

    @property
    def root_steps(self) -> int:
        return _lib.hhs_globals_root_steps(self.h)

    def evolve(self, sys: DeviceParticles, t_end: float, dt_root: float,
               snapshot_path: str | None = None, snapshot_every: int = 0, stream=None):
        """hhs_evolve_f64: root steps until t_end, snapshots at root-step
        boundaries (every `snapshot_every` steps and after the last)."""
        s = sys.struct()
        path = None if snapshot_path is None else os.fsencode(snapshot_path)
        check(_lib.hhs_evolve_f64(self.h, C.byref(s), t_end, dt_root, path, snapshot_every,
                                  _stream(stream)))

    def save_snapshot(self, path: str, sys: DeviceParticles, dt_root: float, n=None,
                      stream=None):
        s = sys.struct()
        check(_lib.hhs_snapshot_write_f64(os.fsencode(path), self.h, C.byref(s),
                                          sys.n if n is None else n, dt_root, _stream(stream)))

    def load_snapshot(self, path: str, sys: DeviceParticles, stream=None) -> dict:
        """Loads particles into `sys` and this globals' simtime and root-step
        count (eta, eps and modes must match the snapshot's)."""
        s = sys.struct()
        h = SnapshotHeader()
        check(_lib.hhs_snapshot_read_f64(os.fsencode(path), C.byref(s), self.h, C.byref(h),
                                         _stream(stream)))
        return h.as_dict()


def read_snapshot_header(path: str) -> dict:
    h = SnapshotHeader()
    check(_lib.hhs_snapshot_read_header(os.fsencode(path), C.byref(h)))
    return h.as_dict()


class HoldIntegrator:
    """AMUSE-style front end: numpy arrays in and out, `evolve_model(t)` in
    root steps of dt_root, snapshots and resume.  Device memory comes from
    the C ABI (hhs_malloc); PyTorch is not used.

    timestep_mode: TIMESTEP_EXACT / _FAST / _FP32; kick_precision: a KICK_*
    mode or -1 (process-wide setting).  KICK_F64_EXACT with TIMESTEP_EXACT
    integrates bit-identically with the FP64 CPU HOLD."""

    _FIELDS = (("pos", 24), ("vel", 24), ("acc", 24), ("timestep", 8), ("m", 8), ("id", 4),
               ("parent_id", 4))

    def __init__(self, pos, vel, m, eta=0.01, eps=0.0, dt_root=1.0 / 64, max_level=48,
                 timestep_mode=TIMESTEP_EXACT, kick_precision=-1, ids=None, _load=None):
        import numpy as np
        self.dt_root = float(dt_root)
        self.eps = float(eps)
        self._ptrs = {}
        if _load is not None:
            hdr = read_snapshot_header(_load)
            n = int(hdr["n"])
            eta, self.eps, self.dt_root = hdr["eta"], hdr["eps"], hdr["dt_root"]
            timestep_mode, kick_precision = hdr["timestep_mode"], hdr["kick_precision"]
        else:
            n = len(m)
        self.n = n
        try:
            for f, b in self._FIELDS:
                p = C.c_void_p()
                check(_lib.hhs_malloc(C.byref(p), max(b * n, 1)))
                self._ptrs[f] = p
            self.globals = Globals(n, eta, self.eps, max_level, timestep_mode, kick_precision)
        except Exception:
            self.close()
            raise
        if _load is not None:
            self.globals.load_snapshot(_load, self, stream=0)
            return
        ids = np.arange(n, dtype=np.int32) if ids is None else ids
        for f, a, dt in (("pos", pos, np.float64), ("vel", vel, np.float64),
                         ("m", m, np.float64), ("id", ids, np.int32)):
            a = np.ascontiguousarray(a, dtype=dt)
            check(_lib.hhs_memcpy_h2d(self._ptrs[f], a.ctypes.data, a.nbytes))

    @classmethod
    def load(cls, path: str, max_level: int = 48) -> "HoldIntegrator":
        """Resume from a snapshot: same n, eta, eps, modes, dt_root and
        model time as the run that wrote it."""
        return cls(None, None, None, max_level=max_level, _load=path)

    def struct(self) -> ParticlesF64:
        p = self._ptrs
        return ParticlesF64(p["pos"], p["vel"], p["acc"], p["timestep"], p["m"], p["id"],
                            p["parent_id"], self.n)

    @property
    def model_time(self) -> float:
        return self.globals.root_steps * self.dt_root

    def evolve_model(self, t_end: float, snapshot_path: str | None = None,
                     snapshot_every: int = 0):
        self.globals.evolve(self, t_end, self.dt_root, snapshot_path, snapshot_every, stream=0)

    def save(self, path: str):
        self.globals.save_snapshot(path, self, self.dt_root, stream=0)

    def get_state(self):
        """(pos[n,3], vel[n,3], m[n], id[n]) copied to the host."""
        import numpy as np
        out = (np.empty((self.n, 3)), np.empty((self.n, 3)), np.empty(self.n),
               np.empty(self.n, dtype=np.int32))
        check(_lib.hhs_stream_synchronize(None))
        for f, a in zip(("pos", "vel", "m", "id"), out):
            check(_lib.hhs_memcpy_d2h(a.ctypes.data, self._ptrs[f], a.nbytes))
        return out

    def energy(self):
        """(kinetic, potential) in FP64 on the device."""
        ke, pe = C.c_double(), C.c_double()
        p = self._ptrs
        check(_lib.hhs_energy_f64(p["pos"], p["vel"], p["m"], self.n, self.eps * self.eps,
                                  C.byref(ke), C.byref(pe), None))
        return ke.value, pe.value

    def close(self):
        g = getattr(self, "globals", None)
        if g is not None:
            g.close()
        for p in getattr(self, "_ptrs", {}).values():
            _lib.hhs_free(p)
        self._ptrs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ic_plummer(n: int, seed: int, mass_model: int = 0, virialize: bool = True,
               eps2: float = 0.0):
    """Host-side Plummer sphere (numpy arrays pos[n,3], vel[n,3], m[n])."""
    import numpy as np
    pos = np.zeros((n, 3), dtype=np.float64)
    vel = np.zeros((n, 3), dtype=np.float64)
    m = np.zeros(n, dtype=np.float64)
    check(_lib.hhs_ic_plummer(n, seed, mass_model, 1 if virialize else 0, eps2,
                              pos.ctypes.data, vel.ctypes.data, m.ctypes.data))
    return pos, vel, m


RESULT_DTYPE = [("e0", "f8"), ("e1", "f8"), ("t", "f8"), ("calls", "i8"),
                ("interactions", "i8"), ("ts_pairs", "i8"), ("max_level", "i4"),
                ("status", "i4")]


class Ensemble:
    """A batch of independent systems integrated by one persistent launch
    (hhs_ensemble_run_f64).  `systems` is a list of (pos[n,3], vel[n,3],
    m[n]) host arrays; t_end and dt_root are scalars or per-system
    sequences."""

    def __init__(self, systems, t_end, dt_root, eta, eps, max_level=0, max_calls=0,
                 wave_team_max_n=0, fast_timesteps=False, exact_decisions=True):
        import numpy as np
        import torch
        self.n_systems = len(systems)
        sizes = np.array([len(s[2]) for s in systems], dtype=np.int64)
        self.offsets = np.zeros(self.n_systems + 1, dtype=np.int32)
        np.cumsum(sizes, out=self.offsets[1:])
        ntot = int(self.offsets[-1])
        pos = np.zeros((max(ntot, 1), 3))
        vel = np.zeros((max(ntot, 1), 3))
        m = np.zeros(max(ntot, 1))
        for k, (p, v, mm) in enumerate(systems):
            a, b = self.offsets[k], self.offsets[k + 1]
            pos[a:b], vel[a:b], m[a:b] = p, v, mm
        dev = "cuda"
        self.pos = torch.as_tensor(pos).to(dev)
        self.vel = torch.as_tensor(vel).to(dev)
        self.m = torch.as_tensor(m).to(dev)
        self.off_dev = torch.as_tensor(self.offsets).to(dev)
        self.t_end = torch.as_tensor(np.broadcast_to(np.asarray(t_end, float),
                                                     (self.n_systems,)).copy()).to(dev)
        self.dt_root = torch.as_tensor(np.broadcast_to(np.asarray(dt_root, float),
                                                       (self.n_systems,)).copy()).to(dev)
        self.results_dev = torch.zeros(max(self.n_systems, 1) * C.sizeof(SystemResult),
                                       dtype=torch.uint8, device=dev)
        self.ws_bytes = int(_lib.hhs_ensemble_workspace_bytes(self.n_systems, ntot))
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=dev)
        self.desc = EnsembleDesc(self.n_systems, self.off_dev.data_ptr(), self.pos.data_ptr(),
                                 self.vel.data_ptr(), self.m.data_ptr(), self.t_end.data_ptr(),
                                 self.dt_root.data_ptr(), eta, eps, max_level, max_calls,
                                 wave_team_max_n, self.results_dev.data_ptr(),
                                 _ts_mode(fast_timesteps), None, 0 if exact_decisions else 1)

    def run(self, stream=None):
        check(_lib.hhs_ensemble_run_f64(C.byref(self.desc),
                                        self.offsets.ctypes.data_as(C.c_void_p),
                                        C.c_void_p(self.ws.data_ptr()), self.ws_bytes,
                                        _stream(stream)))

    def results(self):
        import numpy as np
        raw = self.results_dev.cpu().numpy().tobytes()
        return np.frombuffer(raw, dtype=np.dtype(RESULT_DTYPE))[:self.n_systems].copy()

    def state(self, k):
        a, b = int(self.offsets[k]), int(self.offsets[k + 1])
        return self.pos[a:b].cpu().numpy(), self.vel[a:b].cpu().numpy()


class Server:
    """Continuous-batching HOLD server (hhs_server_*).  submit() copies the
    host arrays and returns a ticket; wait()/poll() report timings on the
    server clock; fetch() returns the final (pos, vel)."""

    def __init__(self, eta, eps, hbm_budget_bytes=8 << 30, small_max_n=0, max_n=0,
                 max_batch=0, batch_streams=0, big_workers=0, max_level=0,
                 fast_timesteps=False):
        cfg = ServerConfig(eta, eps, hbm_budget_bytes, small_max_n, max_n, max_batch,
                           batch_streams, big_workers, max_level, _ts_mode(fast_timesteps))
        h = C.c_void_p()
        check(_lib.hhs_server_create(C.byref(h), C.byref(cfg)))
        self.h = h
        self.sizes = {}

    def close(self):
        if self.h:
            check(_lib.hhs_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, pos, vel, m, t_end, dt_root) -> int:
        import numpy as np
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        m = np.ascontiguousarray(m, dtype=np.float64)
        t = C.c_longlong()
        check(_lib.hhs_server_submit(self.h, len(m), pos.ctypes.data, vel.ctypes.data,
                                     m.ctypes.data, t_end, dt_root, C.byref(t)))
        self.sizes[t.value] = len(m)
        return t.value

    @staticmethod
    def _info(ri):
        r = ri.result
        return {"t_submit": ri.t_submit, "t_admit": ri.t_admit, "t_start": ri.t_start,
                "t_done": ri.t_done, "path": ri.path, "state": ri.state,
                "result": {"e0": r.e0, "e1": r.e1, "t": r.t, "calls": r.calls,
                           "interactions": r.interactions, "ts_pairs": r.ts_pairs,
                           "max_level": r.max_level, "status": r.status}}

    def poll(self, ticket: int) -> dict:
        ri = RequestInfo()
        check(_lib.hhs_server_poll(self.h, ticket, C.byref(ri)))
        return self._info(ri)

    def wait(self, ticket: int) -> dict:
        ri = RequestInfo()
        _lib.hhs_server_wait(self.h, ticket, C.byref(ri))
        return self._info(ri)

    def fetch(self, ticket: int):
        import numpy as np
        n = self.sizes[ticket]
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        check(_lib.hhs_server_fetch(self.h, ticket, pos.ctypes.data, vel.ctypes.data))
        return pos, vel

    def release(self, ticket: int):
        check(_lib.hhs_server_release(self.h, ticket))
        self.sizes.pop(ticket, None)

    def stats(self) -> dict:
        st = ServerStats()
        check(_lib.hhs_server_stats(self.h, C.byref(st)))
        return {"now": st.now, "used_bytes": st.used_bytes, "peak_bytes": st.peak_bytes,
                "pending": st.pending, "running_batches": st.running_batches,
                "running_big": st.running_big}


KICK_MIXED32, KICK_F64_RSQ32, KICK_F64, KICK_F64_EXACT = 0, 1, 2, 3


def kick_set_precision(mode: int):
    check(_lib.hhs_kick_set_precision(mode))


def kick_get_precision() -> int:
    return int(_lib.hhs_kick_get_precision())
