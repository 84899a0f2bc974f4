"""hhs_split_f64_noalloc (the split with a library-owned static scratch and
no allocation) against hhs_split_f64 with a caller workspace: every output
bit for bit -- slow and fast records, parent ids, predicate, scanned
predicate and the partition -- on two calls in a row, the second smaller
(a multi-chunk split, then a single-chunk one), and a lone fast particle."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _set(hhs, torch, n, seed, pivot_frac):
    g = torch.Generator(device="cuda").manual_seed(seed)
    p = hhs.DeviceParticles.empty(n)
    for f in ("pos", "vel", "m"):
        t = getattr(p, f)
        t.copy_(torch.rand(t.shape, generator=g, dtype=torch.float64, device="cuda"))
    p.timestep.copy_(torch.rand(n, generator=g, dtype=torch.float64, device="cuda"))
    p.id.copy_(torch.arange(n, dtype=torch.int32, device="cuda") * 3 + 1)
    return p, float(pivot_frac)


def _outputs(hhs, torch, total, pivot, noalloc):
    n = total.n
    slow, fast = hhs.DeviceParticles.empty(n), hhs.DeviceParticles.empty(n)
    for d in (slow, fast):  # a sentinel in every slot, so unwritten slots compare too
        for f in ("pos", "vel", "acc", "timestep", "m"):
            getattr(d, f).fill_(-7.25)
        d.id.fill_(-3)
        d.parent_id.fill_(-5)
    pred = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    scan = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    part = torch.full((2,), -9, dtype=torch.int32, device="cuda")
    t, s, f = total.struct(), slow.struct(), fast.struct()
    lib = hhs.raw()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        rc = lib.hhs_split_f64_noalloc(C.byref(t), C.byref(s), C.byref(f),
                                       C.c_void_p(pred.data_ptr()), C.c_void_p(scan.data_ptr()),
                                       pivot, n, C.c_void_p(part.data_ptr()), stream)
    else:
        ws = hhs.ScanWorkspace(n)
        rc = lib.hhs_split_f64(C.byref(t), C.byref(s), C.byref(f), C.c_void_p(pred.data_ptr()),
                               C.c_void_p(scan.data_ptr()), pivot, n,
                               C.c_void_p(part.data_ptr()), ws.ptr, ws.nbytes, stream)
    assert rc == 0, hhs.last_error()
    torch.cuda.synchronize()
    out = {"part": part.cpu().numpy(), "pred": pred.cpu().numpy(), "scan": scan.cpu().numpy()}
    for name, d in (("slow", slow), ("fast", fast)):
        for fld in ("pos", "vel", "timestep", "m", "id", "parent_id"):
            out[f"{name}.{fld}"] = getattr(d, fld).cpu().numpy()
    return out


def _assert_same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert a[k].tobytes() == b[k].tobytes(), k


@pytest.mark.parametrize("sizes", [(300000, 1500)])
def test_split_noalloc_matches_split(hhs, torch_cuda, sizes):
    torch = torch_cuda
    for k, n in enumerate(sizes):  # two calls in a row, the second smaller
        total, pivot = _set(hhs, torch, n, 100 + k, 0.4)
        ref = _outputs(hhs, torch, total, pivot, noalloc=False)
        got = _outputs(hhs, torch, total, pivot, noalloc=True)
        _assert_same(ref, got)
        assert int(ref["part"][0]) + int(ref["part"][1]) == n and int(ref["part"][1]) > 0


def test_split_noalloc_lone_fast_particle(hhs, torch_cuda):
    torch = torch_cuda
    n = 777
    total, _ = _set(hhs, torch, n, 5, 0.0)
    ts = total.timestep.cpu().numpy()
    pivot = float(np.sort(ts)[0])  # exactly one timestep <= pivot
    ref = _outputs(hhs, torch, total, pivot, noalloc=False)
    got = _outputs(hhs, torch, total, pivot, noalloc=True)
    _assert_same(ref, got)
    assert ref["part"].tolist() == [n, 0]  # the lone fast particle stays slow
