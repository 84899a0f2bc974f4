"""Scan and split parity: bit-exact against the oracle and closed forms.

Size matrices follow the reference's own scan tests
(tests/test_scan_multi_block.cu:191-217, tests/test_scan_unrolled.cu:191-215):
512, 1023, 2048, 2151, 4090, 524288, 588021, 1048576, 2097152, plus the
edges 0, 1, 2, 3 and a tile boundary.
"""
import numpy as np
import pytest

from conftest import plummer

pytestmark = pytest.mark.gpu

REF_SIZES = [512, 1023, 2048, 2151, 4090, 524288, 588021, 1048576, 2097152]


@pytest.fixture(scope="module")
def ws(hhs, torch_cuda):
    return hhs.ScanWorkspace(1 << 22)


@pytest.mark.parametrize("n", [1, 2, 3, 255, 256, 257, 65536, 65537] + REF_SIZES)
def test_scan_predicates_bit_exact(hhs, torch_cuda, ws, n):
    torch = torch_cuda
    rng = np.random.default_rng(n)
    x = (rng.random(n) * 10.0 <= 5.0).astype(np.int32)
    inp = torch.as_tensor(x).cuda()
    out = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    hhs.scan_exclusive(inp, out, ws)
    torch.cuda.synchronize()
    ref = np.concatenate([[0], np.cumsum(x)[:-1]]).astype(np.int32)
    assert np.array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("n", [1, 100, 4090, 588021])
def test_scan_general_ints(hhs, torch_cuda, ws, n):
    torch = torch_cuda
    x = np.random.default_rng(5).integers(-1000, 1000, n).astype(np.int32)
    inp = torch.as_tensor(x).cuda()
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    hhs.scan_exclusive(inp, out, ws)
    torch.cuda.synchronize()
    ref = np.concatenate([[0], np.cumsum(x.astype(np.int64))[:-1]]).astype(np.int32)
    assert np.array_equal(out.cpu().numpy(), ref)


def test_scan_repeated_calls_reuse_workspace(hhs, torch_cuda, ws):
    torch = torch_cuda
    for k, n in enumerate([5000, 300000, 17, 300000, 1 << 20]):
        x = np.random.default_rng(k).integers(0, 3, n).astype(np.int32)
        inp = torch.as_tensor(x).cuda()
        out = torch.empty(n, dtype=torch.int32, device="cuda")
        hhs.scan_exclusive(inp, out, ws)
        torch.cuda.synchronize()
        ref = np.concatenate([[0], np.cumsum(x)[:-1]]).astype(np.int32)
        assert np.array_equal(out.cpu().numpy(), ref), f"call {k} n={n}"


@pytest.mark.parametrize("n", [512, 4090, 2048, 2151] + REF_SIZES[5:])
def test_split_values_reference_matrix(hhs, oracle, torch_cuda, ws, n):
    """The reference's property checks (sizes sum to N; every fast value is
    <= pivot; every slow value is > pivot) plus bit-exact stable order and
    counts equal to the CPU partition (src/cpu_scan.h)."""
    torch = torch_cuda
    vals = (np.random.default_rng(n).random(n) * 10.0).astype(np.float32)
    pivot = 5.0
    t = torch.as_tensor(vals).cuda()
    slow = torch.empty_like(t)
    fast = torch.empty_like(t)
    part = torch.zeros(2, dtype=torch.int32, device="cuda")
    hhs.split_values(t, slow, fast, pivot, part, ws)
    torch.cuda.synchronize()
    ns, nf = part.cpu().numpy().tolist()
    assert ns + nf == n
    s, f = slow.cpu().numpy()[:ns], fast.cpu().numpy()[:nf]
    assert np.all(f <= pivot) and np.all(s > pivot)
    assert np.array_equal(f, vals[vals <= pivot]) and np.array_equal(s, vals[vals > pivot])
    _, cs, cf, rc = oracle.cpu_split_f32(vals, pivot)
    assert rc == 0 and (cs, cf) == (ns, nf)


def _split_case(hhs, oracle, torch, ws, pos, vel, m, ts, pivot):
    n = len(m)
    host = oracle.HostParticles(n, pos, vel, m, ts)
    o_slow, o_fast, o_ns, o_nf, o_pred, o_scan = oracle.split(host, pivot)
    total = hhs.DeviceParticles.from_host(pos, vel, m, ts)
    slow = hhs.DeviceParticles.empty(n)
    fast = hhs.DeviceParticles.empty(n)
    part = torch.zeros(2, dtype=torch.int32, device="cuda")
    pred = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    scan = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    hhs.split(total, slow, fast, pivot, part, ws, predicate=pred, scanned=scan)
    torch.cuda.synchronize()
    ns, nf = part.cpu().numpy().tolist()
    assert (ns, nf) == (o_ns, o_nf)
    assert np.array_equal(pred.cpu().numpy(), o_pred)
    assert np.array_equal(scan.cpu().numpy(), o_scan)
    for dev, ref, k in ((slow, o_slow, ns), (fast, o_fast, nf)):
        for f in ("pos", "vel", "m", "timestep", "id", "parent_id"):
            a = getattr(dev, f).cpu().numpy()[:k]
            b = getattr(ref, f)[:k]
            # bit patterns, so NaN timesteps compare equal to themselves
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f"field {f}"
    return ns, nf


@pytest.mark.parametrize("n", [1, 2, 3, 64, 1000, 1024, 4097, 65536, 131072, 200003])
def test_split_particles_bit_exact(hhs, oracle, torch_cuda, ws, n):
    rng = np.random.default_rng(n)
    pos, vel = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    m = rng.random(n)
    ts = 2.0 ** rng.uniform(-12, 0, n)
    _split_case(hhs, oracle, torch_cuda, ws, pos, vel, m, ts, 2.0 ** -6)


def test_split_lone_fast_particle_stays_slow(hhs, oracle, torch_cuda, ws):
    """src/split.cuh:321-324: N_fast == 1 becomes N_fast = 0, all slow."""
    for n, k in ((5, 0), (5, 4), (1000, 617), (70000, 69999)):
        rng = np.random.default_rng(n + k)
        pos, vel, m = rng.standard_normal((n, 3)), rng.standard_normal((n, 3)), rng.random(n)
        ts = np.ones(n)
        ts[k] = 1e-3
        ns, nf = _split_case(hhs, oracle, torch_cuda, ws, pos, vel, m, ts, 0.5)
        assert (ns, nf) == (n, 0)


def test_split_all_fast_all_slow_nan(hhs, oracle, torch_cuda, ws):
    n = 3000
    rng = np.random.default_rng(1)
    pos, vel, m = rng.standard_normal((n, 3)), rng.standard_normal((n, 3)), rng.random(n)
    assert _split_case(hhs, oracle, torch_cuda, ws, pos, vel, m, np.full(n, 0.1), 1.0) == (0, n)
    assert _split_case(hhs, oracle, torch_cuda, ws, pos, vel, m, np.full(n, 2.0), 1.0) == (n, 0)
    ts = np.full(n, 0.1)
    ts[::7] = np.nan  # NaN <= pivot is false: slow
    ns, nf = _split_case(hhs, oracle, torch_cuda, ws, pos, vel, m, ts, 1.0)
    assert ns == len(ts[::7])


def test_split_empty(hhs, torch_cuda, ws):
    torch = torch_cuda
    total = hhs.DeviceParticles.empty(4)
    part = torch.full((2,), 9, dtype=torch.int32, device="cuda")
    hhs.split(total, hhs.DeviceParticles.empty(4), hhs.DeviceParticles.empty(4), 1.0, part, ws, n=0)
    torch.cuda.synchronize()
    assert part.cpu().numpy().tolist() == [0, 0]


def test_split_plummer_timesteps(hhs, oracle, torch_cuda, ws):
    """Split of real timesteps at every power-of-two pivot."""
    n = 1024
    pos, vel, m = plummer(hhs, n, seed=1)
    ts = oracle.timesteps(pos, vel, m, 0.01, 0.0)
    for k in range(0, 14):
        _split_case(hhs, oracle, torch_cuda, ws, pos, vel, m, ts, 2.0 ** -k)


def test_split_f32(hhs, torch_cuda, ws):
    torch = torch_cuda
    n = 5000
    rng = np.random.default_rng(3)
    pos, vel, m = rng.standard_normal((n, 3)), rng.standard_normal((n, 3)), rng.random(n)
    ts = rng.random(n)
    total = hhs.DeviceParticles.from_host(pos, vel, m, ts, dtype=torch.float32)
    slow = hhs.DeviceParticles.empty(n, torch.float32)
    fast = hhs.DeviceParticles.empty(n, torch.float32)
    part = torch.zeros(2, dtype=torch.int32, device="cuda")
    hhs.split(total, slow, fast, 0.25, part, ws)
    torch.cuda.synchronize()
    ts32 = ts.astype(np.float32)
    ns, nf = part.cpu().numpy().tolist()
    sel = ts32 <= np.float32(0.25)
    assert nf == sel.sum()
    assert np.array_equal(fast.parent_id.cpu().numpy()[:nf], np.nonzero(sel)[0])
    assert np.array_equal(slow.parent_id.cpu().numpy()[:ns], np.nonzero(~sel)[0])


def test_scan_lookback_epoch_wrap(hhs, torch_cuda):
    """The look-back scan's status words carry a 31-bit call epoch kept in the
    workspace header (word 0); before it wraps, the call's last tile clears
    every status word.  Set the epoch just below the wrap from the host and
    run scans of several sizes across it: every result stays exact and the
    epoch restarts at 1.  Scans above 2^18 elements take the reduce-then-scan,
    whose chunk counts lie beyond every look-back status word: interleaved,
    they leave the epoch alone."""
    torch = torch_cuda
    ws = hhs.ScanWorkspace(1 << 21)
    hdr = ws.buf[:4].view(torch.int32)
    hdr.copy_(torch.tensor([2 ** 31 - 3], dtype=torch.int32))
    torch.cuda.synchronize()
    for k, n in enumerate([200000, 5000, 700000, 100, 1 << 18, 150001, 1 << 21, 4096]):
        x = np.random.default_rng(40 + k).integers(0, 2, n).astype(np.int32)
        inp = torch.as_tensor(x).cuda()
        out = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        hhs.scan_exclusive(inp, out, ws)
        torch.cuda.synchronize()
        ref = np.concatenate([[0], np.cumsum(x)[:-1]]).astype(np.int32)
        assert np.array_equal(out.cpu().numpy(), ref), f"call {k} n={n}"
    # look-back calls 0-1 used epochs 2^31-2 and 2^31-1 (the last one
    # clears), then calls 3, 4, 5, 7 used 1..4; calls 2 and 6 left it alone
    assert int(ws.buf[:4].view(torch.int32).cpu()[0]) == 4


@pytest.mark.parametrize("n", [1, 1000, 4096, 300001])
def test_scan_in_place(hhs, torch_cuda, ws, n):
    """in == out: the in-place scan (reduce-then-scan path) is exact too."""
    torch = torch_cuda
    x = np.random.default_rng(n).integers(-50, 50, n).astype(np.int32)
    buf = torch.as_tensor(x).cuda()
    hhs.scan_exclusive(buf, buf, ws)
    torch.cuda.synchronize()
    ref = np.concatenate([[0], np.cumsum(x.astype(np.int64))[:-1]]).astype(np.int32)
    assert np.array_equal(buf.cpu().numpy(), ref)
