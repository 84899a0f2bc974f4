"""Snapshot file format checks that need no GPU: header layout and the
validation hhs_snapshot_read_header applies (magic, version, sizes)."""
import ctypes as C
import struct

import numpy as np
import pytest


def _header(n, magic=b"HHSSNAP\0", version=1, header_bytes=128, fnv=0):
    h = struct.pack("<8sIIqqddddiiQ", magic, version, header_bytes, n, 3, 0.75, 0.25, 0.01,
                    1e-3, 0, -1, fnv)
    return h + bytes(128 - len(h))


def _fnv1a(b: bytes) -> int:
    h = 0xcbf29ce484222325
    for x in b:
        h = ((h ^ x) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_header_is_128_bytes(hhs):
    assert C.sizeof(hhs.SnapshotHeader) == 128
    assert hhs.SnapshotHeader.payload_fnv1a.offset == 72


def test_read_header_accepts_a_whole_file(hhs, tmp_path):
    n = 5
    payload = np.arange(7 * n, dtype=np.float64).tobytes() + np.arange(n, dtype=np.int32).tobytes()
    assert len(payload) == 60 * n
    p = tmp_path / "ok.snap"
    p.write_bytes(_header(n, fnv=_fnv1a(payload)) + payload)
    h = hhs.read_snapshot_header(str(p))
    assert h["n"] == n and h["root_steps"] == 3 and h["simtime"] == 0.75
    assert h["dt_root"] == 0.25 and h["eta"] == 0.01 and h["eps"] == 1e-3
    assert h["timestep_mode"] == 0 and h["kick_precision"] == -1
    assert h["payload_fnv1a"] == _fnv1a(payload)


@pytest.mark.parametrize("case", ["magic", "version", "header", "short", "long", "empty"])
def test_read_header_rejects_damaged_files(hhs, tmp_path, case):
    n = 4
    payload = bytes(60 * n)
    blob = {
        "magic": _header(n, magic=b"HHSSNAQ\0") + payload,
        "version": _header(n, version=2) + payload,
        "header": _header(n, header_bytes=64) + payload,
        "short": _header(n) + payload[:-1],
        "long": _header(n) + payload + b"x",
        "empty": b"",
    }[case]
    p = tmp_path / f"{case}.snap"
    p.write_bytes(blob)
    with pytest.raises(hhs.HHSError) as ei:
        hhs.read_snapshot_header(str(p))
    assert ei.value.code == hhs.HHS_ERR_FORMAT


def test_read_header_missing_file(hhs, tmp_path):
    with pytest.raises(hhs.HHSError) as ei:
        hhs.read_snapshot_header(str(tmp_path / "none.snap"))
    assert ei.value.code == hhs.HHS_ERR_IO


def test_evolve_and_snapshot_argument_checks(hhs):
    lib = hhs.raw()
    assert lib.hhs_evolve_f64(None, None, 1.0, 0.5, None, 0, None) == hhs.HHS_ERR_INVALID
    assert lib.hhs_snapshot_write_f64(b"x", None, None, 0, 1.0, None) == hhs.HHS_ERR_INVALID
    assert lib.hhs_snapshot_read_f64(b"x", None, None, None, None) == hhs.HHS_ERR_INVALID
