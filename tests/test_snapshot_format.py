"""Snapshot file format round trip (host side; the device copy path is in
tests/test_gpu_snapshots.py)."""
import numpy as np
import pytest


def test_snapshot_file_round_trip(tmp_path):
    from atum_hip.snapshots import read_snapshot, write_snapshot
    a = np.random.default_rng(1).standard_normal((5, 123))
    p = tmp_path / "s.atm"
    write_snapshot(p, a, {"step": 7, "time": 0.25, "law": "euler"})
    b, head = read_snapshot(p)
    np.testing.assert_array_equal(a, b)
    assert head["step"] == 7 and head["time"] == 0.25 and head["law"] == "euler"
    assert head["shape"] == [5, 123] and head["dtype"] == "<f8"
    i = np.arange(10, dtype=np.int32)
    write_snapshot(p, i)
    np.testing.assert_array_equal(read_snapshot(p)[0], i)


def test_snapshot_rejects_other_files(tmp_path):
    from atum_hip.snapshots import read_snapshot
    p = tmp_path / "x.bin"
    p.write_bytes(b"not a snapshot")
    with pytest.raises(ValueError):
        read_snapshot(p)
