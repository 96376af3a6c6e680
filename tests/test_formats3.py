"""The 3-D binary step format (code/snapshot/warp_weno5_3d_rk3/binary_io.py):
header (nz, ny, nx: uint32, time: float64, native byte order) + the
(nz, ny, nx, 5) FP64 primitive array; truncated files and trailing bytes are
errors; compare_steps reports bitwise equality and ULP distance."""
from __future__ import annotations

import struct

import numpy as np
import pytest


def test_step_round_trip(tmp_path):
    from weno_sr_publishing_amd.euler3d import read_step, write_step
    rng = np.random.default_rng(0)
    a = rng.normal(size=(3, 4, 5, 5))
    p = tmp_path / "step_0001.bin"
    write_step(p, 1.25e-5, a)
    raw = p.read_bytes()
    assert struct.unpack_from("=IIId", raw) == (3, 4, 5, 1.25e-5)
    assert len(raw) == struct.calcsize("=IIId") + a.size * 8
    t, b = read_step(p)
    assert t == 1.25e-5 and np.array_equal(a, b)


def test_step_errors(tmp_path):
    from weno_sr_publishing_amd.euler3d import read_step, write_step
    a = np.zeros((2, 2, 2, 5))
    p = tmp_path / "s.bin"
    write_step(p, 0.0, a)
    p.write_bytes(p.read_bytes()[:-8])
    with pytest.raises(ValueError, match="truncated payload"):
        read_step(p)
    write_step(p, 0.0, a)
    p.write_bytes(p.read_bytes() + b"\0")
    with pytest.raises(ValueError, match="trailing"):
        read_step(p)
    p.write_bytes(b"\0" * 7)
    with pytest.raises(ValueError, match="truncated step header"):
        read_step(p)
    with pytest.raises(ValueError):
        write_step(p, 0.0, np.zeros((2, 2, 5)))


def test_compare_steps(tmp_path):
    from weno_sr_publishing_amd.euler3d import compare_steps, write_step
    a = np.linspace(1.0, 2.0, 2 * 3 * 4 * 5).reshape(2, 3, 4, 5)
    b = a.copy()
    b[1, 2, 3, 4] = np.nextafter(b[1, 2, 3, 4], 10.0)
    write_step(tmp_path / "a.bin", 1.0, a)
    write_step(tmp_path / "b.bin", 1.0, b)
    write_step(tmp_path / "c.bin", 1.0, a)
    same = compare_steps(tmp_path / "c.bin", tmp_path / "a.bin")
    assert same["file_bitwise_identical"] and same["array_bitwise_identical"] and same["max_ulp"] == 0
    diff = compare_steps(tmp_path / "b.bin", tmp_path / "a.bin")
    assert not diff["array_bitwise_identical"] and diff["unequal_count"] == 1 and diff["max_ulp"] == 1
    assert diff["header"]["time_bitwise_equal"]


def test_shockbubble_config():
    from weno_sr_publishing_amd.euler3d import ShockBubbleConfig
    c = ShockBubbleConfig()
    assert c.padded_shape == (94, 94, 230, 5)
    assert c.dx == 0.225 / 224 and c.dy == 0.089 / 88
    with pytest.raises(ValueError):
        ShockBubbleConfig(nx=4).validate()
    with pytest.raises(ValueError):
        ShockBubbleConfig(gamma=1.3).validate()
