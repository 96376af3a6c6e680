"""The C-ABI library loads and exports every entry point include/wsr.h declares
(no compute call: runs without a GPU)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from weno_sr_publishing_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "wsr.h"


def declared_functions() -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(wsr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "wsr_euler2d_stage" in names and "wsr_euler2d_loop_steps" in names
    assert set(names) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_config_struct_layout():
    assert ctypes.sizeof(_lib.Euler2DConfig) == 8 * 4 + 5 * 8 + 8
    assert _lib.Euler2DConfig.dx.offset == 32
    assert _lib.Euler2DConfig.mlp.offset == 72


def test_workspace_size_and_argument_checks():
    L = _lib.lib()
    assert L.wsr_euler2d_workspace_bytes(400, 400, 2) == 0
    n = L.wsr_euler2d_workspace_bytes(400, 400, 3)
    assert n >= (406 * 406 + 400 * 401 * 2) * 4 * 8
    assert L.wsr_version().startswith(b"wsr-hip")
    cfg = _lib.Euler2DConfig(nx=8, ny=8, ghost=2, scheme=0, riemann=0, boundary=0, dx=0.1, dy=0.1, gamma=1.4)
    # rejected before anything is enqueued on a device
    assert L.wsr_euler2d_stage(ctypes.byref(cfg), 1, 1, 1, 0.1, 1, 0.0, 1, 1 << 20, None) == -1
    cfg.ghost = 3
    cfg.scheme = 2  # SR32 without parameters
    assert L.wsr_euler2d_stage(ctypes.byref(cfg), 1, 1, 1, 0.1, 1, 0.0, 1, 1 << 20, None) == -2
    cfg.scheme = 0
    a = 1 << 12     # stand-in addresses: never dereferenced, the checks return first
    assert L.wsr_euler2d_stage(ctypes.byref(cfg), a, a, a, 0.1, 1, 0.0, a, 16, None) == -3
    # the kernels move 32-byte records as 16-byte halves: misaligned pointers are refused
    assert L.wsr_euler2d_stage(ctypes.byref(cfg), a + 8, a, a, 0.1, 1, 0.0, a, 16, None) == -1
    assert L.wsr_euler2d_stage(ctypes.byref(cfg), a, a, a, 0.1, 1, 0.0, a + 8, 16, None) == -1
    cfg.boundary = 7
    assert L.wsr_euler2d_stage(ctypes.byref(cfg), 1, 1, 1, 0.1, 1, 0.0, 1, 1 << 20, None) == -1


def test_adapters_reject_cpu_device():
    from weno_sr_publishing_amd.solvers import _adapter
    from weno_sr_publishing_amd.params import Params
    import numpy as np
    p = Params(nx=4, ny=4)
    with pytest.raises(ValueError):
        _adapter.allocate_forward(np.zeros(p.padded_shape), p, "cpu")
