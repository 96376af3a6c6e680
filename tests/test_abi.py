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


def test_initial_state_argument_checks():
    """The device initial-state builders refuse bad arguments before any launch."""
    L = _lib.lib()
    states = (ctypes.c_double * 16)(*([1.0] * 16))
    nodes = (ctypes.c_double * 2)(-0.5, 0.5)
    weights = (ctypes.c_double * 2)(1.0, 1.0)
    a = 1 << 12
    assert L.wsr_init_quadrant(None, 8, 8, 3, 1, states, 0.0, 1.0, 0.0, 1.0, 1.4, nodes, weights, 2, None) == -1
    assert L.wsr_init_quadrant(a, 0, 8, 3, 1, states, 0.0, 1.0, 0.0, 1.0, 1.4, nodes, weights, 2, None) == -1
    # quadrature orders 1..16 only
    assert L.wsr_init_quadrant(a, 8, 8, 3, 1, states, 0.0, 1.0, 0.0, 1.0, 1.4, nodes, weights, 17, None) == -1
    assert L.wsr_init_sod(a, 8, 8, 3, 3, 1.0, 1.0, 1.4, nodes, weights, 2, None) == -1
    assert L.wsr_init_double_mach(a, 8, 8, 4, 4.0, 1.0, 1.4, None, weights, 2, None) == -1
    assert L.wsr_init_vortex(a, 8, -1, 4, 0.0, -10.0, 10.0, -10.0, 10.0, 1.4, nodes, weights, 2, None) == -1
    lengths = (ctypes.c_double * 3)(1.0, 1.0, 1.0)
    assert L.wsr_init_shockbubble3d(a, 8, 8, 8, 3, 2, lengths, lengths, 0.1, 1.4, nodes, weights, 2, None) == -1
