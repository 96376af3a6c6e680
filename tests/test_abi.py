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


def test_beta_protocol_argument_checks():
    """The WENO7 beta-provider entry points refuse bad arguments before any launch."""
    L = _lib.lib()
    a = 1 << 12
    cfg = _lib.Euler7Config(nx=8, ny=8, ghost=4, scheme=_lib.SCHEMES["js"], riemann=0, boundary=0, eno_cutoff=0,
                            flags=0, loop_mode=0, reserved=0, dx=0.1, dy=0.1, gamma=1.4, cfl=0.4, t_end=1.0)
    c = ctypes.byref(cfg)
    assert L.wsr_beta7_workspace_bytes() >= 908 * 8
    # the fill calls need an SR scheme and its parameters
    assert L.wsr_beta7_fill_normal(c, a, a, 1, a, 1 << 20, None) == -1
    cfg.scheme = _lib.SCHEMES["sr64"]
    assert L.wsr_beta7_fill_normal(c, a, a, 1, a, 1 << 20, None) == -2
    params = (ctypes.c_double * 908)()
    cfg.mlp = ctypes.cast(params, ctypes.c_void_p)
    assert L.wsr_beta7_fill_normal(c, a, a, 3, a, 1 << 20, None) == -1          # direction
    assert L.wsr_beta7_fill_normal(c, a, a, 1, a, 16, None) == -3               # workspace
    assert L.wsr_beta7_fill_cross(c, a, a, a + 8, 2, a, 1 << 20, None) == -1     # alignment
    assert L.wsr_euler7_stage_ext(c, a, a, a, None, 1, None) == -1
    assert L.wsr_euler7_flux_ext(c, a, a, a, a, 1, 3, 0.1, 0, None) == -1       # Gauss location
    cfg.riemann = 9
    assert L.wsr_euler7_flux_ext(c, a, a, a, a, 1, 1, 0.1, 0, None) == -1
    cfg.riemann = 0
    assert L.wsr_euler7_rhs_from_flux(c, a, a, a, None, a, None) == -1
    inputs = (ctypes.c_void_p * 2)(a, a)
    assert L.wsr_euler7_shu_combine(c, 5, 0.1, inputs, a, None) == -1
    assert L.wsr_euler7_shu_combine(c, 2, 0.1, (ctypes.c_void_p * 4)(a, a, None, a), a, None) == -1
    assert L.wsr_euler7_max_speed(c, a, a, 64, ctypes.byref(ctypes.c_double()), None) == -3
    cfg.ghost = 3
    assert L.wsr_euler7_boundary(c, a, 0.0, None) == -1
