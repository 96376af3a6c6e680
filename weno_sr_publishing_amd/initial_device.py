"""Initial states built on the GPU (``wsr_init_*``), equal to the host builders of
``initial.py`` (bit for bit for the piecewise-constant problems; the vortex up to
the last bits of exp and pow).  Each returns a torch FP64 tensor on ``device`` in
the layout the solvers take, so that the large reference runs (1200 x 1200 line-cut
references, 1200 x 300 double Mach, the 224 x 88 x 88 shock-bubble) never build
their initial data on the host.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .initial import _GAUSS3_NODES, _GAUSS3_WEIGHTS, CASES
from .params import GAUSS15_W, GAUSS15_XI


def _dev(device) -> torch.device:
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError("the device initial states are built on the GPU")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


def _quad(nodes, weights):
    x = np.ascontiguousarray(nodes, dtype=np.float64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    return x, w, x.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p), int(x.size)


def _stream(dev) -> int:
    return torch.cuda.current_stream(dev).cuda_stream


def _quadrant(nx, ny, ghost, wrap, definition, quadrature, x_min, x_max, y_min, y_max, gamma, dev):
    states = np.ascontiguousarray(np.array(definition["states"], dtype=np.float64).reshape(16))
    x, w, px, pw, n = _quad(*np.polynomial.legendre.leggauss(quadrature))
    u = torch.empty((ny + 2 * ghost, nx + 2 * ghost, 4), dtype=torch.float64, device=dev)
    rc = _lib.lib().wsr_init_quadrant(u.data_ptr(), nx, ny, ghost, int(wrap), states.ctypes.data_as(ctypes.c_void_p),
                                      x_min, x_max, y_min, y_max, gamma, px, pw, n, _stream(dev))
    _lib.check(rc, "wsr_init_quadrant")
    return u


def quadrant_initial(nx: int, ny: int, case: str = "c4", ghost: int = 3, quadrature: int = 15,
                     device="cuda") -> torch.Tensor:
    """initial.make_quadrant_initial on the device (interior averages, periodic ghosts)."""
    return _quadrant(nx, ny, ghost, True, CASES[case], quadrature, 0.0, 1.0, 0.0, 1.0, 1.4, _dev(device))


def quadrant_state(nx: int, ny: int, ghost: int, definition: dict, quadrature: int = 15, x_min: float = 0.0,
                   x_max: float = 1.0, y_min: float = 0.0, y_max: float = 1.0, gamma: float = 1.4,
                   device="cuda") -> torch.Tensor:
    """initial.make_quadrant_state on the device (every padded cell its own average)."""
    return _quadrant(nx, ny, ghost, False, definition, quadrature, x_min, x_max, y_min, y_max, gamma, _dev(device))


def sod_state(params, axis: str = "x", device="cuda") -> torch.Tensor:
    """initial.make_sod_state on the device."""
    dev = _dev(device)
    x, w, px, pw, n = _quad(GAUSS15_XI, GAUSS15_W)
    u = torch.empty(params.padded_shape, dtype=torch.float64, device=dev)
    rc = _lib.lib().wsr_init_sod(u.data_ptr(), params.nx, params.ny, params.ghost, 2 if axis == "y" else 1,
                                 params.x_length, params.y_length, params.gamma, px, pw, n, _stream(dev))
    _lib.check(rc, "wsr_init_sod")
    return u


def double_mach_state(nx: int = 1200, ny: int = 300, ghost: int = 4, x_length: float = 4.0, y_length: float = 1.0,
                      gamma: float = 1.4, device="cuda") -> torch.Tensor:
    """initial.make_double_mach_state (15-point table) on the device."""
    dev = _dev(device)
    x, w, px, pw, n = _quad(GAUSS15_XI, GAUSS15_W)
    u = torch.empty((ny + 2 * ghost, nx + 2 * ghost, 4), dtype=torch.float64, device=dev)
    rc = _lib.lib().wsr_init_double_mach(u.data_ptr(), nx, ny, ghost, x_length, y_length, gamma, px, pw, n,
                                         _stream(dev))
    _lib.check(rc, "wsr_init_double_mach")
    return u


def vortex_state(nx: int, ny: int, t: float = 0.0, quadrature: int = 15, ghost: int = 4, x_min: float = -10.0,
                 x_max: float = 10.0, y_min: float = -10.0, y_max: float = 10.0, gamma: float = 1.4,
                 device="cuda") -> torch.Tensor:
    """initial.make_vortex_state on the device."""
    dev = _dev(device)
    x, w, px, pw, n = _quad(*np.polynomial.legendre.leggauss(quadrature))
    u = torch.empty((ny + 2 * ghost, nx + 2 * ghost, 4), dtype=torch.float64, device=dev)
    rc = _lib.lib().wsr_init_vortex(u.data_ptr(), nx, ny, ghost, t, x_min, x_max, y_min, y_max, gamma, px, pw, n,
                                    _stream(dev))
    _lib.check(rc, "wsr_init_vortex")
    return u


def _bubble(nx, ny, nz, ghost, kind, lengths, shock, ratio, gamma, dev):
    x, w, px, pw, n = _quad(_GAUSS3_NODES, _GAUSS3_WEIGHTS)
    L = np.ascontiguousarray(lengths, dtype=np.float64)
    S = np.ascontiguousarray(shock, dtype=np.float64)
    q = torch.empty((nz + 2 * ghost, ny + 2 * ghost, nx + 2 * ghost, 5), dtype=torch.float64, device=dev)
    rc = _lib.lib().wsr_init_shockbubble3d(q.data_ptr(), nx, ny, nz, ghost, kind, L.ctypes.data_as(ctypes.c_void_p),
                                           S.ctypes.data_as(ctypes.c_void_p), ratio, gamma, px, pw, n, _stream(dev))
    _lib.check(rc, "wsr_init_shockbubble3d")
    return q


def shockbubble_bench_state(nx: int = 224, ny: int = 88, nz: int = 88, ghost: int = 3, mach: float = 3.0,
                            density_ratio: float = 0.138, gamma: float = 1.4, device="cuda") -> torch.Tensor:
    """initial.make_shockbubble_bench_state on the device (the same post-shock values)."""
    rho_post = ((gamma + 1.0) * mach * mach) / ((gamma - 1.0) * mach * mach + 2.0)
    p_post = (2.0 * gamma * mach * mach - (gamma - 1.0)) / (gamma + 1.0)
    u_post = mach * np.sqrt(gamma) * (1.0 - 1.0 / rho_post)
    return _bubble(nx, ny, nz, ghost, 0, (float(nx), float(ny), float(nz)), (rho_post, u_post, p_post),
                   density_ratio, gamma, _dev(device))


def shockbubble_ma3_state(nx: int = 224, ny: int = 88, nz: int = 88, ghost: int = 3,
                          lengths=(0.225, 0.089, 0.089), gamma: float = 1.4, device="cuda") -> torch.Tensor:
    """initial.make_shockbubble_ma3_state on the device."""
    mach, rho_air, p_air = 3.0, 1.29, 101325.0
    rho_post = (((gamma + 1.0) * mach * mach) / ((gamma - 1.0) * mach * mach + 2.0)) * rho_air
    p_post = ((((2.0 * gamma) * mach * mach) - (gamma - 1.0)) / (gamma + 1.0)) * p_air
    return _bubble(nx, ny, nz, ghost, 1, lengths, (rho_post, 736.911, p_post), 0.214, gamma, _dev(device))
