"""Host-side state metrics used by the runners.

Restates code/snapshot/for_paper_results/common.py:78-121 (primitive, interior,
state_health, rho_errors, conservative_block_average) and
code/snapshot/warp_weno5_helpers.py:165-224 (primitive_from_conserved,
compute_dt, interior_stats).
"""
from __future__ import annotations

import numpy as np

GAMMA = 1.4


def primitive(conserved: np.ndarray, gamma: float = GAMMA) -> np.ndarray:
    rho = conserved[..., 0]
    safe_rho = np.maximum(rho, 1.0e-300)
    u = conserved[..., 1] / safe_rho
    v = conserved[..., 2] / safe_rho
    p = (gamma - 1.0) * (conserved[..., 3] - 0.5 * rho * (u * u + v * v))
    return np.stack((rho, u, v, p), axis=-1)


def interior(state: np.ndarray, ghost: int, nx: int, ny: int) -> np.ndarray:
    return state[ghost: ghost + ny, ghost: ghost + nx, :]


def state_health(state: np.ndarray, ghost: int, nx: int, ny: int) -> dict:
    q = interior(state, ghost, nx, ny)
    pri = primitive(q)
    finite = np.isfinite(q).all() and np.isfinite(pri).all()
    return {
        "complete": bool(finite and np.min(pri[..., 0]) > 0.0 and np.min(pri[..., 3]) > 0.0),
        "nan_count": int(np.count_nonzero(~np.isfinite(q))),
        "rho_min": float(np.nanmin(pri[..., 0])),
        "rho_max": float(np.nanmax(pri[..., 0])),
        "p_min": float(np.nanmin(pri[..., 3])),
        "p_max": float(np.nanmax(pri[..., 3])),
    }


def rho_errors(numerical: np.ndarray, reference: np.ndarray) -> dict:
    delta = numerical - reference
    return {
        "rho_l1": float(np.mean(np.abs(delta))),
        "rho_l2": float(np.sqrt(np.mean(delta * delta))),
        "rho_linf": float(np.max(np.abs(delta))),
    }


def conservative_block_average(state: np.ndarray, target_n: int) -> np.ndarray:
    if state.shape[0] != state.shape[1] or state.shape[0] % target_n:
        raise ValueError(f"cannot block-average shape {state.shape} to {target_n}x{target_n}")
    factor = state.shape[0] // target_n
    return state.reshape(target_n, factor, target_n, factor, state.shape[-1]).mean(axis=(1, 3))


def primitive_from_conserved(u: np.ndarray, gamma: float) -> np.ndarray:
    rho = np.maximum(u[..., 0], 1.0e-15)
    vel_x = u[..., 1] / rho
    vel_y = u[..., 2] / rho
    p = (gamma - 1.0) * (u[..., 3] - 0.5 * rho * (vel_x * vel_x + vel_y * vel_y))
    return np.stack([rho, vel_x, vel_y, p], axis=-1)


def compute_dt(u_host: np.ndarray, params) -> float:
    """Host CFL rule of warp_weno5_helpers.py:174-186."""
    gc = params.ghost
    inner = u_host[gc: gc + params.ny, gc: gc + params.nx, :]
    pri = primitive_from_conserved(inner, params.gamma)
    rho = np.maximum(pri[..., 0], 1.0e-15)
    p = np.maximum(pri[..., 3], 1.0e-15)
    sound = np.sqrt(params.gamma * p / rho)
    speed = np.maximum(np.abs(pri[..., 1]) + sound, np.abs(pri[..., 2]) + sound)
    max_speed = float(np.max(speed))
    if max_speed < 1.0e-15:
        return 1.0e10
    return params.cfl * min(params.dx, params.dy) / max_speed


def interior_stats(u_host: np.ndarray, params) -> dict:
    gc = params.ghost
    inner = u_host[gc: gc + params.ny, gc: gc + params.nx, :]
    pri = primitive_from_conserved(inner, params.gamma)
    mass = float(np.sum(inner[..., 0]) * params.dx * params.dy)
    return {
        "mass": mass,
        "rho_min": float(np.min(pri[..., 0])),
        "rho_max": float(np.max(pri[..., 0])),
        "p_min": float(np.min(pri[..., 3])),
        "p_max": float(np.max(pri[..., 3])),
        "nan_count": float(np.isnan(inner).sum()),
        "rho_neg": float(np.sum(pri[..., 0] <= 0.0)),
        "p_neg": float(np.sum(pri[..., 3] <= 0.0)),
    }
