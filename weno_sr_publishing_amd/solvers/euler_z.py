"""WENO-Z drivers of code/snapshot/weno_z_borges_p2_results/euler_z.py:12-74 on the HIP steps.

``run_weno5``: WENO5-Z (p = 2, tau = |beta0 - beta2|, epsilon = h^2) with characteristic
normal and conservative Gauss transverse reconstruction, HLLC, SSPRK3
(weno5_hllc_z_min.py:375-417).  ``run_weno7``: WENO7-Z (p = 3, tau = |-beta0 - 3 beta1 +
3 beta2 + beta3|, epsilon = h^3, the ``weight_kind = 1`` weights) with Shu RK4 and the
downwind operator (weno7_point_rk4_z.py:547-631).
"""
from __future__ import annotations

from typing import Any

import numpy as np

from ..common import interior as _interior
from ..params import Params
from . import _adapter, point_rk4

__all__ = ["make_weno5_params", "make_weno7_params", "run_weno5", "run_weno7", "interior", "primitive"]


def make_weno5_params(nx: int, ny: int, x_length: float, y_length: float, cfl: float, t_end: float) -> Params:
    return Params(nx=nx, ny=ny, x_length=x_length, y_length=y_length, cfl=cfl, t_end=t_end)


def make_weno7_params(nx: int, ny: int, x_min: float, x_max: float, y_min: float, y_max: float, cfl: float,
                      t_end: float):
    from ..euler7 import Params7
    return Params7(nx=nx, ny=ny, x_min=x_min, x_max=x_max, y_min=y_min, y_max=y_max, cfl=cfl, t_end=t_end)


def run_weno5(initial: np.ndarray, params: Params, *, device: str, boundary: str,
              report_interval: int = 0) -> tuple[np.ndarray, dict[str, Any]]:
    def report(step: int, t: float, dt: float, stats: dict) -> None:
        print(f"weno5_z_p2 step={step:05d} t={t:.8e} dt={dt:.4e} "
              f"rho=[{stats['rho_min']:.5e},{stats['rho_max']:.5e}] "
              f"p=[{stats['p_min']:.5e},{stats['p_max']:.5e}] nan={int(stats['nan_count'])}", flush=True)

    final, dts, steps, t = _adapter.run_to_time(initial, params, params.t_end, device, None, boundary, "z",
                                                report_interval, report if report_interval else None)
    return final, {
        "method": "weno5_z_p2", "steps": steps, "t": t,
        "dt_min": min(dts) if dts else 0.0,
        "dt_max": max(dts) if dts else 0.0,
        "dt_mean": float(np.mean(dts)) if dts else 0.0,
    }


def run_weno7(initial: np.ndarray, params, *, device: str, boundary: str, report_interval: int = 0,
              weight_kind: int = 1, unit_gauss_h: bool = False) -> tuple[np.ndarray, dict[str, Any]]:
    """``unit_gauss_h``: evaluate the transverse Gauss-point Z weights with h = 1 (epsilon = 240)
    instead of the transverse spacing that weno7_point_rk4_z.py:473-500 passes; the
    reference's recorded 2-D WENO7-Z results (the C.4-C.6 line-cut tables) follow h = 1."""
    if weight_kind != 1:
        raise ValueError("only the WENO7-Z weights of weight_kind = 1 (p = 3, epsilon = h^3) are built")
    from ..euler7 import Euler7Solver
    point_rk4._check("hllc", True)
    solver = Euler7Solver(params, "z", "hllc", boundary, loop="point", device=device, unit_gauss_h=unit_gauss_h)
    return point_rk4._run(solver, initial, params, report_interval, 10_000_000, "Z-RK4 ")


def interior(final: np.ndarray, ghost: int, nx: int, ny: int) -> np.ndarray:
    return _interior(final, ghost, nx, ny)


def primitive(conserved: np.ndarray, gamma: float = 1.4) -> np.ndarray:
    rho = conserved[..., 0]
    u = conserved[..., 1] / rho
    v = conserved[..., 2] / rho
    p = (gamma - 1.0) * (conserved[..., 3] - 0.5 * rho * (u * u + v * v))
    return np.stack((rho, u, v, p), axis=-1)
