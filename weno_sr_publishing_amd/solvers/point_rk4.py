"""WENO7 + Shu RK4 entry points of code/snapshot/weno7_point_rk4_shu/point_rk4.py
(``Params`` :26-55, ``cell_average_state`` :81-117, ``interior_stats``,
``run_case`` :541-564, ``run_from_initial`` :567-616) on the HIP WENO7 step.

The reference loop ``while t < t_end and step < max_steps`` with dt clipped to
t_end, a stop at dt <= 0 and health checks at report steps runs on the device
(``Euler7Solver`` loop mode "point"); the host synchronises at report steps.
"""
from __future__ import annotations

import numpy as np

from ..common import interior_stats
from ..euler7 import Euler7Solver, Params7 as Params
from ..initial import make_vortex_state

__all__ = ["Params", "cell_average_state", "interior_stats", "run_case", "run_from_initial", "summary_for"]


def cell_average_state(params: Params, t: float, quadrature: int) -> np.ndarray:
    """point_rk4.py:81-117: isentropic-vortex cell averages on the ghost-4 grid."""
    return make_vortex_state(params.nx, params.ny, t, quadrature=quadrature, ghost=params.ghost,
                             x_min=params.x_min, x_max=params.x_max, y_min=params.y_min, y_max=params.y_max,
                             gamma=params.gamma)


def _check(riemann_solver: str, characteristic: bool) -> None:
    if riemann_solver not in ("hllc", "force", "evilin"):
        raise ValueError(f"unknown riemann_solver {riemann_solver!r}")
    if not characteristic:
        raise ValueError("only the characteristic WENO7 reconstruction is built (characteristic=True)")


def summary_for(final: np.ndarray, params: Params, dts: list[float], steps: int, t: float) -> dict:
    return {
        "steps": int(steps),
        "t": float(t),
        "dt_min": float(np.min(dts)) if dts else 0.0,
        "dt_max": float(np.max(dts)) if dts else 0.0,
        "dt_mean": float(np.mean(dts)) if dts else 0.0,
        **interior_stats(final, params),
    }


def _printer(prefix: str, params: Params):
    def report(step: int, t: float, dt: float, s: dict) -> None:
        print(f"{prefix}N={params.nx} step={step:05d} t={t:.8e} dt={dt:.8e} "
              f"rho=[{s['rho_min']:.6e},{s['rho_max']:.6e}] p=[{s['p_min']:.6e},{s['p_max']:.6e}] "
              f"nan={int(s['nan_count'])}", flush=True)
    return report


def run_from_initial(u0: np.ndarray, params: Params, *, device: str, riemann_solver: str, characteristic: bool,
                     report_interval: int = 0, max_steps: int = 10_000_000,
                     boundary: str = "periodic") -> tuple[np.ndarray, dict]:
    """WENO7-JS + Shu RK4 from a padded initial state; returns (final state, summary)."""
    _check(riemann_solver, characteristic)
    solver = Euler7Solver(params, "js", riemann_solver, boundary, loop="point", device=device)
    return _run(solver, u0, params, report_interval, max_steps, "")


def _run(solver: Euler7Solver, u0, params: Params, report_interval: int, max_steps: int, prefix: str):
    u = solver.new_state(np.asarray(u0, dtype=np.float64))
    report = _printer(prefix, params) if report_interval > 0 else None
    dts, steps, t = solver.run(u, params.t_end, report_interval=report_interval, report=report,
                               check_health=report_interval > 0, max_steps=max_steps)
    final = u.cpu().numpy()
    return final, summary_for(final, params, dts, steps, t)


def run_case(params: Params, *, device: str, init_quadrature: int, riemann_solver: str, characteristic: bool,
             report_interval: int = 0, max_steps: int = 10_000_000,
             boundary: str = "periodic") -> tuple[np.ndarray, dict]:
    u0 = cell_average_state(params, 0.0, init_quadrature)
    return run_from_initial(u0, params, device=device, riemann_solver=riemann_solver,
                            characteristic=characteristic, report_interval=report_interval, max_steps=max_steps,
                            boundary=boundary)
