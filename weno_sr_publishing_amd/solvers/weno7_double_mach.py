"""WENO7 double-Mach adapter of code/snapshot/for_paper_results/solvers/weno7_double_mach.py
(``run_to_time`` :456-501): WENO7-JS (beta_model None) or WENO7-SR, HLLC,
characteristic reconstruction, Shu RK4 with the downwind operator, the
time-dependent double-Mach boundary, on the HIP WENO7 step.

Loop semantics as the reference: ``while t < t_end - 1e-14``, dt = min(CFL dt,
t_end - t), stop on a non-finite or non-positive dt; stats at step 1, every
``report_interval`` steps and at the end; a NaN or a non-positive density or
pressure at a report step stops the run.
"""
from __future__ import annotations

from typing import Callable

import numpy as np

from .. import beta7
from ..euler7 import Euler7Solver, Params7 as Params, double_mach_params7
from .point_rk4_mlp import TorchWeno7PointBeta

__all__ = ["Params", "double_mach_params7", "run_to_time"]


def run_to_time(initial: np.ndarray, params: Params, t_end: float, device: str,
                beta_model: TorchWeno7PointBeta | None, *, report_interval: int = 100,
                report: Callable[[int, float, float, dict], None] | None = None,
                riemann_solver: str = "hllc") -> tuple[np.ndarray, list[float], int, float]:
    """WENO7-JS (beta_model None) or WENO7-SR; this library's provider runs on the fused
    step, any other beta provider through the reference's protocol (weno7_double_mach.py:
    275-453 compute_rhs_mlp / _launch_shu_rk4_step on the HIP consumer kernels)."""
    if beta_model is not None and not beta7.is_fusable(beta_model):
        return beta7.run_double_mach(initial, params, t_end, device, beta_model, report_interval=report_interval,
                                     report=report, riemann=riemann_solver)
    scheme = "js" if beta_model is None else beta_model.scheme
    mlp = None if beta_model is None else beta_model.mlp
    solver = Euler7Solver(params, scheme, riemann_solver, "double-mach", mlp, loop="double-mach", device=device)
    u = solver.new_state(np.asarray(initial, dtype=np.float64))
    dts, steps, t = solver.run(u, t_end, report_interval=report_interval, report=report, check_health=True)
    return u.cpu().numpy(), dts, steps, t
