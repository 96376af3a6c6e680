"""WENO7-SR entry points of
code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/point_rk4_mlp.py
(``TorchWeno7PointBeta``, ``run_from_initial_mlp`` :303-362) on the HIP WENO7 step.

The reference's beta model is a torch module that fills the smoothness
indicators of every stencil from the 6-24-16-16-4 closure; here it is the
closure's parameters, evaluated inside the HIP sweep kernels.
"""
from __future__ import annotations

import numpy as np

from ..checkpoints import Mlp7Params, WENO7_SR_FP64, load_weno7_checkpoint
from ..euler7 import Euler7Solver, Params7 as Params
from .point_rk4 import _printer, summary_for


class TorchWeno7PointBeta:
    """Holder of a WENO7-SR checkpoint (point_rk4_mlp.py TorchWeno7PointBeta(model_path,
    device, gamma)): ``precision`` "fp64" (the reference's closure) or "fp32"."""

    def __init__(self, model_path=WENO7_SR_FP64, device: str = "cuda", gamma: float = 1.4,
                 precision: str = "fp64"):
        self.model_path, self.device, self.gamma = model_path, device, gamma
        params = model_path if isinstance(model_path, Mlp7Params) else load_weno7_checkpoint(model_path)
        self.mlp = params if precision == params.precision else params.as_precision(precision)

    @property
    def scheme(self) -> str:
        return "sr64" if self.mlp.precision == "fp64" else "sr32"


def run_from_initial_mlp(u0, params: Params, *, device: str, riemann_solver: str,
                         beta_model: TorchWeno7PointBeta, report_interval: int = 0, max_steps: int = 10_000_000,
                         boundary: str = "periodic", eno_cutoff: bool = False) -> tuple[np.ndarray, dict]:
    if riemann_solver not in ("hllc", "force", "evilin"):
        raise ValueError(f"unknown riemann_solver {riemann_solver!r}")
    solver = Euler7Solver(params, beta_model.scheme, riemann_solver, boundary, beta_model.mlp,
                          eno_cutoff=eno_cutoff, loop="point", device=device)
    u = solver.new_state(np.asarray(u0, dtype=np.float64))
    report = _printer("MLP-RK4 ", params) if report_interval > 0 else None
    dts, steps, t = solver.run(u, params.t_end, report_interval=report_interval, report=report,
                               check_health=report_interval > 0, max_steps=max_steps)
    final = u.cpu().numpy()
    return final, summary_for(final, params, dts, steps, t)
