"""WENO7-SR entry points of
code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/point_rk4_mlp.py
(``TorchWeno7PointBeta`` :22-56, ``allocate_arrays_mlp`` :188-191, ``compute_rhs_mlp``
:194-273, ``launch_shu_rk4_step_mlp`` :275-301, ``run_from_initial_mlp`` :303-362) on the
HIP WENO7 step.

``TorchWeno7PointBeta`` is the reference's beta provider: its four fill methods write
the smoothness-indicator arrays of the 6-24-16-16-4 closure (HIP kernels, FP64 as the
reference, or FP32).  ``run_from_initial_mlp`` runs this provider on the fused HIP step,
where the closure is evaluated inside the sweep and no beta array exists; any other
provider (an object with the four methods that writes ``arrays``' torch tensors) runs
through the reference's unfused protocol (``compute_rhs_mlp`` /
``launch_shu_rk4_step_mlp`` on the HIP consumer kernels), as does this one with
``fused=False``.
"""
from __future__ import annotations

import numpy as np

from .. import beta7
from ..beta7 import allocate_arrays_mlp
from ..checkpoints import Mlp7Params, WENO7_SR_FP64, load_weno7_checkpoint
from ..euler7 import Euler7Solver, Params7 as Params
from .point_rk4 import _printer, summary_for

__all__ = ["Params", "TorchWeno7PointBeta", "allocate_arrays_mlp", "compute_rhs_mlp", "launch_shu_rk4_step_mlp",
           "run_from_initial_mlp"]


class TorchWeno7PointBeta(beta7.HipWeno7Beta):
    """The reference's provider (point_rk4_mlp.py:22-56 TorchWeno7PointBeta(model_path,
    device, gamma)) from a WENO7-SR checkpoint; ``precision`` "fp64" (the reference's
    closure) or "fp32"."""

    def __init__(self, model_path=WENO7_SR_FP64, device: str = "cuda", gamma: float = 1.4,
                 precision: str = "fp64"):
        self.model_path = model_path
        params = model_path if isinstance(model_path, Mlp7Params) else load_weno7_checkpoint(model_path)
        super().__init__(params if precision == params.precision else params.as_precision(precision), device, gamma)


def compute_rhs_mlp(arrays: dict, params: Params, src_name: str, rhs_name: str, dt: float, device=None, *,
                    solver_kind, beta_model, reverse_upwind: bool, boundary: str = "periodic",
                    eno_cutoff: bool = False) -> None:
    """point_rk4_mlp.py:194-273: solver_kind 1 is HLLC, any other EVILIN (a name, "hllc",
    "evilin" or "force", is accepted as well)."""
    del device
    beta7.compute_rhs(arrays, params, src_name, rhs_name, dt, beta_model=beta_model, reverse_upwind=reverse_upwind,
                      riemann=beta7.riemann_name(solver_kind), boundary=boundary, eno_cutoff=eno_cutoff)


def launch_shu_rk4_step_mlp(arrays: dict, params: Params, dt: float, device=None, *, riemann_solver: str,
                            beta_model, boundary: str = "periodic", eno_cutoff: bool = False) -> None:
    """point_rk4_mlp.py:275-301 (solver kind 1 for "hllc", EVILIN otherwise)."""
    del device
    riemann = riemann_solver if riemann_solver in ("hllc", "evilin", "force") else "evilin"
    beta7.shu_rk4_step(arrays, params, dt, beta_model=beta_model, riemann=riemann, boundary=boundary,
                       eno_cutoff=eno_cutoff)


def run_from_initial_mlp(u0, params: Params, *, device: str, riemann_solver: str, beta_model,
                         report_interval: int = 0, max_steps: int = 10_000_000, boundary: str = "periodic",
                         eno_cutoff: bool = False, fused: bool | None = None) -> tuple[np.ndarray, dict]:
    """WENO7-SR + Shu RK4 from a padded initial state: (final state, summary).  ``fused``
    None picks the fused step for this library's provider and the protocol otherwise."""
    if riemann_solver not in ("hllc", "force", "evilin"):
        raise ValueError(f"unknown riemann_solver {riemann_solver!r}")
    use_fused = beta7.is_fusable(beta_model) if fused is None else bool(fused)
    if use_fused and not beta7.is_fusable(beta_model):
        raise ValueError("the fused step evaluates this library's closure: pass a TorchWeno7PointBeta")
    report = _printer("MLP-RK4 ", params) if report_interval > 0 else None
    if not use_fused:
        final, dts, steps, t, stopped = beta7.run_point(
            u0, params, device=device, riemann=riemann_solver, beta_model=beta_model,
            report_interval=report_interval, max_steps=max_steps, boundary=boundary, eno_cutoff=eno_cutoff,
            report=report)
        if stopped:
            print("MLP-RK4 failure: NaN/negative rho/p detected, stopping early", flush=True)
        return final, summary_for(final, params, dts, steps, t)
    solver = Euler7Solver(params, beta_model.scheme, riemann_solver, boundary, beta_model.mlp,
                          eno_cutoff=eno_cutoff, loop="point", device=device)
    u = solver.new_state(np.asarray(u0, dtype=np.float64))
    dts, steps, t = solver.run(u, params.t_end, report_interval=report_interval, report=report,
                               check_health=report_interval > 0, max_steps=max_steps)
    final = u.cpu().numpy()
    return final, summary_for(final, params, dts, steps, t)
