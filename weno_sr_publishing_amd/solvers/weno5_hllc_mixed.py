"""WENO5-SR (FP32 MLP, FP64 state) + HLLC adapter with the entry points of
code/snapshot/for_paper_results/solvers/weno5_hllc_mixed.py:121-212.

Boundaries: "transmissive", "periodic", "double-mach".  MLP parameters loaded
with ``precision="bf16"`` select the BF16 closure instead.
"""
from __future__ import annotations

from typing import Callable

from . import _adapter
from ..euler2d import scheme_for
from ..params import Params

allocate_forward = _adapter.allocate_forward


def _scheme(mlp_params) -> str:
    scheme = scheme_for(mlp_params)
    return "sr32" if scheme == "sr64" else scheme


def launch_stage(u_stage, u0, u_out, stage: dict, params: Params, dt: float, rk: int, device,
                 mlp_params, boundary: str, boundary_time: float = 0.0) -> None:
    _adapter.launch_stage(u_stage, u0, u_out, stage, params, dt, rk, device, mlp_params, boundary,
                          _scheme(mlp_params), boundary_time, False)


def advance_step(arrays: dict, params: Params, dt: float, device, mlp_params, boundary: str,
                 time: float = 0.0) -> None:
    """Three stage launches (boundary times t, t + dt, t + dt/2), then u0 <- u3."""
    _adapter.advance_step(arrays, params, dt, device, mlp_params, boundary, _scheme(mlp_params), time)


compute_dt = _adapter.compute_dt


def run_to_time(u0, params: Params, t_end: float, device, mlp_params, boundary: str,
                report_interval: int = 0, report: Callable | None = None):
    return _adapter.run_to_time(u0, params, t_end, device, mlp_params, boundary, _scheme(mlp_params),
                                report_interval, report)
