"""WENO5-JS + HLLC adapter with the entry points of
code/snapshot/for_paper_results/solvers/weno5_hllc.py:390-506.

``mlp_params=None`` selects WENO5-JS; FP64 MLP parameters select the inline
FP64 WENO5-SR closure, as in the reference (weno5_hllc.py:430-452).
Boundaries: "transmissive" and "periodic" (weno5_hllc.py:405-411).
"""
from __future__ import annotations

from typing import Callable

from . import _adapter
from ..euler2d import scheme_for
from ..params import Params

allocate_forward = _adapter.allocate_forward


def _check_boundary(boundary: str) -> None:
    if boundary not in ("transmissive", "periodic"):
        raise ValueError(f"unsupported boundary {boundary!r}")


def launch_stage(u_stage, u0, u_out, stage: dict, params: Params, dt: float, rk: int, device,
                 mlp_params, boundary: str, eno_cutoff: bool = False) -> None:
    _check_boundary(boundary)
    _adapter.launch_stage(u_stage, u0, u_out, stage, params, dt, rk, device, mlp_params, boundary,
                          scheme_for(mlp_params), 0.0, eno_cutoff)


def advance_step(arrays: dict, params: Params, dt: float, device, mlp_params, boundary: str) -> None:
    """weno5_hllc.py:375-386: three stage launches, then u0 <- u3."""
    _check_boundary(boundary)
    _adapter.advance_step(arrays, params, dt, device, mlp_params, boundary, scheme_for(mlp_params))


compute_dt = _adapter.compute_dt


def run_to_time(u0, params: Params, t_end: float, device, mlp_params, boundary: str,
                report_interval: int = 0, report: Callable | None = None):
    _check_boundary(boundary)
    return _adapter.run_to_time(u0, params, t_end, device, mlp_params, boundary, scheme_for(mlp_params),
                                report_interval, report)
