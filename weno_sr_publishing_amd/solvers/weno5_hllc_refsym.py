"""WENO5-SR (FP64 reflection-symmetric MLP) + HLLC adapter with the entry points of
code/snapshot/teacherfree_lab_weno5_v20_distance_balanced/weno5_hllc_refsym.py:298-544.

Boundaries: "transmissive", "periodic", "double-mach" (time-dependent).
``run_to_time`` checks the interior at report steps (step 1, every
``report_interval`` steps, last step) and stops on NaN or non-positive density
or pressure, like weno5_hllc_refsym.py:528-542.
"""
from __future__ import annotations

from typing import Callable

from . import _adapter
from ..euler2d import scheme_for
from ..params import Params

allocate_forward = _adapter.allocate_forward


def launch_stage(u_stage, u0, u_out, stage: dict, params: Params, dt: float, rk: int, device,
                 mlp_params, boundary: str, boundary_time: float, eno_cutoff: bool) -> None:
    _adapter.launch_stage(u_stage, u0, u_out, stage, params, dt, rk, device, mlp_params, boundary,
                          scheme_for(mlp_params), boundary_time, eno_cutoff)


def advance_step(arrays: dict, params: Params, dt: float, device, mlp_params, boundary: str,
                 time: float = 0.0, eno_cutoff: bool = False) -> None:
    """weno5_hllc_refsym.py:486-526: three stage launches with the boundary times
    t, t + dt, t + dt/2, then u0 <- u3."""
    _adapter.advance_step(arrays, params, dt, device, mlp_params, boundary, scheme_for(mlp_params), time,
                          eno_cutoff)


compute_dt = _adapter.compute_dt


def run_to_time(initial, params: Params, t_end: float, device, mlp_params, boundary: str,
                eno_cutoff: bool = False, report_interval: int = 100,
                report: Callable | None = None):
    return _adapter.run_to_time(initial, params, t_end, device, mlp_params, boundary,
                                scheme_for(mlp_params), report_interval, report,
                                check_health=True, eno_cutoff=eno_cutoff)
