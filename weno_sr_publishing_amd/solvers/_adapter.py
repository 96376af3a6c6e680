"""Shared plumbing of the three WENO5 adapters (allocation, stage launch, time loop)."""
from __future__ import annotations

from typing import Callable

import numpy as np
import torch

from .. import _lib
from ..euler2d import Euler2DSolver
from ..params import Params


def _device(device) -> torch.device:
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError(f"the WENO5 step runs on the GPU only (device {device!r})")
    return dev


def allocate_forward(u0: np.ndarray, params: Params, device) -> dict:
    """State buffers u0..u3 (torch FP64 on the GPU) and one workspace per stage."""
    dev = _device(device)
    host = np.ascontiguousarray(u0, dtype=np.float64)
    if host.shape != params.padded_shape:
        raise ValueError(f"initial state must have shape {params.padded_shape}, got {host.shape}")
    nbytes = int(_lib.lib().wsr_euler2d_workspace_bytes(params.nx, params.ny, params.ghost))
    arrays = {"u0": torch.from_numpy(host).to(dev)}
    for name in ("u1", "u2", "u3"):
        arrays[name] = torch.zeros(params.padded_shape, dtype=torch.float64, device=dev)
    for name in ("s1", "s2", "s3"):
        arrays[name] = {"workspace": torch.empty(nbytes, dtype=torch.uint8, device=dev), "_solvers": {}}
    return arrays


def solver_for_stage(stage: dict, params: Params, scheme: str, mlp_params, boundary: str,
                     eno_cutoff: bool, device) -> Euler2DSolver:
    if boundary not in _lib.BOUNDARY:
        raise ValueError(f"unsupported boundary {boundary!r}")
    key = (params, scheme, boundary, bool(eno_cutoff), id(mlp_params))
    cache = stage.setdefault("_solvers", {})
    solver = cache.get(key)
    if solver is None:
        solver = Euler2DSolver(params, scheme, "hllc", boundary, mlp_params, eno_cutoff, _device(device),
                               workspace=stage.get("workspace"))
        cache[key] = solver
        stage["workspace"] = solver.workspace
    return solver


def launch_stage(u_stage, u0, u_out, stage: dict, params: Params, dt: float, rk: int, device,
                 mlp_params, boundary: str, scheme: str, boundary_time: float = 0.0,
                 eno_cutoff: bool = False) -> None:
    solver = solver_for_stage(stage, params, scheme, mlp_params, boundary, eno_cutoff, device)
    solver.stage(u_stage, u0, u_out, dt, rk, boundary_time)


def run_to_time(u0: np.ndarray, params: Params, t_end: float, device, mlp_params, boundary: str,
                scheme: str, report_interval: int = 0, report: Callable | None = None,
                check_health: bool = False, eno_cutoff: bool = False):
    if boundary not in _lib.BOUNDARY:
        raise ValueError(f"unsupported boundary {boundary!r}")
    solver = Euler2DSolver(params, scheme, "hllc", boundary, mlp_params, eno_cutoff, _device(device))
    u = solver.new_state(u0)
    dts, steps, t = solver.run(u, t_end, report_interval=report_interval, report=report,
                               check_health=check_health)
    return u.cpu().numpy(), dts, steps, t
