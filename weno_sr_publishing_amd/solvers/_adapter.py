"""Shared plumbing of the three WENO5 adapters (allocation, stage launch, time loop)."""
from __future__ import annotations

import hashlib
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
    """State buffers u0..u3 (torch FP64 on the GPU), ``speed`` (nx * ny FP64, the
    reference's per-cell max-speed array; here ``compute_dt`` reduces on the device and
    leaves it untouched) and one workspace per stage (weno5_hllc.py:361-372)."""
    dev = _device(device)
    host = np.ascontiguousarray(u0, dtype=np.float64)
    if host.shape != params.padded_shape:
        raise ValueError(f"initial state must have shape {params.padded_shape}, got {host.shape}")
    nbytes = int(_lib.lib().wsr_euler2d_workspace_bytes(params.nx, params.ny, params.ghost))
    arrays = {"u0": torch.from_numpy(host).to(dev)}
    for name in ("u1", "u2", "u3"):
        arrays[name] = torch.zeros(params.padded_shape, dtype=torch.float64, device=dev)
    arrays["speed"] = torch.zeros(params.nx * params.ny, dtype=torch.float64, device=dev)
    for name in ("s1", "s2", "s3"):
        arrays[name] = {"workspace": torch.empty(nbytes, dtype=torch.uint8, device=dev), "_solvers": {}}
    return arrays


def _mlp_digest(mlp_params) -> str | None:
    """Identity of the closure parameters by value: a solver built for one set of
    weights is never reused for another (a freed and reallocated dict, or weights
    changed in place)."""
    if mlp_params is None:
        return None
    packed = np.ascontiguousarray(mlp_params.packed)
    return hashlib.sha1(packed.tobytes() + str(packed.dtype).encode() +
                        str(getattr(mlp_params, "precision", "")).encode()).hexdigest()


def solver_for_stage(stage: dict, params: Params, scheme: str, mlp_params, boundary: str,
                     eno_cutoff: bool, device) -> Euler2DSolver:
    if boundary not in _lib.BOUNDARY:
        raise ValueError(f"unsupported boundary {boundary!r}")
    key = (params, scheme, boundary, bool(eno_cutoff), _mlp_digest(mlp_params))
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


def compute_dt(u: torch.Tensor, params: Params, device, stage: dict | None = None) -> float:
    """warp_weno5_helpers.py:189-207 compute_dt_from_warp_array: cfl min(dx, dy) / max
    speed (1e10 when the maximum is below 1e-15), the maximum taken on the device."""
    probe = Euler2DSolver(params, "js", "hllc", "transmissive", None, False, _device(device),
                          workspace=stage.get("workspace") if stage is not None else None)
    return probe.compute_dt(u)


def advance_step(arrays: dict, params: Params, dt: float, device, mlp_params, boundary: str, scheme: str,
                 time: float = 0.0, eno_cutoff: bool = False) -> None:
    """One SSPRK3 step through the three stage launches (weno5_hllc.py:375-386, the
    double-Mach boundary times t, t + dt, t + dt/2 of weno5_hllc_refsym.py:486-526);
    the new state is arrays["u0"] afterwards."""
    times = (time, time + dt, time + 0.5 * dt)
    launch_stage(arrays["u0"], arrays["u0"], arrays["u1"], arrays["s1"], params, dt, 1, device, mlp_params,
                 boundary, scheme, times[0], eno_cutoff)
    launch_stage(arrays["u1"], arrays["u0"], arrays["u2"], arrays["s2"], params, dt, 2, device, mlp_params,
                 boundary, scheme, times[1], eno_cutoff)
    launch_stage(arrays["u2"], arrays["u0"], arrays["u3"], arrays["s3"], params, dt, 3, device, mlp_params,
                 boundary, scheme, times[2], eno_cutoff)
    arrays["u0"], arrays["u3"] = arrays["u3"], arrays["u0"]


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
