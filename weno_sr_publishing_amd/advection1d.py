"""1-D periodic scalar advection with the WENO5 closures on the GPU.

Entry points mirror the reference's torch rollout
(code/snapshot/teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/rk3_advection.py):
``balanced_velocities`` (:27-33), ``ssprk3_step_signed`` (:118-131),
``cfl_schedule`` (:149-168) and ``rollout_distance`` (:171-186); ``integrate_ssprk3``
and ``run_gste`` follow code/snapshot/for_paper_results/run_gste.py:103-118,179-264.
A "model" argument is ``None``/"classical" (WENO5-JS) or an ``MlpParams``.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .checkpoints import MlpParams
from .initial import GSTE_LENGTH, gste_cell_averages


def _scheme(model) -> str:
    if model is None or (isinstance(model, str) and model == "classical"):
        return "js"
    if isinstance(model, MlpParams):
        return {"fp64": "sr64", "fp32": "sr32", "bf16": "srbf16"}[model.precision]
    raise TypeError(f"unsupported model {type(model).__name__}")


class Adv1DSolver:
    """Batched (B, N) periodic rollouts; one workgroup per trajectory, all steps in one launch."""

    def __init__(self, n: int, dxinv: float, model=None, device="cuda"):
        self.n = int(n)
        self.dxinv = float(dxinv)
        self.scheme = _scheme(model)
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("the 1-D rollout runs on the GPU only")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self._mlp_host = None
        if self.scheme != "js":
            dtype = np.float64 if self.scheme == "sr64" else np.float32
            self._mlp_host = np.ascontiguousarray(model.packed.astype(dtype))
        self._lib = _lib.lib()

    def _cfg(self, batch: int) -> _lib.Adv1DConfig:
        return _lib.Adv1DConfig(batch=batch, n=self.n, scheme=_lib.SCHEMES[self.scheme], reserved=0,
                                dxinv=self.dxinv,
                                mlp=self._mlp_host.ctypes.data if self._mlp_host is not None else None)

    def rollout(self, u: torch.Tensor, velocities: torch.Tensor, dts, checkpoints: bool = False):
        """Advance ``u`` (B, N) FP64 in place; returns the (B, S+1, N) checkpoints if asked."""
        if u.dtype != torch.float64 or u.dim() != 2 or u.shape[1] != self.n or not u.is_contiguous():
            raise ValueError(f"state must be a contiguous float64 (B, {self.n}) tensor")
        if u.device != self.device:
            raise ValueError("state must be on the solver's device")
        batch = u.shape[0]
        vel = torch.as_tensor(velocities, dtype=torch.float64, device=self.device).reshape(-1)
        if vel.numel() == 1:
            vel = vel.expand(batch)
        vel = vel.contiguous()
        if vel.numel() != batch:
            raise ValueError("one velocity per trajectory")
        dts_t = torch.as_tensor(np.asarray(dts, dtype=np.float64), device=self.device).reshape(-1).contiguous()
        ck = None
        if checkpoints:
            ck = torch.empty((batch, dts_t.numel() + 1, self.n), dtype=torch.float64, device=self.device)
        cfg = self._cfg(batch)
        rc = self._lib.wsr_adv1d_rollout(ctypes.byref(cfg), u.data_ptr(), vel.data_ptr(), dts_t.data_ptr(),
                                         dts_t.numel(), ck.data_ptr() if ck is not None else None,
                                         torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check(rc, "wsr_adv1d_rollout")
        return ck


def balanced_velocities(batch: int, device) -> torch.Tensor:
    if batch < 2 or batch % 2 != 0:
        raise ValueError("bidirectional batches must be positive and even")
    v = torch.ones(batch, device=device, dtype=torch.float64)
    v[batch // 2:] = -1.0
    return v


def ssprk3_step_signed(model, state: torch.Tensor, dt: float, dxinv: float, velocities) -> torch.Tensor:
    out = state.detach().clone().contiguous()
    Adv1DSolver(state.shape[1], dxinv, model, state.device).rollout(out, velocities, [dt])
    return out


def cfl_schedule(distance_cells: float, cfl: float) -> tuple[float, ...]:
    """Full CFL steps plus one shortened step at the exact distance (rk3_advection.py:149-168)."""
    if distance_cells <= 0.0 or cfl <= 0.0:
        raise ValueError("distance_cells and cfl must be positive")
    full_steps = int(math.floor(distance_cells / cfl))
    remainder = distance_cells - full_steps * cfl
    tolerance = 2.0e-13 * max(1.0, distance_cells)
    if remainder < tolerance:
        remainder = 0.0
    elif cfl - remainder < tolerance:
        full_steps += 1
        remainder = 0.0
    schedule = [float(cfl)] * full_steps
    if remainder > 0.0:
        schedule.append(float(remainder))
    if not schedule:
        schedule.append(float(distance_cells))
    if abs(sum(schedule) - distance_cells) > 5.0e-12 * max(1.0, distance_cells):
        raise RuntimeError("CFL schedule does not reach the requested distance")
    return tuple(schedule)


def rollout_distance(model, initial: torch.Tensor, distance_cells: float, cfl: float, velocities) -> torch.Tensor:
    n = initial.shape[1]
    dx = 1.0 / float(n)
    out = initial.detach().clone().contiguous()
    dts = [c * dx for c in cfl_schedule(distance_cells, cfl)]
    Adv1DSolver(n, 1.0 / dx, model, initial.device).rollout(out, velocities, dts)
    return out


def integrate_ssprk3(model, u0: np.ndarray, cfl_limit: float, t_end: float, device="cuda",
                     length: float = GSTE_LENGTH) -> tuple[np.ndarray, dict]:
    nx = u0.size
    dx = length / nx
    steps = int(math.ceil(t_end / (cfl_limit * dx)))
    dt = t_end / steps
    u = torch.as_tensor(np.asarray(u0, dtype=np.float64), device=device).reshape(1, nx).contiguous()
    Adv1DSolver(nx, 1.0 / dx, model, device).rollout(u, [1.0], [dt] * steps)
    return u[0].cpu().numpy(), {"steps": steps, "dt": dt, "cfl": dt / dx, "t": t_end}


def run_gste(model, nx: int = 200, t_end: float = 10.0, cfl: float = 0.6, quadrature: int = 15,
             device="cuda") -> dict:
    """One GSTE row: L1/L2/Linf/TV against the exact translated cell averages."""
    _, u0 = gste_cell_averages(nx, 0.0, quadrature)
    _, exact = gste_cell_averages(nx, t_end, quadrature)
    final, meta = integrate_ssprk3(model, u0, cfl, t_end, device)
    diff = final - exact
    return {**meta, "l1": float(np.mean(np.abs(diff))), "l2": float(np.sqrt(np.mean(diff * diff))),
            "linf": float(np.max(np.abs(diff))), "tv": float(np.sum(np.abs(final - np.roll(final, 1)))),
            "min": float(np.min(final)), "max": float(np.max(final)),
            "complete": bool(np.all(np.isfinite(final))), "final": final}
