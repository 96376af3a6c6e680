"""Differentiable closure training on solver-generated 1-D trajectories.

The forward rollout and its adjoint are hand-written HIP kernels
(``wsr_adv1d_train_forward`` / ``wsr_adv1d_train_adjoint``); torch only holds the
parameters and runs Adam.  The reference does the same training with torch
autograd through checkpointed chunks of the rollout
(code/snapshot/teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/fast_losses.py:412-648,
train.py:449-575).

Loss of one batch: mean over trajectories and cells of (u_T - target)^2 where u_T
is the state after the whole step schedule.
"""
from __future__ import annotations

import ctypes
import json
import math

import numpy as np
import torch

from . import _lib
from .advection1d import cfl_schedule
from .checkpoints import MlpParams, REQUIRED, save_mlp_params

SHAPES = {"w1": (5, 10), "b1": (10,), "w2": (10, 6), "b2": (6,), "w3": (6, 6), "b3": (6,),
          "w4": (6, 3), "b4": (3,)}


def init_params(seed: int = 41) -> dict:
    """Initial MLP: random hidden layers, zero final layer (the linear-d operator),
    as weno5_core.SharedBadnessMLP initialises (weno5_core.py:29-43)."""
    gen = torch.Generator(device="cpu")
    gen.manual_seed(seed)
    p = {}
    p["w1"] = (torch.randn((5, 10), generator=gen, dtype=torch.float64) * math.sqrt(1.0 / 5)).numpy()
    p["b1"] = np.zeros(10)
    p["w2"] = (torch.randn((10, 6), generator=gen, dtype=torch.float64) * math.sqrt(1.0 / 10)).numpy()
    p["b2"] = np.zeros(6)
    p["w3"] = (torch.randn((6, 6), generator=gen, dtype=torch.float64) * math.sqrt(1.0 / 6)).numpy()
    p["b3"] = np.zeros(6)
    p["w4"] = np.zeros((6, 3))
    p["b4"] = np.zeros(3)
    return p


def flatten(params: dict) -> np.ndarray:
    return np.concatenate([np.asarray(params[n]).reshape(-1) for n in REQUIRED])


def unflatten(flat) -> dict:
    flat = np.asarray(flat)
    out, o = {}, 0
    for n in REQUIRED:
        size = int(np.prod(SHAPES[n]))
        out[n] = flat[o:o + size].reshape(SHAPES[n])
        o += size
    return out


class ClosureTrainer:
    """Adam on the 189 closure parameters with gradients from the hand-written adjoint."""

    def __init__(self, params: dict | MlpParams | None = None, precision: str = "fp32", lr: float = 1e-3,
                 device="cuda", seed: int = 41):
        if precision not in ("fp32", "fp64"):
            raise ValueError("training runs with the FP32 or FP64 closure")
        self.precision = precision
        self.device = torch.device(device)
        self.dtype = torch.float32 if precision == "fp32" else torch.float64
        base = params if params is not None else init_params(seed)
        flat = flatten({k: np.asarray(v).reshape(SHAPES[k]) for k, v in base.items() if k in SHAPES})
        self.param = torch.tensor(flat, dtype=self.dtype, device=self.device)
        self.param.grad = torch.zeros_like(self.param)
        self.opt = torch.optim.Adam([self.param], lr=lr)
        self._lib = _lib.lib()
        self.scheme = _lib.SCHEMES["sr32" if precision == "fp32" else "sr64"]

    def _cfg(self, batch: int, n: int, dxinv: float) -> _lib.Adv1DConfig:
        return _lib.Adv1DConfig(batch=batch, n=n, scheme=self.scheme, reserved=0, dxinv=dxinv, mlp=None)

    def loss_and_grad(self, u0: torch.Tensor, velocities: torch.Tensor, target: torch.Tensor, dts,
                      return_state_grad: bool = False, events: dict | None = None):
        """Forward rollout, loss and parameter gradient (summed over the batch).
        ``events``, if given, receives CUDA event pairs recorded on the launch
        stream around the forward and adjoint kernels ("forward", "adjoint")."""
        batch, n = u0.shape
        dxinv = float(n)
        u = u0.detach().to(self.device, torch.float64).contiguous().clone()
        vel = torch.as_tensor(velocities, dtype=torch.float64, device=self.device).reshape(-1).contiguous()
        dts_t = torch.as_tensor(np.asarray(dts, dtype=np.float64), device=self.device).contiguous()
        steps = dts_t.numel()
        ck = torch.empty((batch, steps + 1, n), dtype=torch.float64, device=self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        cfg = self._cfg(batch, n, dxinv)
        p = self.param.detach().contiguous()

        def mark(name):
            if events is not None:
                ev = torch.cuda.Event(enable_timing=True)
                ev.record(torch.cuda.current_stream(self.device))
                events.setdefault(name, []).append(ev)

        mark("forward")
        rc = self._lib.wsr_adv1d_train_forward(ctypes.byref(cfg), u.data_ptr(), vel.data_ptr(), dts_t.data_ptr(),
                                               steps, ck.data_ptr(), p.data_ptr(), stream)
        mark("forward")
        _lib.check(rc, "wsr_adv1d_train_forward")
        tgt = target.to(self.device, torch.float64)
        diff = u - tgt
        loss = torch.mean(diff * diff)
        lam = (2.0 / diff.numel()) * diff
        grad_b = torch.empty((batch, 189), dtype=torch.float64, device=self.device)
        lam0 = torch.empty((batch, n), dtype=torch.float64, device=self.device) if return_state_grad else None
        lam = lam.contiguous()
        mark("adjoint")
        rc = self._lib.wsr_adv1d_train_adjoint(ctypes.byref(cfg), ck.data_ptr(), lam.data_ptr(),
                                               vel.data_ptr(), dts_t.data_ptr(), steps, p.data_ptr(),
                                               grad_b.data_ptr(), lam0.data_ptr() if lam0 is not None else None,
                                               stream)
        mark("adjoint")
        _lib.check(rc, "wsr_adv1d_train_adjoint")
        grad = grad_b.sum(dim=0)
        if return_state_grad:
            return loss, grad, u, lam0
        return loss, grad, u

    def step(self, u0, velocities, target, dts) -> float:
        loss, grad, _ = self.loss_and_grad(u0, velocities, target, dts)
        self.param.grad.copy_(grad.to(self.dtype))
        self.opt.step()
        return float(loss)

    def params(self) -> dict:
        return unflatten(self.param.detach().cpu().numpy())

    def mlp(self) -> MlpParams:
        arrays = {k: v[None, ...] if v.ndim == 2 else v[None, :] for k, v in self.params().items()}
        return MlpParams(arrays, self.precision)

    def save(self, path, meta: dict | None = None) -> None:
        full = {"mlp_architecture": "shared_direct_beta_ratio_5_10_6_6_3_power2_normdelta_scale",
                "training": "hand-written adjoint of the SSPRK3 rollout (wsr_adv1d_train_adjoint)"}
        full.update(meta or {})
        save_mlp_params(path, {k: np.asarray(v, dtype=np.float32 if self.precision == "fp32" else np.float64)
                               for k, v in self.params().items()}, full)


def distance_schedule(n: int, distance_cells: float, cfl: float) -> list[float]:
    """dt per step for a rollout of ``distance_cells`` cells on a unit periodic grid."""
    dx = 1.0 / float(n)
    return [c * dx for c in cfl_schedule(distance_cells, cfl)]
