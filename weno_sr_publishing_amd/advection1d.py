"""1-D periodic scalar advection with the WENO5 closures on the GPU.

Entry points mirror the reference's torch rollout
(code/snapshot/teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/rk3_advection.py):
``balanced_velocities`` (:27-33), ``ssprk3_step_signed`` (:118-131),
``cfl_schedule`` (:149-168) and ``rollout_distance`` (:171-186); ``integrate_ssprk3``
and ``run_gste`` follow code/snapshot/for_paper_results/run_gste.py:103-118,179-264.

A "model" argument is ``None``/"classical" (WENO5-JS), an ``MlpParams``, or a torch
module carrying the closure parameters ``w1 .. b4`` (``model.load_checkpoint``, or the
reference's own closure modules).  With a module, ``ssprk3_step_signed`` and
``rollout_distance`` are differentiable: the rollout is one ``torch.autograd.Function``
whose forward is the HIP rollout kernel (every step's state checkpointed) and whose
backward is the hand-written adjoint kernel (each step recomputed from its checkpoint,
then reversed through the stages and the MLP), giving the gradients of the initial
state and of the module's parameters, as the reference's loss.backward() through its
checkpointed torch rollout does (fast_losses.py:412-648).  Classical and
``MlpParams`` models are not differentiable here: a state that requires grad raises.
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


CLOSURE_NAMES = ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4")


def _closure_module(model) -> bool:
    return isinstance(model, torch.nn.Module) and all(hasattr(model, n) for n in CLOSURE_NAMES)


class _Rollout1D(torch.autograd.Function):
    """u_T = rollout(u_0; theta) over the step schedule ``dts`` (SSPRK3, signed speed
    per trajectory); backward: the adjoint kernel with lambda_T = dL/du_T."""

    @staticmethod
    def forward(ctx, u0, flat, velocities, dts, scheme, dxinv):
        L = _lib.lib()
        batch, n = u0.shape
        u = u0.detach().to(torch.float64).contiguous().clone()
        p = flat.detach().contiguous()
        ck = torch.empty((batch, dts.numel() + 1, n), dtype=torch.float64, device=u.device)
        cfg = _lib.Adv1DConfig(batch=batch, n=n, scheme=_lib.SCHEMES[scheme], reserved=0, dxinv=float(dxinv),
                               mlp=None)
        stream = torch.cuda.current_stream(u.device).cuda_stream
        rc = L.wsr_adv1d_train_forward(ctypes.byref(cfg), u.data_ptr(), velocities.data_ptr(), dts.data_ptr(),
                                       dts.numel(), ck.data_ptr(), p.data_ptr(), stream)
        _lib.check(rc, "wsr_adv1d_train_forward")
        ctx.save_for_backward(ck, p, velocities, dts)
        ctx.cfg = cfg
        return u

    @staticmethod
    def backward(ctx, grad_u):
        ck, p, velocities, dts = ctx.saved_tensors
        cfg = ctx.cfg
        batch, n = cfg.batch, cfg.n
        lam = grad_u.to(torch.float64).contiguous()
        grad_b = torch.empty((batch, 189), dtype=torch.float64, device=lam.device)
        lam0 = torch.empty((batch, n), dtype=torch.float64, device=lam.device) if ctx.needs_input_grad[0] else None
        stream = torch.cuda.current_stream(lam.device).cuda_stream
        rc = _lib.lib().wsr_adv1d_train_adjoint(ctypes.byref(cfg), ck.data_ptr(), lam.data_ptr(),
                                                velocities.data_ptr(), dts.data_ptr(), dts.numel(), p.data_ptr(),
                                                grad_b.data_ptr(), lam0.data_ptr() if lam0 is not None else None,
                                                stream)
        _lib.check(rc, "wsr_adv1d_train_adjoint")
        grad_p = grad_b.sum(dim=0).to(p.dtype) if ctx.needs_input_grad[1] else None
        return lam0, grad_p, None, None, None, None


def _rollout(model, state: torch.Tensor, dts, dxinv: float, velocities) -> torch.Tensor:
    """The rollout for any model; differentiable for a closure module."""
    if state.dim() != 2:
        raise ValueError("state must be a (batch, n) tensor")
    if state.device.type != "cuda":
        raise ValueError("the 1-D rollout runs on the GPU: pass the state on a cuda device")
    batch, n = state.shape
    vel = torch.as_tensor(velocities, dtype=torch.float64, device=state.device).reshape(-1)
    if vel.numel() == 1:
        vel = vel.expand(batch)
    vel = vel.contiguous()
    if vel.numel() != batch:
        raise ValueError("one velocity per trajectory")
    if _closure_module(model):
        flat = torch.cat([getattr(model, name).reshape(-1) for name in CLOSURE_NAMES])
        if flat.numel() != 189:
            raise ValueError("the closure module must be the 5-10-6-6-3 MLP (189 parameters)")
        if flat.dtype not in (torch.float32, torch.float64):
            raise ValueError("the closure parameters must be float32 or float64")
        scheme = "sr32" if flat.dtype == torch.float32 else "sr64"
        dts_t = torch.as_tensor(np.asarray(dts, dtype=np.float64), device=state.device).reshape(-1).contiguous()
        return _Rollout1D.apply(state, flat.to(state.device), vel, dts_t, scheme, float(dxinv))
    if state.requires_grad or (isinstance(model, torch.nn.Module) and any(p.requires_grad for p in model.parameters())):
        raise ValueError("gradients through the 1-D rollout need a closure module with parameters w1 .. b4 "
                         "(model.load_checkpoint); classical and MlpParams models are not differentiable here")
    out = state.detach().to(torch.float64).clone().contiguous()
    Adv1DSolver(n, dxinv, model, state.device).rollout(out, vel, dts)
    return out


def balanced_velocities(batch: int, device) -> torch.Tensor:
    if batch < 2 or batch % 2 != 0:
        raise ValueError("bidirectional batches must be positive and even")
    v = torch.ones(batch, device=device, dtype=torch.float64)
    v[batch // 2:] = -1.0
    return v


def ssprk3_step_signed(model, state: torch.Tensor, dt: float, dxinv: float, velocities) -> torch.Tensor:
    """One signed SSPRK3 step (rk3_advection.py:118-131); differentiable for a closure module."""
    return _rollout(model, state, [dt], dxinv, velocities)


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
    """Exactly ``distance_cells`` cell widths with a shortened last step (rk3_advection.py:171-186),
    as one differentiable rollout for a closure module."""
    n = initial.shape[1]
    dx = 1.0 / float(n)
    dts = [c * dx for c in cfl_schedule(distance_cells, cfl)]
    return _rollout(model, initial, dts, 1.0 / dx, velocities)


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
