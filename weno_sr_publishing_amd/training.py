"""Differentiable closure training on solver-generated 1-D trajectories.

The forward rollout and its adjoint are hand-written HIP kernels
(``wsr_adv1d_train_forward`` / ``wsr_adv1d_train_adjoint``); torch only holds the
parameters and runs Adam.  The reference does the same training with torch
autograd through checkpointed chunks of the rollout
(code/snapshot/teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/fast_losses.py:412-648,
train.py:449-575).

Loss of one batch: mean over trajectories and cells of (u_T - target)^2 where u_T
is the state after the whole step schedule.  Other losses go through the
differentiable drop-ins (``advection1d.rollout_distance`` with a closure module and
``loss.backward()``), which use the same kernels.

Resume state (train.py:47-58, 202-229, 290-318 of the reference): ``save_state``
writes the parameters, the Adam moments and step count, the training step, the batch
generator's state and a recipe signature to one ``.npz`` (arrays plus a JSON record;
no pickled objects) through a temporary file and ``os.replace``; ``resume`` refuses a
file whose recipe or signature differs from the caller's.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import tempfile
from pathlib import Path

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

    # ------------------------------------------------------------ resume state
    def save_state(self, path, step: int, signature: dict, generator: torch.Generator | None = None) -> None:
        """Atomic resume file: parameters, Adam state, step, generator state, signature."""
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        st = self.opt.state.get(self.param, {})
        group = self.opt.param_groups[0]
        meta = {"format_version": 1, "recipe": RECIPE, "step": int(step), "precision": self.precision,
                "training_signature": signature, "adam": {"lr": group["lr"], "betas": list(group["betas"]),
                                                          "eps": group["eps"], "weight_decay": group["weight_decay"]},
                "adam_step": float(st["step"]) if "step" in st else 0.0}
        arrays = {"param": self.param.detach().cpu().numpy(),
                  "meta_json": np.array(json.dumps(meta, sort_keys=True), dtype=np.str_)}
        if "exp_avg" in st:
            arrays["exp_avg"] = st["exp_avg"].detach().cpu().numpy()
            arrays["exp_avg_sq"] = st["exp_avg_sq"].detach().cpu().numpy()
        if generator is not None:
            arrays["generator_state"] = generator.get_state().numpy()
        handle, tmp = tempfile.mkstemp(prefix=path.name + ".", suffix=".tmp", dir=path.parent)
        os.close(handle)
        try:
            with open(tmp, "wb") as f:
                np.savez(f, **arrays)
            os.replace(tmp, path)
        finally:
            if os.path.exists(tmp):
                os.unlink(tmp)

    def resume(self, path, signature: dict, generator: torch.Generator | None = None) -> int:
        """Restore a save_state file; returns the step it was saved at."""
        with np.load(path, allow_pickle=False) as data:
            meta = json.loads(str(data["meta_json"]))
            if meta.get("recipe") != RECIPE:
                raise ValueError("incompatible resume recipe")
            if meta.get("training_signature") != json.loads(json.dumps(signature, sort_keys=True)):
                raise ValueError("resume training signature mismatch")
            if meta.get("precision") != self.precision:
                raise ValueError("resume precision mismatch")
            param = np.array(data["param"])
            moments = (np.array(data["exp_avg"]), np.array(data["exp_avg_sq"])) if "exp_avg" in data.files else None
            gen_state = np.array(data["generator_state"]) if "generator_state" in data.files else None
        with torch.no_grad():
            self.param.copy_(torch.as_tensor(param, device=self.device))
        group = self.opt.param_groups[0]
        group["lr"] = meta["adam"]["lr"]
        group["betas"] = tuple(meta["adam"]["betas"])
        group["eps"] = meta["adam"]["eps"]
        group["weight_decay"] = meta["adam"]["weight_decay"]
        self.opt.state.clear()
        if moments is not None:
            self.opt.state[self.param] = {
                "step": torch.tensor(meta["adam_step"], dtype=torch.float32),
                "exp_avg": torch.as_tensor(moments[0], device=self.device),
                "exp_avg_sq": torch.as_tensor(moments[1], device=self.device)}
        if generator is not None:
            if gen_state is None:
                raise ValueError("the resume file holds no generator state")
            generator.set_state(torch.as_tensor(gen_state, dtype=torch.uint8))
        return int(meta["step"])

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


RECIPE = "weno5_sr_closure_distance_rollout_mse_hand_written_adjoint"


def training_signature(n: int, batch: int, distance_cells: float, cfl: float, precision: str, lr: float,
                       seed: int) -> dict:
    """The recipe values a resumed run must share with the run that saved the state."""
    return {"n": int(n), "batch": int(batch), "distance_cells": float(distance_cells), "cfl": float(cfl),
            "precision": precision, "lr": float(lr), "seed": int(seed)}


def train(trainer: ClosureTrainer, n: int, batch: int, distance_cells: float, cfl: float, steps: int,
          seed: int = 83, state_path=None, save_every: int = 0, resume_from=None) -> list[float]:
    """Training loop on solver-generated batches: each step draws a batch seed from a
    torch generator seeded with ``seed`` (random profiles, a = +1 / -1 balanced,
    target = the exact shifted averages after ``distance_cells`` cells), then one Adam
    step.  With ``state_path`` the resume state is written every ``save_every`` steps
    and at the end; ``resume_from`` continues a saved run (same signature required).
    Returns the losses of the steps run here."""
    from .profiles import make_batch, shifted_target
    if float(distance_cells) != int(distance_cells):
        raise ValueError("the exact shifted target needs a whole number of cells")
    sig = training_signature(n, batch, distance_cells, cfl, trainer.precision,
                             trainer.opt.param_groups[0]["lr"], seed)
    gen = torch.Generator(device="cpu")
    gen.manual_seed(seed)
    start = 0
    if resume_from is not None:
        start = trainer.resume(resume_from, sig, gen)
    dts = distance_schedule(n, distance_cells, cfl)
    losses = []
    for step in range(start + 1, steps + 1):
        batch_seed = int(torch.randint(0, 2 ** 31 - 1, (1,), generator=gen))
        u0, vel = make_batch(batch, n, batch_seed)
        target = shifted_target(u0, vel, int(distance_cells))
        losses.append(trainer.step(torch.tensor(u0), torch.tensor(vel), torch.tensor(target), dts))
        if state_path is not None and ((save_every and step % save_every == 0) or step == steps):
            trainer.save_state(state_path, step, sig, gen)
    return losses


def distance_schedule(n: int, distance_cells: float, cfl: float) -> list[float]:
    """dt per step for a rollout of ``distance_cells`` cells on a unit periodic grid."""
    dx = 1.0 / float(n)
    return [c * dx for c in cfl_schedule(distance_cells, cfl)]
