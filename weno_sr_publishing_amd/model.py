"""The 5-10-6-6-3 WENO5-SR closure as a torch module, and its checkpoint loader.

Mirrors the reference's callable closure models ``model(features) -> ratios``:
the FP32 model of
code/snapshot/teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/model.py:29-80
and the FP64 model of
code/snapshot/teacherfree_lab_weno5_v12_reflection_sym/v12_model.py:14-51 (over
teacherfree_lab_weno5_v9_exact_fvm/v9_model.py:16-37), with ``load_checkpoint``
(model.py:93-115, v12_model.py:61-80).

The module holds the trainable parameters ``w1 .. b4`` (``w1`` is (5, 10): the
layers compute ``h @ w + b``).  The solvers read them: ``advection1d.rollout_distance``
and ``ssprk3_step_signed`` evaluate the closure, and its adjoint, in the HIP kernels,
and the 2-D solvers take ``to_mlp_params()``.  ``forward`` evaluates the closure on a
batch of feature vectors with torch operations, for callers that use the model as a
function of features; no solver calls it.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .checkpoints import REQUIRED, MlpParams

NAMES = tuple(REQUIRED)                    # w1, b1, w2, b2, w3, b3, w4, b4
SHAPES = {"w1": (5, 10), "b1": (10,), "w2": (10, 6), "b2": (6,), "w3": (6, 6), "b3": (6,),
          "w4": (6, 3), "b4": (3,)}
SCALE_LOG10_MIN = -4.0


def remap_scale_feature(features: torch.Tensor) -> torch.Tensor:
    """v9_model.py:16-28: the scale channel arrives as (log10 r + 16) / 16 and the MLP
    sees clamp((log10 r + 4) / 4, 0, 1)."""
    log10_relative = 16.0 * features[..., 4] - 16.0
    new_scale = torch.clamp((log10_relative - SCALE_LOG10_MIN) / (-SCALE_LOG10_MIN), 0.0, 1.0)
    return torch.cat((features[..., :4], new_scale.unsqueeze(-1)), dim=-1)


class ReflectionSymmetricClosure(torch.nn.Module):
    """0.5 [M(x) + P3 M(P5 x)] with M the 5-10-6-6-3 swish MLP, head softmax(6 tanh(z / 6)).

    ``precision`` "fp32" keeps the parameters and the MLP arithmetic in float32 (the
    features are remapped in the caller's dtype, cast to float32 before the first layer,
    and the ratios cast back, as model.py:58-79); "fp64" runs it all in float64.
    The initialisation follows weno5_core.SharedBadnessMLP (seeded normal hidden
    layers, zero final layer: the linear-weight operator).
    """

    def __init__(self, precision: str = "fp32", seed: int = 41):
        super().__init__()
        if precision not in ("fp32", "fp64"):
            raise ValueError("the closure module is FP32 or FP64")
        self.precision = precision
        gen = torch.Generator(device="cpu")
        gen.manual_seed(seed)
        dtype = torch.float32 if precision == "fp32" else torch.float64
        fan_in = {"w1": 5, "w2": 10, "w3": 6}
        for name in NAMES:
            if name in fan_in:
                value = torch.randn(SHAPES[name], generator=gen, dtype=torch.float64) * math.sqrt(1.0 / fan_in[name])
            else:
                value = torch.zeros(SHAPES[name], dtype=torch.float64)
            self.register_parameter(name, torch.nn.Parameter(value.to(dtype)))

    @staticmethod
    def reflect_features(features: torch.Tensor) -> torch.Tensor:
        return features[..., [2, 1, 0, 3, 4]]

    @staticmethod
    def reflect_ratios(ratios: torch.Tensor) -> torch.Tensor:
        return ratios[..., [2, 1, 0]]

    def raw_forward(self, features: torch.Tensor) -> torch.Tensor:
        out_dtype = features.dtype
        h = remap_scale_feature(features).to(self.w1.dtype)
        for w, b in ((self.w1, self.b1), (self.w2, self.b2), (self.w3, self.b3)):
            h = h @ w + b
            h = h * torch.sigmoid(h)
        raw = h @ self.w4 + self.b4
        return torch.softmax(6.0 * torch.tanh(raw / 6.0), dim=-1).to(out_dtype)

    def forward(self, features: torch.Tensor) -> torch.Tensor:
        direct = self.raw_forward(features)
        reflected = self.raw_forward(self.reflect_features(features))
        return 0.5 * (direct + self.reflect_ratios(reflected))

    def flat_parameters(self) -> torch.Tensor:
        """The 189 parameters in the kernels' order (w1 row-major, b1, ..., b4); keeps
        the autograd link to the module's parameters."""
        return torch.cat([getattr(self, n).reshape(-1) for n in NAMES])

    def to_mlp_params(self) -> MlpParams:
        """A detached copy in the 2-D solvers' parameter record."""
        arrays = {n: getattr(self, n).detach().cpu().double().numpy()[None, ...] for n in NAMES}
        return MlpParams(arrays, self.precision)


@torch.no_grad()
def reflection_defect(model: torch.nn.Module, features: torch.Tensor) -> float:
    """max |M(P x) - P M(x)| (model.py:83-90)."""
    left = model(model.reflect_features(features))
    right = model.reflect_ratios(model(features))
    return float(torch.max(torch.abs(left - right)))


def load_checkpoint(path, device: torch.device | str = "cpu", precision: str | None = None
                    ) -> ReflectionSymmetricClosure:
    """model.py:93-115 / v12_model.py:61-80: the arrays ``w1 .. b4`` with their leading
    singleton axis, the same missing / shape errors (ValueError), read without
    unpickling.  ``precision`` defaults to the stored dtype (float32 -> "fp32")."""
    with np.load(path, allow_pickle=False) as data:
        arrays = {n: np.array(data[n]) for n in NAMES if n in data.files}
    missing = [n for n in NAMES if n not in arrays]
    if missing:
        raise ValueError(f"{path} is missing arrays: {missing}")
    wrong = {n: tuple(np.shape(arrays[n])) for n in NAMES if tuple(np.shape(arrays[n])) != (1,) + SHAPES[n]}
    if wrong:
        raise ValueError(f"{path} has incompatible WENO5 MLP shapes: {wrong}")
    if precision is None:
        precision = "fp32" if np.asarray(arrays["w1"]).dtype == np.float32 else "fp64"
    model = ReflectionSymmetricClosure(precision, seed=0)
    with torch.no_grad():
        for n in NAMES:
            getattr(model, n).copy_(torch.as_tensor(np.asarray(arrays[n])[0]))
    return model.to(device)
