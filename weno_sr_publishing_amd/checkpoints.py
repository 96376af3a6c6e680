"""Closure checkpoints: the reference's ``.npz`` layout.

Layout (models/MODEL_CARD.md; writers code/snapshot/teacherfree_lab_weno5/weno5_core.py:75-104):
``w1 (1,5,10) b1 (1,10) w2 (1,10,6) b2 (1,6) w3 (1,6,6) b3 (1,6) w4 (1,6,3) b4 (1,3)``
plus a 0-d unicode ``meta_json``.  The loader mirrors
code/snapshot/run_weno5_circle_mlp_compare.py:49-81 (FP64) and
code/snapshot/teacherfree_lab_weno5_mlp_f32/warp_mlp_f32/run_weno5_circle_mlp_compare_mlp_f32.py:52-84
(FP32): same required names, same shape check and error text, same metadata
warning.  Files are read with ``allow_pickle=False``.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

REQUIRED = ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4")
EXPECTED_SHAPES = {
    "w1": (1, 5, 10), "b1": (1, 10), "w2": (1, 10, 6), "b2": (1, 6),
    "w3": (1, 6, 6), "b3": (1, 6), "w4": (1, 6, 3), "b4": (1, 3),
}
MLP5_COUNT = 189
CHECKPOINT_DIR = Path(__file__).resolve().parent / "checkpoints"
WENO5_SR_FP64 = CHECKPOINT_DIR / "weno5_sr_fp64_step012250.npz"
WENO5_SR_FP32 = CHECKPOINT_DIR / "weno5_sr_fp32_step016500.npz"
WENO7_SR_FP64 = CHECKPOINT_DIR / "weno7_sr_fp64_step016750.npz"


class MlpParams(dict):
    """``{"w1": ..., "b4": ...}`` plus the packed parameter block the kernels take.

    ``packed`` is a contiguous host array of the 189 parameters in the order
    w1 (row-major k,o), b1, w2, b2, w3, b3, w4, b4 and in the closure dtype, built
    from the current arrays each time it is read (weights changed in place reach the
    next solver that packs them).  ``precision`` is "fp64", "fp32" or "bf16".
    """

    def __init__(self, arrays: dict, precision: str, meta: dict | None = None):
        super().__init__(arrays)
        if precision not in ("fp64", "fp32", "bf16"):
            raise ValueError(f"unknown closure precision {precision!r}")
        self.precision = precision
        self.meta = meta or {}
        if self.packed.size != MLP5_COUNT:
            raise ValueError(f"WENO5 MLP must have {MLP5_COUNT} parameters, got {self.packed.size}")

    @property
    def packed(self) -> np.ndarray:
        dtype = np.float64 if self.precision == "fp64" else np.float32
        flat = np.concatenate([np.asarray(self[n]).reshape(-1) for n in REQUIRED]).astype(dtype)
        return np.ascontiguousarray(flat)

    def as_precision(self, precision: str) -> "MlpParams":
        """Same parameters evaluated at another closure precision (fp64 -> fp32 casts etc.)."""
        return MlpParams({k: np.asarray(v) for k, v in self.items()}, precision, self.meta)


def _read(model_path) -> tuple[dict, dict]:
    data = np.load(model_path, allow_pickle=False)
    missing = [name for name in REQUIRED if name not in data.files]
    if missing:
        raise ValueError(f"Model {model_path} is missing arrays: {missing}")
    wrong = {name: data[name].shape for name, shape in EXPECTED_SHAPES.items() if data[name].shape != shape}
    if wrong:
        raise ValueError(
            f"Model {model_path} uses old/incompatible MLP shapes: {wrong}. "
            "Retrain with the shared direct beta-ratio 5->10->6->6->3 WENO5-NN architecture."
        )
    meta: dict = {}
    if "meta_json" in data.files:
        text = str(data["meta_json"])
        if "shared_direct_beta_ratio_5_10_6_6_3" not in text:
            print(
                "warning: model metadata does not say it was trained with the "
                "shared direct beta-ratio 5->10->6->6->3 WENO5-NN architecture; retrain before trusting this comparison."
            )
        try:
            meta = json.loads(text)
        except json.JSONDecodeError:
            meta = {"raw": text}
    arrays = {name: np.array(data[name]) for name in REQUIRED}
    return arrays, meta


def load_mlp_params(model_path, device: str = "cuda", precision: str | None = None) -> MlpParams:
    """Load a WENO5-SR checkpoint.

    ``precision`` defaults to the checkpoint's own dtype (FP64 checkpoints give
    the FP64 closure, FP32 checkpoints the FP32 closure); pass "bf16" for the
    BF16 closure.  ``device`` is accepted for signature compatibility: the
    parameters stay on the host until a solver call stages them into its
    device workspace.
    """
    del device
    arrays, meta = _read(model_path)
    if precision is None:
        precision = "fp64" if arrays["w1"].dtype == np.float64 else "fp32"
    return MlpParams(arrays, precision, meta)


def load_f64(model_path=WENO5_SR_FP64, device: str = "cuda") -> MlpParams:
    return load_mlp_params(model_path, device, "fp64")


def load_f32(model_path=WENO5_SR_FP32, device: str = "cuda") -> MlpParams:
    return load_mlp_params(model_path, device, "fp32")


def save_mlp_params(path, params: dict, meta: dict | None = None) -> None:
    """Write the reference layout (leading singleton axis + ``meta_json``)."""
    payload = {}
    for name in REQUIRED:
        arr = np.asarray(params[name])
        if arr.shape != EXPECTED_SHAPES[name]:
            arr = arr.reshape(EXPECTED_SHAPES[name])
        payload[name] = arr
    payload["meta_json"] = np.array(json.dumps(meta or {}, sort_keys=True))
    np.savez(path, **payload)


# ------------------------------------------------------------------ WENO7 closure
# code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/weno7_core.py:324-423:
# w1 (1,6,h1) b1 (1,h1) w2 (1,h1,h2) b2 (1,h2) w3 (1,h2,h3) b3 (1,h3) w4 (1,h3,4) b4 (1,4)
W7_INPUTS, W7_OUTPUTS, W7_HIDDEN = 6, 4, (24, 16, 16)
MLP7_COUNT = 908


def weno7_checkpoint_shapes(hidden=W7_HIDDEN) -> dict:
    h1, h2, h3 = hidden
    return {"w1": (1, W7_INPUTS, h1), "b1": (1, h1), "w2": (1, h1, h2), "b2": (1, h2),
            "w3": (1, h2, h3), "b3": (1, h3), "w4": (1, h3, W7_OUTPUTS), "b4": (1, W7_OUTPUTS)}


class Mlp7Params(dict):
    """WENO7 closure parameters plus the packed 908-parameter block (order w1
    (row-major), b1, ..., b4) in the closure dtype ("fp64" or "fp32")."""

    def __init__(self, arrays: dict, precision: str, meta: dict | None = None):
        super().__init__(arrays)
        if precision not in ("fp64", "fp32"):
            raise ValueError(f"unknown WENO7 closure precision {precision!r}")
        self.precision = precision
        self.meta = meta or {}
        dtype = np.float64 if precision == "fp64" else np.float32
        flat = np.concatenate([np.asarray(arrays[n]).reshape(-1) for n in REQUIRED]).astype(dtype)
        if flat.size != MLP7_COUNT:
            raise ValueError(f"WENO7 MLP must have {MLP7_COUNT} parameters (6-24-16-16-4), got {flat.size}")
        self.packed = np.ascontiguousarray(flat)

    def as_precision(self, precision: str) -> "Mlp7Params":
        return Mlp7Params({k: np.asarray(v) for k, v in self.items()}, precision, self.meta)


def load_weno7_checkpoint(path, device: str = "cuda", precision: str = "fp64") -> Mlp7Params:
    """weno7_core.py:382-423 load_checkpoint: same required names, input / output /
    hidden shape checks and error texts; read with ``allow_pickle=False``."""
    del device
    data = np.load(path, allow_pickle=False)
    missing = [name for name in REQUIRED if name not in data.files]
    if missing:
        raise ValueError(f"{path} is missing arrays: {missing}")
    if data["w1"].ndim != 3 or data["w1"].shape[1] != W7_INPUTS:
        raise ValueError(f"{path} has incompatible input shape {data['w1'].shape}")
    if data["w4"].ndim != 3 or data["w4"].shape[2] != W7_OUTPUTS:
        raise ValueError(f"{path} has incompatible output shape {data['w4'].shape}")
    hidden = (int(data["w1"].shape[2]), int(data["w2"].shape[2]), int(data["w3"].shape[2]))
    expected = weno7_checkpoint_shapes(hidden)
    wrong = {name: data[name].shape for name, shape in expected.items() if data[name].shape != shape}
    if wrong:
        raise ValueError(f"{path} has incompatible WENO7 MLP shapes: {wrong}")
    if hidden != W7_HIDDEN:
        raise ValueError(f"{path}: the GPU closure is built for hidden layers {W7_HIDDEN}, got {hidden}")
    meta: dict = {}
    if "meta_json" in data.files:
        try:
            meta = json.loads(str(data["meta_json"]))
        except json.JSONDecodeError:
            meta = {"raw": str(data["meta_json"])}
    arrays = {name: np.array(data[name], dtype=np.float64) for name in REQUIRED}
    return Mlp7Params(arrays, precision, meta)


def save_weno7_checkpoint(path, params: dict, meta: dict | None = None) -> None:
    """weno7_core.py:324-379 save_checkpoint layout (FP64, leading singleton axis,
    ``meta_json`` with the architecture tag)."""
    payload = {name: np.asarray(params[name], dtype=np.float64).reshape(shape)
               for name, shape in weno7_checkpoint_shapes().items()}
    full = dict(meta or {})
    full.update({"mlp_architecture": "reflection_sym_direct_beta_ratio_6_24_16_16_4", "hidden": list(W7_HIDDEN),
                 "precision": "float64", "reflection_formula": "0.5*(M(x)+P4*M(P6*x))",
                 "deployment_requires_reflection_symmetrization": True})
    payload["meta_json"] = np.array(json.dumps(full, sort_keys=True), dtype=np.str_)
    np.savez(path, **payload)
