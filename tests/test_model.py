"""The closure as a torch module (model.load_checkpoint, model(features) -> ratios)
against the release verifier's golden outputs (scripts/verify_release.py:157-222,
carried in tests/golden/mlp_golden.json), and the refusals of the 1-D drop-ins.
No GPU needed."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import golden


@pytest.mark.parametrize("label,path_name", [("WENO5 FP64", "WENO5_SR_FP64"), ("WENO5 FP32", "WENO5_SR_FP32")])
def test_closure_module_reproduces_release_golden_vectors(label, path_name):
    from weno_sr_publishing_amd import checkpoints
    from weno_sr_publishing_amd.model import load_checkpoint
    g = golden("mlp_golden.json")
    model = load_checkpoint(getattr(checkpoints, path_name))
    assert model.precision == ("fp64" if "FP64" in label else "fp32")
    with torch.no_grad():
        out = model(torch.tensor(g["features5"], dtype=torch.float64)).numpy()
    np.testing.assert_allclose(out, np.array(g["expected"][label]), rtol=0, atol=g["atol"][label])


def test_closure_module_is_reflection_equivariant_and_starts_linear():
    from weno_sr_publishing_amd import checkpoints
    from weno_sr_publishing_amd.model import ReflectionSymmetricClosure, load_checkpoint, reflection_defect
    feats = torch.rand(512, 5, dtype=torch.float64, generator=torch.Generator().manual_seed(3))
    assert reflection_defect(load_checkpoint(checkpoints.WENO5_SR_FP64), feats) <= 1e-14
    # zero final layer: equal ratios (weno5_core.SharedBadnessMLP initialisation)
    fresh = ReflectionSymmetricClosure("fp64", seed=41)
    with torch.no_grad():
        r = fresh(feats)
    assert torch.allclose(r, torch.full_like(r, 1.0 / 3.0), atol=1e-15)


def test_load_checkpoint_errors(tmp_path):
    from weno_sr_publishing_amd import checkpoints
    from weno_sr_publishing_amd.model import load_checkpoint
    with np.load(checkpoints.WENO5_SR_FP64, allow_pickle=False) as data:
        arrays = {k: data[k] for k in data.files}
    bad = dict(arrays)
    bad.pop("b3")
    np.savez(tmp_path / "missing.npz", **bad)
    with pytest.raises(ValueError, match="missing arrays"):
        load_checkpoint(tmp_path / "missing.npz")
    bad = dict(arrays)
    bad["w2"] = np.zeros((1, 6, 10))
    np.savez(tmp_path / "shape.npz", **bad)
    with pytest.raises(ValueError, match="incompatible WENO5 MLP shapes"):
        load_checkpoint(tmp_path / "shape.npz")


def test_rollout_refuses_host_state_and_non_differentiable_models():
    from weno_sr_publishing_amd import advection1d
    u = torch.zeros((2, 16), dtype=torch.float64)
    with pytest.raises(ValueError, match="GPU"):
        advection1d.ssprk3_step_signed(None, u, 0.01, 16.0, torch.tensor([1.0, -1.0]))


def test_resume_file_is_plain_arrays_and_refuses_other_recipes(tmp_path):
    """The resume file loads without unpickling; a signature or recipe mismatch raises
    before any state is touched (no GPU call: the trainer's parameters live on the CPU
    here)."""
    import json
    from weno_sr_publishing_amd.training import ClosureTrainer, training_signature
    tr = ClosureTrainer(None, "fp64", lr=1e-3, device="cpu")
    sig = training_signature(64, 16, 8, 0.6, "fp64", 1e-3, 83)
    gen = torch.Generator().manual_seed(83)
    tr.save_state(tmp_path / "s.npz", 5, sig, gen)
    with np.load(tmp_path / "s.npz", allow_pickle=False) as data:
        meta = json.loads(str(data["meta_json"]))
        assert meta["step"] == 5 and meta["training_signature"] == sig
        assert data["param"].shape == (189,) and data["generator_state"].dtype == np.uint8
    other = ClosureTrainer(None, "fp64", lr=1e-3, device="cpu", seed=3)
    before = other.param.clone()
    with pytest.raises(ValueError, match="signature mismatch"):
        other.resume(tmp_path / "s.npz", {**sig, "batch": 32})
    assert torch.equal(other.param, before)
    assert other.resume(tmp_path / "s.npz", sig, torch.Generator()) == 5
    assert torch.equal(other.param, tr.param)
