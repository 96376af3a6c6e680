"""Hand-written adjoint of the 1-D rollout (wsr_adv1d_train_forward / _adjoint).

Checked against (1) the reference's own torch-autograd gradients of the same loss
(tests/golden/grad_fixture.npz, written by tests/golden/make_grad_fixture.py from
the reference's rk3_advection / model / v12 modules), (2) central finite
differences of the CPU oracle's FP64 forward, and (3) the chain rule on the state
(adjoint of the initial state vs finite differences).
Tolerances: FP64 closure 1e-8 relative (max-norm over the 189 gradients); FP32
closure 2e-3 relative (FP32 MLP with the GPU's exp2 / reciprocal against torch's
FP32 autograd).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

FIX = Path(__file__).resolve().parent / "golden" / "grad_fixture.npz"
NAMES = ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4")


def fixture():
    return np.load(FIX, allow_pickle=False)


def flat_grad(fx, prefix):
    return np.concatenate([fx[f"{prefix}_grad_{n}"].reshape(-1) for n in NAMES])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adjoint_matches_reference_autograd(mlp64, mlp32, gpu_device, precision):
    from weno_sr_publishing_amd.training import ClosureTrainer
    fx = fixture()
    mlp = mlp64 if precision == "fp64" else mlp32
    tr = ClosureTrainer({k: np.asarray(v)[0] for k, v in mlp.items()}, precision, device=gpu_device)
    loss, grad, final = tr.loss_and_grad(torch.tensor(fx["u0"]), torch.tensor(fx["velocities"]),
                                         torch.tensor(fx["target"]), fx["dts"])
    ref = flat_grad(fx, precision)
    g = grad.cpu().numpy()
    rel = np.max(np.abs(g - ref)) / np.max(np.abs(ref))
    fin = np.max(np.abs(final.cpu().numpy() - fx[f"{precision}_final"]))
    print(f"adjoint {precision}: grad rel {rel:.3e}, final-state diff {fin:.3e}, loss {float(loss)!r} "
          f"vs {float(fx[f'{precision}_loss'])!r}")
    if precision == "fp64":
        assert fin <= 1e-13 and rel <= 1e-8
        assert abs(float(loss) - float(fx["fp64_loss"])) <= 1e-14
    else:
        assert fin <= 1e-6 and rel <= 2e-3


def test_adjoint_matches_oracle_finite_differences(oracle, mlp64, gpu_device):
    """Central differences of the oracle's FP64 loss for 24 parameters and the
    initial-state adjoint for 6 cells."""
    from weno_sr_publishing_amd.training import ClosureTrainer, flatten, unflatten
    fx = fixture()
    base = {k: np.asarray(v)[0] for k, v in mlp64.items()}
    tr = ClosureTrainer(base, "fp64", device=gpu_device)
    _, grad, _, lam0 = tr.loss_and_grad(torch.tensor(fx["u0"]), torch.tensor(fx["velocities"]),
                                        torch.tensor(fx["target"]), fx["dts"], return_state_grad=True)
    g = grad.cpu().numpy()
    flat = flatten(base)
    n = int(fx["n"])

    def loss_of(f, u0=fx["u0"]):
        p = {k: v[None, ...] if v.ndim == 2 else v[None, :] for k, v in unflatten(f).items()}
        fin = oracle.adv1d_run(u0, fx["velocities"], fx["dts"], n, "sr64", p)
        return float(np.mean((fin - fx["target"]) ** 2))

    rng = np.random.default_rng(2)
    picks = sorted(set(rng.choice(189, size=24, replace=False).tolist()) | {0, 55, 130, 170, 187})
    for p in picks:
        h = 1e-6 * max(1.0, abs(flat[p]))
        fp, fm = flat.copy(), flat.copy()
        fp[p] += h
        fm[p] -= h
        fd = (loss_of(fp) - loss_of(fm)) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * max(1e-3, np.max(np.abs(g))), (p, fd, g[p])
    lam = lam0.cpu().numpy()
    ref_u = fx["fp64_grad_u0"]
    assert np.max(np.abs(lam - ref_u)) <= 1e-9 * np.max(np.abs(ref_u))
    # finite differences only where every stencil around the cell is smooth: a
    # perturbation inside a constant region switches the plateau branch on or off
    for (b, i) in ((0, 3), (1, 10), (0, 20), (1, 25), (1, 4)):
        h = 1e-6
        up, um = fx["u0"].copy(), fx["u0"].copy()
        up[b, i] += h
        um[b, i] -= h
        fd = (loss_of(flat, up) - loss_of(flat, um)) / (2 * h)
        assert abs(fd - lam[b, i]) <= 1e-6 * max(1e-3, np.max(np.abs(lam))), (b, i, fd, lam[b, i])


def test_training_step_reduces_loss(gpu_device):
    """A few Adam steps on a 64-trajectory batch of solver-generated data lower the loss."""
    from weno_sr_publishing_amd.profiles import make_batch, shifted_target
    from weno_sr_publishing_amd.training import ClosureTrainer, distance_schedule
    n = 128
    u0, vel = make_batch(64, n, seed=7)
    target = shifted_target(u0, vel, 16)
    dts = distance_schedule(n, 16.0, 0.6)
    tr = ClosureTrainer(None, "fp32", lr=3e-3, device=gpu_device, seed=41)
    losses = [tr.step(torch.tensor(u0), torch.tensor(vel), torch.tensor(target), dts) for _ in range(8)]
    print("losses", losses)
    assert np.all(np.isfinite(losses)) and losses[-1] < losses[0]


def test_rollout_distance_backward_matches_reference_autograd_fp32(gpu_device):
    """mse(rollout_distance(model, u, 3.7, 0.6, v), target).backward() with the FP32
    closure module: the parameter gradients of the reference's own autograd
    (rk3_advection.rollout_distance with model.ReflectionSymmetricBadnessMLPFloat32)."""
    from weno_sr_publishing_amd.advection1d import rollout_distance
    from weno_sr_publishing_amd.checkpoints import WENO5_SR_FP32
    from weno_sr_publishing_amd.model import NAMES, load_checkpoint
    fx = fixture()
    model = load_checkpoint(WENO5_SR_FP32, gpu_device)
    u = torch.tensor(fx["u0"], device=gpu_device)
    v = torch.tensor(fx["velocities"], device=gpu_device)
    final = rollout_distance(model, u, float(fx["distance"]), float(fx["cfl"]), v)
    loss = torch.mean((final - torch.tensor(fx["target"], device=gpu_device)) ** 2)
    loss.backward()
    got = np.concatenate([getattr(model, n).grad.double().cpu().numpy().reshape(-1) for n in NAMES])
    ref = flat_grad(fx, "fp32")
    rel = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    fin = np.max(np.abs(final.detach().cpu().numpy() - fx["fp32_final"]))
    print(f"autograd fp32: grad rel {rel:.3e}, final diff {fin:.3e}", flush=True)
    assert fin <= 1e-6 and rel <= 1e-5


def test_chained_steps_backward_matches_reference_autograd_fp64(gpu_device):
    """Seven chained ssprk3_step_signed calls (one autograd node each) with the FP64
    closure module and an initial state that requires grad: the reference's
    v12_losses.ssprk3_step_signed autograd gradients of the parameters and of u0."""
    from weno_sr_publishing_amd.advection1d import ssprk3_step_signed
    from weno_sr_publishing_amd.checkpoints import WENO5_SR_FP64
    from weno_sr_publishing_amd.model import NAMES, load_checkpoint
    fx = fixture()
    n = int(fx["n"])
    model = load_checkpoint(WENO5_SR_FP64, gpu_device)
    u0 = torch.tensor(fx["u0"], device=gpu_device, requires_grad=True)
    v = torch.tensor(fx["velocities"], device=gpu_device)
    state = u0
    for dt in fx["dts"]:
        state = ssprk3_step_signed(model, state, float(dt), float(n), v)
    loss = torch.mean((state - torch.tensor(fx["target"], device=gpu_device)) ** 2)
    loss.backward()
    got = np.concatenate([getattr(model, name).grad.cpu().numpy().reshape(-1) for name in NAMES])
    ref = flat_grad(fx, "fp64")
    rel = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    ref_u = fx["fp64_grad_u0"]
    rel_u = np.max(np.abs(u0.grad.cpu().numpy() - ref_u)) / np.max(np.abs(ref_u))
    print(f"autograd fp64 chained: grad rel {rel:.3e}, u0 grad rel {rel_u:.3e}, loss {loss.item()!r}", flush=True)
    assert abs(loss.item() - float(fx["fp64_loss"])) <= 1e-14
    assert rel <= 1e-8 and rel_u <= 1e-9


def test_autograd_gradient_at_training_length_matches_oracle_differences(oracle, gpu_device):
    """The training workload's rollout length (N = 256, 64 cells at CFL 0.6 = 107 steps),
    a 16-trajectory slice: the autograd gradient of the FP64 closure module along three
    random parameter directions against central differences of the oracle's FP64
    forward rollout."""
    from weno_sr_publishing_amd.advection1d import cfl_schedule, rollout_distance
    from weno_sr_publishing_amd.checkpoints import WENO5_SR_FP64, MlpParams
    from weno_sr_publishing_amd.model import NAMES, SHAPES, load_checkpoint
    from weno_sr_publishing_amd.profiles import make_batch, shifted_target
    n, batch, distance = 256, 16, 64.0
    u0, vel = make_batch(batch, n, seed=83)
    target = shifted_target(u0, vel, 64)
    dts = [c / n for c in cfl_schedule(distance, 0.6)]
    assert len(dts) == 107
    model = load_checkpoint(WENO5_SR_FP64, gpu_device)
    final = rollout_distance(model, torch.tensor(u0, device=gpu_device), distance, 0.6,
                             torch.tensor(vel, device=gpu_device))
    loss = torch.mean((final - torch.tensor(target, device=gpu_device)) ** 2)
    loss.backward()
    g = np.concatenate([getattr(model, name).grad.cpu().numpy().reshape(-1) for name in NAMES])
    flat = np.concatenate([getattr(model, name).detach().cpu().numpy().reshape(-1) for name in NAMES])

    def loss_of(f):
        arrays, o = {}, 0
        for name in NAMES:
            size = int(np.prod(SHAPES[name]))
            arrays[name] = f[o:o + size].reshape((1,) + SHAPES[name])
            o += size
        fin = oracle.adv1d_run(u0, vel, dts, n, "sr64", MlpParams(arrays, "fp64"))
        return float(np.mean((fin - target) ** 2))

    assert abs(loss_of(flat) - float(loss)) <= 1e-13 * float(loss)
    rng = np.random.default_rng(29)
    for _ in range(3):
        d = rng.normal(size=189)
        d /= np.linalg.norm(d)
        h = 1e-6
        fd = (loss_of(flat + h * d) - loss_of(flat - h * d)) / (2 * h)
        ad = float(g @ d)
        print(f"directional derivative: autograd {ad:.10e}, central difference {fd:.10e}", flush=True)
        assert abs(ad - fd) <= 1e-5 * np.linalg.norm(g)


def test_training_resume_is_bit_identical(gpu_device, tmp_path):
    """Save the resume state after 2 of 4 steps, drop the trainer, resume in a fresh one
    (built from another seed: everything must come from the file): the losses of steps
    3 and 4 and the final parameters equal the uninterrupted run bit for bit.  A resume
    with another recipe signature is refused; the save leaves no temporary file."""
    from weno_sr_publishing_amd.training import ClosureTrainer, train
    kw = dict(n=64, batch=16, distance_cells=8, cfl=0.6, seed=83)
    full = ClosureTrainer(None, "fp32", lr=3e-3, device=gpu_device, seed=41)
    losses_full = train(full, steps=4, **kw)
    first = ClosureTrainer(None, "fp32", lr=3e-3, device=gpu_device, seed=41)
    state = tmp_path / "state.npz"
    losses_a = train(first, steps=2, state_path=state, **kw)
    del first
    second = ClosureTrainer(None, "fp32", lr=3e-3, device=gpu_device, seed=7)
    losses_b = train(second, steps=4, resume_from=state, **kw)
    print("losses", losses_full, losses_a + losses_b, flush=True)
    assert losses_a + losses_b == losses_full
    assert torch.equal(second.param, full.param)
    assert sorted(p.name for p in tmp_path.iterdir()) == ["state.npz"]
    with pytest.raises(ValueError, match="signature mismatch"):
        train(ClosureTrainer(None, "fp32", lr=3e-3, device=gpu_device), steps=5, resume_from=state,
              **{**kw, "cfl": 0.5})
