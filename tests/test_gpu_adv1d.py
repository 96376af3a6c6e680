"""1-D advection path on the GPU (C ABI wsr_adv1d_rollout) against the CPU oracle
and the reference's recorded GSTE table.

Tolerances: max |gpu - oracle| / max |oracle| <= 1e-12 for the FP64 closures after
20 steps; FP32 closure 2e-6, BF16 closure 2e-4 (GPU exp2 / reciprocal vs libm).
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from conftest import golden

pytestmark = pytest.mark.gpu

TOL = {"js": 1e-12, "sr64": 1e-12, "sr32": 2e-6, "srbf16": 2e-4}


def profiles(batch: int, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = (np.arange(n) + 0.5) / n
    out = np.empty((batch, n))
    for b in range(batch):
        k = rng.integers(1, 8)
        out[b] = (np.sin(2 * np.pi * k * x + rng.uniform(0, 6)) * rng.normal()
                  + (np.abs(x - rng.uniform(0.2, 0.8)) < rng.uniform(0.05, 0.2)) * rng.normal()
                  + np.exp(-((x - rng.uniform(0, 1)) / rng.uniform(0.02, 0.2)) ** 2))
        if b % 4 == 3:
            out[b, : n // 3] = 0.3   # a plateau
    return out


def model_for(scheme, mlp64, mlp32):
    return {"js": None, "sr64": mlp64, "sr32": mlp32, "srbf16": mlp32.as_precision("bf16")}[scheme]


@pytest.mark.parametrize("scheme", ["js", "sr64", "sr32", "srbf16"])
def test_batch_rollout_parity(oracle, mlp64, mlp32, gpu_device, scheme):
    from weno_sr_publishing_amd.advection1d import Adv1DSolver
    n, batch, steps = 256, 16, 20
    u0 = profiles(batch, n, 5)
    vel = np.where(np.arange(batch) % 2 == 0, 1.0, -1.0)
    dts = [0.6 / n] * (steps - 1) + [0.17 / n]
    model = model_for(scheme, mlp64, mlp32)
    u = torch.tensor(u0, device=gpu_device)
    ck = Adv1DSolver(n, n, model, gpu_device).rollout(u, torch.tensor(vel), dts, checkpoints=True)
    ref = oracle.adv1d_run(u0, vel, dts, n, scheme, model)
    got = u.cpu().numpy()
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"adv1d parity {scheme}: {err:.3e}")
    assert err <= TOL[scheme]
    ckh = ck.cpu().numpy()
    assert np.array_equal(ckh[:, 0], u0) and np.array_equal(ckh[:, -1], got)
    one = oracle.adv1d_run(u0, vel, dts[:1], n, scheme, model)
    assert np.max(np.abs(ckh[:, 1] - one)) / np.max(np.abs(one)) <= TOL[scheme]


@pytest.mark.parametrize("key,scheme", [("weno5_js", "js"), ("weno5_sr_f64", "sr64"), ("weno5_sr_f32", "sr32")])
def test_gste_against_reference_table(mlp64, mlp32, gpu_device, key, scheme):
    """N=200, t=10, CFL 0.6, 1667 steps on the GPU vs the recorded metrics."""
    from weno_sr_publishing_amd.advection1d import run_gste
    ref = {r["method"]: r for r in golden("gste_metrics.json")["rows"]}[key]
    row = run_gste(model_for(scheme, mlp64, mlp32), nx=200, t_end=10.0, cfl=0.6, device=gpu_device)
    assert row["steps"] == 1667
    tol = 1e-10 if scheme != "sr32" else 1e-6
    for k in ("l1", "l2", "linf", "tv"):
        rel = abs(row[k] / float(ref[k]) - 1.0)
        print(f"gste {key} {k}: {row[k]!r} vs {ref[k]} rel {rel:.2e}")
        assert rel <= tol


def test_gste_against_oracle_final_state(oracle, mlp64, gpu_device):
    from weno_sr_publishing_amd.advection1d import run_gste
    from weno_sr_publishing_amd.initial import gste_cell_averages
    row = run_gste(mlp64, device=gpu_device)
    _, u0 = gste_cell_averages(200, 0.0, 15)
    ref = oracle.adv1d_run(u0[None, :], [1.0], [row["dt"]] * row["steps"], 100.0, "sr64", mlp64)[0]
    err = np.max(np.abs(row["final"] - ref))
    print(f"gste sr64 final-state diff vs oracle after 1667 steps: {err:.3e}")
    assert err <= 1e-10
