"""Random periodic profiles on [0, 1) with exact cell averages (training data).

Families (the reference's profile generator, code/snapshot/
teacherfree_lab_weno5_v20_distance_balanced_mlp_f32_fast/fvm_profiles.py:12-425, draws
the same kinds): Fourier sums with wavenumbers k <= 8 and unit-normal
amplitudes, Gaussians of width U[0.02, 0.2], square, triangle and semi-ellipse
pulses, and composites of two or three of these.  Every cell average is the exact
integral of the profile over the cell (closed forms), so a rollout over an
integer number of cells has the rolled initial averages as its exact target.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

KINDS = ("fourier", "gaussian", "square", "triangle", "ellipse", "composite")


def _antiderivative(kind: str, x: np.ndarray, c: float, w: float) -> np.ndarray:
    """Antiderivative of one non-periodic pulse centred at c with half-width w."""
    t = (x - c) / w
    if kind == "gaussian":
        return 0.5 * math.sqrt(math.pi) * w * erf(t)
    if kind == "square":
        return w * np.clip(t, -1.0, 1.0)
    if kind == "triangle":
        tc = np.clip(t, -1.0, 1.0)
        return w * (tc - 0.5 * np.sign(tc) * tc * tc)
    if kind == "ellipse":
        tc = np.clip(t, -1.0, 1.0)
        return 0.5 * w * (tc * np.sqrt(1.0 - tc * tc) + np.arcsin(tc))
    raise ValueError(kind)


def _pulse_averages(kind: str, edges: np.ndarray, c: float, w: float) -> np.ndarray:
    h = edges[1] - edges[0]
    total = np.zeros(edges.size - 1)
    for shift in (-1.0, 0.0, 1.0):   # periodic images (w <= 0.25)
        F = _antiderivative(kind, edges, c + shift, w)
        total += F[1:] - F[:-1]
    return total / h


def _fourier_averages(edges: np.ndarray, rng) -> np.ndarray:
    h = edges[1] - edges[0]
    kmax = int(rng.integers(1, 9))
    out = np.zeros(edges.size - 1)
    for k in range(1, kmax + 1):
        a, b = rng.normal(), rng.normal()
        wk = 2.0 * math.pi * k
        S = np.sin(wk * edges)
        C = np.cos(wk * edges)
        out += (a * (S[1:] - S[:-1]) - b * (C[1:] - C[:-1])) / (wk * h)
    return out


def profile_averages(kind: str, n: int, rng) -> np.ndarray:
    edges = np.arange(n + 1, dtype=np.float64) / n
    if kind == "fourier":
        return _fourier_averages(edges, rng)
    if kind == "gaussian":
        return _pulse_averages("gaussian", edges, rng.uniform(0, 1), rng.uniform(0.02, 0.2))
    if kind in ("square", "triangle", "ellipse"):
        return _pulse_averages(kind, edges, rng.uniform(0, 1), rng.uniform(0.03, 0.25))
    if kind == "composite":
        parts = rng.choice(["fourier", "gaussian", "square", "triangle", "ellipse"], size=int(rng.integers(2, 4)))
        return sum(rng.normal() * profile_averages(str(p), n, rng) for p in parts)
    raise ValueError(kind)


def make_batch(batch: int, n: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """(B, N) exact cell averages and balanced velocities a = +1 (first half), -1."""
    if batch % 2:
        raise ValueError("bidirectional batches must be even")
    rng = np.random.default_rng(seed)
    kinds = rng.choice(KINDS, size=batch)
    u = np.stack([profile_averages(str(k), n, rng) for k in kinds])
    vel = np.ones(batch)
    vel[batch // 2:] = -1.0
    return u, vel


def shifted_target(u0: np.ndarray, velocities: np.ndarray, cells: int) -> np.ndarray:
    """Exact averages after advecting by an integer number of cells: q(x - a t)."""
    out = np.empty_like(u0)
    for b in range(u0.shape[0]):
        out[b] = np.roll(u0[b], int(round(velocities[b])) * cells)
    return out
