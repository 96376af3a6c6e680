"""Exact self-similar solution of the 1-D Euler Riemann problem (ideal gas).

Host-side numpy, used by the shock-tube runners to score a run against exact
finite-volume averages.  Same entry points as
code/snapshot/for_paper_results/exact_riemann.py (``solve_star`` :40-80,
``sample_similarity`` :83-160, ``wave_speeds`` :163-203, ``sample_points``
:206-226, ``cell_average_conserved`` :241-283) and the Sod helpers the paper's
Sod table uses (``exact_sod_primitive``, pretrain_weno7_offline.py:922-937;
``sod_cell_average_conserved``, weno_z_borges_p2_results/run_riemann_1d.py:47-63).

The method is the textbook exact solver (Toro, *Riemann Solvers and Numerical
Methods for Fluid Dynamics*, ch. 4): the star pressure is the root of
f_L(p) + f_R(p) + (u_R - u_L) = 0 with the shock / rarefaction branches of f_K,
found by Newton iteration from the primitive-variable (PVRS) guess; the star
velocity is (u_L + u_R + f_R - f_L) / 2; the solution at x/t is sampled from the
left / right wave fans.
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np

Primitive = tuple[float, float, float]

SOD_LEFT: Primitive = (1.0, 0.0, 1.0)
SOD_RIGHT: Primitive = (0.125, 0.0, 0.1)


def _f_and_slope(p: float, rho: float, pk: float, ck: float, gamma: float) -> tuple[float, float]:
    """Toro (4.6)-(4.7): velocity jump across the K-wave for star pressure p and its derivative."""
    if p > pk:  # shock
        a_k = 2.0 / ((gamma + 1.0) * rho)
        b_k = (gamma - 1.0) / (gamma + 1.0) * pk
        s = np.sqrt(a_k / (p + b_k))
        return float((p - pk) * s), float(s * (1.0 - 0.5 * (p - pk) / (p + b_k)))
    # rarefaction
    z = (gamma - 1.0) / (2.0 * gamma)
    r = p / pk
    return (float(2.0 * ck / (gamma - 1.0) * (r ** z - 1.0)),
            float(r ** (-(gamma + 1.0) / (2.0 * gamma)) / (rho * ck)))


@lru_cache(maxsize=64)
def solve_star(left: Primitive, right: Primitive, gamma: float) -> tuple[float, float]:
    """(p*, u*) of the Riemann problem; raises if Newton does not converge in 100 steps."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)
    # PVRS guess (Toro 4.47), kept positive
    p = max(float(0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr)), np.finfo(np.float64).tiny)
    for _ in range(100):
        fl, dl = _f_and_slope(p, rl, pl, cl, gamma)
        fr, dr = _f_and_slope(p, rr, pr, cr, gamma)
        p_new = p - (fl + fr + ur - ul) / (dl + dr)
        if not np.isfinite(p_new) or p_new <= 0.0:
            p_new = 0.5 * p
        converged = abs(p_new - p) <= 1.0e-13 * 0.5 * (p_new + p)
        p = p_new
        if converged:
            break
    else:
        raise RuntimeError("exact Riemann pressure solve did not converge")
    fl, _ = _f_and_slope(p, rl, pl, cl, gamma)
    fr, _ = _f_and_slope(p, rr, pr, cr, gamma)
    return float(p), float(0.5 * (ul + ur + fr - fl))


def _shock_density(rho: float, ratio: float, gamma: float) -> float:
    g = (gamma - 1.0) / (gamma + 1.0)
    return float(rho * (ratio + g) / (g * ratio + 1.0))


def sample_similarity(s: float, left: Primitive, right: Primitive, gamma: float) -> Primitive:
    """Primitive (rho, u, p) at x/t = s (Toro 4.4, the sampling procedure)."""
    p_star, u_star = solve_star(left, right, gamma)
    gm1, gp1 = gamma - 1.0, gamma + 1.0
    if s <= u_star:
        rho, u, p = left
        c = np.sqrt(gamma * p / rho)
        sign = 1.0
    else:
        rho, u, p = right
        c = np.sqrt(gamma * p / rho)
        sign = -1.0
    # mirror the right fan onto the left one: s -> -s, u -> -u
    s_m, u_m, u_star_m = sign * s, sign * u, sign * u_star
    if p_star > p:  # shock
        shock = u_m - c * np.sqrt(gp1 / (2.0 * gamma) * (p_star / p) + gm1 / (2.0 * gamma))
        if s_m <= shock:
            return (rho, u, p)
        return (_shock_density(rho, p_star / p, gamma), u_star, p_star)
    c_star = c * (p_star / p) ** (gm1 / (2.0 * gamma))
    head, tail = u_m - c, u_star_m - c_star
    if s_m <= head:
        return (rho, u, p)
    if s_m >= tail:
        return (float(rho * (p_star / p) ** (1.0 / gamma)), u_star, p_star)
    # inside the fan
    u_f = 2.0 / gp1 * (c + 0.5 * gm1 * u_m + s_m)
    c_f = 2.0 / gp1 * (c + 0.5 * gm1 * (u_m - s_m))
    return (float(rho * (c_f / c) ** (2.0 / gm1)), float(sign * u_f),
            float(p * (c_f / c) ** (2.0 * gamma / gm1)))


def wave_speeds(left: Primitive, right: Primitive, gamma: float) -> tuple[float, ...]:
    """Sorted speeds of every wave edge: shocks, rarefaction heads / tails and the contact."""
    p_star, u_star = solve_star(left, right, gamma)
    gm1, gp1 = gamma - 1.0, gamma + 1.0
    out = [u_star]
    for (rho, u, p), sign in ((left, 1.0), (right, -1.0)):
        c = np.sqrt(gamma * p / rho)
        if p_star > p:
            out.append(u - sign * c * np.sqrt(gp1 / (2.0 * gamma) * (p_star / p) + gm1 / (2.0 * gamma)))
        else:
            c_star = c * (p_star / p) ** (gm1 / (2.0 * gamma))
            out.extend((u - sign * c, u_star - sign * c_star))
    return tuple(sorted(float(v) for v in out))


def sample_points(x, time: float, discontinuity: float, left: Primitive, right: Primitive,
                  gamma: float = 1.4) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Pointwise exact primitives at the coordinates x and time t."""
    x = np.asarray(x, dtype=np.float64)
    if time <= 0.0:
        mask = x < discontinuity
        return tuple(np.where(mask, left[k], right[k]).astype(np.float64) for k in range(3))
    vals = np.array([sample_similarity((float(c) - discontinuity) / time, left, right, gamma) for c in x.ravel()],
                    dtype=np.float64).reshape(x.shape + (3,))
    return vals[..., 0], vals[..., 1], vals[..., 2]


def _conserved4(prim: Primitive, gamma: float) -> np.ndarray:
    rho, u, p = prim
    return np.array([rho, rho * u, 0.0, p / (gamma - 1.0) + 0.5 * rho * u * u], dtype=np.float64)


def cell_average_conserved(centers, width: float, time: float, discontinuity: float, left: Primitive,
                           right: Primitive, gamma: float = 1.4, quadrature: int = 20) -> np.ndarray:
    """Exact cell averages (rho, rho u, 0, E): Gauss-Legendre on every smooth piece of each cell
    (the cell is split at the wave edges that cross it)."""
    nodes, weights = np.polynomial.legendre.leggauss(quadrature)
    if time > 0.0:
        edges_all = np.array([discontinuity + time * s for s in wave_speeds(left, right, gamma)])
    else:
        edges_all = np.array([discontinuity])
    centers = np.asarray(centers, dtype=np.float64)
    out = np.zeros((centers.size, 4), dtype=np.float64)
    for n, c in enumerate(centers):
        lo, hi = c - 0.5 * width, c + 0.5 * width
        cuts = np.concatenate(([lo], edges_all[(edges_all > lo) & (edges_all < hi)], [hi]))
        acc = np.zeros(4, dtype=np.float64)
        for a, b in zip(cuts[:-1], cuts[1:]):
            mid, half = 0.5 * (a + b), 0.5 * (b - a)
            for xi, wi in zip(nodes, weights):
                xq = mid + half * float(xi)
                if time <= 0.0:
                    prim = left if xq < discontinuity else right
                else:
                    prim = sample_similarity((xq - discontinuity) / time, left, right, gamma)
                acc += half * float(wi) * _conserved4(prim, gamma)
        out[n] = acc / width
    return out


def exact_sod_primitive(x, t: float, gamma: float) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Sod (1, 0, 1) | (0.125, 0, 0.1) with the discontinuity at x = 0."""
    return sample_points(x, t, 0.0, SOD_LEFT, SOD_RIGHT, gamma)


def sod_cell_average_conserved(centers, dx: float, t: float, discontinuity: float,
                               quadrature: int = 15, gamma: float = 1.4) -> np.ndarray:
    """Unsplit Gauss-Legendre cell averages of the exact Sod solution, components (rho, rho u, 0, E),
    accumulated node by node as run_riemann_1d.py:47-63 does."""
    xi, wi = np.polynomial.legendre.leggauss(quadrature)
    centers = np.asarray(centers, dtype=np.float64)
    state = np.zeros((centers.size, 4), dtype=np.float64)
    for node, weight in zip(xi, wi):
        xq = centers + 0.5 * dx * float(node) - discontinuity
        rho, vel, p = exact_sod_primitive(xq, t, gamma)
        state[:, 0] += 0.5 * float(weight) * rho
        state[:, 1] += 0.5 * float(weight) * rho * vel
        state[:, 3] += 0.5 * float(weight) * (p / (gamma - 1.0) + 0.5 * rho * vel * vel)
    return state
