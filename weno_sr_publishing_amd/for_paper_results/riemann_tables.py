"""Error tables of the Sod and Lax tubes against exact finite-volume averages
(code/snapshot/weno_z_borges_p2_results/plot_riemann_1d_zoom.py:95-140 ``write_errors`` and
plot_lax_paper.py:30-151; the plots are not reproduced).

``density_errors`` scores the transverse-mean density (Sod table columns
``rho_l1_mean, rho_l1_integral, rho_l2, rho_linf``); ``primitive_errors`` scores density,
velocity and pressure of the transverse-mean conserved state (Lax table columns).
"""
from __future__ import annotations

import csv
from pathlib import Path

import numpy as np

from .. import exact_riemann
from . import config
from .run_riemann_1d import PROBLEMS, Problem

LABELS = {**{k: m.label for k, m in config.METHODS.items()}, **config.Z_LABELS}


def centers(problem: Problem) -> tuple[np.ndarray, float]:
    dx = (problem.x_max - problem.x_min) / problem.nx
    return problem.x_min + (np.arange(problem.nx) + 0.5) * dx, dx


def exact_density(problem: Problem) -> np.ndarray:
    xc, dx = centers(problem)
    if problem.key == "sod":
        return exact_riemann.sod_cell_average_conserved(xc, dx, problem.t_end, problem.x0, quadrature=15)[:, 0]
    return exact_riemann.cell_average_conserved(xc, dx, problem.t_end, problem.x0, problem.left,
                                                problem.right, quadrature=15)[:, 0]


def density_errors(problem: Problem, results: dict[str, tuple[np.ndarray, np.ndarray]]) -> list[dict]:
    """results: method -> (cell centres, transverse-mean density)."""
    xc, dx = centers(problem)
    exact = exact_density(problem)
    rows = []
    for method, (x, rho) in results.items():
        if not np.array_equal(x, xc):
            raise RuntimeError(f"{method}: numerical grid does not match exact FV grid")
        error = rho - exact
        rows.append({
            "problem": problem.key, "method": method, "label": LABELS[method], "nx": problem.nx,
            "t_end": problem.t_end, "cfl": problem.cfl, "exact_reference": "15-point Gauss exact cell averages",
            "rho_l1_mean": float(np.mean(np.abs(error))), "rho_l1_integral": float(dx * np.sum(np.abs(error))),
            "rho_l2": float(np.sqrt(np.mean(error * error))), "rho_linf": float(np.max(np.abs(error))),
        })
    return rows


def primitive_fields(conserved: np.ndarray, gamma: float = 1.4) -> dict[str, np.ndarray]:
    rho = conserved[:, 0]
    velocity = conserved[:, 1] / rho
    transverse = conserved[:, 2] / rho
    pressure = (gamma - 1.0) * (conserved[:, 3] - 0.5 * rho * (velocity * velocity + transverse * transverse))
    if not np.all(np.isfinite(conserved)) or np.any(rho <= 0.0) or np.any(pressure <= 0.0):
        raise ValueError("invalid state; refusing repair")
    return {"rho": rho, "velocity": velocity, "pressure": pressure}


def primitive_errors(problem: Problem, results: dict[str, tuple[np.ndarray, np.ndarray]]) -> list[dict]:
    """results: method -> (cell centres, transverse-mean conserved state (nx, 4))."""
    xc, dx = centers(problem)
    exact = exact_riemann.cell_average_conserved(xc, dx, problem.t_end, problem.x0, problem.left,
                                                 problem.right, quadrature=15)
    exact_p = primitive_fields(exact)
    rows = []
    for method, (x, conserved) in results.items():
        if not np.array_equal(x, xc):
            raise RuntimeError(f"{method}: grid does not match the formal {problem.key} grid")
        values = primitive_fields(conserved)
        row = {"problem": problem.key, "method": method, "label": LABELS[method], "nx": problem.nx,
               "t_end": problem.t_end, "cfl": problem.cfl, "reference": "15-point Gauss exact cell averages"}
        for field in ("rho", "velocity", "pressure"):
            delta = values[field] - exact_p[field]
            row.update({
                f"{field}_l1_mean": float(np.mean(np.abs(delta))),
                f"{field}_l1_integral": float(dx * np.sum(np.abs(delta))),
                f"{field}_l2_mean": float(np.sqrt(np.mean(delta * delta))),
                f"{field}_linf": float(np.max(np.abs(delta))),
            })
        rows.append(row)
    return rows


def write_csv(path: Path, rows: list[dict]) -> Path:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("w", newline="", encoding="utf-8") as stream:
        writer = csv.DictWriter(stream, fieldnames=list(rows[0]))
        writer.writeheader()
        writer.writerows(rows)
    return path


def problem(key: str, **overrides) -> Problem:
    from dataclasses import replace
    return replace(PROBLEMS[key], **overrides)
