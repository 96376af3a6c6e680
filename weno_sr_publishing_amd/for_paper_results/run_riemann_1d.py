"""WENO-Z rows of the Sod and Lax tables
(code/snapshot/weno_z_borges_p2_results/run_riemann_1d.py:23-182).

WENO5-Z (p = 2, epsilon = dx^2) with SSPRK3 and WENO7-Z (p = 3, epsilon = dx^3) with
Shu RK4, characteristic reconstruction, HLLC, transmissive / outflow ghost cells; the
row scores the primitive density of the transverse-mean state against the exact
finite-volume averages.
"""
from __future__ import annotations

import argparse
import csv
import json
import time
from dataclasses import dataclass, replace
from pathlib import Path

import numpy as np

from .. import exact_riemann
from ..initial import make_sod_state
from ..solvers import euler_z
from . import config
from .run_sod import make_weno7_sod_state


@dataclass(frozen=True)
class Problem:
    key: str
    x_min: float
    x_max: float
    x0: float
    t_end: float
    nx: int
    ny: int
    cfl: float
    left: tuple[float, float, float]
    right: tuple[float, float, float]


PROBLEMS = {
    "sod": Problem("sod", 0.0, 1.0, 0.5, 0.2, 51, 11, 0.4, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)),
    "lax": Problem("lax", -5.0, 5.0, 0.0, 1.3, 200, 10, 0.8, (0.445, 0.698, 3.528), (0.5, 0.0, 0.571)),
}


def initial_state(params, problem: Problem, order: int) -> np.ndarray:
    """run_riemann_1d.py:66-85."""
    if problem.key == "sod":
        if order == 5:
            return make_sod_state(params, "x")
        return make_weno7_sod_state(params, 15, problem.x0)
    g = params.ghost
    dx = (problem.x_max - problem.x_min) / problem.nx
    centers = problem.x_min + (np.arange(problem.nx + 2 * g, dtype=np.float64) - g + 0.5) * dx
    cell = exact_riemann.cell_average_conserved(centers, dx, 0.0, problem.x0, problem.left, problem.right,
                                                quadrature=15)
    return np.broadcast_to(cell[None, :, :], (problem.ny + 2 * g, problem.nx + 2 * g, 4)).copy()


def make_params(problem: Problem, order: int):
    dx = (problem.x_max - problem.x_min) / problem.nx
    y_length = problem.ny * dx
    if order == 5:
        return euler_z.make_weno5_params(problem.nx, problem.ny, problem.x_max - problem.x_min, y_length,
                                         problem.cfl, problem.t_end)
    return euler_z.make_weno7_params(problem.nx, problem.ny, problem.x_min, problem.x_max, 0.0, y_length,
                                     problem.cfl, problem.t_end)


def exact_conserved(problem: Problem, x: np.ndarray, dx: float) -> np.ndarray:
    if problem.key == "sod":
        return exact_riemann.sod_cell_average_conserved(x, dx, problem.t_end, problem.x0, quadrature=15)
    return exact_riemann.cell_average_conserved(x, dx, problem.t_end, problem.x0, problem.left, problem.right,
                                                quadrature=15)


def run(problem_key: str, device: str = "cuda", report_interval: int = 0, cfl: float | None = None,
        nx: int | None = None, ny: int | None = None, out_tag: str | None = None, write: bool = True) -> list[dict]:
    problem = PROBLEMS[problem_key]
    if cfl is not None:
        problem = replace(problem, cfl=cfl)
    if nx is not None:
        problem = replace(problem, nx=nx)
    if ny is not None:
        problem = replace(problem, ny=ny)
    out = config.RAW / f"riemann_1d/{problem.key}"
    if out_tag:
        out = out / out_tag
    if write:
        out.mkdir(parents=True, exist_ok=True)
    rows = []
    for order in (5, 7):
        dx = (problem.x_max - problem.x_min) / problem.nx
        params = make_params(problem, order)
        run_fn = euler_z.run_weno5 if order == 5 else euler_z.run_weno7
        boundary = "transmissive" if order == 5 else "outflow"
        initial = initial_state(params, problem, order)
        started = time.perf_counter()
        final, summary = run_fn(initial, params, device=device, boundary=boundary, report_interval=report_interval)
        elapsed = time.perf_counter() - started
        state = euler_z.interior(final, params.ghost, problem.nx, problem.ny)
        conserved = state.mean(axis=0)
        primitive = euler_z.primitive(conserved)
        x = problem.x_min + (np.arange(problem.nx) + 0.5) * dx
        exact = exact_conserved(problem, x, dx)
        exact_primitive = euler_z.primitive(exact)
        delta = primitive[:, 0] - exact_primitive[:, 0]
        power = 2 if order == 5 else 3
        key = f"weno{order}_z_p{power}"
        row = {
            "problem": problem.key, "method": key, "nx": problem.nx, "ny": problem.ny, "cfl": problem.cfl,
            "t_end": problem.t_end, "steps": int(summary["steps"]), "t": float(summary["t"]),
            "rho_l1": float(np.mean(np.abs(delta))), "rho_l2": float(np.sqrt(np.mean(delta * delta))),
            "rho_linf": float(np.max(np.abs(delta))), "rho_min": float(primitive[:, 0].min()),
            "p_min": float(primitive[:, 3].min()), "nan_count": int(np.isnan(state).sum()), "seconds": elapsed,
            "weno_z_p": power,
            "weno_z_tau": "abs(beta0-beta2)" if order == 5 else "abs(-beta0-3*beta1+3*beta2+beta3)",
            "weno_z_epsilon": float(dx ** power), "weno_z_epsilon_convention": f"paper epsilon=dx^{power}",
        }
        if write:
            np.savez(out / f"{key}.npz", x=x, state=state, conserved_1d=conserved, rho=primitive[:, 0],
                     velocity=primitive[:, 1], pressure=primitive[:, 3], exact_conserved=exact,
                     metadata_json=np.array(json.dumps(row, sort_keys=True)))
        row["_x"], row["_conserved_1d"] = x, conserved
        rows.append(row)
        print(json.dumps({k: v for k, v in row.items() if not k.startswith("_")}, sort_keys=True), flush=True)
    if write:
        with (out / "metrics.csv").open("w", newline="") as stream:
            fields = [k for k in rows[0] if not k.startswith("_")]
            writer = csv.DictWriter(stream, fieldnames=fields, extrasaction="ignore")
            writer.writeheader()
            writer.writerows(rows)
    return rows


def main() -> None:
    parser = argparse.ArgumentParser()
    parser.add_argument("--problem", choices=tuple(PROBLEMS), required=True)
    parser.add_argument("--device", default="cuda")
    parser.add_argument("--report-interval", type=int, default=0)
    parser.add_argument("--cfl", type=float, default=None)
    parser.add_argument("--nx", type=int, default=None)
    parser.add_argument("--ny", type=int, default=None)
    parser.add_argument("--out-tag", default=None)
    args = parser.parse_args()
    run(args.problem, args.device, args.report_interval, args.cfl, args.nx, args.ny, args.out_tag)


if __name__ == "__main__":
    main()
