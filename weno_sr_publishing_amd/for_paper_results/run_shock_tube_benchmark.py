"""Lax and left-half Woodward-Colella shock tubes
(code/snapshot/for_paper_results/run_shock_tube_benchmark.py:20-252).

The initial state is the exact finite-volume average of the two constant states,
including the cut interface cell; every method runs with transmissive ghost cells,
characteristic reconstruction and HLLC.  The result keeps the transverse mean of the
conserved state (``conserved_1d``) and its primitives.
"""
from __future__ import annotations

import argparse
import json
import time
from dataclasses import dataclass
from pathlib import Path
from typing import NamedTuple

import numpy as np

from ..common import interior, state_health
from ..solvers import euler_methods
from . import config
from ._io import write_json


@dataclass(frozen=True)
class ShockTube:
    key: str
    title: str
    x_min: float
    x_max: float
    discontinuity: float
    t_end: float
    left: tuple[float, float, float]
    right: tuple[float, float, float]


BENCHMARKS = {
    "lax": ShockTube("lax", "Lax shock tube", -5.0, 5.0, 0.0, 1.3, (0.445, 0.698, 3.528), (0.500, 0.000, 0.571)),
    "woodward_colella_half": ShockTube("woodward_colella_half", "Left half of the Woodward--Colella blast wave",
                                       0.0, 1.0, 0.5, 0.012, (1.0, 0.0, 1000.0), (1.0, 0.0, 0.01)),
}


class ShockTubeResult(NamedTuple):
    metadata: dict
    x: np.ndarray
    conserved_1d: np.ndarray
    state: np.ndarray


def _conserved(rho: float, u: float, p: float, gamma: float) -> np.ndarray:
    # warp_weno5_helpers.py:90-92 with zero transverse velocity
    energy = 0.5 * rho * (u * u + 0.0 * 0.0) + p / (gamma - 1.0)
    return np.array([rho, rho * u, rho * 0.0, energy], dtype=np.float64)


def make_initial_state(params, problem: ShockTube) -> np.ndarray:
    """Length-weighted average of the two states in every cell (run_shock_tube_benchmark.py:57-78)."""
    g = params.ghost
    nx_total, ny_total = params.nx + 2 * g, params.ny + 2 * g
    dx = (problem.x_max - problem.x_min) / float(params.nx)
    x_left = problem.x_min + (np.arange(nx_total, dtype=np.float64) - g) * dx
    x_right = x_left + dx
    left_length = np.clip(np.minimum(x_right, problem.discontinuity) - x_left, 0.0, dx)
    right_length = dx - left_length
    cl = _conserved(*problem.left, params.gamma)
    cr = _conserved(*problem.right, params.gamma)
    cell = (left_length[:, None] * cl[None, :] + right_length[:, None] * cr[None, :]) / dx
    return np.broadcast_to(cell[None, :, :], (ny_total, nx_total, 4)).copy()


def primitive_1d(conserved: np.ndarray, gamma: float) -> tuple[np.ndarray, ...]:
    """(rho, u, p) of a transverse-mean state; raises instead of repairing an invalid state."""
    rho = conserved[:, 0]
    if not np.all(np.isfinite(conserved)) or np.any(rho <= 0.0):
        raise RuntimeError("invalid conserved state; refusing diagnostic state repair")
    velocity = conserved[:, 1] / rho
    transverse = conserved[:, 2] / rho
    pressure = (gamma - 1.0) * (conserved[:, 3] - 0.5 * rho * (velocity * velocity + transverse * transverse))
    if np.any(pressure <= 0.0) or not np.all(np.isfinite(pressure)):
        raise RuntimeError("invalid pressure; refusing diagnostic state repair")
    return rho, velocity, pressure


def make_params(problem: ShockTube, method: str, nx: int, ny: int, cfl: float):
    x_length = problem.x_max - problem.x_min
    dx = x_length / float(nx)
    y_length = ny * dx
    if config.METHODS[method].family.startswith("weno5"):
        return euler_methods.make_weno5_params(nx, ny, x_length, y_length, cfl, problem.t_end,
                                               mixed=method == "weno5_sr_f32")
    return euler_methods.make_weno7_params(nx, ny, problem.x_min, problem.x_max, 0.0, y_length, cfl, problem.t_end)


def run(benchmark: str, method: str, nx: int = 200, ny: int = 10, cfl: float = 0.8, device: str = "cuda",
        report_interval: int = 100, out_dir: Path | None = None, write: bool = True) -> "ShockTubeResult":
    """One method on one shock tube; writes ``<method>.npz`` / ``.json`` like the reference and
    returns the metadata with the cell centres and the transverse-mean conserved state."""
    if nx <= 0 or ny < 10:
        raise ValueError("shock-tube tests require nx > 0 and ny >= 10")
    if not 0.0 < cfl <= 0.8:
        raise ValueError(f"shock-tube CFL must lie in (0, 0.8]; got {cfl}")
    if method not in config.METHODS:
        raise ValueError(f"unknown method {method!r}")
    config.validate_models()
    problem = BENCHMARKS[benchmark]
    meth = config.METHODS[method]
    dx = (problem.x_max - problem.x_min) / float(nx)
    params = make_params(problem, method, nx, ny, cfl)
    out_dir = Path(out_dir) if out_dir else config.RAW / "shock_tubes_cfl08" / problem.key / f"N{nx}x{ny}"
    initial = make_initial_state(params, problem)
    print(f"shock_tube_start benchmark={problem.key} method={method} grid={nx}x{ny} "
          f"domain=[{problem.x_min},{problem.x_max}] x0={problem.discontinuity} cfl={cfl} t_end={problem.t_end} "
          f"integrator={meth.time_integrator} space=characteristic solver=hllc boundary=transmissive cutoff=False",
          flush=True)
    started = time.perf_counter()
    run_fn = euler_methods.run_weno5 if meth.family.startswith("weno5") else euler_methods.run_weno7
    final, summary = run_fn(method, initial, params, device=device, boundary="transmissive",
                            report_interval=report_interval)
    solve_seconds = time.perf_counter() - started
    health = state_health(final, params.ghost, nx, ny)
    complete_time = abs(float(summary["t"]) - problem.t_end) < 1.0e-12
    state = interior(final, params.ghost, nx, ny)
    conserved_1d = state.mean(axis=0)
    diagnostic_error = None
    try:
        rho, velocity, pressure = primitive_1d(conserved_1d, params.gamma)
    except RuntimeError as exc:
        diagnostic_error = str(exc)
        rho = velocity = pressure = None
    x = problem.x_min + (np.arange(nx, dtype=np.float64) + 0.5) * dx
    metadata = {
        "benchmark": problem.key, "title": problem.title, "method": method, "label": meth.label,
        "model": str(meth.model) if meth.model else None, "nx": nx, "ny": ny,
        "x_min": problem.x_min, "x_max": problem.x_max, "discontinuity": problem.discontinuity,
        "left_primitive": list(problem.left), "right_primitive": list(problem.right), "dx": dx, "cfl": cfl,
        "t_end": problem.t_end, "t": float(summary["t"]), "steps": int(summary["steps"]),
        "dt_min": float(summary["dt_min"]), "dt_max": float(summary["dt_max"]), "dt_mean": float(summary["dt_mean"]),
        "initialization": "exact_finite_volume_cell_average", "riemann_solver": "hllc",
        "weno_space": "characteristic", "time_integrator": meth.time_integrator, "boundary": "transmissive",
        "eno_cutoff": False, "reflection_symmetrized_mlp": "_sr_" in method, "state_repair": False,
        "solve_seconds": solve_seconds, "complete_time": complete_time, "diagnostic_error": diagnostic_error,
        **health,
    }
    metadata["complete"] = bool(metadata["complete"] and complete_time and diagnostic_error is None)
    if diagnostic_error is not None:
        if write:
            write_json(out_dir / f"{method}.failed.json", metadata)
        raise RuntimeError(f"shock-tube validation failed without state repair: {metadata}")
    if write:
        out_dir.mkdir(parents=True, exist_ok=True)
        np.savez(out_dir / f"{method}.npz", state=state, conserved_1d=conserved_1d, rho=rho, velocity=velocity,
                 pressure=pressure, x=x, metadata_json=np.array(json.dumps(metadata, sort_keys=True)))
        write_json(out_dir / f"{method}.json", metadata)
    if not metadata["complete"]:
        raise RuntimeError(f"shock-tube validation failed: {metadata}")
    return ShockTubeResult(metadata, x, conserved_1d, state)


def main() -> None:
    parser = argparse.ArgumentParser()
    parser.add_argument("--benchmark", choices=tuple(BENCHMARKS), required=True)
    parser.add_argument("--method", choices=config.EULER_METHODS, required=True)
    parser.add_argument("--nx", type=int, default=200)
    parser.add_argument("--ny", type=int, default=10)
    parser.add_argument("--cfl", type=float, default=0.8)
    parser.add_argument("--device", default="cuda")
    parser.add_argument("--report-interval", type=int, default=100)
    parser.add_argument("--out-dir", type=Path, default=None)
    args = parser.parse_args()
    config.ensure_output_dirs()
    run(args.benchmark, args.method, args.nx, args.ny, args.cfl, args.device, args.report_interval, args.out_dir)


if __name__ == "__main__":
    main()
