"""Sod shock tube for the five paper methods (code/snapshot/for_paper_results/run_sod.py:18-136).

WENO5 methods start from the Sod strip of warp_sod_validation.py (centred on 0 with
``x_length = 1``), WENO7 methods from the 15-point Gauss-Legendre cell averages on the
ghost-4 grid of [0, 1] with the discontinuity at 0.5; both run with clamped
(transmissive / "outflow") ghost cells.  Each row carries the density errors of the
transverse-mean density against the exact finite-volume averages.
"""
from __future__ import annotations

import argparse
import csv
import json

import numpy as np

from .. import exact_riemann
from ..common import interior, rho_errors, state_health
from ..initial import make_sod_state
from ..solvers import euler_methods
from . import config
from ._io import write_json

GAMMA = 1.4


def primitive_to_conserved_1d(rho, velocity, pressure, gamma: float) -> np.ndarray:
    return np.stack([rho, rho * velocity, pressure / (gamma - 1.0) + 0.5 * rho * velocity * velocity], axis=-1)


def exact_sod_cell_averages(centers: np.ndarray, dx: float, t: float, gamma: float, quadrature: int,
                            discontinuity: float = 0.5) -> np.ndarray:
    """(rho, rho u, E) averages of the exact Sod solution by unsplit Gauss-Legendre (run_sod.py:18-33)."""
    xi, wi = np.polynomial.legendre.leggauss(quadrature)
    state = np.zeros((centers.size, 3), dtype=np.float64)
    for node, weight in zip(xi, wi):
        xq = centers + 0.5 * dx * float(node) - discontinuity
        rho, velocity, pressure = exact_riemann.exact_sod_primitive(xq, t, gamma)
        state += 0.5 * float(weight) * primitive_to_conserved_1d(rho, velocity, pressure, gamma)
    return state


def make_weno7_sod_state(params, quadrature: int, discontinuity: float = 0.5) -> np.ndarray:
    """Ghost-4 WENO7 initial state, uniform in y (run_sod.py:36-47)."""
    g = params.ghost
    centers = params.x_min + (np.arange(params.nx + 2 * g) - g + 0.5) * params.dx
    state1d = exact_sod_cell_averages(centers, params.dx, 0.0, params.gamma, quadrature, discontinuity)
    state = np.zeros((params.ny + 2 * g, params.nx + 2 * g, 4), dtype=np.float64)
    state[..., 0] = state1d[None, :, 0]
    state[..., 1] = state1d[None, :, 1]
    state[..., 3] = state1d[None, :, 2]
    return state


def initial_state(method: str, nx: int, ny: int, cfl: float, t_end: float, quadrature: int = 15):
    """(params, padded initial state, boundary) of one method (run_sod.py:79-102)."""
    if method.startswith("weno5"):
        params = euler_methods.make_weno5_params(nx, ny, 1.0, ny / nx, cfl, t_end, method == "weno5_sr_f32")
        return params, make_sod_state(params, "x"), "transmissive"
    params = euler_methods.make_weno7_params(nx, ny, 0.0, 1.0, 0.0, ny / nx, cfl, t_end)
    return params, make_weno7_sod_state(params, quadrature, 0.5), "outflow"


def run(nx: int = 100, ny: int = 10, cfl: float = 0.4, t_end: float = 0.25, device: str = "cuda",
        report_interval: int = 0, init_quadrature: int = 15, out_tag: str | None = None,
        methods=config.EULER_METHODS, write: bool = True) -> list[dict]:
    if init_quadrature != 15:
        raise ValueError("paper Sod tests require 15-point Gauss-Legendre initialisation")
    config.validate_models()
    out_dir = config.RAW / "sod" / out_tag if out_tag else config.RAW / "sod"
    if write:
        out_dir.mkdir(parents=True, exist_ok=True)
    discontinuity = 0.5
    dx = 1.0 / nx
    x = (np.arange(nx) + 0.5) * dx
    exact_rho = exact_sod_cell_averages(x, dx, t_end, GAMMA, init_quadrature, discontinuity)[:, 0]
    rows: list[dict] = []
    for key in methods:
        params, initial, boundary = initial_state(key, nx, ny, cfl, t_end, init_quadrature)
        run_fn = euler_methods.run_weno5 if key.startswith("weno5") else euler_methods.run_weno7
        final, summary = run_fn(key, initial, params, device=device, boundary=boundary,
                                report_interval=report_interval)
        state = interior(final, params.ghost, nx, ny)
        rho = state[..., 0].mean(axis=0)
        health = state_health(final, params.ghost, nx, ny)
        complete_time = abs(float(summary["t"]) - t_end) < 1.0e-12
        row = {
            "method": key, "nx": nx, "ny": ny, "init_quadrature": init_quadrature,
            "x_min": 0.0, "x_max": 1.0, "discontinuity": discontinuity,
            **rho_errors(rho, exact_rho), **health,
            "complete_time": complete_time, "steps": int(summary["steps"]), "t": float(summary["t"]),
        }
        row["complete"] = bool(row["complete"] and complete_time)
        if write:
            np.savez(out_dir / f"{key}.npz", x=x, rho=rho, exact_rho=exact_rho, state=state,
                     metadata_json=np.array(json.dumps(row, sort_keys=True)))
            write_json(out_dir / f"{key}.json", row)
        rows.append(row)
        print(row, flush=True)
    if write:
        fields = ["method", "nx", "ny", "init_quadrature", "complete", "steps", "t", "rho_l1", "rho_l2", "rho_linf",
                  "rho_min", "rho_max", "p_min", "p_max", "nan_count"]
        with (out_dir / "metrics.csv").open("w", newline="") as stream:
            writer = csv.DictWriter(stream, fieldnames=fields, extrasaction="ignore")
            writer.writeheader()
            writer.writerows(rows)
        write_json(out_dir / "summary.json", {"configuration": {"nx": nx, "ny": ny, "cfl": cfl, "t_end": t_end,
                                                                "device": device, "init_quadrature": init_quadrature,
                                                                "out_tag": out_tag}, "rows": rows})
    return rows


def main() -> None:
    parser = argparse.ArgumentParser()
    parser.add_argument("--nx", type=int, default=100)
    parser.add_argument("--ny", type=int, default=10)
    parser.add_argument("--cfl", type=float, default=0.4)
    parser.add_argument("--t-end", type=float, default=0.25)
    parser.add_argument("--device", default="cuda")
    parser.add_argument("--report-interval", type=int, default=0)
    parser.add_argument("--init-quadrature", type=int, default=15)
    parser.add_argument("--out-tag", default=None)
    args = parser.parse_args()
    config.ensure_output_dirs()
    run(args.nx, args.ny, args.cfl, args.t_end, args.device, args.report_interval, args.init_quadrature, args.out_tag)


if __name__ == "__main__":
    main()
