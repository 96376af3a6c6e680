"""Periodic isentropic-vortex convergence study
(code/snapshot/for_paper_results/run_vortex.py:17-169, and the WENO-Z rows of
weno_z_borges_p2_results/run_vortex.py:45-140).

Grid N x N on [-10, 10]^2, 15 x 15 Gauss-Legendre cell averages on the ghost-4 WENO7
grid; WENO5 methods take the interior re-padded periodically to ghost 3.  Every method
runs to t = 2 with periodic ghost cells; rows carry the density errors against the exact
cell averages at t = 2 (vortex centre moved by (2, 2)) and the observed orders.
"""
from __future__ import annotations

import argparse
import csv
import json

import numpy as np

from ..common import interior, rho_errors, state_health
from ..initial import pad_periodic
from ..solvers import euler_methods, euler_z, point_rk4
from . import config
from ._io import write_json

Z_METHODS = ("weno5_z_p2", "weno7_z_p3")
ALL_METHODS = config.EULER_METHODS + Z_METHODS
INTEGRATOR = {"weno5_z_p2": "SSPRK3", "weno7_z_p3": "four-stage fourth-order downwind TVD-RK"}


def add_orders(rows: list[dict]) -> None:
    """Observed order between consecutive grids of one method (run_vortex.py:37-50)."""
    for method in {row["method"] for row in rows}:
        subset = sorted((row for row in rows if row["method"] == method), key=lambda row: row["N"])
        for index, row in enumerate(subset):
            for norm in ("rho_l1", "rho_l2"):
                if index == 0 or row[norm] <= 0.0 or subset[index - 1][norm] <= 0.0:
                    row[f"{norm}_order"] = float("nan")
                else:
                    prev = subset[index - 1]
                    row[f"{norm}_order"] = float(np.log(prev[norm] / row[norm]) / np.log(row["N"] / prev["N"]))


def run_one(key: str, n: int, cfl: float = 0.4, t_end: float = 2.0, quadrature: int = 15, device: str = "cuda",
            report_interval: int = 0) -> tuple[dict, np.ndarray, np.ndarray]:
    """(row, final interior state, exact interior state) of one method on one grid."""
    p7 = euler_methods.make_weno7_params(n, n, -10.0, 10.0, -10.0, 10.0, cfl, t_end)
    initial7 = point_rk4.cell_average_state(p7, 0.0, quadrature)
    exact_interior = interior(point_rk4.cell_average_state(p7, t_end, quadrature), p7.ghost, n, n)
    if key.startswith("weno5"):
        params = euler_methods.make_weno5_params(n, n, 20.0, 20.0, cfl, t_end, key == "weno5_sr_f32")
        initial = pad_periodic(interior(initial7, p7.ghost, n, n), params.ghost)
        if key == "weno5_z_p2":
            final, summary = euler_z.run_weno5(initial, params, device=device, boundary="periodic",
                                               report_interval=report_interval)
        else:
            final, summary = euler_methods.run_weno5(key, initial, params, device=device, boundary="periodic",
                                                     report_interval=report_interval)
    else:
        params = p7
        if key == "weno7_z_p3":
            final, summary = euler_z.run_weno7(initial7.copy(), params, device=device, boundary="periodic",
                                               report_interval=report_interval, weight_kind=1)
        else:
            final, summary = euler_methods.run_weno7(key, initial7.copy(), params, device=device,
                                                     boundary="periodic", report_interval=report_interval)
    state = interior(final, params.ghost, n, n)
    health = state_health(final, params.ghost, n, n)
    complete_time = abs(float(summary["t"]) - t_end) < 1.0e-12
    integrator = config.METHODS[key].time_integrator if key in config.METHODS else INTEGRATOR[key]
    row = {
        "method": key, "N": n, "dx": 20.0 / n, "domain": "[-10,10]^2", "cfl": cfl, "t_end": t_end,
        "quadrature": quadrature, "initialization": "15x15_Gauss_finite_volume_cell_average",
        "exact_solution": "periodic_isentropic_vortex_center_(t,t)", "boundary": "periodic",
        "riemann_solver": "hllc", "weno_space": "characteristic", "eno_cutoff": False,
        "time_integrator": integrator, **rho_errors(state[..., 0], exact_interior[..., 0]), **health,
        "complete_time": complete_time, "steps": int(summary["steps"]), "t": float(summary["t"]),
    }
    row["complete"] = bool(row["complete"] and complete_time)
    return row, state, exact_interior


def run(grids=(25, 50, 100, 200), cfl: float = 0.4, t_end: float = 2.0, device: str = "cuda",
        quadrature: int = 15, report_interval: int = 0, out_tag: str = "vortex", methods=config.EULER_METHODS,
        write: bool = True) -> list[dict]:
    if quadrature != 15:
        raise ValueError("formal vortex tests require 15x15 Gauss-Legendre averages")
    if t_end != 2.0:
        raise ValueError("formal vortex tests require t_end=2")
    if cfl <= 0.0:
        raise ValueError("CFL must be positive")
    unknown = [m for m in methods if m not in ALL_METHODS]
    if unknown:
        raise ValueError(f"unknown method(s): {unknown}")
    config.validate_models()
    out_dir = config.RAW / out_tag
    if write:
        out_dir.mkdir(parents=True, exist_ok=True)
    rows: list[dict] = []
    for n in grids:
        for key in methods:
            row, state, exact = run_one(key, n, cfl, t_end, quadrature, device, report_interval)
            if write:
                np.savez(out_dir / f"{key}_N{n}.npz", state=state, exact=exact,
                         metadata_json=np.array(json.dumps(row, sort_keys=True)))
                write_json(out_dir / f"{key}_N{n}.json", row)
            rows.append(row)
            print(row, flush=True)
    order = {key: i for i, key in enumerate(ALL_METHODS)}
    rows.sort(key=lambda row: (order[row["method"]], int(row["N"])))
    add_orders(rows)
    if write:
        fields = ["method", "N", "dx", "cfl", "t_end", "quadrature", "complete", "steps", "t", "rho_l1",
                  "rho_l1_order", "rho_l2", "rho_l2_order", "rho_linf", "rho_min", "rho_max", "p_min", "p_max",
                  "nan_count"]
        with (out_dir / "metrics.csv").open("w", newline="") as stream:
            writer = csv.DictWriter(stream, fieldnames=fields, extrasaction="ignore")
            writer.writeheader()
            writer.writerows(rows)
        write_json(out_dir / "summary.json", {"configuration": {"grids": list(grids), "cfl": cfl, "t_end": t_end,
                                                                "quadrature": quadrature, "device": device},
                                              "rows": rows})
    return rows


def main() -> None:
    parser = argparse.ArgumentParser()
    parser.add_argument("--grids", default="25,50,100,200")
    parser.add_argument("--cfl", type=float, default=0.4)
    parser.add_argument("--t-end", type=float, default=2.0)
    parser.add_argument("--device", default="cuda")
    parser.add_argument("--quadrature", type=int, default=15)
    parser.add_argument("--report-interval", type=int, default=0)
    parser.add_argument("--out-tag", default="vortex")
    parser.add_argument("--methods", default=",".join(config.EULER_METHODS))
    args = parser.parse_args()
    grids = [int(v) for v in args.grids.split(",") if v.strip()]
    if not grids or any(g <= 0 for g in grids):
        raise ValueError("--grids requires positive comma-separated integers")
    config.ensure_output_dirs()
    run(grids, args.cfl, args.t_end, args.device, args.quadrature, args.report_interval, args.out_tag,
        [m.strip() for m in args.methods.split(",") if m.strip()])


if __name__ == "__main__":
    main()
