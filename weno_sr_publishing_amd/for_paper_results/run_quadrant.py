"""Four-quadrant 2-D Riemann problems C.3-C.6 for the paper methods
(code/snapshot/for_paper_results/run_quadrant.py:19-223, and the WENO7-Z run of
weno_z_borges_p2_results/run_quadrant.py:19-120).

The initial state is the 15 x 15 Gauss-Legendre cell average on the ghost-4 WENO7 grid
of the unit square; WENO5 methods take its interior re-padded periodically to ghost 3
and run with transmissive ghost cells, WENO7 methods run on it with clamped ("outflow")
ghost cells.  Rows carry the state health and step counts; states are written as
``<method>.npz`` (``state`` = interior, ``metadata_json``).
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from ..common import interior, state_health
from ..initial import CASES, make_quadrant_state as _quadrant_cells, pad_periodic
from ..solvers import euler_methods, euler_z
from . import config
from ._io import write_json

Z_METHOD = "weno7_z_p3"


def make_quadrant_state(params, definition: dict, quadrature: int, device: str | None = None) -> np.ndarray:
    """Cell averages on the padded grid of ``params`` (run_quadrant.py:87-103).  With a
    GPU ``device`` they are computed there (``initial_device.quadrant_state``, equal bit
    for bit) and returned as a host array."""
    if device is not None and torch.device(device).type == "cuda":
        from ..initial_device import quadrant_state
        state = quadrant_state(params.nx, params.ny, params.ghost, definition, quadrature, params.x_min,
                               params.x_max, params.y_min, params.y_max, params.gamma, device=device)
        return state.cpu().numpy()
    return _quadrant_cells(params.nx, params.ny, params.ghost, definition, quadrature, params.x_min,
                           params.x_max, params.y_min, params.y_max, params.gamma)


def run_method(key: str, case: str, nx: int, ny: int, cfl: float, t_end: float, device: str = "cuda",
               report_interval: int = 0, init_quadrature: int = 15,
               z7_unit_gauss_h: bool = False) -> tuple[np.ndarray, dict, float]:
    """(padded final state, solver summary, wall seconds) of one method on one case;
    ``z7_unit_gauss_h`` selects the transverse WENO7-Z weight convention of the reference's
    recorded results (see ``solvers.euler_z.run_weno7``)."""
    definition = CASES[case]
    p7 = euler_methods.make_weno7_params(nx, ny, 0.0, 1.0, 0.0, 1.0, cfl, t_end)
    initial7 = make_quadrant_state(p7, definition, init_quadrature, device)
    started = time.perf_counter()
    if key.startswith("weno5"):
        params = euler_methods.make_weno5_params(nx, ny, 1.0, 1.0, cfl, t_end, key == "weno5_sr_f32")
        initial = pad_periodic(interior(initial7, p7.ghost, nx, ny), params.ghost)
        final, summary = euler_methods.run_weno5(key, initial, params, device=device, boundary="transmissive",
                                                 report_interval=report_interval)
    elif key == Z_METHOD:
        final, summary = euler_z.run_weno7(initial7.copy(), p7, device=device, boundary="outflow",
                                           report_interval=report_interval, unit_gauss_h=z7_unit_gauss_h)
    else:
        final, summary = euler_methods.run_weno7(key, initial7.copy(), p7, device=device, boundary="outflow",
                                                 report_interval=report_interval)
    return final, summary, time.perf_counter() - started


def run(case: str, nx: int = 400, ny: int = 400, cfl: float = 0.4, t_end: float | None = None,
        init_quadrature: int = 15, methods=config.EULER_METHODS, reference_only: bool = False,
        device: str = "cuda", report_interval: int = 50, write: bool = True,
        z7_unit_gauss_h: bool = False) -> tuple[list[dict], dict]:
    """Rows (one per method) and the interior states by method."""
    if nx != ny:
        raise ValueError("paper quadrant tests require square meshes")
    if init_quadrature != 15:
        raise ValueError("paper Riemann tests require 15-point initialisation")
    unknown = [m for m in methods if m not in config.METHODS and m != Z_METHOD]
    if unknown:
        raise ValueError(f"unknown method(s): {unknown}")
    config.validate_models()
    definition = CASES[case]
    t_end = definition["t_end"] if t_end is None else t_end
    methods = ["weno7_js"] if reference_only else list(methods)
    suffix = "reference1200" if reference_only else f"N{nx}"
    out_dir = config.RAW / case / suffix
    if write:
        out_dir.mkdir(parents=True, exist_ok=True)
    rows, states, failed = [], {}, []
    for key in methods:
        final, summary, elapsed = run_method(key, case, nx, ny, cfl, t_end, device, report_interval,
                                             init_quadrature, z7_unit_gauss_h)
        ghost = 3 if key.startswith("weno5") else 4
        state = interior(final, ghost, nx, ny)
        health = state_health(final, ghost, nx, ny)
        complete_time = abs(float(summary["t"]) - t_end) < 1.0e-12
        label = config.METHODS[key].label if key in config.METHODS else "WENO7-Z-RK4"
        row = {
            "method": key, "case": case, "configuration": definition["label"], "nx": nx, "ny": ny, "cfl": cfl,
            "t_end": t_end, "t": float(summary["t"]), "steps": int(summary["steps"]),
            "wall_seconds_including_setup": elapsed, **health, "complete_time": complete_time,
            "riemann_solver": "hllc", "weno_space": "characteristic", "eno_cutoff": False, "label": label,
        }
        row["complete"] = bool(row["complete"] and complete_time)
        if write:
            np.savez(out_dir / f"{key}.npz", state=state, metadata_json=np.array(json.dumps(row, sort_keys=True)))
            write_json(out_dir / f"{key}.json", row)
        rows.append(row)
        states[key] = state
        print(row, flush=True)
        if not row["complete"]:
            failed.append(key)
    if write:
        write_json(out_dir / "summary.json", {"configuration": {"case": case, "nx": nx, "ny": ny, "cfl": cfl,
                                                                "t_end": t_end, "device": device},
                                              "rows": rows})
    if failed:
        raise RuntimeError("formal validation failed for: " + ", ".join(failed))
    return rows, states


def main() -> None:
    parser = argparse.ArgumentParser()
    parser.add_argument("--case", choices=CASES, required=True)
    parser.add_argument("--nx", type=int, default=400)
    parser.add_argument("--ny", type=int, default=400)
    parser.add_argument("--cfl", type=float, default=0.4)
    parser.add_argument("--t-end", type=float, default=None)
    parser.add_argument("--init-quadrature", type=int, default=15)
    parser.add_argument("--methods", default=",".join(config.EULER_METHODS))
    parser.add_argument("--reference-only", action="store_true")
    parser.add_argument("--device", default="cuda")
    parser.add_argument("--report-interval", type=int, default=50)
    parser.add_argument("--z7-unit-gauss-h", action="store_true",
                        help="WENO7-Z transverse weights with h = 1, as the reference's recorded results")
    args = parser.parse_args()
    config.ensure_output_dirs()
    methods = [m.strip() for m in args.methods.split(",") if m.strip()]
    run(args.case, args.nx, args.ny, args.cfl, args.t_end, args.init_quadrature, methods, args.reference_only,
        args.device, args.report_interval, z7_unit_gauss_h=args.z7_unit_gauss_h)


if __name__ == "__main__":
    main()
