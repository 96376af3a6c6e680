"""Double Mach reflection at 1200 x 300 to t = 0.2 for the paper methods
(code/snapshot/for_paper_results/run_double_mach.py:31-260; the WENO5-JS field of
run_double_mach_compare.py:253-300 and the WENO-Z fields of
weno_z_borges_p2_results/run_double_mach.py), and the density line cuts of
weno_z_borges_p2_results/plot_double_mach.py:274-320.

Every method starts from the 15 x 15 cell averages of the Mach-10 shock on
[0, 4] x [0, 1] (run_double_mach_compare.py:80-88), runs with the time-dependent
double-Mach boundary, HLLC and characteristic reconstruction: WENO5 with SSPRK3,
WENO7 with Shu RK4 and the downwind operator.
"""
from __future__ import annotations

import argparse
import json

import numpy as np
import torch

from .. import checkpoints
from ..common import interior, state_health
from ..euler7 import Euler7Solver
from ..initial import make_double_mach_state
from ..initial_device import double_mach_state
from ..solvers import euler_methods, weno5_hllc_mixed, weno5_hllc_refsym
from ..solvers.point_rk4_mlp import TorchWeno7PointBeta
from ..solvers import weno7_double_mach
from . import config
from ._io import write_json

METHODS = ("weno5_js", "weno5_z_p2", "weno5_sr_f64", "weno5_sr_f32", "weno7_js", "weno7_z_p3", "weno7_sr_f64")
CUT_Y = (0.2, 0.4, 0.6)


def _initial(nx: int, ny: int, ghost: int, device) -> np.ndarray:
    """run_double_mach_compare.py:69-88 cell averages, computed on a GPU ``device``
    (equal bit for bit to ``initial.make_double_mach_state``)."""
    if torch.device(device).type == "cuda":
        return double_mach_state(nx, ny, ghost, device=device).cpu().numpy()
    return make_double_mach_state(nx, ny, ghost=ghost)


def run_method(key: str, nx: int = 1200, ny: int = 300, cfl: float = 0.4, t_end: float = 0.2,
               device: str = "cuda", report_interval: int = 0, z7_unit_gauss_h: bool = False):
    """(padded final state, dts, steps, t, ghost) of one method."""
    if key.startswith("weno5"):
        params = euler_methods.make_weno5_params(nx, ny, 4.0, 1.0, cfl, t_end, key == "weno5_sr_f32")
        initial = _initial(nx, ny, 3, device)
        from ..solvers import _adapter
        if key == "weno5_js":
            # the reference's WENO5-JS field comes from run_double_mach_compare.py (its
            # classical WENO5 path); weno5_hllc.py itself has no double-Mach boundary
            out = _adapter.run_to_time(initial, params, t_end, device, None, "double-mach", "js", report_interval)
        elif key == "weno5_sr_f64":
            out = weno5_hllc_refsym.run_to_time(initial, params, t_end, device,
                                                checkpoints.load_mlp_params(config.METHODS[key].model, device),
                                                "double-mach", report_interval=report_interval)
        elif key == "weno5_sr_f32":
            out = weno5_hllc_mixed.run_to_time(initial, params, t_end, device,
                                               checkpoints.load_mlp_params(config.METHODS[key].model, device),
                                               "double-mach", report_interval=report_interval)
        else:
            out = _adapter.run_to_time(initial, params, t_end, device, None, "double-mach", "z", report_interval)
        final, dts, steps, t = out
        return final, dts, steps, t, 3
    params = weno7_double_mach.Params(nx=nx, ny=ny, x_min=0.0, x_max=4.0, y_min=0.0, y_max=1.0, cfl=cfl, t_end=t_end)
    initial = _initial(nx, ny, 4, device)
    if key == "weno7_z_p3":
        solver = Euler7Solver(params, "z", "hllc", "double-mach", loop="double-mach", device=device,
                              unit_gauss_h=z7_unit_gauss_h)
        u = solver.new_state(initial)
        dts, steps, t = solver.run(u, t_end, report_interval=report_interval, check_health=True)
        return u.cpu().numpy(), dts, steps, t, 4
    beta = None if key == "weno7_js" else TorchWeno7PointBeta(config.METHODS[key].model, device, params.gamma)
    final, dts, steps, t = weno7_double_mach.run_to_time(initial, params, t_end, device, beta,
                                                         report_interval=report_interval)
    return final, dts, steps, t, 4


def run(methods=METHODS, nx: int = 1200, ny: int = 300, cfl: float = 0.4, t_end: float = 0.2,
        device: str = "cuda", report_interval: int = 0, write: bool = True,
        z7_unit_gauss_h: bool = False) -> tuple[list[dict], dict]:
    """Rows and interior states by method."""
    unknown = [m for m in methods if m not in METHODS]
    if unknown:
        raise ValueError(f"unsupported double-Mach method(s): {sorted(unknown)}")
    config.validate_models()
    out_dir = config.RAW / "double_mach"
    rows, states = [], {}
    for key in methods:
        final, dts, steps, t, ghost = run_method(key, nx, ny, cfl, t_end, device, report_interval, z7_unit_gauss_h)
        state = interior(final, ghost, nx, ny)
        health = state_health(final, ghost, nx, ny)
        complete_time = abs(float(t) - t_end) < 1.0e-12
        row = {"method": key, "nx": nx, "ny": ny, "cfl": cfl, "t_end": t_end, "t": float(t), "steps": int(steps),
               "dt_min": float(min(dts)) if dts else 0.0, "dt_max": float(max(dts)) if dts else 0.0,
               "dt_mean": float(np.mean(dts)) if dts else 0.0, "complete_time": complete_time,
               "riemann_solver": "hllc", "weno_space": "characteristic", "eno_cutoff": False,
               "time_integrator": "SSP-RK4" if key.startswith("weno7") else "SSPRK3", **health}
        row["complete"] = bool(row["complete"] and complete_time)
        if write:
            out_dir.mkdir(parents=True, exist_ok=True)
            np.savez(out_dir / f"{key}.npz", state=state, metadata_json=np.array(json.dumps(row, sort_keys=True)))
            write_json(out_dir / f"{key}.json", row)
        rows.append(row)
        states[key] = state
        print(row, flush=True)
    return rows, states


def linecut_rows(states: dict[str, np.ndarray], reference_method: str = "weno7_js") -> list[dict]:
    """Density along the rows nearest y = 0.2, 0.4, 0.6 against the WENO7-JS field."""
    ny = next(iter(states.values())).shape[0]
    y = (np.arange(ny) + 0.5) / ny
    reference = states[reference_method][..., 0]
    rows = []
    for requested in CUT_Y:
        index = int(np.argmin(np.abs(y - requested)))
        for method in METHODS:
            if method not in states:
                continue
            error = states[method][index, :, 0] - reference[index]
            rows.append({"requested_y": requested, "actual_y": float(y[index]), "method": method,
                         "status": "complete", "l1_vs_weno7_js": float(np.mean(np.abs(error))),
                         "l2_vs_weno7_js": float(np.sqrt(np.mean(error * error))),
                         "linf_vs_weno7_js": float(np.max(np.abs(error)))})
    return rows


def main() -> None:
    parser = argparse.ArgumentParser()
    parser.add_argument("--nx", type=int, default=1200)
    parser.add_argument("--ny", type=int, default=300)
    parser.add_argument("--cfl", type=float, default=0.4)
    parser.add_argument("--t-end", type=float, default=0.2)
    parser.add_argument("--init-quadrature", type=int, default=15)
    parser.add_argument("--device", default="cuda")
    parser.add_argument("--report-interval", type=int, default=100)
    # run_double_mach.py:63-68: both providers evaluate the same FP64 checkpoint (here the
    # same HIP closure), so the choice does not change the run
    parser.add_argument("--weno7-beta-backend", choices=("warp", "torch"), default="warp")
    parser.add_argument("--methods", default="weno5_sr_f64,weno5_sr_f32,weno7_js,weno7_sr_f64")
    parser.add_argument("--z7-unit-gauss-h", action="store_true")
    args = parser.parse_args()
    if args.init_quadrature != 15:
        raise ValueError("paper double-Mach tests require 15-point initialisation")
    if (args.nx, args.ny) != (1200, 300):
        raise ValueError("paper double-Mach comparison is fixed at 1200x300")
    config.ensure_output_dirs()
    run([m.strip() for m in args.methods.split(",") if m.strip()], args.nx, args.ny, args.cfl, args.t_end,
        args.device, args.report_interval, z7_unit_gauss_h=args.z7_unit_gauss_h)


if __name__ == "__main__":
    main()
