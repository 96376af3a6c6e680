"""Method dispatch of code/snapshot/for_paper_results/solvers/euler_methods.py:27-96
(WENO5 methods on the 2-D WENO5 step, WENO7 methods on the WENO7 + Shu RK4 step)."""
from __future__ import annotations

from typing import Any

import numpy as np

from .. import checkpoints
from ..params import Params
from . import weno5_hllc, weno5_hllc_mixed, weno5_hllc_refsym

WENO5_METHODS = ("weno5_js", "weno5_sr_f64", "weno5_sr_f32", "weno5_sr_bf16", "weno5_z")
WENO7_METHODS = ("weno7_js", "weno7_sr_f64")


def make_weno5_params(nx: int, ny: int, x_length: float, y_length: float, cfl: float, t_end: float,
                      mixed: bool = False) -> Params:
    del mixed  # the reference's mixed Params class is identical
    return Params(nx=nx, ny=ny, x_length=x_length, y_length=y_length, cfl=cfl, t_end=t_end)


def make_weno7_params(nx: int, ny: int, x_min: float, x_max: float, y_min: float, y_max: float, cfl: float,
                      t_end: float):
    from ..euler7 import Params7
    return Params7(nx=nx, ny=ny, x_min=x_min, x_max=x_max, y_min=y_min, y_max=y_max, cfl=cfl, t_end=t_end)


def run_weno7(method_key: str, initial: np.ndarray, params, *, device: str, boundary: str,
              report_interval: int = 0) -> tuple[np.ndarray, dict[str, Any]]:
    """euler_methods.py:81-96: WENO7-JS or WENO7-SR (FP64 closure), HLLC, characteristic."""
    from . import point_rk4, point_rk4_mlp
    if method_key == "weno7_js":
        return point_rk4.run_from_initial(initial, params, device=device, riemann_solver="hllc",
                                          characteristic=True, boundary=boundary, report_interval=report_interval)
    if method_key == "weno7_sr_f64":
        beta = point_rk4_mlp.TorchWeno7PointBeta(checkpoints.WENO7_SR_FP64, device, params.gamma)
        return point_rk4_mlp.run_from_initial_mlp(initial, params, device=device, riemann_solver="hllc",
                                                  beta_model=beta, boundary=boundary, eno_cutoff=False,
                                                  report_interval=report_interval)
    raise ValueError(f"not a WENO7 paper method: {method_key}")


def run_weno5(method_key: str, initial: np.ndarray, params: Params, *, device: str, boundary: str,
              report_interval: int = 0) -> tuple[np.ndarray, dict[str, Any]]:
    from . import _adapter
    if method_key == "weno5_js":
        final, dts, steps, t = weno5_hllc.run_to_time(initial, params, params.t_end, device, None, boundary,
                                                      report_interval=report_interval,
                                                      report=_printer(method_key) if report_interval else None)
    elif method_key == "weno5_sr_f64":
        mlp = checkpoints.load_f64(device=device)
        final, dts, steps, t = weno5_hllc_refsym.run_to_time(
            initial, params, params.t_end, device, mlp, boundary, report_interval=report_interval,
            report=_printer(method_key) if report_interval else None)
    elif method_key in ("weno5_sr_f32", "weno5_sr_bf16"):
        mlp = checkpoints.load_mlp_params(checkpoints.WENO5_SR_FP32, device,
                                          "fp32" if method_key == "weno5_sr_f32" else "bf16")
        final, dts, steps, t = weno5_hllc_mixed.run_to_time(
            initial, params, params.t_end, device, mlp, boundary, report_interval=report_interval,
            report=_printer(method_key) if report_interval else None)
    elif method_key == "weno5_z":
        final, dts, steps, t = _adapter.run_to_time(initial, params, params.t_end, device, None, boundary, "z",
                                                    report_interval,
                                                    _printer(method_key) if report_interval else None)
    else:
        raise ValueError(f"not a WENO5 paper method: {method_key}")
    return final, {
        "method": method_key,
        "riemann_solver": "hllc",
        "weno_space": "characteristic",
        "eno_cutoff": False,
        "steps": steps,
        "t": t,
        "dt_min": min(dts) if dts else 0.0,
        "dt_max": max(dts) if dts else 0.0,
        "dt_mean": float(np.mean(dts)) if dts else 0.0,
    }


def _printer(method_key: str):
    def report(step: int, t: float, dt: float, stats: dict) -> None:
        print(f"{method_key} step={step:05d} t={t:.8e} dt={dt:.4e} "
              f"rho=[{stats['rho_min']:.5e},{stats['rho_max']:.5e}] "
              f"p=[{stats['p_min']:.5e},{stats['p_max']:.5e}] nan={int(stats['nan_count'])}", flush=True)
    return report
