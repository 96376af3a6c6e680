"""GSTE long-time advection rows (code/snapshot/for_paper_results/run_gste.py:38-118,
179-264: ``cell_averages``, ``integrate_ssprk3``, ``run_torch_methods``) for the WENO5
methods the 1-D HIP rollout provides (WENO5-JS, WENO5-SR FP64 / FP32).

Each row: steps, dt, CFL, t, L1 / L2 / Linf against the exact translated cell averages,
total variation, extrema; the final profile goes to ``<method>.npz``.
"""
from __future__ import annotations

import argparse
import json
from pathlib import Path

import numpy as np

from .. import checkpoints
from ..advection1d import integrate_ssprk3
from ..initial import gste_cell_averages
from . import config

GSTE_METHODS = ("weno5_js", "weno5_sr_f64", "weno5_sr_f32")


def cell_averages(nx: int, t: float, quadrature: int = 15) -> tuple[np.ndarray, np.ndarray]:
    return gste_cell_averages(nx, t, quadrature)


def model_for(method: str, device: str = "cuda"):
    if method == "weno5_js":
        return "classical"
    if method == "weno5_sr_f64":
        return checkpoints.load_mlp_params(config.METHODS[method].model, device, "fp64")
    if method == "weno5_sr_f32":
        return checkpoints.load_mlp_params(config.METHODS[method].model, device, "fp32")
    raise ValueError(f"the 1-D HIP rollout provides {GSTE_METHODS}, not {method!r}")


def run_methods(nx: int = 200, t_end: float = 10.0, device: str = "cuda", out_dir: Path | None = None,
                quadrature: int = 15, methods=GSTE_METHODS, weno5_cfl: float = 0.6) -> list[dict]:
    x, u0 = cell_averages(nx, 0.0, quadrature)
    _, exact = cell_averages(nx, t_end, quadrature)
    rows = []
    for key in methods:
        model = model_for(key, device)
        final, meta = integrate_ssprk3(model, u0, weno5_cfl, t_end, device)
        diff = final - exact
        row = {
            "method": key, **meta,
            "l1": float(np.mean(np.abs(diff))), "l2": float(np.sqrt(np.mean(diff * diff))),
            "linf": float(np.max(np.abs(diff))), "tv": float(np.sum(np.abs(final - np.roll(final, 1)))),
            "min": float(np.min(final)), "max": float(np.max(final)),
            "complete": bool(np.all(np.isfinite(final))), "eno_cutoff": False, "eno_cutoff_threshold": None,
            "model": str(config.METHODS[key].model) if config.METHODS[key].model else None,
        }
        if out_dir is not None:
            Path(out_dir).mkdir(parents=True, exist_ok=True)
            np.savez(Path(out_dir) / f"{key}.npz", x=x, initial=u0, final=final, exact=exact,
                     metadata_json=np.array(json.dumps(row, sort_keys=True)))
        rows.append(row)
        print(row, flush=True)
    return rows


def main() -> None:
    parser = argparse.ArgumentParser()
    parser.add_argument("--nx", type=int, default=200)
    parser.add_argument("--t-end", type=float, default=10.0)
    parser.add_argument("--device", default="cuda")
    parser.add_argument("--quadrature", type=int, default=15)
    parser.add_argument("--methods", default=",".join(GSTE_METHODS))
    parser.add_argument("--weno5-cfl", type=float, default=0.6)
    args = parser.parse_args()
    config.ensure_output_dirs()
    run_methods(args.nx, args.t_end, args.device, config.RAW / "gste", args.quadrature,
                [m.strip() for m in args.methods.split(",") if m.strip()], args.weno5_cfl)


if __name__ == "__main__":
    main()
