"""Wall time of the WENO5 paths on C.4 400 x 400 (code/snapshot/for_paper_results/run_weno5_timing.py:20-182).

Same protocol as the reference: one untimed warm-up trajectory, then ``repeats`` timed
calls of the drop-in ``run_to_time`` adapters (host state in, host state out, device time
loop), each checked for a complete trajectory; rows carry mean / std seconds, cost
relative to WENO5-JS and the closure overhead, and a LaTeX table like the reference's.
"""
from __future__ import annotations

import argparse
import statistics
import time

from .. import checkpoints
from ..common import interior, state_health
from ..initial import CASES, pad_periodic
from ..solvers import euler_methods, weno5_hllc, weno5_hllc_mixed, weno5_hllc_refsym
from . import config
from ._io import write_json
from .run_quadrant import make_quadrant_state

VARIANTS = (
    ("weno5_js", "WENO5-JS", False, weno5_hllc, None),
    ("weno5_sr_f64", "FP64 MLP", False, weno5_hllc_refsym, "fp64"),
    ("weno5_sr_f32", "FP32 MLP", True, weno5_hllc_mixed, "fp32"),
)


def latex_table(rows: list[dict]) -> str:
    lines = [
        "\\begin{table}[t]", "\\centering",
        "\\caption{WENO5 wall time on the $400^2$ Riemann configuration C.4. "
        "Compilation, model loading, initialisation, diagnostics, and output are excluded.}",
        "\\label{tab:weno5_precision_timing}", "\\begin{tabular}{llrrrr}", "\\toprule",
        "Problem & Method & Mean (s) & Std. dev. (s) & Cost/JS & Overhead \\\\", "\\midrule",
    ]
    for i, row in enumerate(rows):
        if i and row["case"] != rows[i - 1]["case"]:
            lines.append("\\midrule")
        lines.append(f"{row['case_label']} & {row['label']} & {row['mean_seconds']:.3f} & {row['std_seconds']:.3f} & "
                     f"{row['cost_vs_classical']:.2f}$\\times$ & {row['overhead_percent']:.1f}\\% \\\\")
    lines.extend(["\\bottomrule", "\\end{tabular}", "\\end{table}"])
    return "\n".join(lines) + "\n"


def run(repeats: int = 3, device: str = "cuda", methods=("weno5_js", "weno5_sr_f64", "weno5_sr_f32"),
        nx: int = 400, write: bool = True) -> list[dict]:
    if repeats < 1:
        raise ValueError("--repeats must be positive")
    selected = set(methods)
    known = {v[0] for v in VARIANTS}
    if not selected or selected - known:
        raise ValueError(f"invalid --methods selection: {sorted(selected - known)}")
    if "weno5_js" not in selected:
        raise ValueError("the cost ratios need the weno5_js row")
    config.validate_models()
    case, case_label, definition = "c4", "C.4", CASES["c4"]
    t_end = float(definition["t_end"])
    p7 = euler_methods.make_weno7_params(nx, nx, 0.0, 1.0, 0.0, 1.0, 0.4, t_end)
    initial_interior = interior(make_quadrant_state(p7, definition, 15, device), p7.ghost, nx, nx)
    rows = []
    for key, label, mixed, adapter, precision in VARIANTS:
        if key not in selected:
            continue
        params = euler_methods.make_weno5_params(nx, nx, 1.0, 1.0, 0.4, t_end, mixed)
        initial = pad_periodic(initial_interior, params.ghost)
        mlp = None if precision is None else checkpoints.load_mlp_params(config.METHODS[key].model, device,
                                                                         precision)
        warm, _, _, warm_t = adapter.run_to_time(initial, params, t_end, device, mlp, "transmissive",
                                                 report_interval=0, report=None)
        health = state_health(warm, params.ghost, nx, nx)
        if not health["complete"] or abs(warm_t - t_end) >= 1.0e-12:
            raise RuntimeError(f"warm-up failed for {case}/{key}: {health}, t={warm_t}")
        samples = []
        for rep in range(repeats):
            started = time.perf_counter()
            final, _, _, final_t = adapter.run_to_time(initial, params, t_end, device, mlp, "transmissive",
                                                       report_interval=0, report=None)
            elapsed = time.perf_counter() - started
            health = state_health(final, params.ghost, nx, nx)
            if not health["complete"] or abs(final_t - t_end) >= 1.0e-12:
                raise RuntimeError(f"timed repeat {rep + 1} failed for {case}/{key}: {health}, t={final_t}")
            samples.append(elapsed)
            print(f"{case} {key} repeat={rep + 1} seconds={elapsed:.6f}", flush=True)
        rows.append({"case": case, "case_label": case_label, "method": key, "label": label,
                     "samples_seconds": samples, "mean_seconds": statistics.fmean(samples),
                     "std_seconds": statistics.stdev(samples) if len(samples) > 1 else 0.0, "complete": True,
                     **health})
    classical = next(r["mean_seconds"] for r in rows if r["method"] == "weno5_js")
    for row in rows:
        row["cost_vs_classical"] = row["mean_seconds"] / classical
        row["overhead_percent"] = 100.0 * (row["cost_vs_classical"] - 1.0)
    if write:
        out = config.RAW / "timing"
        out.mkdir(parents=True, exist_ok=True)
        write_json(out / "summary.json", {"rows": rows})
        config.TABLES.mkdir(parents=True, exist_ok=True)
        (config.TABLES / "weno5_precision_timing.tex").write_text(latex_table(rows), encoding="utf-8")
    return rows


def main() -> None:
    parser = argparse.ArgumentParser()
    parser.add_argument("--repeats", type=int, default=3)
    parser.add_argument("--device", default="cuda")
    parser.add_argument("--methods", default="weno5_js,weno5_sr_f64,weno5_sr_f32")
    args = parser.parse_args()
    config.ensure_output_dirs()
    run(args.repeats, args.device, [m.strip() for m in args.methods.split(",") if m.strip()])


if __name__ == "__main__":
    main()
