"""Methods, checkpoints and output roots (code/snapshot/for_paper_results/config.py:31-93).

The checkpoints are the reference's selected models, carried byte-identical in
``weno_sr_publishing_amd/checkpoints``.  Outputs go under ``$WSR_PAPER_OUT``
(default ``./paper_results``) instead of the package directory.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from pathlib import Path

from .. import checkpoints

ROOT = Path(os.environ.get("WSR_PAPER_OUT", "paper_results")).resolve()
RAW = ROOT / "raw"
FIGURES = ROOT / "figures"
TABLES = ROOT / "tables"
LOGS = ROOT / "logs"


@dataclass(frozen=True)
class Method:
    key: str
    label: str
    family: str
    time_integrator: str
    model: Path | None
    mlp_precision: str | None
    color: str
    linestyle: str


METHODS = {
    "weno5_js": Method("weno5_js", "WENO5-JS-RK3", "weno5", "SSPRK3", None, None, "#4D4D4D", "--"),
    "weno5_sr_f64": Method("weno5_sr_f64", "WENO5-SR-RK3", "weno5", "SSPRK3", checkpoints.WENO5_SR_FP64,
                           "float64", "#0072B2", "-"),
    "weno5_sr_f32": Method("weno5_sr_f32", "WENO5-SR-FP32-RK3", "weno5_mixed", "SSPRK3", checkpoints.WENO5_SR_FP32,
                           "float32", "#009E73", "-."),
    "weno7_js": Method("weno7_js", "WENO7-JS-RK4", "weno7", "SSP-RK4", None, None, "#D55E00", "--"),
    "weno7_sr_f64": Method("weno7_sr_f64", "WENO7-SR-RK4", "weno7", "SSP-RK4", checkpoints.WENO7_SR_FP64,
                           "float64", "#CC79A7", "-"),
}

EULER_METHODS = tuple(METHODS)

# the WENO-Z rows of the shock-tube tables (weno_z_borges_p2_results/plot_riemann_1d_zoom.py:22-30)
Z_LABELS = {"weno5_z_p2": "WENO5-Z-RK3", "weno7_z_p3": "WENO7-Z-RK4"}


def ensure_output_dirs() -> None:
    for path in (RAW, FIGURES, TABLES, LOGS):
        path.mkdir(parents=True, exist_ok=True)


def validate_models() -> None:
    missing = [str(m.model) for m in METHODS.values() if m.model is not None and not Path(m.model).is_file()]
    if missing:
        raise FileNotFoundError("missing selected checkpoint(s):\n" + "\n".join(missing))
