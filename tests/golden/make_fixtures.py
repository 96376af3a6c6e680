"""Write the golden fixtures under tests/golden/ from data files of the reference checkout.

Run in a container that has the reference at /root/reference.  Only recorded
numbers are extracted (JSON values, CSV rows and the literal vectors of the
release verifier); nothing of the reference is imported or executed.

Sources (relative to the reference checkout):
  code/snapshot/for_paper_results/verification_smoke_euler.json  (40x4 Sod smoke, 2 steps)
  scripts/verify_release.py:157-222 (fixed feature vectors and expected MLP outputs)
  code/snapshot/for_paper_results/verification_hllc.json         (recorded HLLC agreement)
  experiments/01_gste_long_time_advection/tables/gste_metrics.csv (1-D GSTE metrics)
  experiments/05_isentropic_vortex/tables/vortex_convergence.md (2-D vortex L1/L2, 9 digits)
  experiments/14_mixed_precision_timing/tables/weno5_precision_timing.tex (timings)
  experiments/02_sod_shock_tube/tables/sod_metrics.csv          (Sod N=100 density errors, 7 methods)
  experiments/03_lax_shock_tube/tables/lax_metrics.csv          (Lax N=200 primitive errors, 7 methods)
  experiments/06-09_riemann_c3..c6/tables/linecuts.csv          (C.3-C.6 density line cuts, 6 methods)
  experiments/10_double_mach_reflection/tables/linecuts.csv     (double Mach line cuts vs WENO7-JS, 7 methods)
"""
from __future__ import annotations

import ast
import csv
import json
import re
from pathlib import Path

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent


def smoke() -> None:
    data = json.loads((REF / "code/snapshot/for_paper_results/verification_smoke_euler.json").read_text())
    keep = {k: v for k, v in data.items() if k.startswith("weno5")}
    (OUT / "smoke_euler_weno5.json").write_text(json.dumps(
        {"source": "code/snapshot/for_paper_results/verification_smoke_euler.json",
         "case": {"nx": 40, "ny": 4, "x_length": 1.0, "y_length": 0.1, "cfl": 0.4, "t_end": 0.01,
                  "boundary": "transmissive", "initial": "Sod strip along x centred on 0"},
         "methods": keep}, indent=2, sort_keys=True) + "\n")


def mlp_golden() -> None:
    src = (REF / "scripts/verify_release.py").read_text()
    def grab(name: str) -> list:
        m = re.search(name + r"\s*=\s*torch\.tensor\(\s*(\[.*?\]),\s*dtype", src, re.S)
        return ast.literal_eval(m.group(1))
    f5 = grab("features5")
    f7 = grab("features7")
    outs = {}
    for label in ("WENO5 FP64", "WENO5 FP32", "WENO7 FP64"):
        start = src.index('"' + label + '": np.array(', src.index("expected_outputs"))
        start = src.index("[", start)
        depth, end = 0, start
        for end in range(start, len(src)):
            depth += {"[": 1, "]": -1}.get(src[end], 0)
            if depth == 0:
                break
        outs[label] = ast.literal_eval(src[start:end + 1])
    (OUT / "mlp_golden.json").write_text(json.dumps(
        {"source": "scripts/verify_release.py:157-234",
         "note": "inputs are old-style features (scale channel (log10+16)/16); the reference remaps "
                 "the scale channel to (log10+4)/4 before the MLP and reflection-averages the output",
         "features5": f5, "features7": f7, "expected": outs,
         "atol": {"WENO5 FP64": 5e-15, "WENO5 FP32": 2e-7, "WENO7 FP64": 5e-15},
         "reflection_defect_limit": {"WENO5 FP64": 1e-14, "WENO5 FP32": 2e-7, "WENO7 FP64": 1e-14}},
        indent=2) + "\n")


def hllc() -> None:
    data = json.loads((REF / "code/snapshot/for_paper_results/verification_hllc.json").read_text())
    (OUT / "hllc_recorded.json").write_text(json.dumps(
        {"source": "code/snapshot/for_paper_results/verification_hllc.json", "recorded": data},
        indent=2, sort_keys=True) + "\n")


def gste() -> None:
    path = REF / "experiments/01_gste_long_time_advection/tables/gste_metrics.csv"
    rows = list(csv.DictReader(path.open()))
    (OUT / "gste_metrics.json").write_text(json.dumps(
        {"source": "experiments/01_gste_long_time_advection/tables/gste_metrics.csv", "rows": rows},
        indent=2) + "\n")


def vortex() -> None:
    path = REF / "experiments/05_isentropic_vortex/tables/vortex_convergence.md"
    rows = []
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[1].isdigit():
            rows.append({"method": cells[0], "N": int(cells[1]), "rho_l1": cells[2], "rho_l2": cells[4]})
    (OUT / "vortex_convergence.json").write_text(json.dumps(
        {"source": "experiments/05_isentropic_vortex/tables/vortex_convergence.md", "rows": rows},
        indent=2) + "\n")


def timing() -> None:
    path = REF / "experiments/14_mixed_precision_timing/tables/weno5_precision_timing.tex"
    (OUT / "timing_table.json").write_text(json.dumps(
        {"source": "experiments/14_mixed_precision_timing/tables/weno5_precision_timing.tex",
         "rows": [line.strip() for line in path.read_text().splitlines() if "&" in line]},
        indent=2) + "\n")


def shock_tubes() -> None:
    out = {}
    for key, rel in (("sod", "experiments/02_sod_shock_tube/tables/sod_metrics.csv"),
                     ("lax", "experiments/03_lax_shock_tube/tables/lax_metrics.csv")):
        out[key] = {"source": rel, "rows": list(csv.DictReader((REF / rel).open()))}
    (OUT / "shock_tube_metrics.json").write_text(json.dumps(out, indent=2) + "\n")


def linecuts() -> None:
    out = {}
    for case, folder in (("c3", "06_riemann_c3"), ("c4", "07_riemann_c4"), ("c5", "08_riemann_c5"),
                         ("c6", "09_riemann_c6")):
        rel = f"experiments/{folder}/tables/linecuts.csv"
        out[case] = {"source": rel, "rows": list(csv.DictReader((REF / rel).open()))}
    rel = "experiments/10_double_mach_reflection/tables/linecuts.csv"
    out["double_mach"] = {"source": rel, "rows": list(csv.DictReader((REF / rel).open()))}
    (OUT / "riemann_linecuts.json").write_text(json.dumps(out, indent=2) + "\n")


if __name__ == "__main__":
    linecuts()
    shock_tubes()
    smoke()
    mlp_golden()
    hllc()
    gste()
    vortex()
    timing()
    print("fixtures written to", OUT)
