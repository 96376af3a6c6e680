"""The vortex convergence runner on the HIP steps reproduces the reference's table
(experiments/05_isentropic_vortex/tables/vortex_convergence.md, 9 significant digits,
carried in tests/golden/vortex_convergence.json) for all seven methods and all four grids.
"""
from __future__ import annotations

import json
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROWS = json.loads((Path(__file__).parent / "golden" / "vortex_convergence.json").read_text())["rows"]
KEYS = {"WENO5-JS-RK3": "weno5_js", "WENO5-SR-RK3": "weno5_sr_f64", "WENO5-SR-FP32-RK3": "weno5_sr_f32",
        "WENO5-Z-RK3": "weno5_z_p2", "WENO7-JS-RK4": "weno7_js", "WENO7-SR-RK4": "weno7_sr_f64",
        "WENO7-Z-RK4": "weno7_z_p3"}
# the table's 9 significant digits bound the FP64 comparison; the FP32 closure adds
# ulp-level differences of the GPU's FP32 transcendentals against the reference's build
BAR = {k: 6e-9 for k in KEYS.values()}
BAR["weno5_sr_f32"] = 2e-6


@pytest.mark.parametrize("label", list(KEYS))
def test_vortex_table_on_gpu(gpu_device, label):
    from weno_sr_publishing_amd.for_paper_results import run_vortex
    key = KEYS[label]
    want = {r["N"]: r for r in ROWS if r["method"] == label}
    assert sorted(want) == [25, 50, 100, 200]
    print(f"run_vortex {key} grids 25,50,100,200", flush=True)
    rows = run_vortex.run(grids=(25, 50, 100, 200), methods=(key,), write=False)
    for row in rows:
        assert row["complete"] and row["t"] == 2.0
        for norm in ("rho_l1", "rho_l2"):
            dev = abs(row[norm] / float(want[row["N"]][norm]) - 1.0)
            assert dev <= BAR[key], (key, row["N"], norm, row[norm], want[row["N"]][norm])
    orders = [r["rho_l1_order"] for r in rows[1:]]
    assert all(o > 2.0 for o in orders), orders
