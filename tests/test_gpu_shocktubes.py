"""The paper shock-tube runners on the HIP steps reproduce the reference's tables.

run_sod / run_shock_tube_benchmark / run_riemann_1d (same entry points as the
reference's for_paper_results and weno_z_borges_p2_results modules) run the Sod
(N=100, CFL 0.8, t=0.2) and Lax (N=200, CFL 0.8, t=1.3) tubes for all seven methods on
the GPU; their error rows must meet the 16-digit tables the reference's own GPU runs
wrote (tests/golden/shock_tube_metrics.json).
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((Path(__file__).parent / "golden" / "shock_tube_metrics.json").read_text())
# relative bars: FP64 closures; the WENO7-JS indicator cancellation; the FP32 closure (GPU
# FP32 transcendentals against the reference's FP32 build, ulp-level per evaluation)
BAR = {"weno5_js": 1e-11, "weno5_z_p2": 1e-11, "weno5_sr_f64": 1e-11, "weno5_sr_f32": 2e-6,
       "weno7_js": 5e-9, "weno7_z_p3": 1e-11, "weno7_sr_f64": 1e-10}


def table(problem: str) -> dict:
    return {r["method"]: r for r in GOLDEN[problem]["rows"]}


def _close(got: float, want: str, bar: float) -> bool:
    return abs(got / float(want) - 1.0) <= bar


def test_sod_table_on_gpu(gpu_device):
    from weno_sr_publishing_amd.for_paper_results import run_riemann_1d, run_sod
    ref = table("sod")
    print("run_sod N=100 ny=10 cfl=0.8 t=0.2 (5 methods)", flush=True)
    rows = run_sod.run(nx=100, ny=10, cfl=0.8, t_end=0.2, device="cuda", write=False)
    print("run_riemann_1d sod N=100 (WENO5-Z, WENO7-Z)", flush=True)
    rows += run_riemann_1d.run("sod", device="cuda", cfl=0.8, nx=100, ny=10, write=False)
    assert {r["method"] for r in rows} == set(ref)
    for r in rows:
        m = r["method"]
        assert r["complete"] if "complete" in r else r["nan_count"] == 0
        assert r["t"] == pytest.approx(0.2, abs=1e-12)
        for mine, theirs in (("rho_l1", "rho_l1_mean"), ("rho_l2", "rho_l2"), ("rho_linf", "rho_linf")):
            assert _close(r[mine], ref[m][theirs], BAR[m]), (m, mine, r[mine], ref[m][theirs])


def test_lax_table_on_gpu(gpu_device):
    from weno_sr_publishing_amd.for_paper_results import riemann_tables as rt
    from weno_sr_publishing_amd.for_paper_results import run_riemann_1d, run_shock_tube_benchmark as stb
    ref = table("lax")
    results = {}
    for method in ("weno5_js", "weno5_sr_f64", "weno5_sr_f32", "weno7_js", "weno7_sr_f64"):
        print(f"run_shock_tube_benchmark lax {method} N=200x10 cfl=0.8", flush=True)
        res = stb.run("lax", method, nx=200, ny=10, cfl=0.8, device="cuda", report_interval=0, write=False)
        assert res.metadata["complete"] and res.metadata["steps"] > 100
        results[method] = (res.x, res.conserved_1d)
    print("run_riemann_1d lax (WENO5-Z, WENO7-Z)", flush=True)
    for row in run_riemann_1d.run("lax", device="cuda", write=False):
        results[row["method"]] = (row["_x"], row["_conserved_1d"])
    rows = rt.primitive_errors(run_riemann_1d.PROBLEMS["lax"], results)
    for r in rows:
        m = r["method"]
        for field in ("rho", "velocity", "pressure"):
            for stat in ("l1_mean", "l1_integral", "l2_mean", "linf"):
                key = f"{field}_{stat}"
                assert _close(r[key], ref[m][key], 10 * BAR[m]), (m, key, r[key], ref[m][key])


def test_runner_writes_reference_files(gpu_device, tmp_path, monkeypatch):
    from weno_sr_publishing_amd.for_paper_results import config, run_shock_tube_benchmark as stb
    from weno_sr_publishing_amd.for_paper_results import run_sod
    monkeypatch.setattr(config, "RAW", tmp_path / "raw")
    print("run_sod weno5_js with files", flush=True)
    rows = run_sod.run(nx=100, ny=10, cfl=0.8, t_end=0.2, device="cuda", methods=("weno5_js",), out_tag="N100")
    out = tmp_path / "raw" / "sod" / "N100"
    with np.load(out / "weno5_js.npz", allow_pickle=False) as data:
        assert data["state"].shape == (10, 100, 4)
        meta = json.loads(str(data["metadata_json"]))
    assert meta["rho_l1"] == rows[0]["rho_l1"] and meta["steps"] == rows[0]["steps"]
    assert (out / "weno5_js.json").is_file() and (out / "metrics.csv").is_file() and (out / "summary.json").is_file()
    res = stb.run("lax", "weno5_js", device="cuda", report_interval=0, out_dir=tmp_path / "lax")
    with np.load(tmp_path / "lax" / "weno5_js.npz", allow_pickle=False) as data:
        np.testing.assert_array_equal(data["conserved_1d"], res.conserved_1d)
        assert json.loads(str(data["metadata_json"]))["benchmark"] == "lax"


def test_shock_tube_rejects_bad_arguments(gpu_device):
    from weno_sr_publishing_amd.for_paper_results import run_shock_tube_benchmark as stb
    with pytest.raises(ValueError):
        stb.run("lax", "weno5_js", ny=4, write=False)
    with pytest.raises(ValueError):
        stb.run("lax", "weno5_js", cfl=0.9, write=False)
    with pytest.raises(ValueError):
        stb.run("lax", "weno9", write=False)
