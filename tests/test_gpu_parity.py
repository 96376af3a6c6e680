"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances (max |gpu - oracle| / max |oracle| over the interior):
  * all-FP64 closures (WENO5-JS, WENO5-Z, WENO5-SR FP64): 1e-12 per stage and per step;
  * FP32 closure: 2e-6, BF16 closure: 2e-4 per stage -- the MLP runs with the GPU's
    exp2 / reciprocal instructions and fused multiply-adds, the oracle with libm
    expf and unfused FP32 arithmetic, so ratios differ at the FP32 rounding level.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from weno_sr_publishing_amd.common import state_health
from weno_sr_publishing_amd.initial import make_quadrant_initial, make_sod_state
from weno_sr_publishing_amd.params import Params
from conftest import golden

pytestmark = pytest.mark.gpu

GAMMA = 1.4
TOL = {"js": 1e-12, "z": 1e-12, "sr64": 1e-12, "sr32": 2e-6, "srbf16": 2e-4}


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    a = a[3:-3, 3:-3]
    b = b[3:-3, 3:-3]
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))


def rich_state(nx: int, ny: int, seed: int = 0) -> np.ndarray:
    """Smooth waves, two jumps and a plateau region: exercises every branch."""
    rng = np.random.default_rng(seed)
    x = (np.arange(-3, nx + 3) + 0.5) / nx
    y = (np.arange(-3, ny + 3) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X + 0.4) * np.cos(2 * np.pi * Y) + 0.6 * (X + 0.7 * Y > 0.8)
    rho = np.where((X < 0.25) & (Y < 0.3), 0.8, rho)
    u = 0.4 * np.cos(2 * np.pi * Y) + 0.2 * (X > 0.6)
    v = -0.3 * np.sin(2 * np.pi * X) + 0.1 * rng.normal(size=X.shape) * (Y > 0.5)
    p = 1.0 + 0.2 * np.cos(2 * np.pi * (X - Y)) + 0.5 * (Y > 0.4)
    u = np.where((X < 0.25) & (Y < 0.3), 0.0, u)
    v = np.where((X < 0.25) & (Y < 0.3), 0.0, v)
    p = np.where((X < 0.25) & (Y < 0.3), 1.0, p)
    E = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)
    return np.ascontiguousarray(np.stack([rho, rho * u, rho * v, E], axis=-1))


def mlp_for(scheme, mlp64, mlp32):
    if scheme == "sr64":
        return mlp64
    if scheme in ("sr32", "srbf16"):
        return mlp32.as_precision("bf16") if scheme == "srbf16" else mlp32
    return None


def solver(params, scheme, mlp, boundary="transmissive", riemann="hllc", device="cuda:0"):
    from weno_sr_publishing_amd.euler2d import Euler2DSolver
    return Euler2DSolver(params, scheme, riemann, boundary, mlp, False, device)


@pytest.mark.parametrize("scheme", ["js", "z", "sr64", "sr32", "srbf16"])
@pytest.mark.parametrize("boundary,riemann", [("transmissive", "hllc"), ("periodic", "hllc"),
                                              ("double-mach", "hllc"), ("transmissive", "force")])
def test_stage_parity(oracle, mlp64, mlp32, gpu_device, scheme, boundary, riemann):
    nx, ny = 37, 29
    p = Params(nx=nx, ny=ny, x_length=1.0, y_length=0.8, cfl=0.4)
    us = rich_state(nx, ny, 1)
    u0 = rich_state(nx, ny, 2)
    mlp = mlp_for(scheme, mlp64, mlp32)
    s = solver(p, scheme, mlp, boundary, riemann)
    dt, t = 2.0e-3, 0.013
    g_us, g_u0 = s.new_state(us), s.new_state(u0)
    out = s.new_state()
    worst = 0.0
    for rk in (1, 2, 3):
        s.stage(g_us, g_u0, out, dt, rk, t)
        torch.cuda.synchronize()
        ref = oracle.stage2d(us, u0, nx, ny, p.dx, p.dy, GAMMA, dt, rk, scheme=scheme, riemann=riemann,
                             boundary=boundary, boundary_time=t, mlp=mlp)
        err = rel_err(out.cpu().numpy(), ref)
        worst = max(worst, err)
    print(f"stage parity {scheme} {boundary} {riemann}: {worst:.3e}")
    assert worst <= TOL[scheme]


@pytest.mark.parametrize("scheme", ["js", "sr64", "sr32"])
def test_sod_smoke_against_reference_record(mlp64, mlp32, gpu_device, scheme):
    """GPU run of the 40x4 Sod smoke against the reference's recorded values."""
    from weno_sr_publishing_amd.solvers import weno5_hllc, weno5_hllc_mixed, weno5_hllc_refsym
    fx = golden("smoke_euler_weno5.json")
    key = {"js": "weno5_js", "sr64": "weno5_sr_f64", "sr32": "weno5_sr_f32"}[scheme]
    case, ref = fx["case"], fx["methods"][key]
    p = Params(nx=case["nx"], ny=case["ny"], x_length=case["x_length"], y_length=case["y_length"],
               cfl=case["cfl"], t_end=case["t_end"])
    u0 = make_sod_state(p)
    if scheme == "js":
        final, dts, steps, t = weno5_hllc.run_to_time(u0, p, p.t_end, "cuda:0", None, "transmissive")
    elif scheme == "sr64":
        final, dts, steps, t = weno5_hllc_refsym.run_to_time(u0, p, p.t_end, "cuda:0", mlp64, "transmissive")
    else:
        final, dts, steps, t = weno5_hllc_mixed.run_to_time(u0, p, p.t_end, "cuda:0", mlp32, "transmissive")
    assert steps == 2 and abs(t - 0.01) < 1e-15
    assert abs(min(dts) - ref["dt_min"]) < 1e-17 and abs(max(dts) - ref["dt_max"]) < 1e-17
    h = state_health(final, 3, p.nx, p.ny)
    tol = 1e-14 if scheme != "sr32" else 1e-9
    for k in ("rho_min", "rho_max", "p_min", "p_max"):
        assert abs(h[k] - ref[k]) <= tol, (k, h[k], ref[k])


@pytest.mark.parametrize("scheme", ["js", "sr64", "sr32"])
def test_device_loop_matches_oracle_run(oracle, mlp64, mlp32, gpu_device, scheme):
    """Device-resident dt and time loop vs the oracle's run_to_time (C.4, 48x40)."""
    nx, ny = 48, 40
    p = Params(nx=nx, ny=ny, x_length=1.0, y_length=1.0, cfl=0.4, t_end=0.02)
    u0 = make_quadrant_initial(nx, ny, "c4")
    mlp = mlp_for(scheme, mlp64, mlp32)
    s = solver(p, scheme, mlp)
    u = s.new_state(u0)
    dts, steps, t = s.run(u, p.t_end, chunk=5)
    ref, rdts, rsteps, rt = oracle.run2d(u0, nx, ny, p.dx, p.dy, GAMMA, p.cfl, p.t_end, scheme=scheme, mlp=mlp)
    assert steps == rsteps and abs(t - rt) <= 1e-15
    np.testing.assert_allclose(dts, rdts, rtol=1e-12 if scheme != "sr32" else 1e-6)
    err = rel_err(u.cpu().numpy(), ref)
    print(f"loop parity {scheme}: steps {steps} err {err:.3e}")
    assert err <= (1e-12 if scheme != "sr32" else 1e-5)


def test_max_speed_and_dt(oracle, gpu_device):
    nx, ny = 64, 50
    p = Params(nx=nx, ny=ny, x_length=1.0, y_length=1.0, cfl=0.4)
    us = rich_state(nx, ny, 3)
    s = solver(p, "js", None)
    g = s.new_state(us)
    assert s.max_speed(g) == oracle.max_speed2d(us, nx, ny, GAMMA)
    assert s.compute_dt(g) == oracle.dt2d(us, nx, ny, p.dx, p.dy, GAMMA, p.cfl)
    # The kernel's max(a, b) = a > b ? a : b (warp_weno5_helpers.py:886-896 uses wp.max):
    # a NaN density or energy is floored to 1e-15 and one NaN velocity loses against the
    # other one, so only a cell with both momenta NaN gives a NaN speed, which np.max
    # then propagates and the loop stops.
    bad = us.copy()
    bad[10, 10, 1] = bad[10, 10, 2] = np.nan
    assert np.isnan(s.max_speed(s.new_state(bad)))
    assert np.isnan(oracle.max_speed2d(bad, nx, ny, GAMMA))
    for comp in (0, 1, 3):
        bad = us.copy()
        bad[10, 10, comp] = np.nan
        assert s.max_speed(s.new_state(bad)) == oracle.max_speed2d(bad, nx, ny, GAMMA)


def test_loop_stops_on_nan(gpu_device):
    """A NaN state yields a non-finite dt: the loop stops at once (run_to_time's break)."""
    nx, ny = 16, 16
    p = Params(nx=nx, ny=ny, x_length=1.0, y_length=1.0, cfl=0.4, t_end=0.1)
    u0 = make_quadrant_initial(nx, ny, "c4")
    u0[8, 8, 1] = u0[8, 8, 2] = np.nan
    s = solver(p, "js", None)
    u = s.new_state(u0)
    dts, steps, t = s.run(u, p.t_end)
    assert steps == 0 and t == 0.0 and dts == []


@pytest.mark.parametrize("scheme", ["js", "sr32"])
def test_closure_counters(mlp32, gpu_device, scheme):
    """Counting does not change the result; the counts are consistent."""
    nx, ny = 48, 40
    p = Params(nx=nx, ny=ny, x_length=1.0, y_length=1.0, cfl=0.4, t_end=0.02)
    u0 = make_quadrant_initial(nx, ny, "c4")
    mlp = mlp_for(scheme, None, mlp32)
    plain = solver(p, scheme, mlp)
    ua = plain.new_state(u0)
    plain.run(ua, p.t_end)
    counting = solver(p, scheme, mlp)
    counting.set_counting(True)
    ub = counting.new_state(u0)
    counting.run(ub, p.t_end)
    assert torch.equal(ua, ub)
    c = counting.counters()
    print(f"counters {scheme}: {c}")
    assert c[0] > 0 and c[2] > 0
    assert c[4] % 64 == 0 and c[5] % 64 == 0
    if scheme == "js":
        assert c[1] == c[3] == c[4] == c[5] == 0
    else:
        assert 0 < c[1] <= c[4] <= c[0] and 0 < c[3] <= c[5]


@pytest.mark.parametrize("scheme", ["js", "sr32"])
def test_full_size_one_step_parity(oracle, mlp32, gpu_device, scheme):
    """400x400 C.4 (the benchmark workload): one SSPRK3 step vs the oracle."""
    nx = ny = 400
    p = Params(nx=nx, ny=ny, x_length=1.0, y_length=1.0, cfl=0.4, t_end=0.25)
    u0 = make_quadrant_initial(nx, ny, "c4")
    mlp = mlp32 if scheme == "sr32" else None
    s = solver(p, scheme, mlp)
    u = s.new_state(u0)
    dts, steps, t = s.run(u, p.t_end, max_steps=1)
    ref, rdts, _, _ = oracle.run2d(u0, nx, ny, p.dx, p.dy, GAMMA, p.cfl, p.t_end, scheme=scheme, mlp=mlp,
                                   max_steps=1)
    assert steps == 1 and dts[0] == rdts[0]
    err = rel_err(u.cpu().numpy(), ref)
    print(f"400^2 step parity {scheme}: {err:.3e}")
    assert err <= TOL[scheme]


@pytest.mark.parametrize("scheme", ["js", "sr32"])
def test_full_size_periodic_conservation_and_symmetry(mlp32, gpu_device, scheme):
    """400x400: periodic runs conserve every conserved total; C.4 stays symmetric
    under the transpose (x <-> y with the momenta swapped)."""
    nx = ny = 400
    p = Params(nx=nx, ny=ny, x_length=1.0, y_length=1.0, cfl=0.4, t_end=1.0)
    u0 = make_quadrant_initial(nx, ny, "c4")
    mlp = mlp32 if scheme == "sr32" else None
    s = solver(p, scheme, mlp, boundary="periodic")
    u = s.new_state(u0)
    s.run(u, p.t_end, max_steps=20)
    a = u.cpu().numpy()[3:-3, 3:-3]
    b = u0[3:-3, 3:-3]
    # telescoping flux differences conserve to rounding (~1e-16 per cell update)
    np.testing.assert_allclose(a.sum(axis=(0, 1)), b.sum(axis=(0, 1)), rtol=1e-11)
    s2 = solver(p, scheme, mlp, boundary="transmissive")
    u = s2.new_state(u0)
    s2.run(u, p.t_end, max_steps=20)
    a = u.cpu().numpy()[3:-3, 3:-3]
    at = a.transpose(1, 0, 2)[..., [0, 2, 1, 3]]
    sym = np.max(np.abs(a - at)) / np.max(np.abs(a))
    print(f"transpose symmetry {scheme}: {sym:.3e}")
    assert sym <= (1e-11 if scheme == "js" else 1e-4)


@pytest.mark.parametrize("scheme", ["sr64", "sr32", "srbf16"])
def test_full_size_developed_flow_one_step_parity(oracle, mlp64, mlp32, gpu_device, scheme):
    """400x400 C.4 from the developed flow at t = 0.125 (the GPU's own trajectory for the
    closure; 31-41 % of stencils run the closure there), one SSPRK3 step on the GPU and in
    the oracle from the same state: the whole field within the closure's bar."""
    nx = ny = 400
    p = Params(nx=nx, ny=ny, x_length=1.0, y_length=1.0, cfl=0.4, t_end=0.25)
    mlp = {"sr64": mlp64, "sr32": mlp32, "srbf16": mlp32.as_precision("bf16")}[scheme]
    s = solver(p, scheme, mlp)
    u = s.new_state(make_quadrant_initial(nx, ny, "c4"))
    s.run(u, 0.125)
    start = u.cpu().numpy()
    dts, steps, t = s.run(u, p.t_end, max_steps=1)
    ref, rdts, _, _ = oracle.run2d(start, nx, ny, p.dx, p.dy, GAMMA, p.cfl, p.t_end, scheme=scheme, mlp=mlp,
                                   max_steps=1)
    assert steps == 1 and abs(dts[0] - rdts[0]) <= 1e-15 * rdts[0]
    err = rel_err(u.cpu().numpy(), ref)
    print(f"400^2 developed-flow step parity {scheme}: {err:.3e}", flush=True)
    assert err <= TOL[scheme]
