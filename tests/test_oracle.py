"""Pins the CPU oracle against the reference's recorded outputs and against
properties the mathematics fixes.  Runs without a GPU."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import golden
from weno_sr_publishing_amd.common import state_health
from weno_sr_publishing_amd.initial import make_sod_state, make_quadrant_initial
from weno_sr_publishing_amd.params import GAUSS15_W, GAUSS15_XI, Params

GAMMA = 1.4


def prim_to_con(rho, u, v, p, gamma=GAMMA):
    return np.array([rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)])


# ------------------------------------------------------------------ full step
SMOKE_TOL = {"weno5_js": 1e-15, "weno5_sr_f64": 1e-15, "weno5_sr_f32": 1e-12}


@pytest.mark.parametrize("key", ["weno5_js", "weno5_sr_f64", "weno5_sr_f32"])
def test_sod_smoke_matches_reference_record(oracle, mlp64, mlp32, key):
    """40x4 Sod, 2 steps: the reference's recorded 16-digit extrema and dt
    (for_paper_results/verification_smoke_euler.json)."""
    fx = golden("smoke_euler_weno5.json")
    case, ref = fx["case"], fx["methods"][key]
    p = Params(nx=case["nx"], ny=case["ny"], x_length=case["x_length"], y_length=case["y_length"],
               cfl=case["cfl"], t_end=case["t_end"])
    scheme, mlp = {"weno5_js": ("js", None), "weno5_sr_f64": ("sr64", mlp64),
                   "weno5_sr_f32": ("sr32", mlp32)}[key]
    u, dts, steps, t = oracle.run2d(make_sod_state(p), p.nx, p.ny, p.dx, p.dy, p.gamma, p.cfl, p.t_end,
                                    scheme=scheme, mlp=mlp)
    assert steps == ref["steps"] and abs(t - ref["t"]) < 1e-15
    assert abs(min(dts) - ref["dt_min"]) < 1e-17 and abs(max(dts) - ref["dt_max"]) < 1e-17
    h = state_health(u, 3, p.nx, p.ny)
    for k in ("rho_min", "rho_max", "p_min", "p_max"):
        assert abs(h[k] - ref[k]) <= SMOKE_TOL[key], (k, h[k], ref[k])


# -------------------------------------------------------------------- the MLP
def _remap(features):
    f = np.array(features, dtype=np.float64)
    log10_rel = 16.0 * f[..., 4] - 16.0
    f[..., 4] = np.clip((log10_rel + 4.0) / 4.0, 0.0, 1.0)
    return f


def _sym(oracle, feats, scheme, mlp):
    direct = oracle.mlp_raw(feats, scheme, mlp)
    refl = oracle.mlp_raw(feats[:, [2, 1, 0, 3, 4]], scheme, mlp)
    return 0.5 * (direct + refl[:, ::-1])


@pytest.mark.parametrize("label,scheme", [("WENO5 FP64", "sr64"), ("WENO5 FP32", "sr32")])
def test_mlp_matches_release_golden_vectors(oracle, mlp64, mlp32, label, scheme):
    """scripts/verify_release.py:157-222 fixed-stencil inference signature."""
    g = golden("mlp_golden.json")
    feats = _remap(g["features5"])
    out = _sym(oracle, feats, scheme, mlp64 if scheme == "sr64" else mlp32)
    np.testing.assert_allclose(out, np.array(g["expected"][label]), rtol=0, atol=g["atol"][label])


@pytest.mark.parametrize("scheme", ["sr64", "sr32"])
def test_reflection_equivariance(oracle, mlp64, mlp32, scheme):
    rng = np.random.default_rng(7)
    feats = rng.uniform(0, 1, size=(256, 5))
    mlp = mlp64 if scheme == "sr64" else mlp32
    lhs = _sym(oracle, feats[:, [2, 1, 0, 3, 4]], scheme, mlp)
    rhs = _sym(oracle, feats, scheme, mlp)[:, ::-1]
    assert np.max(np.abs(lhs - rhs)) <= (1e-14 if scheme == "sr64" else 2e-7)


def test_zero_final_layer_gives_linear_weights(oracle, mlp64):
    """self_test.py:50-60: w4 = 0, b4 = 0 => equal ratios => omega = d exactly."""
    zero = {k: np.array(v) for k, v in mlp64.items()}
    zero["w4"] = np.zeros_like(zero["w4"])
    zero["b4"] = np.zeros_like(zero["b4"])
    rng = np.random.default_rng(3)
    d = {"normal_lr1": [0.1, 0.6, 0.3], "normal_lr2": [0.3, 0.6, 0.1],
         "gauss_lr1": [(210 + math.sqrt(3)) / 1080, 11 / 18, (210 - math.sqrt(3)) / 1080],
         "gauss_lr2": [(210 - math.sqrt(3)) / 1080, 11 / 18, (210 + math.sqrt(3)) / 1080]}
    for _ in range(20):
        q = rng.normal(size=5)
        for head, dd in d.items():
            w = oracle.mlp_weights(q, head, "sr64", zero)
            assert np.max(np.abs(w - np.array(dd))) <= 2e-15


def test_bf16_rounding(oracle):
    assert oracle.round_bf16(1.0) == 1.0
    assert oracle.round_bf16(1.0 + 2.0 ** -8) == 1.0            # tie -> even
    assert oracle.round_bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert oracle.round_bf16(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7


def test_features_and_plateau(oracle):
    assert oracle.plateau([1.0, 1.0, 1.0, 1.0, 1.0])
    assert oracle.plateau([2.0, 2.0, 2.0, 2.0, 2.0 + 1e-14])
    assert not oracle.plateau([0.0, 0.1, 0.2, 0.3, 0.5])
    q = np.array([0.3, -0.2, 0.7, 1.1, 0.4])
    f = oracle.features(q)
    fr = oracle.features(q[::-1])
    np.testing.assert_allclose(fr, f[[2, 1, 0, 3, 4]], rtol=0, atol=1e-15)
    assert max(f[:3]) == 1.0 and 0.0 <= f[3] <= 1.0 and 0.0 <= f[4] <= 1.0
    # scale channel: clip((log10(delta_max / q_scale) + 4) / 4, 0, 1)
    d2 = [q[0] - 2 * q[1] + q[2], q[1] - 2 * q[2] + q[3], q[2] - 2 * q[3] + q[4]]
    deltas = [13 / 12 * abs(d2[0]) + 0.25 * abs(q[0] - 4 * q[1] + 3 * q[2]),
              13 / 12 * abs(d2[1]) + 0.25 * abs(q[1] - q[3]),
              13 / 12 * abs(d2[2]) + 0.25 * abs(3 * q[2] - 4 * q[3] + q[4])]
    scale = min(1.0, max(0.0, (math.log10(max(deltas) / max(1.0, np.max(np.abs(q)))) + 4) / 4))
    assert abs(f[4] - scale) < 1e-15


# ------------------------------------------------------ stencil coefficients
def _cell_avg_poly(coeffs, a, b):
    """Exact average of sum c_k x^k over [a, b]."""
    return sum(c * (b ** (k + 1) - a ** (k + 1)) / (k + 1) for k, c in enumerate(coeffs)) / (b - a)


def _poly(coeffs, x):
    return sum(c * x ** k for k, c in enumerate(coeffs))


@pytest.mark.parametrize("head,point,deg_cand,tol", [
    ("normal_lr1", 0.5, 2, 1e-13), ("normal_lr2", -0.5, 2, 1e-13),
    ("gauss_lr1", -0.5 / math.sqrt(3), 2, 2e-14), ("gauss_lr2", 0.5 / math.sqrt(3), 2, 2e-14)])
def test_candidates_exact_for_quadratics(oracle, head, point, deg_cand, tol):
    """W5:248-281: every candidate is exact for degree <= 2 at its point; the
    linear-weight combination is exact for degree <= 4 (Gauss constants are the
    reference's 15-digit truncations, hence the looser tolerance)."""
    rng = np.random.default_rng(11)
    d = {"normal_lr1": [0.1, 0.6, 0.3], "normal_lr2": [0.3, 0.6, 0.1],
         "gauss_lr1": [(210 + math.sqrt(3)) / 1080, 11 / 18, (210 - math.sqrt(3)) / 1080],
         "gauss_lr2": [(210 - math.sqrt(3)) / 1080, 11 / 18, (210 + math.sqrt(3)) / 1080]}[head]
    for deg in range(5):
        c = rng.normal(size=deg + 1)
        q = [_cell_avg_poly(c, m - 2.5, m - 1.5) for m in range(5)]
        s = oracle.candidates(q, head)
        exact = _poly(c, point)
        if deg <= deg_cand:
            assert np.max(np.abs(s - exact)) < tol
        assert abs(np.dot(d, s) - exact) < tol
        # small smooth data (beta << 1e-6) -> the JS weights approach d
        scale = 1e-7
        assert abs(oracle.scalar_value(np.array(q) * scale, head, "js") - scale * exact) < 1e-4 * scale


def test_weno_z_and_js_on_constant_and_jump(oracle):
    for head in ("normal_lr1", "normal_lr2", "gauss_lr1", "gauss_lr2"):
        for scheme in ("js", "z"):
            assert abs(oracle.scalar_value([2.5] * 5, head, scheme, h=0.01) - 2.5) < 1e-14
    # a jump between q2 and q3: the right-face JS value must stay on the left plateau
    v = oracle.scalar_value([1.0, 1.0, 1.0, 0.0, 0.0], "normal_lr1", "js")
    assert abs(v - 1.0) < 1e-6
    v = oracle.scalar_value([1.0, 1.0, 1.0, 0.0, 0.0], "normal_lr1", "z", h=0.01)
    assert abs(v - 1.0) < 1e-3


# -------------------------------------------------------------- Euler algebra
def _random_states(rng, n):
    out = []
    for _ in range(n):
        rho = rng.uniform(0.1, 5.0)
        u, v = rng.normal(scale=2.0, size=2)
        p = rng.uniform(0.05, 10.0)
        out.append(prim_to_con(rho, u, v, p))
    return out


def _hllc_toro(ul, ur, direction, gamma=GAMMA):
    """Independent HLLC in Toro's alternative form (Toro, Riemann Solvers, eq. 10.75):
    F*_K = (S*(S_K U_K - F_K) + S_K (p_K + rho_L (S_L - u_L)(S* - u_L)) D*) / (S_K - S*)."""
    def prim(q):
        rho = q[0]
        u, v = q[1] / rho, q[2] / rho
        return rho, u, v, (gamma - 1) * (q[3] - 0.5 * rho * (u * u + v * v))

    def flux(q, direction):
        rho, u, v, p = prim(q)
        un = u if direction == 1 else v
        f = un * np.asarray(q, dtype=float)
        f[direction] += p
        f[3] += un * p
        return f

    rl, ul_, vl, pl = prim(ul)
    rr, ur_, vr, pr = prim(ur)
    unl, unr = (ul_, ur_) if direction == 1 else (vl, vr)
    al, ar = math.sqrt(gamma * pl / rl), math.sqrt(gamma * pr / rr)
    sl, sr = min(unl - al, unr - ar), max(unl + al, unr + ar)
    fl, fr = flux(ul, direction), flux(ur, direction)
    if sl >= 0:
        return fl
    if sr <= 0:
        return fr
    ss = (pr - pl + rl * unl * (sl - unl) - rr * unr * (sr - unr)) / (rl * (sl - unl) - rr * (sr - unr))
    pLR = 0.5 * (pl + pr + rl * (sl - unl) * (ss - unl) + rr * (sr - unr) * (ss - unr))
    D = np.zeros(4)
    D[direction] = 1.0
    D[3] = ss
    if ss >= 0:
        return (ss * (sl * np.asarray(ul) - fl) + sl * pLR * D) / (sl - ss)
    return (ss * (sr * np.asarray(ur) - fr) + sr * pLR * D) / (sr - ss)


@pytest.mark.parametrize("direction", [1, 2])
def test_hllc_against_independent_form(oracle, direction):
    """verify_hllc.py:62-127 compares against a direct restatement on 512 states; here the
    comparison is against Toro's algebraically different HLLC form (scaled <= 1e-13,
    recorded reference agreement 1.4e-14)."""
    rng = np.random.default_rng(20260711)
    states = _random_states(rng, 1024)
    worst = 0.0
    for a, b in zip(states[::2], states[1::2]):
        f = oracle.hllc(a, b, direction)
        g = _hllc_toro(a, b, direction)
        worst = max(worst, np.max(np.abs(f - g)) / max(1.0, np.max(np.abs(g))))
    assert worst < 1e-13
    rec = golden("hllc_recorded.json")["recorded"]["hllc_max_scaled_diff_vs_direct_numpy"]
    assert rec[f"direction_{direction}"] < 1e-13


@pytest.mark.parametrize("direction", [1, 2])
def test_flux_consistency(oracle, direction):
    rng = np.random.default_rng(5)
    for q in _random_states(rng, 64):
        f = oracle.physical_flux(q, direction)
        np.testing.assert_allclose(oracle.hllc(q, q, direction), f, rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(oracle.force(q, q, direction, 0.3), f, rtol=1e-12, atol=1e-12)


def test_hllc_supersonic_upwinding(oracle):
    ql = prim_to_con(1.0, 5.0, 0.3, 1.0)
    qr = prim_to_con(0.5, 4.0, -0.2, 0.8)
    np.testing.assert_array_equal(oracle.hllc(ql, qr, 1), oracle.physical_flux(ql, 1))
    ql = prim_to_con(1.0, 0.1, -5.0, 1.0)
    qr = prim_to_con(0.5, 0.0, -6.0, 0.8)
    np.testing.assert_array_equal(oracle.hllc(ql, qr, 2), oracle.physical_flux(qr, 2))


@pytest.mark.parametrize("direction", [1, 2])
def test_characteristic_roundtrip(oracle, direction):
    """L R = I for every Roe state (H5:645-748)."""
    rng = np.random.default_rng(9)
    st = _random_states(rng, 80)
    for a, b, q in zip(st[::3], st[1::3], st[2::3]):
        roe = oracle.roe(a, b)
        c, back = oracle.char_roundtrip(q, roe, direction)
        np.testing.assert_allclose(back, q, rtol=1e-12, atol=1e-12 * np.max(np.abs(q)))


def test_roe_average_of_equal_states(oracle):
    q = prim_to_con(1.3, 0.4, -0.7, 2.2)
    np.testing.assert_allclose(oracle.roe(q, q), q, rtol=1e-14)


# ------------------------------------------------------------ whole stages
@pytest.mark.parametrize("scheme", ["js", "sr64", "sr32", "srbf16", "z"])
@pytest.mark.parametrize("boundary", ["transmissive", "periodic"])
def test_uniform_flow_preserved(oracle, mlp64, mlp32, scheme, boundary):
    nx, ny = 12, 9
    q = prim_to_con(1.2, 0.7, -0.4, 2.0)
    u = np.broadcast_to(q, (ny + 6, nx + 6, 4)).copy()
    mlp = mlp64 if scheme == "sr64" else mlp32
    for rk in (1, 2, 3):
        out = oracle.stage2d(u, u, nx, ny, 0.1, 0.12, GAMMA, 0.01, rk, scheme=scheme, boundary=boundary, mlp=mlp)
        inner = out[3:-3, 3:-3]
        assert np.max(np.abs(inner - q)) < 1e-14


def test_periodic_mass_conservation(oracle, mlp32):
    nx, ny = 16, 12
    rng = np.random.default_rng(1)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) + 0.3 * (X > 0.5)
    st = np.stack([rho, rho * 0.5, rho * -0.3, 2.5 + 0.1 * rng.normal(size=rho.shape)], axis=-1)
    u = np.pad(st, ((3, 3), (3, 3), (0, 0)), mode="wrap")
    for scheme in ("js", "sr32"):
        out = oracle.step2d(u, nx, ny, 1 / nx, 1 / ny, GAMMA, 0.004, scheme=scheme, boundary="periodic", mlp=mlp32)
        before = u[3:-3, 3:-3].sum(axis=(0, 1))
        after = out[3:-3, 3:-3].sum(axis=(0, 1))
        np.testing.assert_allclose(after, before, rtol=2e-14)


def test_dt_rule(oracle):
    p = Params(nx=20, ny=10, x_length=1.0, y_length=0.6, cfl=0.4)
    u = make_quadrant_initial(20, 10, "c4")
    from weno_sr_publishing_amd.common import compute_dt
    assert oracle.dt2d(u, p.nx, p.ny, p.dx, p.dy, GAMMA, p.cfl) == compute_dt(u, p)


# --------------------------------------------------------------- quadrature
def test_gauss15_table_integrates_polynomials():
    xi, w = np.array(GAUSS15_XI), np.array(GAUSS15_W)
    assert len(xi) == 15 and np.all(np.diff(xi) > 0)
    for k in range(30):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * xi ** k) - exact) < 5e-15


def test_quadrant_initial_state():
    u = make_quadrant_initial(40, 40, "c4")
    assert u.shape == (46, 46, 4)
    ne = prim_to_con(1.1, 0.0, 0.0, 1.1)
    sw = prim_to_con(1.1, 0.8939, 0.8939, 1.1)
    np.testing.assert_allclose(u[3 + 30, 3 + 30], ne, rtol=1e-14)
    np.testing.assert_allclose(u[3 + 5, 3 + 5], sw, rtol=1e-14)


# --------------------------------------------------------------- 1-D (GSTE)
GSTE_TOL = {"weno5_js": 1e-12, "weno5_sr_f64": 1e-12, "weno5_sr_f32": 5e-7}


@pytest.mark.parametrize("key", ["weno5_js", "weno5_sr_f64", "weno5_sr_f32"])
def test_gste_metrics_match_reference_table(oracle, mlp64, mlp32, key):
    """1-D GSTE, N=200, t=10, CFL 0.6 (1667 SSPRK3 steps): L1/L2/Linf/TV of the
    reference's recorded table (experiments/01_gste_long_time_advection/tables/gste_metrics.csv)."""
    from weno_sr_publishing_amd.initial import gste_cell_averages
    rows = {r["method"]: r for r in golden("gste_metrics.json")["rows"]}
    ref = rows[key]
    nx, t_end, cfl = int(ref["nx"]), float(ref["t_end"]), float(ref["cfl"])
    dx = 2.0 / nx
    steps = int(math.ceil(t_end / (cfl * dx)))
    _, u0 = gste_cell_averages(nx, 0.0, 15)
    _, exact = gste_cell_averages(nx, t_end, 15)
    scheme, mlp = {"weno5_js": ("js", None), "weno5_sr_f64": ("sr64", mlp64), "weno5_sr_f32": ("sr32", mlp32)}[key]
    final = oracle.adv1d_run(u0[None, :], [1.0], [t_end / steps] * steps, 1.0 / dx, scheme, mlp)[0]
    d = final - exact
    got = {"l1": np.mean(np.abs(d)), "l2": np.sqrt(np.mean(d * d)), "linf": np.max(np.abs(d)),
           "tv": np.sum(np.abs(final - np.roll(final, 1)))}
    for k, v in got.items():
        assert abs(v / float(ref[k]) - 1.0) <= GSTE_TOL[key], (k, v, ref[k])


def test_adv1d_reflection_symmetry_and_mass(oracle, mlp32):
    """self_test.py:69-111: periodic mass conservation and bidirectional reflection
    (flip the profile and a -> -a gives the flipped solution)."""
    rng = np.random.default_rng(4)
    n = 64
    x = (np.arange(n) + 0.5) / n
    u = np.sin(2 * np.pi * x) + (x > 0.3) * (x < 0.55) + 0.2 * rng.normal(size=n)
    dts = [0.5 / n] * 20
    for scheme in ("js", "sr32"):
        pos = oracle.adv1d_run(u[None, :], [1.0], dts, n, scheme, mlp32)[0]
        neg = oracle.adv1d_run(u[::-1][None, :], [-1.0], dts, n, scheme, mlp32)[0]
        assert abs(pos.sum() - u.sum()) < 2e-12 * n
        assert np.max(np.abs(pos - neg[::-1])) < 2e-12


def test_adv1d_rollout_matches_reference_torch(oracle, mlp64, mlp32):
    """Final states of the reference's own torch rollouts (tests/golden/grad_fixture.npz,
    rk3_advection.rollout_distance with the FP32 model; v12 ssprk3_step_signed with FP64)."""
    from pathlib import Path
    fx = np.load(Path(__file__).resolve().parent / "golden" / "grad_fixture.npz", allow_pickle=False)
    n = int(fx["n"])
    f64 = oracle.adv1d_run(fx["u0"], fx["velocities"], fx["dts"], n, "sr64", mlp64)
    assert np.max(np.abs(f64 - fx["fp64_final"])) <= 1e-14
    f32 = oracle.adv1d_run(fx["u0"], fx["velocities"], fx["dts"], n, "sr32", mlp32)
    assert np.max(np.abs(f32 - fx["fp32_final"])) <= 1e-6


def test_oracle_fd_gradient_matches_reference_autograd(oracle, mlp64):
    """Central differences of the oracle's FP64 loss against the reference's autograd
    gradient for a handful of parameters (pins the fixture and the oracle together)."""
    from pathlib import Path
    from weno_sr_publishing_amd.training import flatten, unflatten
    fx = np.load(Path(__file__).resolve().parent / "golden" / "grad_fixture.npz", allow_pickle=False)
    n = int(fx["n"])
    base = flatten({k: np.asarray(v)[0] for k, v in mlp64.items()})
    ref = np.concatenate([fx[f"fp64_grad_{k}"].reshape(-1) for k in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4")])

    def loss(f):
        p = {k: v[None, ...] if v.ndim == 2 else v[None, :] for k, v in unflatten(f).items()}
        fin = oracle.adv1d_run(fx["u0"], fx["velocities"], fx["dts"], n, "sr64", p)
        return float(np.mean((fin - fx["target"]) ** 2))

    assert abs(loss(base) - float(fx["fp64_loss"])) <= 1e-14
    for p in (0, 7, 52, 75, 124, 140, 165, 172, 186, 188):
        h = 1e-6 * max(1.0, abs(base[p]))
        fp, fm = base.copy(), base.copy()
        fp[p] += h
        fm[p] -= h
        fd = (loss(fp) - loss(fm)) / (2 * h)
        assert abs(fd - ref[p]) <= 1e-6 * max(1e-3, np.max(np.abs(ref))), (p, fd, ref[p])


# --------------------------------------------------------------------------
# FORCE (warp_weno5_helpers.py:401-435): the oracle against an independent
# restatement of Toro's combined form and a closed form that pins the signs
# and the placement of c = dt/h and 1/c.
def _euler_flux_conserved(U, direction, gamma=1.4):
    rho, mx, my, E = U
    p = (gamma - 1.0) * (E - 0.5 * (mx * mx + my * my) / rho)
    un = (mx if direction == 1 else my) / rho
    out = un * np.asarray(U, dtype=np.float64)
    out[1 if direction == 1 else 2] += p
    out[3] += p * un
    return out


def _random_state(rng, gamma=1.4):
    rho = rng.uniform(0.2, 5.0)
    u, v = rng.normal(scale=1.5, size=2)
    p = rng.uniform(0.1, 10.0)
    return np.array([rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)])


@pytest.mark.parametrize("direction", [1, 2])
def test_force_matches_toro_combined_form(oracle, direction):
    """FORCE = 1/4 [F_L + 2 F(U_RI) + F_R - (h/dt)(U_R - U_L)],
    U_RI = 1/2 (U_L + U_R) - 1/2 (dt/h) (F_R - F_L)  (Toro, Riemann Solvers, 3rd ed., 7.4.2)."""
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(512):
        ul, ur = _random_state(rng), _random_state(rng)
        c = rng.uniform(0.02, 0.3)
        fl, fr = _euler_flux_conserved(ul, direction), _euler_flux_conserved(ur, direction)
        uri = 0.5 * (ul + ur) - 0.5 * c * (fr - fl)
        if uri[0] <= 0.0:
            continue
        want = 0.25 * (fl + 2.0 * _euler_flux_conserved(uri, direction) + fr - (ur - ul) / c)
        got = oracle.force(ul, ur, direction, c)
        scale = max(np.abs(fl).max(), np.abs(fr).max(), np.abs(ur - ul).max() / c)
        worst = max(worst, float(np.max(np.abs(got - want)) / scale))
    assert worst <= 1e-13


@pytest.mark.parametrize("direction", [1, 2])
def test_force_on_a_contact_is_the_linear_advection_flux(oracle, direction):
    """Equal velocity and pressure, different density: along this family the Euler flux
    is affine, F(U) = un U + (0, p, 0, p un) (p on the normal momentum), and FORCE is the
    linear closed form 1/2 un (U_L + U_R) + P - 1/4 (c un^2 + 1/c) (U_R - U_L)."""
    gamma = 1.4
    rng = np.random.default_rng(5)
    for _ in range(64):
        u, v, p = rng.normal(), rng.normal(), rng.uniform(0.5, 3.0)
        rl, rr = rng.uniform(0.3, 3.0, size=2)
        c = rng.uniform(0.05, 0.4)
        mk = lambda rho: np.array([rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)])
        ul, ur = mk(rl), mk(rr)
        un = u if direction == 1 else v
        P = np.zeros(4)
        P[1 if direction == 1 else 2] = p
        P[3] = p * un
        want = 0.5 * un * (ul + ur) + P - 0.25 * (c * un * un + 1.0 / c) * (ur - ul)
        got = oracle.force(ul, ur, direction, c)
        assert np.allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
