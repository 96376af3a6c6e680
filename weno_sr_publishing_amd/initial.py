"""Initial-condition hooks (host numpy, excluded from timing like the reference).

* Four-quadrant 2-D Riemann problems C.3-C.6:
  code/snapshot/for_paper_results/run_quadrant.py:20-106 (states, region rule,
  tensor-product Gauss-Legendre cell averages via numpy's ``leggauss``).
* Planar Sod strip at t=0: code/snapshot/teacherfree_lab_weno5/warp_sod_validation.py:48-62,176-210
  (15-point cell averages with the reference's own quadrature table).
* ``pad_periodic``: code/snapshot/for_paper_results/run_vortex.py:34-35.
* Periodic isentropic vortex on [-10, 10]^2 (WENO7 grid, ghost 4):
  code/snapshot/weno7_point_rk4_shu/point_rk4.py:81-117 (``leggauss`` cell averages).
* Double-Mach reflection on [0, 4] x [0, 1]: code/snapshot/run_double_mach_compare.py:43-88
  (Mach-10 shock at 60 degrees, 15-point reference quadrature table).
"""
from __future__ import annotations

import numpy as np

from .params import GAUSS15_W, GAUSS15_XI

# state order: 1 = upper right, 2 = upper left, 3 = lower left, 4 = lower right
CASES = {
    "c3": {"label": "C.3 (q400)", "t_end": 0.50,
           "states": ((1.5, 0.0, 0.0, 1.5), (0.5323, 1.206, 0.0, 0.3),
                      (0.138, 1.206, 1.206, 0.029), (0.5323, 0.0, 1.206, 0.3))},
    "c4": {"label": "C.4", "t_end": 0.25,
           "states": ((1.1, 0.0, 0.0, 1.1), (0.5065, 0.8939, 0.0, 0.35),
                      (1.1, 0.8939, 0.8939, 1.1), (0.5065, 0.0, 0.8939, 0.35))},
    "c5": {"label": "C.5", "t_end": 0.23,
           "states": ((1.0, -0.75, -0.5, 1.0), (2.0, -0.75, 0.5, 1.0),
                      (1.0, 0.75, 0.5, 1.0), (3.0, 0.75, -0.5, 1.0))},
    "c6": {"label": "C.6", "t_end": 0.30,
           "states": ((1.0, 0.75, -0.5, 1.0), (2.0, 0.75, 0.5, 1.0),
                      (1.0, -0.75, 0.5, 1.0), (3.0, -0.75, -0.5, 1.0))},
}


def conserved_from_primitive(rho, u, v, p, gamma):
    out = np.empty(np.shape(rho) + (4,), dtype=np.float64)
    out[..., 0] = rho
    out[..., 1] = rho * u
    out[..., 2] = rho * v
    out[..., 3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return out


def quadrant_primitive(x, y, states):
    if len(states) != 4:
        raise ValueError("a quadrant problem requires four primitive states")
    rho = np.empty_like(x, dtype=np.float64)
    vx = np.empty_like(x, dtype=np.float64)
    vy = np.empty_like(x, dtype=np.float64)
    pressure = np.empty_like(x, dtype=np.float64)
    regions = ((x >= 0.5) & (y >= 0.5), (x < 0.5) & (y >= 0.5),
               (x < 0.5) & (y < 0.5), (x >= 0.5) & (y < 0.5))
    for mask, (rho_q, vx_q, vy_q, p_q) in zip(regions, states):
        rho[mask] = rho_q
        vx[mask] = vx_q
        vy[mask] = vy_q
        pressure[mask] = p_q
    return rho, vx, vy, pressure


def make_quadrant_state(nx: int, ny: int, ghost: int, definition: dict, quadrature: int = 15,
                        x_min: float = 0.0, x_max: float = 1.0, y_min: float = 0.0,
                        y_max: float = 1.0, gamma: float = 1.4) -> np.ndarray:
    """Tensor-product Gauss-Legendre cell averages on the padded grid."""
    if quadrature < 1:
        raise ValueError("quadrature order must be at least one")
    dx = (x_max - x_min) / nx
    dy = (y_max - y_min) / ny
    xi, wi = np.polynomial.legendre.leggauss(quadrature)
    jj, ii = np.indices((ny + 2 * ghost, nx + 2 * ghost))
    xc = x_min + (ii - ghost + 0.5) * dx
    yc = y_min + (jj - ghost + 0.5) * dy
    state = np.zeros((ny + 2 * ghost, nx + 2 * ghost, 4), dtype=np.float64)
    for sx, wx in zip(xi, wi):
        x = xc + 0.5 * dx * float(sx)
        for sy, wy in zip(xi, wi):
            y = yc + 0.5 * dy * float(sy)
            rho, vx, vy, pressure = quadrant_primitive(x, y, definition["states"])
            state += 0.25 * float(wx) * float(wy) * conserved_from_primitive(rho, vx, vy, pressure, gamma)
    return state


def pad_periodic(interior_state: np.ndarray, ghost: int) -> np.ndarray:
    return np.pad(interior_state, ((ghost, ghost), (ghost, ghost), (0, 0)), mode="wrap")


def make_quadrant_initial(nx: int, ny: int, case: str = "c4", ghost: int = 3,
                          quadrature: int = 15) -> np.ndarray:
    """Initial state exactly as the timing harness builds it
    (code/snapshot/for_paper_results/run_weno5_timing.py:86-104): averages on the
    unit square, interior re-padded periodically to ``ghost`` cells."""
    definition = CASES[case]
    full = make_quadrant_state(nx, ny, 4, definition, quadrature)
    inner = full[4: 4 + ny, 4: 4 + nx, :]
    return pad_periodic(inner, ghost)


def _sod_states():
    return (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)


def _primitive_to_conserved_axis(rho, normal_vel, pressure, gamma, axis):
    energy = 0.5 * rho * normal_vel * normal_vel + pressure / (gamma - 1.0)
    if axis == "y":
        return np.array([rho, 0.0, rho * normal_vel, energy], dtype=np.float64)
    return np.array([rho, rho * normal_vel, 0.0, energy], dtype=np.float64)


def _cell_average_sod0(coord_center: float, h: float, gamma: float, axis: str) -> np.ndarray:
    left, right = _sod_states()
    state = np.zeros(4, dtype=np.float64)
    for xi, weight in zip(GAUSS15_XI, GAUSS15_W):
        coord = coord_center + 0.5 * h * float(xi)
        rho, vel, pressure = left if coord < 0.0 else right
        state += float(weight) * _primitive_to_conserved_axis(rho, vel, pressure, gamma, axis)
    return 0.5 * state


def make_sod_state(params, axis: str = "x") -> np.ndarray:
    """Sod strip centred on 0 at t = 0 (warp_sod_validation.py:191-210)."""
    ny_total, nx_total, _ = params.padded_shape
    u = np.zeros(params.padded_shape, dtype=np.float64)
    if axis == "y":
        y0 = -0.5 * params.y_length
        con = np.array([_cell_average_sod0(y0 + (j - params.ghost + 0.5) * params.dy, params.dy,
                                           params.gamma, axis) for j in range(ny_total)])
        u[:, :, :] = con[:, None, :]
    else:
        x0 = -0.5 * params.x_length
        con = np.array([_cell_average_sod0(x0 + (i - params.ghost + 0.5) * params.dx, params.dx,
                                           params.gamma, axis) for i in range(nx_total)])
        u[:, :, :] = con[None, :, :]
    return u


# ---------------------------------------------------------------- 1-D GSTE
# Gaussian / square / triangle / semi-ellipse profile on [-1, 1]
# (code/snapshot/for_paper_results/run_gste.py:38-100).
GSTE_X_MIN = -1.0
GSTE_X_MAX = 1.0
GSTE_LENGTH = GSTE_X_MAX - GSTE_X_MIN


def _gste_wrap(x: np.ndarray) -> np.ndarray:
    return GSTE_X_MIN + np.mod(x - GSTE_X_MIN, GSTE_LENGTH)


def gste_point(x) -> np.ndarray:
    x = _gste_wrap(np.asarray(x, dtype=np.float64))
    delta = 0.005
    beta = np.log(2.0) / (36.0 * delta * delta)
    z, a, alpha = -0.7, 0.5, 10.0

    def gaussian(center):
        return np.exp(-beta * (x - center) ** 2)

    def ellipse(center):
        return np.sqrt(np.maximum(1.0 - alpha * alpha * (x - center) ** 2, 0.0))

    out = np.zeros_like(x)
    mask = (-0.8 < x) & (x < -0.6)
    out[mask] = (gaussian(z - delta)[mask] + 4.0 * gaussian(z)[mask] + gaussian(z + delta)[mask]) / 6.0
    mask = (-0.4 < x) & (x < -0.2)
    out[mask] = 1.0
    mask = (0.0 < x) & (x < 0.2)
    out[mask] = 1.0 - np.abs(10.0 * (x[mask] - 0.1))
    mask = (0.4 < x) & (x < 0.6)
    out[mask] = (ellipse(a - delta)[mask] + 4.0 * ellipse(a)[mask] + ellipse(a + delta)[mask]) / 6.0
    return out


def gste_cell_averages(nx: int, time: float, quadrature: int = 15) -> tuple[np.ndarray, np.ndarray]:
    """Cell centres and Gauss-Legendre cell averages of the profile advected by ``time``."""
    dx = GSTE_LENGTH / nx
    centers = GSTE_X_MIN + (np.arange(nx) + 0.5) * dx
    xi, weights = np.polynomial.legendre.leggauss(quadrature)
    samples = centers[:, None] + 0.5 * dx * xi[None, :] - time
    averages = 0.5 * np.sum(weights[None, :] * gste_point(samples), axis=1)
    return centers, averages


# ---------------------------------------------------------------- isentropic vortex
def _pdelta(d: np.ndarray, length: float) -> np.ndarray:
    return d - length * np.round(d / length)


def isentropic_vortex_conserved(x, y, t: float, gamma: float = 1.4) -> np.ndarray:
    """point_rk4.py:81-99: vortex of strength 5 advected with (1, 1), period 20."""
    beta, length = 5.0, 20.0
    dx = _pdelta(x - 1.0 * t, length)
    dy = _pdelta(y - 1.0 * t, length)
    r2 = dx * dx + dy * dy
    d_t = -((gamma - 1.0) * beta * beta) / (8.0 * gamma * np.pi * np.pi) * np.exp(1.0 - r2)
    rho = 1.0 + (np.power(1.0 + d_t, 1.0 / (gamma - 1.0)) - 1.0)
    p = 1.0 + (np.power(1.0 + d_t, gamma / (gamma - 1.0)) - 1.0)
    factor = beta / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r2))
    u = 1.0 - factor * dy
    v = 1.0 + factor * dx
    return conserved_from_primitive(rho, u, v, p, gamma)


def make_vortex_state(nx: int, ny: int, t: float = 0.0, quadrature: int = 15, ghost: int = 4,
                      x_min: float = -10.0, x_max: float = 10.0, y_min: float = -10.0, y_max: float = 10.0,
                      gamma: float = 1.4) -> np.ndarray:
    """point_rk4.py:102-117 cell_average_state: tensor-product Gauss-Legendre
    cell averages over the whole padded grid (ghost cells included)."""
    xi, wi = np.polynomial.legendre.leggauss(quadrature)
    dx = (x_max - x_min) / float(nx)
    dy = (y_max - y_min) / float(ny)
    jj, ii = np.indices((ny + 2 * ghost, nx + 2 * ghost))
    xc = x_min + (ii - ghost + 0.5) * dx
    yc = y_min + (jj - ghost + 0.5) * dy
    out = np.zeros((ny + 2 * ghost, nx + 2 * ghost, 4), dtype=np.float64)
    for sx, wx in zip(xi.astype(np.float64), wi.astype(np.float64)):
        x = xc + 0.5 * dx * float(sx)
        for sy, wy in zip(xi.astype(np.float64), wi.astype(np.float64)):
            y = yc + 0.5 * dy * float(sy)
            out += 0.25 * float(wx) * float(wy) * isentropic_vortex_conserved(x, y, t, gamma)
    return out


# ---------------------------------------------------------------- double-Mach reflection
DOUBLE_MACH_POST = (8.0, 8.25 * np.sqrt(3.0) * 0.5, -8.25 * 0.5, 116.5)
DOUBLE_MACH_PRE = (1.4, 0.0, 0.0, 1.0)


def double_mach_conserved(x, y, t: float, gamma: float = 1.4) -> np.ndarray:
    """run_double_mach_compare.py:43-49: post-shock state left of x0 + (y + 20 t)/sqrt(3)."""
    root3 = np.sqrt(3.0)
    shock_x = 1.0 / 6.0 + y / root3 + 20.0 * t / root3
    post = conserved_from_primitive(*DOUBLE_MACH_POST, gamma)
    pre = conserved_from_primitive(*DOUBLE_MACH_PRE, gamma)
    left = (np.asarray(x) < shock_x)[..., None]
    return np.where(left, post, pre)


def make_double_mach_state(nx: int = 1200, ny: int = 300, ghost: int = 4, quadrature: int = 15,
                           x_length: float = 4.0, y_length: float = 1.0, gamma: float = 1.4) -> np.ndarray:
    """run_double_mach_compare.py:69-88 make_initial_state: cell averages (the
    reference's 15-point table) over the padded grid, origin at (0, 0)."""
    if quadrature == 15:
        xi, wi = GAUSS15_XI, GAUSS15_W
    elif quadrature == 1:
        xi, wi = np.array([0.0]), np.array([2.0])
    else:
        raise ValueError("quadrature must be 1 or 15")
    dx = x_length / nx
    dy = y_length / ny
    jj, ii = np.indices((ny + 2 * ghost, nx + 2 * ghost))
    xc = (ii - ghost + 0.5) * dx
    yc = (jj - ghost + 0.5) * dy
    state = np.zeros((ny + 2 * ghost, nx + 2 * ghost, 4), dtype=np.float64)
    for ax, wx in zip(xi, wi):
        x = xc + 0.5 * dx * float(ax)
        for ay, wy in zip(xi, wi):
            y = yc + 0.5 * dy * float(ay)
            state += float(wx) * float(wy) * double_mach_conserved(x, y, 0.0, gamma)
    return 0.25 * state


# ---------------------------------------------------------------- 3-D shock-bubble
_GAUSS3_NODES = (-0.77459666924148340427791481488384306430816650390625, 0.0,
                 0.77459666924148340427791481488384306430816650390625)
_GAUSS3_WEIGHTS = (0.55555555555555546920487586248782463371753692626953125,
                   0.88888888888888917261255073754000477492809295654296875,
                   0.55555555555555546920487586248782463371753692626953125)


def shockbubble_ma3_conserved(x, y, z, gamma: float = 1.4) -> np.ndarray:
    """code/snapshot/warp_weno5_3d_rk3/kernels.py:730-767: Mach-3 shock in air at
    x < 0.005 and a helium sphere (radius 0.025 at (0.035, 0.0445, 0.0445))."""
    mach, rho_air, p_air, rho_he = 3.0, 1.29, 101325.0, 0.214
    rho_post = (((gamma + 1.0) * mach * mach) / ((gamma - 1.0) * mach * mach + 2.0)) * rho_air
    p_post = ((((2.0 * gamma) * mach * mach) - (gamma - 1.0)) / (gamma + 1.0)) * p_air
    x, y, z = np.broadcast_arrays(np.asarray(x, float), np.asarray(y, float), np.asarray(z, float))
    rho = np.full(x.shape, rho_air)
    u = np.zeros(x.shape)
    p = np.full(x.shape, p_air)
    shocked = x < 0.005
    rho = np.where(shocked, rho_post, rho)
    u = np.where(shocked, 736.911, u)
    p = np.where(shocked, p_post, p)
    dxb, dyb, dzb = x - 0.035, y - 0.0445, z - 0.0445
    bubble = dxb * dxb + dyb * dyb + dzb * dzb <= 0.025 * 0.025
    rho = np.where(bubble, rho_he, rho)
    u = np.where(bubble, 0.0, u)
    p = np.where(bubble, p_air, p)
    out = np.empty(x.shape + (5,))
    out[..., 0] = rho
    out[..., 1] = rho * u
    out[..., 2] = 0.0
    out[..., 3] = 0.0
    out[..., 4] = 0.5 * rho * (u * u) + p / (gamma - 1.0)
    return out


def make_shockbubble_ma3_state(nx: int = 224, ny: int = 88, nz: int = 88, ghost: int = 3,
                               lengths=(0.225, 0.089, 0.089)) -> np.ndarray:
    """kernels.py:770-799 initialize_shockbubble_kernel: 3 x 3 x 3 Gauss-Legendre
    cell averages over the padded grid (sum over x, then y, then z nodes)."""
    dx, dy, dz = lengths[0] / nx, lengths[1] / ny, lengths[2] / nz
    kk, jj, ii = np.indices((nz + 2 * ghost, ny + 2 * ghost, nx + 2 * ghost))
    xc = (ii - ghost + 0.5) * dx
    yc = (jj - ghost + 0.5) * dy
    zc = (kk - ghost + 0.5) * dz
    total = np.zeros(xc.shape + (5,))
    for gi in range(3):
        x = xc + 0.5 * dx * _GAUSS3_NODES[gi]
        for gj in range(3):
            y = yc + 0.5 * dy * _GAUSS3_NODES[gj]
            for gk in range(3):
                z = zc + 0.5 * dz * _GAUSS3_NODES[gk]
                w = _GAUSS3_WEIGHTS[gi] * _GAUSS3_WEIGHTS[gj] * _GAUSS3_WEIGHTS[gk]
                total = total + w * shockbubble_ma3_conserved(x, y, z)
    return 0.125 * total


def make_shockbubble_bench_state(nx: int = 224, ny: int = 88, nz: int = 88, ghost: int = 3, mach: float = 3.0,
                                 density_ratio: float = 0.138, gamma: float = 1.4) -> np.ndarray:
    """A planar Mach-3 shock in a uniform gas (rho = 1, p = 1 ahead of it) striking
    a spherical light-gas bubble of radius ny/4 cells with density ratio 0.138
    in pressure equilibrium, on a unit-spacing grid (cell averages by 3 x 3 x 3
    Gauss-Legendre quadrature).  No reference counterpart: the benchmark case."""
    rho_post = ((gamma + 1.0) * mach * mach) / ((gamma - 1.0) * mach * mach + 2.0)
    p_post = (2.0 * gamma * mach * mach - (gamma - 1.0)) / (gamma + 1.0)
    c_pre = np.sqrt(gamma)
    u_post = mach * c_pre * (1.0 - 1.0 / rho_post)
    radius = 0.25 * ny
    centre = (0.5 * ny + 8.0, 0.5 * ny, 0.5 * nz)
    kk, jj, ii = np.indices((nz + 2 * ghost, ny + 2 * ghost, nx + 2 * ghost))
    xc, yc, zc = ii - ghost + 0.5, jj - ghost + 0.5, kk - ghost + 0.5
    total = np.zeros(xc.shape + (5,))
    for gi in range(3):
        x = xc + 0.5 * _GAUSS3_NODES[gi]
        for gj in range(3):
            y = yc + 0.5 * _GAUSS3_NODES[gj]
            for gk in range(3):
                z = zc + 0.5 * _GAUSS3_NODES[gk]
                w = _GAUSS3_WEIGHTS[gi] * _GAUSS3_WEIGHTS[gj] * _GAUSS3_WEIGHTS[gk]
                shocked = x < 4.0
                bubble = (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2 <= radius * radius
                rho = np.where(shocked, rho_post, np.where(bubble, density_ratio, 1.0))
                u = np.where(shocked, u_post, 0.0)
                p = np.where(shocked, p_post, 1.0)
                q = np.zeros(xc.shape + (5,))
                q[..., 0] = rho
                q[..., 1] = rho * u
                q[..., 4] = 0.5 * rho * u * u + p / (gamma - 1.0)
                total = total + w * q
    return 0.125 * total
