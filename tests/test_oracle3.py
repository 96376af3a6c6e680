"""The 3-D part of the CPU oracle (code/snapshot/warp_weno5_3d_rk3) against what
the mathematics and the pinned 2-D path fix:

* L R = I for the 5 x 5 eigenvector matrices of all three directions.
* EVILIN and the five-variable HLLC are consistent (F(u, u) = f(u)).
* A uniform flow is preserved by one SSPRK3 step.
* Reduction to 2-D: a state that does not depend on z (w = 0) advanced with
  HLLC reproduces the 2-D oracle's step (itself pinned to the reference's Sod
  record), since the z-Gauss points of a z-constant field coincide.
* Exchanging two axes (and the matching velocity components) commutes with
  the step for states that do not vary along the third axis.  For general
  states it does not: the reference's z faces run their transverse Gauss
  passes y then x (code/snapshot/warp_weno5_3d_rk3/solver.py:256-270) while
  x faces run y then z, and two nonlinear WENO passes do not commute; the
  full-state exchange is therefore checked to differ only at the size such
  an ordering produces.
"""
from __future__ import annotations

import numpy as np
import pytest

GAMMA = 1.4


def state3(nx, ny, nz, seed=0):
    rng = np.random.default_rng(seed)
    x = (np.arange(-3, nx + 3) + 0.5) / nx
    y = (np.arange(-3, ny + 3) + 0.5) / ny
    z = (np.arange(-3, nz + 3) + 0.5) / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) + 0.5 * (X + 0.6 * Y + 0.3 * Z > 0.9)
    u = 0.3 * np.cos(2 * np.pi * Y) + 0.1 * (X > 0.6)
    v = -0.2 * np.sin(2 * np.pi * X) + 0.05 * rng.normal(size=X.shape)
    w = 0.1 * np.sin(2 * np.pi * Z)
    p = 1.0 + 0.2 * np.cos(2 * np.pi * (X - Y)) + 0.4 * (Y > 0.45)
    E = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return np.ascontiguousarray(np.stack([rho, rho * u, rho * v, rho * w, E], axis=-1))


def test_eigenvectors_are_inverse(oracle):
    rng = np.random.default_rng(1)
    for _ in range(10):
        rho, u, v, w, p = rng.uniform(0.5, 2), rng.normal(), rng.normal(), rng.normal(), rng.uniform(0.5, 2)
        s = np.array([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)])
        for d in (1, 2, 3):
            L, R = oracle.eigen3d(s, d)
            assert np.allclose(L @ R, np.eye(5), atol=1e-12)


@pytest.mark.parametrize("riemann", ["evilin", "hllc"])
def test_3d_fluxes_are_consistent(oracle, riemann):
    rng = np.random.default_rng(2)
    for _ in range(10):
        rho, u, v, w, p = rng.uniform(0.5, 2), rng.normal(), rng.normal(), rng.normal(), rng.uniform(0.5, 2)
        q = np.array([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)])
        for d in (1, 2, 3):
            un = (u, v, w)[d - 1]
            f = np.array([rho * un, rho * un * u, rho * un * v, rho * un * w, un * (q[4] + p)])
            f[d] += p
            assert np.allclose(oracle.flux3d(q, q, d, 0.1, riemann), f, rtol=1e-13, atol=1e-13)


def _prim5(rho, u, v, w, p):
    return np.array([rho, rho * u, rho * v, rho * w, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def _hllc_pairs(rng, count):
    """Left / right primitive pairs spread over the HLLC cases: supersonic either way,
    star state on either side of the contact (normal speeds up to Mach ~3)."""
    for _ in range(count):
        yield (rng.uniform(0.2, 3.0), rng.normal(0, 2.0), rng.normal(), rng.normal(), rng.uniform(0.1, 3.0)), \
              (rng.uniform(0.2, 3.0), rng.normal(0, 2.0), rng.normal(), rng.normal(), rng.uniform(0.1, 3.0))


def test_hllc5_rotational_invariance(oracle):
    """The y and z fluxes are the x flux of the states with the normal velocity
    moved into the x slot (and back): a wrong velocity index in the wave speeds,
    the contact speed or the star momentum breaks it."""
    rng = np.random.default_rng(11)
    for (l, r) in _hllc_pairs(rng, 200):
        for d, perm in ((2, [0, 2, 1, 3, 4]), (3, [0, 3, 2, 1, 4])):
            ql, qr = _prim5(*l), _prim5(*r)
            fd = oracle.flux3d(ql, qr, d, 0.1, "hllc")
            fx = oracle.flux3d(ql[perm], qr[perm], 1, 0.1, "hllc")
            assert np.allclose(fd, fx[perm], rtol=1e-14, atol=1e-14), (d, l, r)


def test_hllc5_carries_tangential_velocity_with_the_mass_flux(oracle):
    """In every HLLC region the tangential momentum fluxes are the mass flux times
    the tangential velocity of the upwind side of the contact (the star states keep
    the side's tangential velocities)."""
    rng = np.random.default_rng(12)
    regions = set()
    for (l, r) in _hllc_pairs(rng, 300):
        ql, qr = _prim5(*l), _prim5(*r)
        f = oracle.flux3d(ql, qr, 1, 0.1, "hllc")
        a_l, a_r = np.sqrt(GAMMA * l[4] / l[0]), np.sqrt(GAMMA * r[4] / r[0])
        s_l, s_r = min(l[1] - a_l, r[1] - a_r), max(l[1] + a_l, r[1] + a_r)
        if s_l >= 0.0 or s_r <= 0.0:
            side = l if s_l >= 0.0 else r
            regions.add("supersonic")
        else:
            num = r[4] - l[4] + l[0] * l[1] * (s_l - l[1]) - r[0] * r[1] * (s_r - r[1])
            s_star = num / (l[0] * (s_l - l[1]) - r[0] * (s_r - r[1]))
            side = l if s_star >= 0.0 else r
            regions.add("star-left" if s_star >= 0.0 else "star-right")
        for c in (2, 3):
            assert f[c] == pytest.approx(f[0] * side[c], rel=1e-12, abs=1e-13), (l, r, c)
    assert regions == {"supersonic", "star-left", "star-right"}


def test_hllc5_resolves_a_stationary_contact(oracle):
    """Equal pressure, zero normal velocity, any density and tangential velocities:
    the flux is the pressure alone (HLLC keeps contacts exact)."""
    rng = np.random.default_rng(13)
    for _ in range(50):
        p = rng.uniform(0.1, 3.0)
        ql = _prim5(rng.uniform(0.2, 3.0), 0.0, rng.normal(), rng.normal(), p)
        qr = _prim5(rng.uniform(0.2, 3.0), 0.0, rng.normal(), rng.normal(), p)
        for d in (1, 2, 3):
            perm = {1: [0, 1, 2, 3, 4], 2: [0, 2, 1, 3, 4], 3: [0, 3, 2, 1, 4]}[d]
            f = oracle.flux3d(ql[perm], qr[perm], d, 0.1, "hllc")
            want = np.zeros(5)
            want[d] = p
            assert np.allclose(f, want, rtol=0, atol=1e-14), (d, f)


@pytest.mark.parametrize("scheme", ["js", "sr64"])
def test_3d_uniform_flow(oracle, mlp64, scheme):
    n = 6
    q = np.zeros((n + 6, n + 6, n + 6, 5))
    q[...] = [1.3, 1.3 * 0.2, -1.3 * 0.1, 1.3 * 0.3, 1.0 / 0.4 + 0.5 * 1.3 * (0.04 + 0.01 + 0.09)]
    out = oracle.step3d(q, n, n, n, 0.1, 0.1, 0.1, 0.01, scheme=scheme, mlp=mlp64 if scheme == "sr64" else None)
    assert np.max(np.abs(out - q)) <= 1e-14


@pytest.mark.parametrize("scheme", ["js", "sr64", "sr32"])
def test_3d_reduces_to_2d_for_z_invariant_states(oracle, mlp64, mlp32, scheme):
    nx, ny, nz = 20, 16, 5
    q2 = state3(nx, ny, 1)[0]                       # (ny+6, nx+6, 5) slice, w = 0 below
    q2[..., 3] = 0.0
    q2[..., 4] = (1.0 + 0.2 * np.cos(np.arange(ny + 6)[:, None] * 0.3)) / 0.4 + \
        0.5 * (q2[..., 1] ** 2 + q2[..., 2] ** 2) / q2[..., 0]
    q3 = np.ascontiguousarray(np.broadcast_to(q2, (nz + 6,) + q2.shape))
    dx, dy, dz, dt = 1.0 / nx, 1.0 / ny, 1.0 / nz, 2e-3
    mlp = {"sr64": mlp64, "sr32": mlp32}.get(scheme)
    out3 = oracle.step3d(q3, nx, ny, nz, dx, dy, dz, dt, scheme=scheme, riemann="hllc", mlp=mlp)
    u2 = np.ascontiguousarray(q2[..., [0, 1, 2, 4]])
    out2 = oracle.step2d(u2, nx, ny, dx, dy, GAMMA, dt, scheme=scheme, riemann="hllc", boundary="transmissive",
                         mlp=mlp)
    a = out3[3:-3, 3:-3, 3:-3][..., [0, 1, 2, 4]]
    b = out2[3:-3, 3:-3][None]
    err = np.max(np.abs(a - b)) / np.max(np.abs(b))
    assert err <= 2e-13, err
    assert np.max(np.abs(out3[3:-3, 3:-3, 3:-3, 3])) <= 1e-13


def _exchange(oracle, q, n, axes, perm, riemann, scheme="js", mlp=None):
    h, dt = 1.0 / n, 1e-3
    out = oracle.step3d(q, n, n, n, h, h, h, dt, scheme=scheme, riemann=riemann, mlp=mlp)
    qt = np.ascontiguousarray(np.swapaxes(q, *axes)[..., perm])
    outt = oracle.step3d(qt, n, n, n, h, h, h, dt, scheme=scheme, riemann=riemann, mlp=mlp)
    back = np.swapaxes(outt, *axes)[..., perm]
    return np.max(np.abs(back[3:-3, 3:-3, 3:-3] - out[3:-3, 3:-3, 3:-3])) / np.max(np.abs(out))


XY = ((1, 2), [0, 2, 1, 3, 4])      # array axes (z, y, x, comp): y <-> x
XZ = ((0, 2), [0, 3, 2, 1, 4])      # z <-> x


@pytest.mark.parametrize("riemann", ["evilin", "hllc"])
@pytest.mark.parametrize("scheme", ["js", "sr64"])
def test_3d_axis_exchange_commutes(oracle, mlp64, riemann, scheme):
    n = 10
    q = state3(n, n, n, 3)
    mlp = mlp64 if scheme == "sr64" else None
    qz = np.ascontiguousarray(np.broadcast_to(q[5:6], q.shape))          # no z dependence, w != 0
    assert _exchange(oracle, qz, n, *XY, riemann, scheme, mlp) <= 1e-14
    qy = np.ascontiguousarray(np.broadcast_to(q[:, 5:6], q.shape))       # no y dependence
    assert _exchange(oracle, qy, n, *XZ, riemann, scheme, mlp) <= 1e-14


def test_3d_transverse_order_is_not_symmetric(oracle):
    """With all three axes active the exchanges differ by the ordering effect
    of the transverse passes (proportional to dt, far above rounding)."""
    n = 10
    q = state3(n, n, n, 3)
    for axes, perm in (XY, XZ):
        err = _exchange(oracle, q, n, axes, perm, "hllc")
        assert 1e-8 <= err <= 1e-4, err


def test_shockbubble_bench_state():
    """The 3-D benchmark's initial state (a planar Mach-3 shock into rho = 1, p = 1,
    striking a light bubble, density ratio 0.138, radius ny/4, in pressure equilibrium):
    the shocked cells satisfy the Rankine-Hugoniot conditions with the quiescent gas at
    the shock speed 3 sqrt(1.4) (mass, momentum and energy fluxes, restated here), the
    unshocked gas has p = 1 inside and outside the bubble, and the bubble's mass deficit
    is (1 - 0.138) (4/3) pi r^3 within the 3x3x3 quadrature's resolution of the sphere."""
    from weno_sr_publishing_amd.initial import make_shockbubble_bench_state
    gamma, mach = 1.4, 3.0
    nx, ny, nz = 112, 44, 44
    q = make_shockbubble_bench_state(nx, ny, nz)[3:-3, 3:-3, 3:-3]
    assert q.shape == (nz, ny, nx, 5)
    rho = q[..., 0]
    vel = q[..., 1:4] / rho[..., None]
    p = (gamma - 1.0) * (q[..., 4] - 0.5 * rho * np.sum(vel * vel, axis=-1))
    post = q[:, :, :3]                                   # cells entirely behind the shock (x < 3)
    assert np.ptp(post, axis=(0, 1, 2)).max() == 0.0
    r2, m2, e2 = post[0, 0, 0, 0], post[0, 0, 0, 1], post[0, 0, 0, 4]
    u2 = m2 / r2
    p2 = (gamma - 1.0) * (e2 - 0.5 * r2 * u2 * u2)
    s = mach * np.sqrt(gamma * 1.0 / 1.0)
    r1, u1, p1 = 1.0, 0.0, 1.0
    e1 = p1 / (gamma - 1.0)
    assert abs(r2 * (u2 - s) - r1 * (u1 - s)) <= 1e-13 * r2 * s
    assert abs(r2 * u2 * (u2 - s) + p2 - (r1 * u1 * (u1 - s) + p1)) <= 1e-13 * p2
    assert abs(e2 * (u2 - s) + p2 * u2 - (e1 * (u1 - s) + p1 * u1)) <= 1e-13 * e2 * s
    ahead = q[:, :, 6:]
    assert np.all(ahead[..., 1:4] == 0.0)
    assert np.max(np.abs(p[:, :, 6:] - 1.0)) <= 1e-14
    assert np.all(vel[..., 1:] == 0.0)
    deficit = float(np.sum(1.0 - ahead[..., 0]))
    radius = 0.25 * ny
    assert abs(deficit / ((1.0 - 0.138) * 4.0 / 3.0 * np.pi * radius ** 3) - 1.0) <= 2e-3
    # quadrature weights sum to 8 up to rounding: the extremes are 0.138 and 1 to 1e-15
    assert abs(ahead[..., 0].min() - 0.138) <= 1e-15 and abs(ahead[..., 0].max() - 1.0) <= 1e-15
