"""Device initial states (``initial_device``) against the host builders of ``initial.py``.

The piecewise-constant problems (quadrants, Sod, double Mach, both shock-bubble cases)
are compared bit for bit: the kernels evaluate each cell average with the host's node
order and expression order and without fused multiply-adds.  The vortex goes through
exp and pow, whose device and host implementations may differ in the last bits; it is
compared to 1e-15 relative to the state's scale.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from weno_sr_publishing_amd import initial, initial_device
from weno_sr_publishing_amd.params import Params

pytestmark = pytest.mark.gpu


def _host(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("case", ["c3", "c4", "c5", "c6"])
@pytest.mark.parametrize("ghost", [3, 4])
def test_quadrant_initial_bitwise(case, ghost):
    nx, ny = 96, 80
    dev = initial_device.quadrant_initial(nx, ny, case=case, ghost=ghost)
    np.testing.assert_array_equal(_host(dev), initial.make_quadrant_initial(nx, ny, case=case, ghost=ghost))


@pytest.mark.parametrize("quadrature", [1, 5, 15])
def test_quadrant_state_bitwise(quadrature):
    nx, ny, ghost = 70, 90, 4
    definition = initial.CASES["c6"]
    dev = initial_device.quadrant_state(nx, ny, ghost, definition, quadrature=quadrature, x_min=-0.25, x_max=1.25)
    host = initial.make_quadrant_state(nx, ny, ghost, definition, quadrature=quadrature, x_min=-0.25, x_max=1.25)
    np.testing.assert_array_equal(_host(dev), host)


@pytest.mark.parametrize("axis", ["x", "y"])
def test_sod_bitwise(axis):
    params = Params(nx=64 if axis == "x" else 8, ny=8 if axis == "x" else 64, ghost=3, x_length=0.3, y_length=0.2)
    np.testing.assert_array_equal(_host(initial_device.sod_state(params, axis)), initial.make_sod_state(params, axis))


def test_double_mach_bitwise():
    nx, ny = 480, 120
    dev = initial_device.double_mach_state(nx, ny)
    np.testing.assert_array_equal(_host(dev), initial.make_double_mach_state(nx, ny))


def test_vortex_close():
    nx, ny = 64, 48
    dev = _host(initial_device.vortex_state(nx, ny, t=3.7, quadrature=7))
    host = initial.make_vortex_state(nx, ny, t=3.7, quadrature=7)
    assert np.max(np.abs(dev - host)) <= 1e-15 * np.max(np.abs(host))
    # the periodic wrap (numpy round: half to even) lands every quadrature point in the same image
    assert np.max(np.abs(dev - host)) < 1e-13


@pytest.mark.parametrize("shape", [(56, 22, 22), (112, 44, 44)])
def test_shockbubble_bench_bitwise(shape):
    nx, ny, nz = shape
    dev = initial_device.shockbubble_bench_state(nx, ny, nz)
    np.testing.assert_array_equal(_host(dev), initial.make_shockbubble_bench_state(nx, ny, nz))


def test_shockbubble_ma3_bitwise():
    nx, ny, nz = 90, 36, 36
    dev = initial_device.shockbubble_ma3_state(nx, ny, nz)
    np.testing.assert_array_equal(_host(dev), initial.make_shockbubble_ma3_state(nx, ny, nz))


def test_full_size_states_shape_and_sums():
    """The full benchmark sizes, built on the device only: shapes, finiteness and the
    totals the piecewise-constant data fix exactly."""
    u = initial_device.double_mach_state()
    assert tuple(u.shape) == (308, 1208, 4)
    assert bool(torch.isfinite(u).all())
    q = initial_device.shockbubble_bench_state()
    assert tuple(q.shape) == (94, 94, 230, 5)
    assert float(q[..., 0].min()) >= 0.138 - 1e-15 and float(q[..., 2:4].abs().max()) == 0.0
    c4 = initial_device.quadrant_initial(1200, 1200, "c4", ghost=3)
    assert tuple(c4.shape) == (1206, 1206, 4)
    # the C.4 quadrants meet on cell faces at 1200^2: every average is one state's density
    # (up to the rounding of the weight sum)
    states = np.array(initial.CASES["c4"]["states"])
    rho = _host(c4[..., 0]).reshape(-1, 1)
    assert np.min(np.abs(rho - states[None, :, 0]), axis=1).max() < 1e-14
