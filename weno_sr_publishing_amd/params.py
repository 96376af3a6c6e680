"""Grid record of the two-dimensional WENO5 solver.

Mirrors ``Params`` of code/snapshot/warp_weno5_helpers.py:66-87 (same field
names, defaults and derived properties) so that reference call sites construct
it unchanged.
"""
from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class Params:
    nx: int = 450
    ny: int = 178
    ghost: int = 3
    gamma: float = 1.4
    x_length: float = 0.225
    y_length: float = 0.089
    cfl: float = 0.228
    t_end: float = 0.0002

    @property
    def dx(self) -> float:
        return self.x_length / float(self.nx)

    @property
    def dy(self) -> float:
        return self.y_length / float(self.ny)

    @property
    def padded_shape(self) -> tuple[int, int, int]:
        return (self.ny + 2 * self.ghost, self.nx + 2 * self.ghost, 4)


# The reference's 15-point Gauss-Legendre table (code/snapshot/warp_weno5_helpers.py:22-62).
# These doubles are not the correctly rounded Gauss-Legendre rule (nor numpy's
# leggauss), so initial data built from them is only bit-identical to the
# reference's when the same doubles are used.  Stored as the non-negative half
# in hexadecimal and mirrored.
_XI_HALF = (
    "0x0.0p+0", "0x1.9c0ba62ef04b6p-3", "0x1.939c69257d6b6p-2", "0x1.245676f08f3a4p-1",
    "0x1.72e6e181ab3c4p-1", "0x1.b248221fffd63p-1", "0x1.dfe24c4f8b447p-1", "0x1.f9da27c32e6d0p-1",
)
_W_HALF = (
    "0x1.9ee1575f9c975p-3", "0x1.96633f1fd02c9p-3", "0x1.7d41fa76dc25dp-3", "0x1.5484f30a86ecdp-3",
    "0x1.1dd73b4963157p-3", "0x1.b6ec9635f1143p-4", "0x1.2038260b5d042p-4", "0x1.f7dc7227a2a01p-6",
)
GAUSS15_XI = tuple([-float.fromhex(v) for v in reversed(_XI_HALF[1:])] + [float.fromhex(v) for v in _XI_HALF])
GAUSS15_W = tuple([float.fromhex(v) for v in reversed(_W_HALF[1:])] + [float.fromhex(v) for v in _W_HALF])
