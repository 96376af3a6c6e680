"""3-D WENO5 entry points of code/snapshot/warp_weno5_3d_rk3 (config.py
ShockBubbleConfig, solver.py Weno5Rk3Solver / run_solver, binary_io.py) on the
HIP 3-D step."""
from __future__ import annotations

from ..euler3d import (Euler3Solver as Weno5Rk3Solver, ShockBubbleConfig, compare_steps, read_step, run_solver,
                       write_step)

__all__ = ["ShockBubbleConfig", "Weno5Rk3Solver", "run_solver", "read_step", "write_step", "compare_steps"]
