"""Device-side driver of the 2-D WENO7 step (Shu's downwind RK4) on one MI355X.

``Euler7Solver`` owns the workspace and calls the ``wsr_euler7_*`` C ABI on
torch's current stream.  States are FP64 in the reference layout
``(ny + 8, nx + 8, 4)`` (ghost 4).

The time loops follow code/snapshot/weno7_point_rk4_shu/point_rk4.py:537-616
(``loop="point"``) and code/snapshot/for_paper_results/solvers/
weno7_double_mach.py:434-470 (``loop="double-mach"``); dt is computed on the
device and the host synchronises only every ``chunk`` steps and at report steps.
Only the interior of the final state is defined (ghost cells of the result
are not those of the reference's in-place boundary fills).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Callable

import numpy as np
import torch

from . import _lib
from .checkpoints import Mlp7Params
from .common import interior_stats
from .euler2d import _state_copy

DEFAULT_CHUNK = 64


@dataclass(frozen=True)
class Params7:
    """point_rk4.py:26-55 Params (ghost 4)."""
    nx: int
    ny: int
    x_min: float = -10.0
    x_max: float = 10.0
    y_min: float = -10.0
    y_max: float = 10.0
    ghost: int = 4
    gamma: float = 1.4
    cfl: float = 0.4
    t_end: float = 2.0

    @property
    def x_length(self) -> float:
        return self.x_max - self.x_min

    @property
    def y_length(self) -> float:
        return self.y_max - self.y_min

    @property
    def dx(self) -> float:
        return self.x_length / float(self.nx)

    @property
    def dy(self) -> float:
        return self.y_length / float(self.ny)

    @property
    def padded_shape(self) -> tuple[int, int, int]:
        return (self.ny + 2 * self.ghost, self.nx + 2 * self.ghost, 4)


def double_mach_params7(nx: int = 1200, ny: int = 300, cfl: float = 0.4, t_end: float = 0.2) -> Params7:
    """The double-Mach grid on [0, 4] x [0, 1] (weno7_double_mach.Params as used by run_double_mach.py)."""
    return Params7(nx=nx, ny=ny, x_min=0.0, x_max=4.0, y_min=0.0, y_max=1.0, cfl=cfl, t_end=t_end)


BOUNDARY7_ALIASES = {"outflow": "transmissive"}


class Euler7Solver:
    def __init__(self, params: Params7, scheme: str = "js", riemann: str = "hllc",
                 boundary: str = "periodic", mlp: Mlp7Params | None = None, eno_cutoff: bool = False,
                 loop: str = "point", z_fallback: bool = False, device: str | torch.device = "cuda",
                 unit_gauss_h: bool = False):
        if params.ghost != 4:
            raise ValueError("the WENO7 step needs ghost = 4")
        if scheme not in ("js", "sr64", "sr32", "z"):
            raise ValueError(f"unknown WENO7 scheme {scheme!r}")
        if riemann not in _lib.RIEMANN:
            raise ValueError(f"unknown Riemann solver {riemann!r}")
        # point_rk4.py:446-455 clamps the ghost cells for every name but "periodic";
        # run_sod.py / run_quadrant.py call it "outflow"
        boundary = BOUNDARY7_ALIASES.get(boundary, boundary)
        if boundary not in _lib.BOUNDARY:
            raise ValueError(f"unsupported boundary {boundary!r}")
        if loop not in _lib.LOOP7:
            raise ValueError(f"unknown time loop {loop!r}")
        self.params, self.scheme, self.riemann, self.boundary = params, scheme, riemann, boundary
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self._mlp_host = None
        if scheme in ("sr64", "sr32"):
            if mlp is None:
                raise ValueError("WENO7-SR needs the 908 closure parameters")
            dtype = np.float64 if scheme == "sr64" else np.float32
            self._mlp_host = np.ascontiguousarray(np.asarray(mlp.packed).astype(dtype))
            if self._mlp_host.size != 908:
                raise ValueError("WENO7-SR needs a 6-24-16-16-4 closure")
        self._lib = _lib.lib()
        flags = (_lib.FLAG_Z_FALLBACK if z_fallback else 0) | (_lib.FLAG_Z7_UNIT_GAUSS_H if unit_gauss_h else 0)
        self.cfg = _lib.Euler7Config(
            nx=params.nx, ny=params.ny, ghost=4, scheme=_lib.SCHEMES[scheme], riemann=_lib.RIEMANN[riemann],
            boundary=_lib.BOUNDARY[boundary], eno_cutoff=1 if eno_cutoff else 0, flags=flags,
            loop_mode=_lib.LOOP7[loop], reserved=0, dx=params.dx, dy=params.dy, gamma=params.gamma,
            cfl=params.cfl, t_end=params.t_end,
            mlp=self._mlp_host.ctypes.data if self._mlp_host is not None else None)
        self.ws_bytes = int(self._lib.wsr_euler7_workspace_bytes(ctypes.byref(self.cfg)))
        if self.ws_bytes == 0:
            raise ValueError("invalid WENO7 configuration")
        self.workspace = torch.empty(self.ws_bytes, dtype=torch.uint8, device=self.device)
        self._dts = None

    # ------------------------------------------------------------- helpers
    def _stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def new_state(self, host: np.ndarray | torch.Tensor) -> torch.Tensor:
        """A solver state copied from a host array or a torch tensor (the device
        initial states of ``initial_device``)."""
        return _state_copy(host, self.params.padded_shape, self.device)

    def _check(self, u: torch.Tensor) -> None:
        if u.dtype != torch.float64 or not u.is_contiguous() or tuple(u.shape) != self.params.padded_shape:
            raise ValueError("state must be a contiguous FP64 tensor of the padded shape")
        if u.device != self.device:
            raise ValueError("state on the wrong device")

    # ------------------------------------------------------------- operators
    def rhs(self, u: torch.Tensor, dt: float = 0.0, time: float = 0.0, reverse: bool = False) -> torch.Tensor:
        """L(u) (or the downwind L~(u)) as a padded array with zero ghost cells."""
        self._check(u)
        out = torch.empty_like(u)
        other = torch.empty_like(u) if reverse else None
        rc = self._lib.wsr_euler7_rhs(ctypes.byref(self.cfg), u.data_ptr(), out.data_ptr(),
                                      other.data_ptr() if other is not None else None, float(dt), float(time),
                                      self.workspace.data_ptr(), self.ws_bytes, self._stream())
        _lib.check(rc, "wsr_euler7_rhs")
        return other if reverse else out

    def step(self, u: torch.Tensor, dt: float, time: float = 0.0) -> None:
        self._check(u)
        rc = self._lib.wsr_euler7_step(ctypes.byref(self.cfg), u.data_ptr(), float(dt), float(time),
                                       self.workspace.data_ptr(), self.ws_bytes, self._stream())
        _lib.check(rc, "wsr_euler7_step")

    # ------------------------------------------------------------- time loop
    def loop_init(self, u: torch.Tensor, t_end: float, max_steps: int = 10_000_000) -> None:
        self._check(u)
        self.cfg.t_end = float(t_end)
        cap = min(max_steps, 2_000_000)
        if self._dts is None or self._dts.numel() < cap:
            self._dts = torch.zeros(cap + 64, dtype=torch.float64, device=self.device)
        rc = self._lib.wsr_euler7_loop_init(ctypes.byref(self.cfg), u.data_ptr(), int(min(max_steps, cap)),
                                            self.workspace.data_ptr(), self.ws_bytes, self._stream())
        _lib.check(rc, "wsr_euler7_loop_init")

    def loop_steps(self, u: torch.Tensor, n: int) -> None:
        self._check(u)
        rc = self._lib.wsr_euler7_loop_steps(ctypes.byref(self.cfg), u.data_ptr(), self._dts.data_ptr(), int(n),
                                             self.workspace.data_ptr(), self.ws_bytes, self._stream())
        _lib.check(rc, "wsr_euler7_loop_steps")

    def loop_steps_timed(self, u: torch.Tensor, n: int) -> tuple[list[float], list[int]]:
        """n (<= 32) steps with HIP events around every sweep launch: ([ms] x 2,
        [launches] x 2) for the L-and-L~ sweeps and the L-only sweeps that did work."""
        self._check(u)
        ms = (ctypes.c_double * 2)()
        cnt = (ctypes.c_int32 * 2)()
        rc = self._lib.wsr_euler7_loop_steps_timed(ctypes.byref(self.cfg), u.data_ptr(), self._dts.data_ptr(),
                                                   int(n), self.workspace.data_ptr(), self.ws_bytes, self._stream(),
                                                   ms, cnt)
        _lib.check(rc, "wsr_euler7_loop_steps_timed")
        return [float(v) for v in ms], [int(v) for v in cnt]

    def loop_status(self) -> tuple[int, float, float, bool]:
        steps, t, dt, done = ctypes.c_int32(0), ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_int32(0)
        rc = self._lib.wsr_euler2d_loop_status(self.workspace.data_ptr(), ctypes.byref(steps), ctypes.byref(t),
                                               ctypes.byref(dt), ctypes.byref(done), self._stream())
        _lib.check(rc, "wsr_euler2d_loop_status")
        return int(steps.value), float(t.value), float(dt.value), bool(done.value)

    def set_counting(self, on: bool) -> None:
        if on:
            self.cfg.flags |= _lib.FLAG_COUNT_CLOSURE
        else:
            self.cfg.flags &= ~_lib.FLAG_COUNT_CLOSURE

    def counters(self) -> list[int]:
        """[normal recons, closures, Gauss recons, closures, lanes of closure waves (normal,
        Gauss), states that fell back to WENO7-Z]: the first six with set_counting(True),
        the last whenever the solver was built with z_fallback."""
        out = (ctypes.c_uint64 * 7)()
        rc = self._lib.wsr_euler2d_counters(self.workspace.data_ptr(), out, 7, self._stream())
        _lib.check(rc, "wsr_euler2d_counters")
        return [int(v) for v in out]

    def fallbacks(self) -> int:
        """Reconstructed states that fell back to WENO7-Z since loop_init (or in the last
        rhs / step call): each normal face state (face, side) and each Gauss-point state
        (face, side, point) counts once."""
        return self.counters()[6]

    def run(self, u: torch.Tensor, t_end: float, report_interval: int = 0,
            report: Callable[[int, float, float, dict], None] | None = None, check_health: bool = False,
            max_steps: int = 10_000_000, chunk: int = DEFAULT_CHUNK) -> tuple[list[float], int, float]:
        """Advance u in place to t_end; returns (dts, steps, t).  ``report`` /
        ``check_health`` as the reference loops (stats of the interior at step
        1, every ``report_interval`` steps and at the end; a NaN or a
        non-positive density or pressure stops the run)."""
        self.loop_init(u, t_end, max_steps)
        last = 0
        while True:
            n = chunk
            if report_interval > 0 or check_health:
                steps_now = self.loop_status()[0]
                targets = [1] + ([((steps_now // report_interval) + 1) * report_interval] if report_interval > 0 else [])
                nxt = min(tt for tt in targets if tt > steps_now) if any(tt > steps_now for tt in targets) else None
                if nxt is not None:
                    n = max(1, min(chunk, nxt - steps_now))
            self.loop_steps(u, n)
            steps, t, dt, done = self.loop_status()
            at_report = steps != last and (steps == 1 or (report_interval > 0 and steps % report_interval == 0)
                                           or done)
            if (report is not None or check_health) and at_report and steps > 0:
                last = steps
                stats = interior_stats(u.cpu().numpy(), self.params)
                if report is not None:
                    report(steps, t, dt, stats)
                if check_health and (stats["nan_count"] or stats["rho_neg"] or stats["p_neg"]):
                    break
            if done:
                break
        steps, t, _, _ = self.loop_status()
        dts = self._dts[:steps].cpu().tolist()
        return dts, steps, t
