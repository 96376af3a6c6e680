"""Device-side driver of the 2-D WENO5 step on one MI355X.

``Euler2DSolver`` owns the workspace (a torch uint8 tensor) and calls the C ABI
of ``libwsr_hip.so`` on torch's current stream.  State tensors are FP64 in the
reference layout ``(ny + 6, nx + 6, 4)``.

The time loop (``run``) follows run_to_time of
code/snapshot/for_paper_results/solvers/weno5_hllc_mixed.py:185-212 and
code/snapshot/teacherfree_lab_weno5_v20_distance_balanced/weno5_hllc_refsym.py:462-544,
with dt computed on the device: the host only synchronises every ``chunk``
steps and at report steps.
"""
from __future__ import annotations

import ctypes
from typing import Callable

import numpy as np
import torch

from . import _lib
from .checkpoints import MlpParams
from .common import interior_stats
from .params import Params

DEFAULT_CHUNK = 64


def _state_copy(state, shape: tuple, device: torch.device) -> torch.Tensor:
    """A contiguous FP64 copy of ``state`` (host array or torch tensor) on ``device``,
    checked against the padded shape."""
    if tuple(state.shape) != tuple(shape):
        raise ValueError(f"initial state must have shape {tuple(shape)}, got {tuple(state.shape)}")
    if isinstance(state, torch.Tensor):
        return state.detach().to(device=device, dtype=torch.float64, copy=True).contiguous()
    return torch.from_numpy(np.ascontiguousarray(state, dtype=np.float64)).to(device)


def scheme_for(mlp: MlpParams | None, closure: str | None = None) -> str:
    """Closure name from the MLP precision ("js" without parameters)."""
    if closure is not None:
        return closure
    if mlp is None:
        return "js"
    return {"fp64": "sr64", "fp32": "sr32", "bf16": "srbf16"}[getattr(mlp, "precision", "fp64")]


class Euler2DSolver:
    def __init__(self, params: Params, scheme: str = "js", riemann: str = "hllc",
                 boundary: str = "transmissive", mlp: MlpParams | None = None,
                 eno_cutoff: bool = False, device: str | torch.device = "cuda",
                 workspace: torch.Tensor | None = None):
        if params.ghost != 3:
            raise ValueError("the WENO5 step needs ghost = 3")
        if boundary not in _lib.BOUNDARY:
            raise ValueError(f"unsupported boundary {boundary!r}")
        if scheme not in _lib.SCHEMES:
            raise ValueError(f"unknown scheme {scheme!r}")
        if riemann not in _lib.RIEMANN:
            raise ValueError(f"unknown Riemann solver {riemann!r}")
        self.params = params
        self.scheme = scheme
        self.riemann = riemann
        self.boundary = boundary
        self.eno_cutoff = bool(eno_cutoff)
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self._mlp_host = None
        if scheme in ("sr64", "sr32", "srbf16"):
            if mlp is None:
                raise ValueError("WENO5-SR needs MLP parameters")
            dtype = np.float64 if scheme == "sr64" else np.float32
            self._mlp_host = np.ascontiguousarray(mlp.packed.astype(dtype))
        L = _lib.lib()
        self._lib = L
        self.ws_bytes = int(L.wsr_euler2d_workspace_bytes(params.nx, params.ny, params.ghost))
        if self.ws_bytes == 0:
            raise ValueError("invalid grid")
        if workspace is not None and workspace.numel() >= self.ws_bytes and workspace.device == self.device:
            self.workspace = workspace
        else:
            self.workspace = torch.empty(self.ws_bytes, dtype=torch.uint8, device=self.device)
        self.cfg = _lib.Euler2DConfig(
            nx=params.nx, ny=params.ny, ghost=params.ghost,
            scheme=_lib.SCHEMES[scheme], riemann=_lib.RIEMANN[riemann],
            boundary=_lib.BOUNDARY[boundary], eno_cutoff=1 if eno_cutoff else 0, flags=0,
            dx=params.dx, dy=params.dy, gamma=params.gamma, cfl=params.cfl, t_end=params.t_end,
            mlp=self._mlp_host.ctypes.data if self._mlp_host is not None else None)
        self._u1 = None
        self._u2 = None
        self._dts = None

    # ------------------------------------------------------------ helpers
    def _stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def _check_state(self, u: torch.Tensor) -> None:
        if u.dtype != torch.float64 or tuple(u.shape) != self.params.padded_shape:
            raise ValueError(f"state must be float64 {self.params.padded_shape}, got {u.dtype} {tuple(u.shape)}")
        if not u.is_contiguous() or u.device != self.device:
            raise ValueError("state must be a contiguous tensor on the solver's device")

    def new_state(self, host: np.ndarray | torch.Tensor | None = None) -> torch.Tensor:
        """A fresh solver state: zeros, a copy of a host array, or a device copy of a
        torch tensor (the device initial states of ``initial_device``)."""
        if host is None:
            return torch.zeros(self.params.padded_shape, dtype=torch.float64, device=self.device)
        return _state_copy(host, self.params.padded_shape, self.device)

    # ------------------------------------------------------------ one stage
    def stage(self, u_stage: torch.Tensor, u0: torch.Tensor, u_out: torch.Tensor, dt: float, rk: int,
              boundary_time: float = 0.0) -> None:
        for t in (u_stage, u0, u_out):
            self._check_state(t)
        rc = self._lib.wsr_euler2d_stage(ctypes.byref(self.cfg), u_stage.data_ptr(), u0.data_ptr(),
                                         u_out.data_ptr(), float(dt), int(rk), float(boundary_time),
                                         self.workspace.data_ptr(), self.ws_bytes, self._stream())
        _lib.check(rc, "wsr_euler2d_stage")

    def stage_noalloc(self, u_stage: torch.Tensor, u0: torch.Tensor, u_out: torch.Tensor, dt: float, rk: int,
                      boundary_time: float = 0.0) -> None:
        """``stage`` through wsr_euler2d_stage_noalloc: the same kernels on the library's
        file-scope device scratch instead of this solver's workspace (for callers that
        cannot allocate; grids up to the scratch size, no closure counting; calls that
        overlap on one GPU race)."""
        for t in (u_stage, u0, u_out):
            self._check_state(t)
        rc = self._lib.wsr_euler2d_stage_noalloc(ctypes.byref(self.cfg), u_stage.data_ptr(), u0.data_ptr(),
                                                 u_out.data_ptr(), float(dt), int(rk), float(boundary_time),
                                                 self._stream())
        _lib.check(rc, "wsr_euler2d_stage_noalloc")

    def max_speed(self, u: torch.Tensor) -> float:
        self._check_state(u)
        out = ctypes.c_double(0.0)
        rc = self._lib.wsr_euler2d_max_speed(ctypes.byref(self.cfg), u.data_ptr(), self.workspace.data_ptr(),
                                             self.ws_bytes, ctypes.byref(out), self._stream())
        _lib.check(rc, "wsr_euler2d_max_speed")
        return float(out.value)

    def compute_dt(self, u: torch.Tensor) -> float:
        """warp_weno5_helpers.py:189-207 compute_dt_from_warp_array."""
        max_speed = self.max_speed(u)
        if max_speed < 1.0e-15:
            return 1.0e10
        return self.params.cfl * min(self.params.dx, self.params.dy) / max_speed

    # ------------------------------------------------------------ time loop
    def _buffers(self, max_steps: int) -> None:
        if self._u1 is None:
            self._u1 = self.new_state()
            self._u2 = self.new_state()
        if self._dts is None or self._dts.numel() < max_steps:
            self._dts = torch.zeros(max(max_steps, 1), dtype=torch.float64, device=self.device)

    def loop_init(self, u: torch.Tensor, t_end: float, max_steps: int) -> None:
        self._check_state(u)
        self.cfg.t_end = float(t_end)
        self._buffers(max_steps)
        rc = self._lib.wsr_euler2d_loop_init(ctypes.byref(self.cfg), u.data_ptr(), int(max_steps),
                                             self.workspace.data_ptr(), self.ws_bytes, self._stream())
        _lib.check(rc, "wsr_euler2d_loop_init")

    def loop_steps(self, u: torch.Tensor, n: int) -> None:
        self._check_state(u)
        rc = self._lib.wsr_euler2d_loop_steps(ctypes.byref(self.cfg), u.data_ptr(), self._u1.data_ptr(),
                                              self._u2.data_ptr(), self._dts.data_ptr(), int(n),
                                              self.workspace.data_ptr(), self.ws_bytes, self._stream())
        _lib.check(rc, "wsr_euler2d_loop_steps")

    def loop_steps_timed(self, u: torch.Tensor, n: int) -> tuple[float, int]:
        """loop_steps with HIP events around every sweep launch; returns (ms, launches)."""
        self._check_state(u)
        ms = ctypes.c_double(0.0)
        launches = ctypes.c_int32(0)
        rc = self._lib.wsr_euler2d_loop_steps_timed(
            ctypes.byref(self.cfg), u.data_ptr(), self._u1.data_ptr(), self._u2.data_ptr(),
            self._dts.data_ptr(), int(n), self.workspace.data_ptr(), self.ws_bytes, self._stream(),
            ctypes.byref(ms), ctypes.byref(launches))
        _lib.check(rc, "wsr_euler2d_loop_steps_timed")
        return float(ms.value), int(launches.value)

    def set_counting(self, on: bool) -> None:
        self.cfg.flags = (self.cfg.flags | _lib.FLAG_COUNT_CLOSURE) if on else (self.cfg.flags & ~_lib.FLAG_COUNT_CLOSURE)

    def counters(self) -> list[int]:
        """[normal recons, normal closures, Gauss recons, Gauss closures,
        lanes of waves running the normal closure, same for Gauss]."""
        out = (ctypes.c_uint64 * 6)()
        rc = self._lib.wsr_euler2d_counters(self.workspace.data_ptr(), out, 6, self._stream())
        _lib.check(rc, "wsr_euler2d_counters")
        return [int(v) for v in out]

    def loop_status(self) -> tuple[int, float, float, bool]:
        steps = ctypes.c_int32(0)
        t = ctypes.c_double(0.0)
        dt = ctypes.c_double(0.0)
        done = ctypes.c_int32(0)
        rc = self._lib.wsr_euler2d_loop_status(self.workspace.data_ptr(), ctypes.byref(steps), ctypes.byref(t),
                                               ctypes.byref(dt), ctypes.byref(done), self._stream())
        _lib.check(rc, "wsr_euler2d_loop_status")
        return int(steps.value), float(t.value), float(dt.value), bool(done.value)

    def dts(self, steps: int) -> list[float]:
        return self._dts[:steps].cpu().tolist()

    def run(self, u: torch.Tensor, t_end: float, report_interval: int = 0,
            report: Callable | None = None, check_health: bool = False,
            max_steps: int = 10_000_000, chunk: int = DEFAULT_CHUNK) -> tuple[list[float], int, float]:
        """Advance ``u`` in place to ``t_end``.  Returns (dts, steps, t).

        Reports (and, with ``check_health``, the NaN / non-positive early exit of
        weno5_hllc_refsym.py:528-542) happen at step 1, every ``report_interval``
        steps and at the final step, exactly as in the reference loop."""
        self.loop_init(u, t_end, max_steps)
        reporting = report is not None or check_health
        steps = 0
        t = 0.0
        last_reported = 0
        while True:
            n = chunk
            if reporting:
                if steps == 0:
                    nxt = 1
                elif report_interval > 0:
                    nxt = (steps // report_interval + 1) * report_interval
                else:
                    nxt = steps + chunk
                n = max(1, min(chunk, nxt - steps))
            self.loop_steps(u, n)
            steps, t, dt, done = self.loop_status()
            if reporting and steps > last_reported:
                at_end = t >= t_end - 1.0e-14
                hit = steps == 1 or (report_interval > 0 and steps % report_interval == 0) or at_end
                if hit:
                    last_reported = steps
                    stats = interior_stats(u.cpu().numpy(), self.params)
                    if report is not None:
                        report(steps, t, dt, stats)
                    if check_health and (stats["nan_count"] or stats["rho_neg"] or stats["p_neg"]):
                        break
            if done:
                break
        return self.dts(steps), steps, t
