"""Device-side driver of the 3-D WENO5 step on one MI355X.

Mirrors the interface of code/snapshot/warp_weno5_3d_rk3 (config.py:6-54
ShockBubbleConfig, solver.py:23-346 Weno5Rk3Solver, solver.py:348-420
run_solver, binary_io.py read_step / write_step / compare_steps).  States are
FP64 conserved (rho, rho u, rho v, rho w, E) in the reference layout
(nz + 6, ny + 6, nx + 6, 5); the step runs in the HIP library through the
``wsr_euler3_*`` C ABI on torch's current stream, with dt computed on the
device and the host synchronising only every ``chunk`` steps.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import math
import struct
import time as _time
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .checkpoints import MlpParams
from .euler2d import _state_copy

DEFAULT_CHUNK = 32
SCHEMES3 = ("js", "sr64", "sr32", "srbf16")
RIEMANN3 = ("evilin", "hllc")


@dataclass(frozen=True)
class ShockBubbleConfig:
    """config.py:6-54 (the Ma = 3 shock-bubble run, SI units)."""
    nx: int = 224
    ny: int = 88
    nz: int = 88
    ghost: int = 3
    x_start: float = 0.0
    x_end: float = 0.225
    y_start: float = 0.0
    y_end: float = 0.089
    z_start: float = 0.0
    z_end: float = 0.089
    gamma: float = 1.4
    cfl: float = 0.25
    t_start: float = 0.0
    t_end: float = 2.0e-5

    @property
    def dx(self) -> float:
        return (self.x_end - self.x_start) / self.nx

    @property
    def dy(self) -> float:
        return (self.y_end - self.y_start) / self.ny

    @property
    def dz(self) -> float:
        return (self.z_end - self.z_start) / self.nz

    @property
    def padded_shape(self) -> tuple[int, int, int, int]:
        g = self.ghost
        return (self.nz + 2 * g, self.ny + 2 * g, self.nx + 2 * g, 5)

    def validate(self) -> None:
        if min(self.nx, self.ny, self.nz) < 5:
            raise ValueError("WENO5 requires at least five cells per direction")
        if self.ghost != 3:
            raise ValueError("the 3-D step requires three ghost cells")
        if self.gamma != 1.4:
            raise ValueError("the 3-D reference hard-codes gamma=1.4")
        if self.cfl <= 0.0:
            raise ValueError("CFL must be positive")
        if self.t_end < self.t_start:
            raise ValueError("t_end must not precede t_start")


def primitive3(q: np.ndarray, gamma: float = 1.4) -> np.ndarray:
    """kernels.py:29-56 conserved -> primitive (rho, u, v, w, p), no floors."""
    rho = q[..., 0]
    u, v, w = q[..., 1] / rho, q[..., 2] / rho, q[..., 3] / rho
    p = (gamma - 1.0) * (q[..., 4] - 0.5 * rho * (u * u + v * v + w * w))
    return np.stack((rho, u, v, w, p), axis=-1)


class Euler3Solver:
    """The 3-D WENO5 / SSPRK3 step (solver.py:23-346) on the HIP library.

    scheme: "js" | "sr64" | "sr32" | "srbf16"; riemann: "evilin" (the
    reference's flux) | "hllc" (five-variable HLLC, no 3-D counterpart in the
    reference).  ``mlp`` holds the 189 closure parameters of an SR scheme.
    """

    def __init__(self, config: ShockBubbleConfig, scheme: str = "js", riemann: str = "evilin",
                 mlp: MlpParams | None = None, eno_cutoff: bool = False, device: str | torch.device = "cuda",
                 check_config: bool = True):
        if check_config:
            config.validate()
        if scheme not in SCHEMES3:
            raise ValueError(f"unknown 3-D scheme {scheme!r}")
        if riemann not in RIEMANN3:
            raise ValueError(f"unknown 3-D Riemann solver {riemann!r}")
        self.config, self.scheme, self.riemann = config, scheme, riemann
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self._mlp_host = None
        if scheme != "js":
            if mlp is None:
                raise ValueError("an SR scheme needs the 189 closure parameters")
            dtype = np.float64 if scheme == "sr64" else np.float32
            self._mlp_host = np.ascontiguousarray(np.asarray(mlp.packed).astype(dtype))
        self._lib = _lib.lib()
        self.cfg = _lib.Euler3Config(
            nx=config.nx, ny=config.ny, nz=config.nz, ghost=config.ghost, scheme=_lib.SCHEMES[scheme],
            riemann=_lib.RIEMANN[riemann], boundary=_lib.BOUNDARY["transmissive"],
            eno_cutoff=1 if eno_cutoff else 0, flags=0, reserved=0, dx=config.dx, dy=config.dy, dz=config.dz,
            gamma=config.gamma, cfl=config.cfl, t_end=config.t_end,
            mlp=self._mlp_host.ctypes.data if self._mlp_host is not None else None)
        self.ws_bytes = int(self._lib.wsr_euler3_workspace_bytes(ctypes.byref(self.cfg)))
        if self.ws_bytes == 0:
            raise ValueError("invalid 3-D configuration")
        self.workspace = torch.empty(self.ws_bytes, dtype=torch.uint8, device=self.device)
        self._dts = None
        self.q: torch.Tensor | None = None
        self.time = float(config.t_start)
        self.step_count = 0
        self.dt_trace: list[dict] = []

    # ------------------------------------------------------------- helpers
    def _stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def new_state(self, host: np.ndarray | torch.Tensor) -> torch.Tensor:
        """A solver state copied from a host array or a torch tensor (the device
        initial states of ``initial_device``)."""
        return _state_copy(host, self.config.padded_shape, self.device)

    def _check(self, q: torch.Tensor) -> None:
        if q.dtype != torch.float64 or not q.is_contiguous() or tuple(q.shape) != self.config.padded_shape:
            raise ValueError("state must be a contiguous FP64 tensor of the padded shape")
        if q.device != self.device:
            raise ValueError("state on the wrong device")

    # ------------------------------------------------------------- reference-style API
    def initialize(self, host_state: np.ndarray | torch.Tensor | None = None) -> None:
        """solver.py:134-150: the Ma = 3 shock-bubble cell averages, built on the device
        (or a given state)."""
        if host_state is None:
            from .initial_device import shockbubble_ma3_state
            c = self.config
            self.q = shockbubble_ma3_state(c.nx, c.ny, c.nz, c.ghost,
                                           (c.x_end - c.x_start, c.y_end - c.y_start, c.z_end - c.z_start),
                                           device=self.device)
        else:
            self.q = self.new_state(host_state)
        self.time = float(self.config.t_start)
        self.step_count = 0
        self.dt_trace = []

    def max_speed(self, q: torch.Tensor | None = None) -> float:
        q = self.q if q is None else q
        self._check(q)
        out = ctypes.c_double(0.0)
        rc = self._lib.wsr_euler3_max_speed(ctypes.byref(self.cfg), q.data_ptr(), self.workspace.data_ptr(),
                                            self.ws_bytes, ctypes.byref(out), self._stream())
        _lib.check(rc, "wsr_euler3_max_speed")
        return float(out.value)

    def step(self, q: torch.Tensor, dt: float) -> None:
        """One SSPRK3 step of q in place with the given dt."""
        self._check(q)
        rc = self._lib.wsr_euler3_step(ctypes.byref(self.cfg), q.data_ptr(), float(dt), self.workspace.data_ptr(),
                                       self.ws_bytes, self._stream())
        _lib.check(rc, "wsr_euler3_step")

    def advance(self) -> bool:
        """solver.py:303-329 (host-side dt, one synchronisation per step)."""
        c = self.config
        if self.time == c.t_end:
            return False
        maximum = self.max_speed()
        raw_dt = 1.0e10 if maximum < 1.0e-15 else c.cfl * min(c.dx, min(c.dy, c.dz)) / maximum
        dt = raw_dt
        if self.time + dt > c.t_end:
            dt = c.t_end - self.time
        start = self.time
        self.step_count += 1
        self.time += dt
        self.step(self.q, dt)
        self.dt_trace.append({"step": self.step_count, "time_start": start, "raw_dt": raw_dt, "dt": dt,
                              "time_end": self.time, "max_speed": maximum})
        return True

    def primitive_host(self) -> np.ndarray:
        c, g = self.config, self.config.ghost
        q = self.q.cpu().numpy()[g: g + c.nz, g: g + c.ny, g: g + c.nx]
        return np.ascontiguousarray(primitive3(q, c.gamma))

    def diagnostics(self) -> dict:
        """solver.py:337-345."""
        values = self.primitive_host()
        return {
            "nan_count": int(np.count_nonzero(~np.isfinite(values))),
            "rho_min": float(np.min(values[..., 0])),
            "rho_max": float(np.max(values[..., 0])),
            "p_min": float(np.min(values[..., 4])),
            "p_max": float(np.max(values[..., 4])),
        }

    # ------------------------------------------------------------- device time loop
    def loop_init(self, q: torch.Tensor, t_end: float, max_steps: int = 10_000_000) -> None:
        self._check(q)
        self.cfg.t_end = float(t_end)
        cap = min(max_steps, 2_000_000)
        if self._dts is None or self._dts.numel() < cap:
            self._dts = torch.zeros(cap + 64, dtype=torch.float64, device=self.device)
        rc = self._lib.wsr_euler3_loop_init(ctypes.byref(self.cfg), q.data_ptr(), int(cap),
                                            self.workspace.data_ptr(), self.ws_bytes, self._stream())
        _lib.check(rc, "wsr_euler3_loop_init")

    def loop_steps(self, q: torch.Tensor, n: int) -> None:
        self._check(q)
        rc = self._lib.wsr_euler3_loop_steps(ctypes.byref(self.cfg), q.data_ptr(), self._dts.data_ptr(), int(n),
                                             self.workspace.data_ptr(), self.ws_bytes, self._stream())
        _lib.check(rc, "wsr_euler3_loop_steps")

    def loop_steps_timed(self, q: torch.Tensor, n: int) -> tuple[list[float], list[int]]:
        """n (<= 32) steps with HIP events around the normal / first-pass / flux
        launches: ([ms] x 3, [launches] x 3) of the launches that did work."""
        self._check(q)
        ms = (ctypes.c_double * 3)()
        cnt = (ctypes.c_int32 * 3)()
        rc = self._lib.wsr_euler3_loop_steps_timed(ctypes.byref(self.cfg), q.data_ptr(), self._dts.data_ptr(),
                                                   int(n), self.workspace.data_ptr(), self.ws_bytes, self._stream(),
                                                   ms, cnt)
        _lib.check(rc, "wsr_euler3_loop_steps_timed")
        return [float(v) for v in ms], [int(v) for v in cnt]

    def set_counting(self, on: bool) -> None:
        if on:
            self.cfg.flags |= _lib.FLAG_COUNT_CLOSURE
        else:
            self.cfg.flags &= ~_lib.FLAG_COUNT_CLOSURE

    def counters(self) -> list[int]:
        """Closure counters since loop_init: [normal recons, of which closure ran,
        Gauss-pair recons, of which closure ran, lanes of normal-closure waves,
        lanes of Gauss-closure waves]."""
        out = (ctypes.c_uint64 * 6)()
        rc = self._lib.wsr_euler2d_counters(self.workspace.data_ptr(), out, 6, self._stream())
        _lib.check(rc, "wsr_euler2d_counters")
        return [int(v) for v in out]

    def loop_status(self) -> tuple[int, float, float, bool]:
        steps, t, dt, done = ctypes.c_int32(0), ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_int32(0)
        rc = self._lib.wsr_euler2d_loop_status(self.workspace.data_ptr(), ctypes.byref(steps), ctypes.byref(t),
                                               ctypes.byref(dt), ctypes.byref(done), self._stream())
        _lib.check(rc, "wsr_euler2d_loop_status")
        return int(steps.value), float(t.value), float(dt.value), bool(done.value)

    def run(self, q: torch.Tensor | None = None, t_end: float | None = None, max_steps: int = 10_000_000,
            chunk: int = DEFAULT_CHUNK) -> tuple[list[float], int, float]:
        """Advance q (default: the solver's state) in place until t == t_end or
        max_steps steps; returns (dts, steps, t)."""
        q = self.q if q is None else q
        t_end = self.config.t_end if t_end is None else t_end
        self.loop_init(q, t_end, max_steps)
        while True:
            self.loop_steps(q, chunk)
            steps, t, _, done = self.loop_status()
            if done or steps >= max_steps:
                break
        steps, t, _, _ = self.loop_status()
        dts = self._dts[:steps].cpu().tolist()
        if q is self.q:
            self.step_count, self.time = steps, t
        return dts, steps, t


# ----------------------------------------------------------------- binary steps
_HEADER = struct.Struct("=IIId")


def write_step(path, time: float, primitive: np.ndarray) -> None:
    """binary_io.py write_step: header (nz, ny, nx: uint32, time: float64), then
    the (nz, ny, nx, 5) FP64 primitive array."""
    primitive = np.ascontiguousarray(primitive, dtype=np.float64)
    if primitive.ndim != 4 or primitive.shape[-1] != 5:
        raise ValueError("primitive must have shape (nz, ny, nx, 5)")
    nz, ny, nx, _ = primitive.shape
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("wb") as f:
        f.write(_HEADER.pack(nz, ny, nx, float(time)))
        f.write(primitive.tobytes())


def read_step(path) -> tuple[float, np.ndarray]:
    """binary_io.py read_step (rejects truncated files and trailing bytes)."""
    data = Path(path).read_bytes()
    if len(data) < _HEADER.size:
        raise ValueError(f"truncated step header: {path}")
    nz, ny, nx, time = _HEADER.unpack_from(data)
    count = nz * ny * nx * 5
    payload = data[_HEADER.size:]
    if len(payload) < 8 * count:
        raise ValueError(f"truncated payload: expected {count}, got {len(payload) // 8}")
    if len(payload) > 8 * count:
        raise ValueError(f"unexpected trailing bytes in {path}")
    return float(time), np.frombuffer(payload, dtype=np.float64).reshape(nz, ny, nx, 5).copy()


def _ordered(values: np.ndarray) -> np.ndarray:
    bits = np.asarray(values, dtype=np.float64).view(np.uint64)
    sign = np.uint64(1) << np.uint64(63)
    return np.where((bits & sign) != 0, ~bits, bits | sign)


def compare_steps(candidate, reference) -> dict:
    """binary_io.py compare_steps: bitwise equality, ULP distance and error norms."""
    tc, uc = read_step(candidate)
    tr, ur = read_step(reference)
    if uc.shape != ur.shape:
        raise ValueError(f"shape mismatch: {uc.shape} != {ur.shape}")
    same = uc.view(np.uint64) == ur.view(np.uint64)
    err = np.abs(uc - ur)
    oa, ob = _ordered(uc), _ordered(ur)
    ulp = np.maximum(oa, ob) - np.minimum(oa, ob)
    denom = float(np.sum(np.abs(ur)))
    sha = lambda p: hashlib.sha256(Path(p).read_bytes()).hexdigest()
    return {
        "candidate": str(candidate), "reference": str(reference),
        "candidate_sha256": sha(candidate), "reference_sha256": sha(reference),
        "file_bitwise_identical": Path(candidate).read_bytes() == Path(reference).read_bytes(),
        "array_bitwise_identical": bool(np.all(same)),
        "header": {"shape_equal": True, "candidate_time": tc, "reference_time": tr,
                   "time_bitwise_equal": struct.pack("=d", tc) == struct.pack("=d", tr)},
        "unequal_count": int(np.count_nonzero(~same)),
        "finite_pair_count": int(np.count_nonzero(np.isfinite(uc) & np.isfinite(ur))),
        "normalized_l1": (float(np.sum(err)) / denom) if denom else (0.0 if not np.any(err) else math.inf),
        "mean_absolute_error": float(np.mean(err)),
        "max_absolute_error": float(np.max(err)),
        "max_ulp": int(np.max(ulp)),
    }


def run_solver(config: ShockBubbleConfig, device: str = "cuda:0", out_dir=None, reference=None,
               stop_step: int | None = None, model=None, eno_cutoff: bool = False, precision: str | None = None,
               riemann: str = "evilin", chunk: int = DEFAULT_CHUNK) -> dict:
    """solver.py:348-420 run_solver: initialise, advance to t_end (or stop_step),
    write step_NNNN.bin, dt_trace and run_manifest.json into out_dir.  ``model``
    selects WENO5-SR (FP64 checkpoint -> sr64, FP32 -> sr32, or ``precision``)."""
    from .checkpoints import load_mlp_params
    mlp = None
    scheme = "js"
    if model is not None:
        mlp = load_mlp_params(model)      # the checkpoint's own dtype picks the closure precision
        scheme = precision or {"fp64": "sr64", "fp32": "sr32", "bf16": "srbf16"}[mlp.precision]
    solver = Euler3Solver(config, scheme=scheme, riemann=riemann, mlp=mlp, eno_cutoff=eno_cutoff, device=device)
    wall = _time.perf_counter()
    solver.initialize()
    initialized = solver.diagnostics()
    max_steps = stop_step if stop_step is not None else 10_000_000
    dts, steps, t = solver.run(max_steps=max_steps, chunk=chunk)
    torch.cuda.synchronize(solver.device)
    elapsed = _time.perf_counter() - wall
    report = {
        "config": asdict(config), "device": str(solver.device),
        "reconstruction": f"weno5-sr-v20-reflection-symmetric-mlp-{scheme}" if mlp is not None else "weno5-js",
        "riemann": riemann, "eno_cutoff": bool(eno_cutoff), "step": steps, "time": t,
        "elapsed_seconds": elapsed, "initial_diagnostics": initialized, "final_diagnostics": solver.diagnostics(),
    }
    if out_dir is not None:
        out_dir = Path(out_dir)
        out_path = out_dir / f"step_{steps:04d}.bin"
        write_step(out_path, t, solver.primitive_host())
        report["output"] = str(out_path)
        clock, rows = 0.0, []
        for n, dt in enumerate(dts, start=1):
            rows.append(f"{n},{clock!r},{dt!r},{clock + dt!r}")
            clock += dt
        (out_dir / "dt_trace.csv").write_text("step,time_start,dt,time_end\n" + "\n".join(rows) + "\n",
                                              encoding="ascii")
        if reference is not None and Path(reference).exists() and t == config.t_end:
            report["comparison"] = compare_steps(out_path, reference)
        (out_dir / "run_manifest.json").write_text(json.dumps(report, indent=2, sort_keys=True) + "\n",
                                                   encoding="ascii")
    return report
