"""Beta-provider protocol of the WENO7-SR point step on the HIP library.

The reference drives WENO7-SR through a *beta provider*: an object with four
methods that fill arrays of smoothness indicators (beta = 4 r of the
reflection-averaged closure, ones on a plateau), which separate kernels then
consume --
code/snapshot/teacherfree_lab_weno7_rk4_distance_balanced_fast/warp_sod/point_rk4_mlp.py
(``TorchWeno7PointBeta`` :22-56, ``allocate_arrays_mlp`` :188-191, ``compute_rhs_mlp``
:194-273, ``launch_shu_rk4_step_mlp`` :275-301, ``run_from_initial_mlp`` :303-362),
code/snapshot/for_paper_results/solvers/weno7_warp_beta.py:562-682
(``WarpWeno7PointBeta``) and the double-Mach variant of
code/snapshot/for_paper_results/solvers/weno7_double_mach.py:275-501.

``HipWeno7Beta`` is this library's closure as such a provider (HIP kernels, FP64 or
FP32 MLP).  ``compute_rhs``, ``shu_rk4_step``, ``run_point`` and ``run_double_mach``
drive the step through any provider -- a caller's own object with the four methods
that writes the torch tensors of ``arrays`` on the current stream -- with the HIP
consumer kernels (``wsr_euler7_stage_ext``, ``wsr_euler7_flux_ext``,
``wsr_euler7_rhs_from_flux``, ``wsr_euler7_shu_combine``).  For this library's own
provider the fused step (``Euler7Solver``) computes the same thing in one sweep per
right-hand side; the drop-in runners choose it unless another provider is given.
"""
from __future__ import annotations

import ctypes
from typing import Callable

import numpy as np
import torch

from . import _lib
from .checkpoints import Mlp7Params
from .common import interior_stats
from .euler2d import _state_copy

NORMAL_BETA_STRIDE = 32
CROSS_BETA_STRIDE = 256
SPEED_WORKSPACE_BYTES = 4096   # WSR_SPEED_WORKSPACE_BYTES

# point_rk4.py:524 / point_rk4_mlp.py:286: solver kind 1 is HLLC, any other EVILIN;
# "force" (the FORCE flux) is this library's addition, as in Euler7Solver
RIEMANN_KIND = {"hllc": 1, "evilin": 0}


def riemann_name(solver_kind: int | str) -> str:
    if isinstance(solver_kind, str):
        if solver_kind not in _lib.RIEMANN:
            raise ValueError(f"unknown riemann_solver {solver_kind!r}")
        return solver_kind
    return "hllc" if int(solver_kind) == 1 else "evilin"


def config(params, *, scheme: str = "js", riemann: str = "hllc", boundary: str = "periodic",
           eno_cutoff: bool = False, mlp_host: np.ndarray | None = None) -> _lib.Euler7Config:
    """The C-ABI record of the WENO7 grid ``params`` (point_rk4.Params) and method."""
    if params.ghost != 4:
        raise ValueError("the WENO7 step needs ghost = 4")
    boundary = {"outflow": "transmissive"}.get(boundary, boundary)
    if boundary not in _lib.BOUNDARY:
        raise ValueError(f"unsupported boundary {boundary!r}")
    return _lib.Euler7Config(
        nx=params.nx, ny=params.ny, ghost=4, scheme=_lib.SCHEMES[scheme], riemann=_lib.RIEMANN[riemann],
        boundary=_lib.BOUNDARY[boundary], eno_cutoff=1 if eno_cutoff else 0, flags=0,
        loop_mode=_lib.LOOP7["point"], reserved=0, dx=params.dx, dy=params.dy, gamma=params.gamma,
        cfl=params.cfl, t_end=params.t_end,
        mlp=mlp_host.ctypes.data if mlp_host is not None else None)


def _stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _expect(t: torch.Tensor, shape: tuple, name: str) -> None:
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_contiguous() \
            or tuple(t.shape) != tuple(shape) or t.device.type != "cuda":
        raise ValueError(f"arrays[{name!r}] must be a contiguous FP64 device tensor of shape {tuple(shape)}")


def array_shapes(params) -> dict[str, tuple]:
    """point_rk4.py:406-432 allocate_arrays + warp_weno7_external_overwrite_fullall.py:334-340."""
    nx, ny = params.nx, params.ny
    padded = params.padded_shape
    shapes = {name: padded for name in ("u", "u1", "u2", "u3", "rhs0", "rhs1", "rhs2", "rhs3", "rhs_t0", "rhs_t1",
                                        "x_l", "x_r", "y_l", "y_r")}
    shapes.update({"fx1": (ny, nx + 1, 4), "fx2": (ny, nx + 1, 4), "fy1": (ny + 1, nx, 4), "fy2": (ny + 1, nx, 4),
                   "speed": (nx * ny,),
                   "beta_x": (ny + 8, nx + 2, NORMAL_BETA_STRIDE), "beta_y": (ny + 2, nx + 8, NORMAL_BETA_STRIDE),
                   "beta_cross_x": (ny, nx + 2, CROSS_BETA_STRIDE),
                   "beta_cross_y": (ny + 2, nx, CROSS_BETA_STRIDE)})
    return shapes


def allocate_arrays_mlp(u0_host, params, device) -> dict[str, torch.Tensor]:
    """point_rk4_mlp.py:188-191: the step's arrays as torch device tensors (zeros),
    ``u`` a copy of the initial state."""
    dev = torch.device(device)
    if dev.type == "cuda" and dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    arrays = {name: torch.zeros(shape, dtype=torch.float64, device=dev)
              for name, shape in array_shapes(params).items() if name != "u"}
    arrays["u"] = _state_copy(u0_host, params.padded_shape, dev)
    return arrays


class HipWeno7Beta:
    """The 6-24-16-16-4 reflection-averaged closure of this library as a beta provider:
    the four fill methods of point_rk4_mlp.py:22-56 / weno7_warp_beta.py:616-682 on
    HIP kernels (``wsr_beta7_fill_normal`` / ``wsr_beta7_fill_cross``), in FP64
    (``mlp.precision`` "fp64", the reference's provider) or FP32."""

    def __init__(self, mlp: Mlp7Params, device="cuda", gamma: float = 1.4):
        self.mlp, self.gamma = mlp, float(gamma)
        self.device = torch.device(device)
        dtype = np.float64 if mlp.precision == "fp64" else np.float32
        self._host = np.ascontiguousarray(np.asarray(mlp.packed).astype(dtype))
        if self._host.size != 908:
            raise ValueError("WENO7-SR needs a 6-24-16-16-4 closure")
        self._ws: dict[torch.device, torch.Tensor] = {}

    @property
    def scheme(self) -> str:
        return "sr64" if self.mlp.precision == "fp64" else "sr32"

    def _workspace(self, dev: torch.device) -> torch.Tensor:
        ws = self._ws.get(dev)
        if ws is None:
            n = int(_lib.lib().wsr_beta7_workspace_bytes())
            ws = self._ws[dev] = torch.empty(n, dtype=torch.uint8, device=dev)
        return ws

    def _normal(self, arrays: dict, params, src_name: str, direction: int, out: str) -> None:
        shapes = array_shapes(params)
        src, beta = arrays[src_name], arrays[out]
        _expect(src, params.padded_shape, src_name)
        _expect(beta, shapes[out], out)
        cfg = config(params, scheme=self.scheme, mlp_host=self._host)
        ws = self._workspace(src.device)
        rc = _lib.lib().wsr_beta7_fill_normal(ctypes.byref(cfg), src.data_ptr(), beta.data_ptr(), direction,
                                              ws.data_ptr(), ws.numel(), _stream(src.device))
        _lib.check(rc, "wsr_beta7_fill_normal")

    def _cross(self, arrays: dict, params, right: str, left: str, direction: int, out: str) -> None:
        shapes = array_shapes(params)
        r, l, beta = arrays[right], arrays[left], arrays[out]
        _expect(r, params.padded_shape, right)
        _expect(l, params.padded_shape, left)
        _expect(beta, shapes[out], out)
        cfg = config(params, scheme=self.scheme, mlp_host=self._host)
        ws = self._workspace(r.device)
        rc = _lib.lib().wsr_beta7_fill_cross(ctypes.byref(cfg), r.data_ptr(), l.data_ptr(), beta.data_ptr(),
                                             direction, ws.data_ptr(), ws.numel(), _stream(r.device))
        _lib.check(rc, "wsr_beta7_fill_cross")

    def fill_normal_x_for(self, arrays: dict, params, src_name: str) -> None:
        self._normal(arrays, params, src_name, 1, "beta_x")

    def fill_normal_y_for(self, arrays: dict, params, src_name: str) -> None:
        self._normal(arrays, params, src_name, 2, "beta_y")

    def fill_cross_x_point(self, arrays: dict, params) -> None:
        self._cross(arrays, params, "x_r", "x_l", 1, "beta_cross_x")

    def fill_cross_y_point(self, arrays: dict, params) -> None:
        self._cross(arrays, params, "y_r", "y_l", 2, "beta_cross_y")


PROTOCOL = ("fill_normal_x_for", "fill_normal_y_for", "fill_cross_x_point", "fill_cross_y_point")


def is_fusable(beta_model) -> bool:
    """True for this library's provider with its own fill methods (the fused step then
    computes the same right-hand sides); False for any other provider."""
    if not isinstance(beta_model, HipWeno7Beta):
        return False
    return all(getattr(type(beta_model), name) is getattr(HipWeno7Beta, name) for name in PROTOCOL)


def _check_provider(beta_model) -> None:
    missing = [name for name in PROTOCOL if not callable(getattr(beta_model, name, None))]
    if missing:
        raise TypeError(f"beta_model lacks the provider methods {missing}")


def compute_rhs(arrays: dict, params, src_name: str, rhs_name: str, dt: float, *, beta_model,
                reverse_upwind: bool, riemann: str = "hllc", boundary: str = "periodic",
                boundary_time: float = 0.0, eno_cutoff: bool = False) -> None:
    """point_rk4_mlp.py:194-273 compute_rhs_mlp (weno7_double_mach.py:275-371 with the
    double-Mach boundary at ``boundary_time``): boundary fill of ``arrays[src_name]`` in
    place, then per direction the provider's normal beta, the normal face values, the
    provider's cross beta and the fluxes of both Gauss locations, and the right-hand side."""
    _check_provider(beta_model)
    shapes = array_shapes(params)
    for name in ("x_l", "x_r", "y_l", "y_r", "fx1", "fx2", "fy1", "fy2", "beta_x", "beta_y", "beta_cross_x",
                 "beta_cross_y"):
        _expect(arrays[name], shapes[name], name)
    src, rhs = arrays[src_name], arrays[rhs_name]
    _expect(src, params.padded_shape, src_name)
    _expect(rhs, params.padded_shape, rhs_name)
    L = _lib.lib()
    cfg = config(params, riemann=riemann, boundary=boundary, eno_cutoff=eno_cutoff)
    c = ctypes.byref(cfg)
    st = _stream(src.device)
    _lib.check(L.wsr_euler7_boundary(c, src.data_ptr(), float(boundary_time), st), "wsr_euler7_boundary")
    rev = 1 if reverse_upwind else 0
    for direction, axis in ((1, "x"), (2, "y")):
        fill_normal = beta_model.fill_normal_x_for if direction == 1 else beta_model.fill_normal_y_for
        fill_cross = beta_model.fill_cross_x_point if direction == 1 else beta_model.fill_cross_y_point
        fill_normal(arrays, params, src_name)
        st = _stream(src.device)
        _lib.check(L.wsr_euler7_stage_ext(c, src.data_ptr(), arrays[f"beta_{axis}"].data_ptr(),
                                          arrays[f"{axis}_l"].data_ptr(), arrays[f"{axis}_r"].data_ptr(),
                                          direction, st), "wsr_euler7_stage_ext")
        fill_cross(arrays, params)
        st = _stream(src.device)
        for loca in (1, 2):
            _lib.check(L.wsr_euler7_flux_ext(c, arrays[f"{axis}_l"].data_ptr(), arrays[f"{axis}_r"].data_ptr(),
                                             arrays[f"beta_cross_{axis}"].data_ptr(),
                                             arrays[f"f{axis}{loca}"].data_ptr(), direction, loca, float(dt), rev,
                                             st), "wsr_euler7_flux_ext")
    _lib.check(L.wsr_euler7_rhs_from_flux(c, rhs.data_ptr(), arrays["fx1"].data_ptr(), arrays["fx2"].data_ptr(),
                                          arrays["fy1"].data_ptr(), arrays["fy2"].data_ptr(),
                                          _stream(src.device)), "wsr_euler7_rhs_from_flux")


# point_rk4.py:315-403: inputs of the four Shu stage combinations and their outputs
_SHU = {1: (("u", "rhs0"), "u1"),
        2: (("u", "rhs_t0", "u1", "rhs1"), "u2"),
        3: (("u", "rhs_t0", "u1", "rhs_t1", "u2", "rhs2"), "u3"),
        4: (("u", "rhs0", "u1", "rhs1", "u2", "u3", "rhs3"), "u")}


def shu_combine(arrays: dict, params, stage: int, dt: float) -> None:
    names, out = _SHU[stage]
    for name in names + (out,):
        _expect(arrays[name], params.padded_shape, name)
    ptrs = (ctypes.c_void_p * len(names))(*[arrays[n].data_ptr() for n in names])
    cfg = config(params)
    rc = _lib.lib().wsr_euler7_shu_combine(ctypes.byref(cfg), int(stage), float(dt), ptrs, arrays[out].data_ptr(),
                                           _stream(arrays[out].device))
    _lib.check(rc, "wsr_euler7_shu_combine")


def shu_rk4_step(arrays: dict, params, dt: float, *, beta_model, riemann: str = "hllc",
                 boundary: str = "periodic", time: float = 0.0, eno_cutoff: bool = False) -> None:
    """point_rk4_mlp.py:275-301 launch_shu_rk4_step_mlp / weno7_double_mach.py:374-453
    _launch_shu_rk4_step: six right-hand sides (L and the downwind L~) and four
    stage combinations; the boundary time advances with the stages (double Mach)."""
    kw = dict(beta_model=beta_model, riemann=riemann, boundary=boundary, eno_cutoff=eno_cutoff)
    compute_rhs(arrays, params, "u", "rhs0", dt, reverse_upwind=False, boundary_time=time, **kw)
    shu_combine(arrays, params, 1, dt)
    compute_rhs(arrays, params, "u", "rhs_t0", dt, reverse_upwind=True, boundary_time=time, **kw)
    compute_rhs(arrays, params, "u1", "rhs1", dt, reverse_upwind=False, boundary_time=time + 0.5 * dt, **kw)
    shu_combine(arrays, params, 2, dt)
    compute_rhs(arrays, params, "u1", "rhs_t1", dt, reverse_upwind=True, boundary_time=time + 0.5 * dt, **kw)
    compute_rhs(arrays, params, "u2", "rhs2", dt, reverse_upwind=False, boundary_time=time + 0.5 * dt, **kw)
    shu_combine(arrays, params, 3, dt)
    compute_rhs(arrays, params, "u3", "rhs3", dt, reverse_upwind=False, boundary_time=time + dt, **kw)
    shu_combine(arrays, params, 4, dt)


class _Speed:
    """point_rk4.py:136-147 compute_dt_from_warp_array on wsr_euler7_max_speed."""

    def __init__(self, params, device: torch.device):
        self.params = params
        self.cfg = config(params)
        self.ws = torch.empty(SPEED_WORKSPACE_BYTES, dtype=torch.uint8, device=device)

    def dt(self, u: torch.Tensor) -> float:
        out = ctypes.c_double(0.0)
        rc = _lib.lib().wsr_euler7_max_speed(ctypes.byref(self.cfg), u.data_ptr(), self.ws.data_ptr(),
                                             self.ws.numel(), ctypes.byref(out), _stream(u.device))
        _lib.check(rc, "wsr_euler7_max_speed")
        max_speed = out.value
        if max_speed < 1.0e-15:
            return 1.0e10
        return self.params.cfl * min(self.params.dx, self.params.dy) / max_speed


def run_point(u0, params, *, device, riemann: str, beta_model, report_interval: int = 0,
              max_steps: int = 10_000_000, boundary: str = "periodic", eno_cutoff: bool = False,
              report: Callable[[int, float, float, dict], None] | None = None):
    """point_rk4_mlp.py:303-362 run_from_initial_mlp through the provider protocol:
    returns (final host state, dts, steps, t, stopped) with stopped set when a report
    step found NaN or a non-positive density or pressure."""
    arrays = allocate_arrays_mlp(u0, params, device)
    speed = _Speed(params, arrays["u"].device)
    t, step, dts, stopped = 0.0, 0, [], False
    while t < params.t_end and step < max_steps:
        dt = speed.dt(arrays["u"])
        if t + dt > params.t_end:
            dt = params.t_end - t
        if dt <= 0.0:
            break
        shu_rk4_step(arrays, params, dt, beta_model=beta_model, riemann=riemann, boundary=boundary,
                     eno_cutoff=eno_cutoff)
        t += dt
        step += 1
        dts.append(float(dt))
        if report_interval > 0 and (step == 1 or step % report_interval == 0 or t >= params.t_end):
            stats = interior_stats(arrays["u"].cpu().numpy(), params)
            if report is not None:
                report(step, t, dt, stats)
            if stats["nan_count"] or stats["rho_neg"] or stats["p_neg"]:
                stopped = True
                break
    return arrays["u"].cpu().numpy(), dts, step, t, stopped


def run_double_mach(initial, params, t_end: float, device, beta_model, *, report_interval: int = 100,
                    report: Callable[[int, float, float, dict], None] | None = None, riemann: str = "hllc"):
    """weno7_double_mach.py:456-501 run_to_time through the provider protocol
    (``while t < t_end - 1e-14``, dt = min(CFL dt, t_end - t), stop on a non-finite or
    non-positive dt or an unhealthy report step): (final host state, dts, steps, t)."""
    arrays = allocate_arrays_mlp(initial, params, device)
    speed = _Speed(params, arrays["u"].device)
    t, step, dts = 0.0, 0, []
    while t < t_end - 1.0e-14:
        dt = min(speed.dt(arrays["u"]), t_end - t)
        if not np.isfinite(dt) or dt <= 0.0:
            break
        shu_rk4_step(arrays, params, dt, beta_model=beta_model, riemann=riemann, boundary="double-mach", time=t)
        t += dt
        step += 1
        dts.append(float(dt))
        if step == 1 or (report_interval > 0 and step % report_interval == 0) or t >= t_end - 1.0e-14:
            stats = interior_stats(arrays["u"].cpu().numpy(), params)
            if report is not None:
                report(step, t, dt, stats)
            if stats["nan_count"] or stats["rho_neg"] or stats["p_neg"]:
                break
    return arrays["u"].cpu().numpy(), dts, step, t
