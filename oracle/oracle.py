"""ctypes front-end of the C oracle (``wsr_oracle.c``) -- test infrastructure only.

The oracle restates the reference's FP64 two-dimensional WENO5 step in plain C
(see the citations inside ``wsr_oracle.c``).  Arrays are numpy float64 in the
reference layout ``(ny + 2*ghost, nx + 2*ghost, 4)``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libwsr_oracle.so"

SCHEMES = {"js": 0, "sr64": 1, "sr32": 2, "srbf16": 3, "z": 4}
RIEMANN = {"hllc": 0, "force": 1}
BOUNDARY = {"transmissive": 0, "periodic": 1, "double-mach": 2}
HEADS = {"normal_lr1": 0, "normal_lr2": 1, "gauss_lr1": 2, "gauss_lr2": 3}

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p
_lib = None


def build(force: bool = False) -> Path:
    """Compile the oracle with its Makefile (no FMA contraction, no fast-math)."""
    opc = HERE / "libwsr_opcount.so"
    stale = not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < (HERE / "wsr_oracle.c").stat().st_mtime
    stale = stale or not opc.exists() or opc.stat().st_mtime < (HERE / "opcount.cpp").stat().st_mtime
    if force or stale:
        subprocess.run(["make", "-s", "-C", str(HERE), "-B" if force else "all"], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.oracle_euler2d_work_doubles.restype = ctypes.c_size_t
        L.oracle_euler2d_work_doubles.argtypes = [_I, _I, _I]
        L.oracle_euler2d_stage.restype = _I
        L.oracle_euler2d_stage.argtypes = [_P, _P, _P, _I, _I, _I, _D, _D, _D, _D, _I, _I, _I, _I, _D, _I, _P, _P]
        L.oracle_euler2d_max_speed.restype = _D
        L.oracle_euler2d_max_speed.argtypes = [_P, _I, _I, _I, _D]
        L.oracle_euler2d_dt.restype = _D
        L.oracle_euler2d_dt.argtypes = [_P, _I, _I, _I, _D, _D, _D, _D]
        L.oracle_euler2d_run.restype = _I
        L.oracle_euler2d_run.argtypes = [_P, _I, _I, _I, _D, _D, _D, _D, _D, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P]
        for name in ("oracle_hllc",):
            getattr(L, name).argtypes = [_P, _P, _I, _D, _P]
        L.oracle_force.argtypes = [_P, _P, _I, _D, _D, _P]
        L.oracle_physical_flux.argtypes = [_P, _I, _D, _P]
        L.oracle_roe.argtypes = [_P, _P, _D, _P]
        L.oracle_char_roundtrip.argtypes = [_P, _P, _I, _D, _P, _P]
        L.oracle_candidates.argtypes = [_P, _I, _P]
        L.oracle_scalar_value.restype = _D
        L.oracle_scalar_value.argtypes = [_P, _I, _I, _P, _I, _D]
        L.oracle_features.argtypes = [_P, _P]
        L.oracle_plateau.restype = _I
        L.oracle_plateau.argtypes = [_P]
        L.oracle_mlp_raw.argtypes = [_P, _I, _I, _P, _P]
        L.oracle_mlp_weights.argtypes = [_P, _I, _I, _P, _I, _P]
        L.oracle_round_bf16.restype = ctypes.c_float
        L.oracle_round_bf16.argtypes = [ctypes.c_float]
        L.oracle_num_threads.restype = _I
        L.oracle_set_num_threads.argtypes = [_I]
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None):
    """Pointer argument that keeps the array alive for the duration of the call
    (``data_as`` holds a reference, so temporaries such as a freshly packed
    parameter block are not freed before the C function reads them)."""
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def pack_mlp(mlp: dict | None, scheme: str) -> np.ndarray | None:
    """Flatten w1,b1,...,b4 (leading singleton axis allowed) in the oracle layout."""
    if mlp is None or scheme in ("js", "z"):
        return None
    dtype = np.float64 if scheme == "sr64" else np.float32
    parts = [np.asarray(mlp[name]).reshape(-1) for name in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4")]
    flat = np.concatenate(parts).astype(dtype)
    if flat.size != 189:
        raise ValueError(f"WENO5 MLP must have 189 parameters, got {flat.size}")
    return np.ascontiguousarray(flat)


def set_threads(n: int) -> None:
    lib().oracle_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def stage2d(u_stage, u0, nx, ny, dx, dy, gamma, dt, rk, scheme="js", riemann="hllc",
            boundary="transmissive", boundary_time=0.0, eno_cutoff=False, mlp=None, ghost=3):
    u_stage = _f64(u_stage)
    u0 = _f64(u0)
    out = np.zeros_like(u_stage)
    packed = pack_mlp(mlp, scheme)
    work = np.zeros(lib().oracle_euler2d_work_doubles(nx, ny, ghost), dtype=np.float64)
    rc = lib().oracle_euler2d_stage(
        _ptr(u_stage), _ptr(u0), _ptr(out), nx, ny, ghost, dx, dy, gamma, dt, rk,
        SCHEMES[scheme], RIEMANN[riemann], BOUNDARY[boundary], boundary_time,
        1 if eno_cutoff else 0, _ptr(packed), _ptr(work))
    if rc:
        raise RuntimeError(f"oracle_euler2d_stage failed: {rc}")
    return out


def max_speed2d(u, nx, ny, gamma, ghost=3) -> float:
    u = _f64(u)
    return float(lib().oracle_euler2d_max_speed(_ptr(u), nx, ny, ghost, gamma))


def dt2d(u, nx, ny, dx, dy, gamma, cfl, ghost=3) -> float:
    u = _f64(u)
    return float(lib().oracle_euler2d_dt(_ptr(u), nx, ny, ghost, dx, dy, gamma, cfl))


def step2d(u, nx, ny, dx, dy, gamma, dt, t=0.0, **kw):
    """One SSPRK3 step (three stages) following A64:486-526."""
    u1 = stage2d(u, u, nx, ny, dx, dy, gamma, dt, 1, boundary_time=t, **kw)
    u2 = stage2d(u1, u, nx, ny, dx, dy, gamma, dt, 2, boundary_time=t + dt, **kw)
    u3 = stage2d(u2, u, nx, ny, dx, dy, gamma, dt, 3, boundary_time=t + 0.5 * dt, **kw)
    return u3


def run2d(u, nx, ny, dx, dy, gamma, cfl, t_end, scheme="js", riemann="hllc",
          boundary="transmissive", eno_cutoff=False, mlp=None, max_steps=1_000_000, ghost=3):
    """run_to_time without reporting: returns (state, dts, steps, t)."""
    u = _f64(u).copy()
    n = u.size
    tmp = np.zeros(3 * n, dtype=np.float64)
    work = np.zeros(lib().oracle_euler2d_work_doubles(nx, ny, ghost), dtype=np.float64)
    cap = max(1, min(max_steps, 1_000_000))
    dts = np.zeros(cap, dtype=np.float64)
    t = ctypes.c_double(0.0)
    packed = pack_mlp(mlp, scheme)
    steps = lib().oracle_euler2d_run(
        _ptr(u), nx, ny, ghost, dx, dy, gamma, cfl, t_end, cap, SCHEMES[scheme],
        RIEMANN[riemann], BOUNDARY[boundary], 1 if eno_cutoff else 0, _ptr(packed),
        _ptr(tmp), _ptr(work), _ptr(dts), ctypes.addressof(t))
    if steps < 0:
        raise RuntimeError(f"oracle_euler2d_run failed: {steps}")
    return u, dts[:steps].tolist(), int(steps), float(t.value)


# ---------------------------------------------------------------- unit level
def hllc(ul, ur, direction, gamma=1.4):
    ul, ur = _f64(ul), _f64(ur)
    out = np.zeros(4)
    lib().oracle_hllc(_ptr(ul), _ptr(ur), direction, gamma, _ptr(out))
    return out


def force(ul, ur, direction, c, gamma=1.4):
    ul, ur = _f64(ul), _f64(ur)
    out = np.zeros(4)
    lib().oracle_force(_ptr(ul), _ptr(ur), direction, c, gamma, _ptr(out))
    return out


def physical_flux(q, direction, gamma=1.4):
    q = _f64(q)
    out = np.zeros(4)
    lib().oracle_physical_flux(_ptr(q), direction, gamma, _ptr(out))
    return out


def roe(left, right, gamma=1.4):
    left, right = _f64(left), _f64(right)
    out = np.zeros(4)
    lib().oracle_roe(_ptr(left), _ptr(right), gamma, _ptr(out))
    return out


def char_roundtrip(q, roe_state, direction, gamma=1.4):
    q, roe_state = _f64(q), _f64(roe_state)
    c = np.zeros(4)
    back = np.zeros(4)
    lib().oracle_char_roundtrip(_ptr(q), _ptr(roe_state), direction, gamma, _ptr(c), _ptr(back))
    return c, back


def candidates(q, head):
    q = _f64(q)
    s = np.zeros(3)
    lib().oracle_candidates(_ptr(q), HEADS[head], _ptr(s))
    return s


def scalar_value(q, head, scheme="js", mlp=None, eno_cutoff=False, h=0.0):
    q = _f64(q)
    packed = pack_mlp(mlp, scheme)
    return float(lib().oracle_scalar_value(_ptr(q), HEADS[head], SCHEMES[scheme], _ptr(packed),
                                           1 if eno_cutoff else 0, h))


def features(q):
    q = _f64(q)
    f = np.zeros(5)
    lib().oracle_features(_ptr(q), _ptr(f))
    return f


def plateau(q) -> bool:
    q = _f64(q)
    return bool(lib().oracle_plateau(_ptr(q)))


def mlp_raw(feats, scheme, mlp):
    feats = _f64(np.atleast_2d(feats))
    out = np.zeros((feats.shape[0], 3))
    packed = pack_mlp(mlp, scheme)
    lib().oracle_mlp_raw(_ptr(feats), feats.shape[0], SCHEMES[scheme], _ptr(packed), _ptr(out))
    return out


def mlp_weights(q, head, scheme, mlp, eno_cutoff=False):
    q = _f64(q)
    w = np.zeros(3)
    packed = pack_mlp(mlp, scheme)
    lib().oracle_mlp_weights(_ptr(q), HEADS[head], SCHEMES[scheme], _ptr(packed),
                             1 if eno_cutoff else 0, _ptr(w))
    return w


def round_bf16(x: float) -> float:
    return float(lib().oracle_round_bf16(float(x)))


os.environ.setdefault("OMP_PROC_BIND", "false")


# ------------------------------------------------------------------ 1-D path
def _lib1d():
    L = lib()
    if not getattr(L, "_adv1d_ready", False):
        L.oracle_adv1d_run.restype = _I
        L.oracle_adv1d_run.argtypes = [_P, _I, _I, _P, _P, _I, _D, _I, _P]
        L.oracle_face_1d.restype = _D
        L.oracle_face_1d.argtypes = [_P, _I, _I, _P]
        L.oracle_features_1d.argtypes = [_P, _P]
        L._adv1d_ready = True
    return L


def adv1d_run(u, velocities, dts, dxinv, scheme="js", mlp=None):
    """Advance a batch (B, N) of periodic 1-D trajectories with the signed SSPRK3
    step (rk3_advection.py:82-131) over the dt schedule ``dts``."""
    u = np.array(np.atleast_2d(u), dtype=np.float64, order="C")
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(velocities, dtype=np.float64), (u.shape[0],)))
    dts = np.ascontiguousarray(dts, dtype=np.float64)
    packed = pack_mlp(mlp, scheme)
    rc = _lib1d().oracle_adv1d_run(_ptr(u), u.shape[0], u.shape[1], _ptr(a), _ptr(dts), dts.size,
                                   float(dxinv), SCHEMES[scheme], _ptr(packed))
    if rc:
        raise RuntimeError(f"oracle_adv1d_run failed: {rc}")
    return u


def face_1d(q, lr, scheme="js", mlp=None):
    q = _f64(q)
    return float(_lib1d().oracle_face_1d(_ptr(q), lr, SCHEMES[scheme], _ptr(pack_mlp(mlp, scheme))))


def features_1d(q):
    q = _f64(q)
    f = np.zeros(5)
    _lib1d().oracle_features_1d(_ptr(q), _ptr(f))
    return f


# ------------------------------------------------------------------ WENO7
RIEMANN7 = {"hllc": 0, "force": 1, "evilin": 2}
W7_NAMES = ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4")


def pack_mlp7(mlp: dict | None, scheme: str) -> np.ndarray | None:
    """Flatten the 6-24-16-16-4 network (908 parameters) in the oracle layout."""
    if mlp is None or scheme in ("js", "z"):
        return None
    dtype = np.float64 if scheme == "sr64" else np.float32
    flat = np.concatenate([np.asarray(mlp[n]).reshape(-1) for n in W7_NAMES]).astype(dtype)
    if flat.size != 908:
        raise ValueError(f"WENO7 MLP must have 908 parameters, got {flat.size}")
    return np.ascontiguousarray(flat)


def _lib7():
    L = lib()
    if not getattr(L, "_w7_ready", False):
        L.oracle_adv1d7_run.restype = _I
        L.oracle_adv1d7_run.argtypes = [_P, _I, _I, _P, _P, _I, _D, _I, _P]
        L.oracle_face7_1d.restype = _D
        L.oracle_face7_1d.argtypes = [_P, _I, _I, _P, _D]
        L.oracle_features7.argtypes = [_P, _P]
        L.oracle_optimal7.argtypes = [_I, _P]
        L.oracle_plateau7.restype = _I
        L.oracle_plateau7.argtypes = [_P]
        L.oracle_candidates7.argtypes = [_P, _I, _P]
        L.oracle_classical_beta7.argtypes = [_P, _P]
        L.oracle_mlp7_raw.argtypes = [_P, _I, _I, _P, _P]
        L.oracle_mlp7_sym.argtypes = [_P, _I, _I, _P, _P]
        L.oracle_beta7_poly.argtypes = [_P, _P]
        L.oracle_beta7_sos.argtypes = [_P, _P]
        L.oracle_scalar7_2d.restype = _D
        L.oracle_scalar7_2d.argtypes = [_P, _I, _I, _I, _P, _I, _D]
        L.oracle_riemann7.argtypes = [_P, _P, _I, _D, _D, _I, _P]
        L.oracle_euler7_rhs.restype = _I
        L.oracle_euler7_rhs.argtypes = [_P, _P, _I, _I, _D, _D, _D, _D, _D, _I, _I, _I, _I, _I, _P]
        L.oracle_euler7_step.restype = _I
        L.oracle_euler7_step.argtypes = [_P, _I, _I, _D, _D, _D, _D, _D, _I, _I, _I, _I, _P]
        L.oracle_euler7_run.restype = _I
        L.oracle_euler7_run.argtypes = [_P, _I, _I, _D, _D, _D, _D, _D, _I, _I, _I, _I, _P, _I, _I, _P, _P, _P]
        L.oracle_guard7.restype = _I
        L.oracle_guard7.argtypes = [_P, _I, _I, _I, _P, _D]
        L.oracle_euler7_fallbacks.restype = ctypes.c_longlong
        L.oracle_euler7_fallbacks.argtypes = []
        L._w7_ready = True
    return L


def adv1d7_run(u, velocities, dts, dxinv, scheme="js", mlp=None):
    """Batch of periodic 1-D trajectories advanced by the signed Shu RK4 step
    (rk4_advection.py:137-195) with WENO7 faces."""
    u = np.array(np.atleast_2d(u), dtype=np.float64, order="C")
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(velocities, dtype=np.float64), (u.shape[0],)))
    dts = np.ascontiguousarray(dts, dtype=np.float64)
    packed = pack_mlp7(mlp, scheme)
    rc = _lib7().oracle_adv1d7_run(_ptr(u), u.shape[0], u.shape[1], _ptr(a), _ptr(dts), dts.size,
                                   float(dxinv), SCHEMES[scheme], _ptr(packed))
    if rc:
        raise RuntimeError(f"oracle_adv1d7_run failed: {rc}")
    return u


def face7_1d(q, lr, scheme="js", mlp=None, h=1.0):
    q = _f64(q)
    return float(_lib7().oracle_face7_1d(_ptr(q), lr, SCHEMES[scheme], _ptr(pack_mlp7(mlp, scheme)), float(h)))


def features7(q):
    q = _f64(q)
    f = np.zeros(6)
    _lib7().oracle_features7(_ptr(q), _ptr(f))
    return f


def optimal7(lr):
    d = np.zeros(4)
    _lib7().oracle_optimal7(int(lr), _ptr(d))
    return d


def plateau7(q) -> bool:
    return bool(_lib7().oracle_plateau7(_ptr(_f64(q))))


def candidates7(q, lr):
    s = np.zeros(4)
    _lib7().oracle_candidates7(_ptr(_f64(q)), int(lr), _ptr(s))
    return s


def classical_beta7(q):
    b = np.zeros(4)
    _lib7().oracle_classical_beta7(_ptr(_f64(q)), _ptr(b))
    return b


def beta7_poly(q):
    b = np.zeros(4)
    _lib7().oracle_beta7_poly(_ptr(_f64(q)), _ptr(b))
    return b


def beta7_sos(q):
    b = np.zeros(4)
    _lib7().oracle_beta7_sos(_ptr(_f64(q)), _ptr(b))
    return b


def mlp7(feats, scheme, mlp, symmetric=True):
    f = _f64(np.atleast_2d(feats))
    out = np.zeros((f.shape[0], 4))
    fn = _lib7().oracle_mlp7_sym if symmetric else _lib7().oracle_mlp7_raw
    fn(_ptr(f), f.shape[0], SCHEMES[scheme], _ptr(pack_mlp7(mlp, scheme)), _ptr(out))
    return out


def scalar7_2d(q, gauss, lr, scheme="js", mlp=None, eno_cutoff=False, h=1.0):
    return float(_lib7().oracle_scalar7_2d(_ptr(_f64(q)), int(gauss), int(lr), SCHEMES[scheme],
                                           _ptr(pack_mlp7(mlp, scheme)), int(eno_cutoff), float(h)))


def riemann7(ul, ur, direction, c, gamma=1.4, riemann="hllc"):
    out = np.zeros(4)
    _lib7().oracle_riemann7(_ptr(_f64(ul)), _ptr(_f64(ur)), direction, float(c), float(gamma), RIEMANN7[riemann],
                            _ptr(out))
    return out


def _eno7(eno_cutoff: bool, z_fallback: bool) -> int:
    """The oracle's WENO7 closure flags: bit 0 ENO cutoff, bit 1 WENO7-Z fallback."""
    return (1 if eno_cutoff else 0) | (2 if z_fallback else 0)


def guard7(q, gauss, lr, scheme, mlp, h) -> bool:
    """The WENO7-Z fallback guard's decision for one stencil (SR closures)."""
    return bool(_lib7().oracle_guard7(_ptr(_f64(q)), int(gauss), int(lr), SCHEMES[scheme],
                                      _ptr(pack_mlp7(mlp, scheme)), float(h)))


def fallbacks7() -> int:
    """Values that fell back to WENO7-Z in the last rhs7_2d / step7_2d / run7_2d call."""
    return int(_lib7().oracle_euler7_fallbacks())


def rhs7_2d(u, nx, ny, dx, dy, gamma, dt, t=0.0, scheme="js", riemann="hllc", boundary="periodic",
            reverse=False, eno_cutoff=False, mlp=None, z_fallback=False):
    """One WENO7 right-hand side L(u) (ghost 4); the interior of the result."""
    u = _f64(u)
    rhs = np.zeros_like(u)
    rc = _lib7().oracle_euler7_rhs(_ptr(u), _ptr(rhs), nx, ny, dx, dy, gamma, dt, t, SCHEMES[scheme],
                                   RIEMANN7[riemann], BOUNDARY[boundary], int(reverse), _eno7(eno_cutoff, z_fallback),
                                   _ptr(pack_mlp7(mlp, scheme)))
    if rc:
        raise RuntimeError(f"oracle_euler7_rhs failed: {rc}")
    return rhs


def step7_2d(u, nx, ny, dx, dy, gamma, dt, t=0.0, scheme="js", riemann="hllc", boundary="periodic",
             eno_cutoff=False, mlp=None, z_fallback=False):
    u = np.array(u, dtype=np.float64, order="C")
    rc = _lib7().oracle_euler7_step(_ptr(u), nx, ny, dx, dy, gamma, dt, t, SCHEMES[scheme], RIEMANN7[riemann],
                                    BOUNDARY[boundary], _eno7(eno_cutoff, z_fallback), _ptr(pack_mlp7(mlp, scheme)))
    if rc:
        raise RuntimeError(f"oracle_euler7_step failed: {rc}")
    return u


def run7_2d(u, nx, ny, dx, dy, gamma, cfl, t_end, scheme="js", riemann="hllc", boundary="periodic",
            eno_cutoff=False, mlp=None, mode="point", max_steps=10_000_000, z_fallback=False):
    """WENO7 + Shu RK4 time loop; mode "point" = point_rk4.run_from_initial,
    "double-mach" = weno7_double_mach.run_to_time.  Returns (u, dts, steps, t)."""
    u = np.array(u, dtype=np.float64, order="C")
    cap = 200000 if max_steps > 200000 else max_steps
    dts = np.zeros(cap)
    steps = ctypes.c_int(0)
    t = ctypes.c_double(0.0)
    rc = _lib7().oracle_euler7_run(_ptr(u), nx, ny, dx, dy, gamma, cfl, t_end, SCHEMES[scheme], RIEMANN7[riemann],
                                   BOUNDARY[boundary], _eno7(eno_cutoff, z_fallback), _ptr(pack_mlp7(mlp, scheme)),
                                   0 if mode == "point" else 1, min(max_steps, cap), _ptr(dts),
                                   ctypes.byref(steps), ctypes.byref(t))
    if rc:
        raise RuntimeError(f"oracle_euler7_run failed: {rc}")
    return u, dts[: steps.value].tolist(), steps.value, t.value


# ------------------------------------------------------------------ 3-D WENO5
def _lib3():
    L = lib()
    if not getattr(L, "_w3_ready", False):
        L.oracle_euler3_work_doubles.restype = ctypes.c_size_t
        L.oracle_euler3_work_doubles.argtypes = [_I, _I, _I]
        L.oracle_euler3_step.restype = _I
        L.oracle_euler3_step.argtypes = [_P, _I, _I, _I, _D, _D, _D, _D, _I, _I, _I, _P]
        L.oracle_euler3_max_speed.restype = _D
        L.oracle_euler3_max_speed.argtypes = [_P, _I, _I, _I]
        L.oracle_euler3_run.restype = _I
        L.oracle_euler3_run.argtypes = [_P, _I, _I, _I, _D, _D, _D, _D, _D, _I, _I, _I, _I, _P, _P, _P, _P]
        L.oracle_flux3.argtypes = [_P, _P, _I, _D, _I, _P]
        L.oracle_eigen3.argtypes = [_P, _I, _P, _P]
        L._w3_ready = True
    return L


def step3d(q, nx, ny, nz, dx, dy, dz, dt, scheme="js", riemann="evilin", eno_cutoff=False, mlp=None):
    """One SSPRK3 step of the padded (nz+6, ny+6, nx+6, 5) state (solver.py:166-210)."""
    q = np.array(q, dtype=np.float64, order="C")
    rc = _lib3().oracle_euler3_step(_ptr(q), nx, ny, nz, dx, dy, dz, dt, SCHEMES[scheme], RIEMANN7[riemann],
                                    int(eno_cutoff), _ptr(pack_mlp(mlp, scheme)))
    if rc:
        raise RuntimeError(f"oracle_euler3_step failed: {rc}")
    return q


def max_speed3d(q, nx, ny, nz) -> float:
    return float(_lib3().oracle_euler3_max_speed(_ptr(_f64(q)), nx, ny, nz))


def run3d(q, nx, ny, nz, dx, dy, dz, cfl, t_end, max_steps=10_000_000, scheme="js", riemann="evilin",
          eno_cutoff=False, mlp=None):
    q = np.array(q, dtype=np.float64, order="C")
    cap = min(max_steps, 100000)
    dts = np.zeros(cap)
    steps = ctypes.c_int(0)
    t = ctypes.c_double(0.0)
    rc = _lib3().oracle_euler3_run(_ptr(q), nx, ny, nz, dx, dy, dz, cfl, t_end, cap, SCHEMES[scheme],
                                   RIEMANN7[riemann], int(eno_cutoff), _ptr(pack_mlp(mlp, scheme)), _ptr(dts),
                                   ctypes.byref(steps), ctypes.byref(t))
    if rc:
        raise RuntimeError(f"oracle_euler3_run failed: {rc}")
    return q, dts[: steps.value].tolist(), steps.value, t.value


def flux3d(ul, ur, direction, dt_h, riemann="evilin"):
    out = np.zeros(5)
    _lib3().oracle_flux3(_ptr(_f64(ul)), _ptr(_f64(ur)), direction, float(dt_h), RIEMANN7[riemann], _ptr(out))
    return out


def eigen3d(state, direction):
    L = np.zeros((5, 5))
    R = np.zeros((5, 5))
    _lib3().oracle_eigen3(_ptr(_f64(state)), direction, _ptr(L), _ptr(R))
    return L, R
