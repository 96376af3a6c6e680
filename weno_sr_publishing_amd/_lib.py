"""ctypes binding of ``libwsr_hip.so`` (the C ABI declared in include/wsr.h).

The HIP library is the only compute path: if it is missing this module raises,
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libwsr_hip.so"

SCHEMES = {"js": 0, "sr64": 1, "sr32": 2, "srbf16": 3, "z": 4}
RIEMANN = {"hllc": 0, "force": 1, "evilin": 2}
BOUNDARY = {"transmissive": 0, "periodic": 1, "double-mach": 2}
ERRORS = {-1: "bad argument", -2: "SR scheme without MLP parameters", -3: "workspace too small",
          -4: "HIP runtime error"}

EXPORTS = (
    "wsr_euler2d_workspace_bytes", "wsr_euler2d_stage", "wsr_euler2d_stage_noalloc", "wsr_euler2d_max_speed",
    "wsr_euler2d_loop_init", "wsr_euler2d_loop_steps", "wsr_euler2d_loop_status", "wsr_version",
    "wsr_euler2d_loop_steps_timed", "wsr_euler2d_counters", "wsr_adv1d_rollout",
    "wsr_adv1d_train_forward", "wsr_adv1d_train_adjoint",
    "wsr_euler7_workspace_bytes", "wsr_euler7_rhs", "wsr_euler7_step", "wsr_euler7_loop_init",
    "wsr_euler7_loop_steps", "wsr_euler7_loop_steps_timed", "wsr_euler7_boundary", "wsr_euler7_max_speed",
    "wsr_beta7_workspace_bytes", "wsr_beta7_fill_normal", "wsr_beta7_fill_cross", "wsr_euler7_stage_ext",
    "wsr_euler7_flux_ext", "wsr_euler7_rhs_from_flux", "wsr_euler7_shu_combine",
    "wsr_euler3_workspace_bytes", "wsr_euler3_step", "wsr_euler3_max_speed", "wsr_euler3_loop_init",
    "wsr_euler3_loop_steps", "wsr_euler3_loop_steps_timed",
    "wsr_init_quadrant", "wsr_init_sod", "wsr_init_double_mach", "wsr_init_vortex", "wsr_init_shockbubble3d",
)
FLAG_COUNT_CLOSURE = 1
FLAG_Z_FALLBACK = 2
NOALLOC_SCRATCH_BYTES = 32 << 20   # WSR_NOALLOC_SCRATCH_BYTES
FLAG_Z7_UNIT_GAUSS_H = 4
LOOP7 = {"point": 0, "double-mach": 1}


class Euler2DConfig(ctypes.Structure):
    """Mirror of ``wsr_euler2d_config``."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("ghost", ctypes.c_int32),
        ("scheme", ctypes.c_int32), ("riemann", ctypes.c_int32), ("boundary", ctypes.c_int32),
        ("eno_cutoff", ctypes.c_int32), ("flags", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("gamma", ctypes.c_double),
        ("cfl", ctypes.c_double), ("t_end", ctypes.c_double),
        ("mlp", ctypes.c_void_p),
    ]


class Euler7Config(ctypes.Structure):
    """Mirror of ``wsr_euler7_config``."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("ghost", ctypes.c_int32),
        ("scheme", ctypes.c_int32), ("riemann", ctypes.c_int32), ("boundary", ctypes.c_int32),
        ("eno_cutoff", ctypes.c_int32), ("flags", ctypes.c_int32),
        ("loop_mode", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("gamma", ctypes.c_double),
        ("cfl", ctypes.c_double), ("t_end", ctypes.c_double),
        ("mlp", ctypes.c_void_p),
    ]


class Euler3Config(ctypes.Structure):
    """Mirror of ``wsr_euler3_config``."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("ghost", ctypes.c_int32),
        ("scheme", ctypes.c_int32), ("riemann", ctypes.c_int32), ("boundary", ctypes.c_int32),
        ("eno_cutoff", ctypes.c_int32), ("flags", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("gamma", ctypes.c_double),
        ("cfl", ctypes.c_double), ("t_end", ctypes.c_double),
        ("mlp", ctypes.c_void_p),
    ]


class Adv1DConfig(ctypes.Structure):
    """Mirror of ``wsr_adv1d_config``."""
    _fields_ = [
        ("batch", ctypes.c_int32), ("n", ctypes.c_int32), ("scheme", ctypes.c_int32),
        ("reserved", ctypes.c_int32), ("dxinv", ctypes.c_double), ("mlp", ctypes.c_void_p),
    ]


_lib = None


class WsrError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise WsrError(f"{what} failed: {ERRORS.get(rc, rc)} ({rc})")


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m weno_sr_publishing_amd.build` "
                "(the HIP library is the only compute path)")
        L = ctypes.CDLL(str(LIB_PATH))
        P, I32, D, SZ = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_size_t
        CFG = ctypes.POINTER(Euler2DConfig)
        L.wsr_euler2d_workspace_bytes.restype = SZ
        L.wsr_euler2d_workspace_bytes.argtypes = [I32, I32, I32]
        L.wsr_euler2d_stage.restype = ctypes.c_int
        L.wsr_euler2d_stage.argtypes = [CFG, P, P, P, D, I32, D, P, SZ, P]
        L.wsr_euler2d_stage_noalloc.restype = ctypes.c_int
        L.wsr_euler2d_stage_noalloc.argtypes = [CFG, P, P, P, D, I32, D, P]
        L.wsr_euler2d_max_speed.restype = ctypes.c_int
        L.wsr_euler2d_max_speed.argtypes = [CFG, P, P, SZ, ctypes.POINTER(D), P]
        L.wsr_euler2d_loop_init.restype = ctypes.c_int
        L.wsr_euler2d_loop_init.argtypes = [CFG, P, I32, P, SZ, P]
        L.wsr_euler2d_loop_steps.restype = ctypes.c_int
        L.wsr_euler2d_loop_steps.argtypes = [CFG, P, P, P, P, I32, P, SZ, P]
        L.wsr_euler2d_loop_status.restype = ctypes.c_int
        L.wsr_euler2d_loop_status.argtypes = [P, ctypes.POINTER(I32), ctypes.POINTER(D),
                                              ctypes.POINTER(D), ctypes.POINTER(I32), P]
        L.wsr_euler2d_loop_steps_timed.restype = ctypes.c_int
        L.wsr_euler2d_loop_steps_timed.argtypes = [CFG, P, P, P, P, I32, P, SZ, P, ctypes.POINTER(D),
                                                   ctypes.POINTER(I32)]
        L.wsr_euler2d_counters.restype = ctypes.c_int
        L.wsr_euler2d_counters.argtypes = [P, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int32, P]
        L.wsr_adv1d_rollout.restype = ctypes.c_int
        L.wsr_adv1d_rollout.argtypes = [ctypes.POINTER(Adv1DConfig), P, P, P, I32, P, P]
        A1 = ctypes.POINTER(Adv1DConfig)
        L.wsr_adv1d_train_forward.restype = ctypes.c_int
        L.wsr_adv1d_train_forward.argtypes = [A1, P, P, P, I32, P, P, P]
        L.wsr_adv1d_train_adjoint.restype = ctypes.c_int
        L.wsr_adv1d_train_adjoint.argtypes = [A1, P, P, P, P, I32, P, P, P, P]
        C7 = ctypes.POINTER(Euler7Config)
        L.wsr_euler7_workspace_bytes.restype = SZ
        L.wsr_euler7_workspace_bytes.argtypes = [C7]
        L.wsr_euler7_rhs.restype = ctypes.c_int
        L.wsr_euler7_rhs.argtypes = [C7, P, P, P, D, D, P, SZ, P]
        L.wsr_euler7_step.restype = ctypes.c_int
        L.wsr_euler7_step.argtypes = [C7, P, D, D, P, SZ, P]
        L.wsr_euler7_loop_init.restype = ctypes.c_int
        L.wsr_euler7_loop_init.argtypes = [C7, P, I32, P, SZ, P]
        L.wsr_euler7_loop_steps.restype = ctypes.c_int
        L.wsr_euler7_loop_steps.argtypes = [C7, P, P, I32, P, SZ, P]
        L.wsr_euler7_loop_steps_timed.restype = ctypes.c_int
        L.wsr_euler7_loop_steps_timed.argtypes = [C7, P, P, I32, P, SZ, P, ctypes.POINTER(D), ctypes.POINTER(I32)]
        L.wsr_euler7_boundary.restype = ctypes.c_int
        L.wsr_euler7_boundary.argtypes = [C7, P, D, P]
        L.wsr_euler7_max_speed.restype = ctypes.c_int
        L.wsr_euler7_max_speed.argtypes = [C7, P, P, SZ, ctypes.POINTER(D), P]
        L.wsr_beta7_workspace_bytes.restype = SZ
        L.wsr_beta7_workspace_bytes.argtypes = []
        L.wsr_beta7_fill_normal.restype = ctypes.c_int
        L.wsr_beta7_fill_normal.argtypes = [C7, P, P, I32, P, SZ, P]
        L.wsr_beta7_fill_cross.restype = ctypes.c_int
        L.wsr_beta7_fill_cross.argtypes = [C7, P, P, P, I32, P, SZ, P]
        L.wsr_euler7_stage_ext.restype = ctypes.c_int
        L.wsr_euler7_stage_ext.argtypes = [C7, P, P, P, P, I32, P]
        L.wsr_euler7_flux_ext.restype = ctypes.c_int
        L.wsr_euler7_flux_ext.argtypes = [C7, P, P, P, P, I32, I32, D, I32, P]
        L.wsr_euler7_rhs_from_flux.restype = ctypes.c_int
        L.wsr_euler7_rhs_from_flux.argtypes = [C7, P, P, P, P, P, P]
        L.wsr_euler7_shu_combine.restype = ctypes.c_int
        L.wsr_euler7_shu_combine.argtypes = [C7, I32, D, ctypes.POINTER(P), P, P]
        C3 = ctypes.POINTER(Euler3Config)
        for name, args in (("wsr_init_quadrant", [P, I32, I32, I32, I32, P, D, D, D, D, D, P, P, I32, P]),
                           ("wsr_init_sod", [P, I32, I32, I32, I32, D, D, D, P, P, I32, P]),
                           ("wsr_init_double_mach", [P, I32, I32, I32, D, D, D, P, P, I32, P]),
                           ("wsr_init_vortex", [P, I32, I32, I32, D, D, D, D, D, D, P, P, I32, P]),
                           ("wsr_init_shockbubble3d", [P, I32, I32, I32, I32, I32, P, P, D, D, P, P, I32, P])):
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = args
        L.wsr_euler3_workspace_bytes.restype = SZ
        L.wsr_euler3_workspace_bytes.argtypes = [C3]
        L.wsr_euler3_step.restype = ctypes.c_int
        L.wsr_euler3_step.argtypes = [C3, P, D, P, SZ, P]
        L.wsr_euler3_max_speed.restype = ctypes.c_int
        L.wsr_euler3_max_speed.argtypes = [C3, P, P, SZ, ctypes.POINTER(D), P]
        L.wsr_euler3_loop_init.restype = ctypes.c_int
        L.wsr_euler3_loop_init.argtypes = [C3, P, I32, P, SZ, P]
        L.wsr_euler3_loop_steps.restype = ctypes.c_int
        L.wsr_euler3_loop_steps.argtypes = [C3, P, P, I32, P, SZ, P]
        L.wsr_euler3_loop_steps_timed.restype = ctypes.c_int
        L.wsr_euler3_loop_steps_timed.argtypes = [C3, P, P, I32, P, SZ, P, ctypes.POINTER(D), ctypes.POINTER(I32)]
        L.wsr_version.restype = ctypes.c_char_p
        _lib = L
    return _lib
