"""ctypes front-end of the op-counting build of the oracle (``opcount.cpp``) --
test infrastructure only.

``opcount.cpp`` compiles ``wsr_oracle.c`` unchanged over a counting scalar and
attributes every operation to the oracle functions active when it ran.  This module
runs single building blocks and whole stages through it and returns the tallies:

* ``block_costs()`` -- operations per call of each building block of the WENO5 step,
  each measured in isolation (one Roe average, one projection, one scalar
  reconstruction of each closure and head, one HLLC flux on the subsonic path, ...);
* ``count_stage(...)`` -- the per-function tallies of one ``oracle_euler2d_stage``.

Operations are counted by kind (add/sub, mul, div, sqrt, exp, log, tanh);
comparisons, min / max, fabs and sign flips count zero.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import oracle as _oracle

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libwsr_opcount.so"
KINDS = ("add", "mul", "div", "sqrt", "exp", "log", "tanh")

_D, _I, _P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            _oracle.build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.opcount_init.restype = _I
        L.opcount_nkinds.restype = _I
        L.opcount_nfunctions.restype = _I
        L.opcount_name.restype = ctypes.c_char_p
        L.opcount_name.argtypes = [_I]
        L.opcount_tally.argtypes = [_I, _P]
        L.opcount_total.argtypes = [_P]
        L.opcount_euler2d_stage.restype = _I
        L.opcount_euler2d_stage.argtypes = [_P, _I, _I, _D, _D, _D, _D, _I, _I, _I, _I, _D, _I, _P]
        L.opcount_roe.argtypes = [_P, _P, _D, _P]
        L.opcount_char_roundtrip.argtypes = [_P, _P, _I, _D, _P, _P]
        L.opcount_scalar.restype = _D
        L.opcount_scalar.argtypes = [_P, _I, _I, _P, _I, _D]
        L.opcount_hllc.argtypes = [_P, _P, _I, _D, _P]
        L.opcount_scalar7.restype = _D
        L.opcount_scalar7.argtypes = [_P, _I, _I, _I, _P, _I, _D]
        if L.opcount_nkinds() != len(KINDS):
            raise RuntimeError("op-counting library and wrapper disagree on the operation kinds")
        _lib = L
    return _lib


@dataclass
class Tally:
    """Calls of one oracle function and the operations run inside it, by kind:
    ``incl`` counts everything below the function, ``excl`` only what no other
    tracked function below it ran."""
    calls: int
    incl: np.ndarray = field(default_factory=lambda: np.zeros(len(KINDS), dtype=np.int64))
    excl: np.ndarray = field(default_factory=lambda: np.zeros(len(KINDS), dtype=np.int64))

    def incl_per_call(self) -> np.ndarray:
        return self.incl // max(self.calls, 1)

    def excl_per_call(self) -> np.ndarray:
        return self.excl // max(self.calls, 1)


def _reset() -> None:
    if lib().opcount_init() <= 0:
        raise RuntimeError("opcount_init failed")


def _tallies() -> tuple[dict[str, Tally], np.ndarray]:
    L = lib()
    nk = len(KINDS)
    buf = (ctypes.c_ulonglong * (1 + 2 * nk))()
    out = {}
    for i in range(L.opcount_nfunctions()):
        L.opcount_tally(i, buf)
        v = np.array(list(buf), dtype=np.int64)
        out[L.opcount_name(i).decode()] = Tally(int(v[0]), v[1:1 + nk].copy(), v[1 + nk:].copy())
    tot = (ctypes.c_ulonglong * nk)()
    L.opcount_total(tot)
    return out, np.array(list(tot), dtype=np.int64)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def count_stage(u, nx, ny, dx, dy, gamma, dt, rk=1, scheme="js", riemann="hllc", boundary="transmissive",
                boundary_time=0.0, eno_cutoff=False, mlp=None) -> tuple[dict[str, Tally], np.ndarray]:
    """Tallies of one counted ``oracle_euler2d_stage`` (ghost 3) and the stage total."""
    if scheme not in ("js", "z", "sr64"):
        raise ValueError("the counting build runs the FP64 arithmetic: js, z or sr64")
    u = _f64(u)
    if u.shape != (ny + 6, nx + 6, 4):
        raise ValueError("state must be the padded (ny + 6, nx + 6, 4) array")
    packed = _oracle.pack_mlp(mlp, scheme)
    _reset()
    rc = lib().opcount_euler2d_stage(u.ctypes.data, nx, ny, dx, dy, gamma, dt, rk, _oracle.SCHEMES[scheme],
                                     _oracle.RIEMANN[riemann], _oracle.BOUNDARY[boundary], boundary_time,
                                     int(eno_cutoff), None if packed is None else packed.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"counted stage failed ({rc})")
    return _tallies()


def count_scalar(q, head, scheme="js", mlp=None, eno_cutoff=False, h=0.0) -> tuple[dict[str, Tally], np.ndarray]:
    q = _f64(q)
    packed = _oracle.pack_mlp(mlp, scheme)
    _reset()
    lib().opcount_scalar(q.ctypes.data, _oracle.HEADS[head], _oracle.SCHEMES[scheme],
                         None if packed is None else packed.ctypes.data, int(eno_cutoff), h)
    return _tallies()


def count_scalar7(q, gauss, lr, scheme="js", mlp=None, eno=0, h=0.01) -> tuple[dict[str, Tally], np.ndarray]:
    """One counted WENO7 scalar reconstruction (``gauss`` 0 normal / 1 Gauss head)."""
    q = _f64(q)
    packed = _oracle.pack_mlp7(mlp, scheme)
    _reset()
    lib().opcount_scalar7(q.ctypes.data, int(gauss), int(lr), _oracle.SCHEMES[scheme],
                          None if packed is None else packed.ctypes.data, int(eno), h)
    return _tallies()


def count_hllc(ul, ur, direction=1, gamma=1.4) -> tuple[dict[str, Tally], np.ndarray]:
    ul, ur, out = _f64(ul), _f64(ur), np.zeros(4)
    _reset()
    lib().opcount_hllc(ul.ctypes.data, ur.ctypes.data, direction, gamma, out.ctypes.data)
    return _tallies()


def count_roe_and_projection(left, right, direction=1, gamma=1.4) -> tuple[dict[str, Tally], np.ndarray]:
    """One Roe average, then one projection of ``left`` onto its fields and back."""
    left, right, roe = _f64(left), _f64(right), np.zeros(4)
    _reset()
    lib().opcount_roe(left.ctypes.data, right.ctypes.data, gamma, roe.ctypes.data)
    c, back = np.zeros(4), np.zeros(4)
    lib().opcount_char_roundtrip(left.ctypes.data, roe.ctypes.data, direction, gamma, c.ctypes.data,
                                 back.ctypes.data)
    return _tallies()


# A smooth, non-plateau stencil (the closure runs), a plateau stencil, and a subsonic
# state pair (HLLC takes its star-state path); the counts of the blocks below do not
# depend on the values beyond these path choices.
SMOOTH_STENCIL = (1.0, 1.3, 1.9, 2.2, 2.3)
PLATEAU_STENCIL = (1.0, 1.0, 1.0, 1.0, 1.0)
SUBSONIC_PAIR = ((1.0, 0.2, 0.1, 2.6), (0.6, 0.1, -0.05, 1.7))


def block_costs(mlp: dict) -> dict[str, np.ndarray]:
    """Operations per call, by kind, of each building block of the WENO5 step.

    ``mlp`` is any 5-10-6-6-3 parameter set (the count does not depend on its values).
    Keys: roe, jacobian, project (con_to_char without the Jacobian values),
    back_project, candidates_normal, candidates_gauss, js_beta, js_combine, z_combine,
    raw_sensors, plateau_test, features (nn_features beyond the raw sensors), mlp_raw,
    reflection_average, sr_weights, weighted_sum, hllc (subsonic path).
    """
    out: dict[str, np.ndarray] = {}
    t, _ = count_roe_and_projection(*SUBSONIC_PAIR)
    out["roe"] = t["roe_average_state"].incl_per_call()
    out["jacobian"] = t["jacobian_state_values"].incl_per_call()
    out["project"] = t["con_to_char_x"].excl_per_call()
    out["back_project"] = t["char_to_con_x"].excl_per_call()
    t2, _ = count_roe_and_projection(*SUBSONIC_PAIR, direction=2)
    if not (np.array_equal(t2["con_to_char_y"].excl_per_call(), out["project"])
            and np.array_equal(t2["char_to_con_y"].excl_per_call(), out["back_project"])):
        raise AssertionError("x and y projections differ in cost")
    for head, key in (("normal_lr1", "candidates_normal"), ("gauss_lr1", "candidates_gauss")):
        t, _ = count_scalar(SMOOTH_STENCIL, head, "js")
        out[key] = t["candidates"].incl_per_call()
        out["js_beta"] = t["js_beta"].incl_per_call()
        out["js_combine"] = t["js_combine"].incl_per_call()
    t, _ = count_scalar(SMOOTH_STENCIL, "normal_lr1", "z", h=0.01)
    out["z_combine"] = t["z_combine"].incl_per_call()
    t, _ = count_scalar(SMOOTH_STENCIL, "normal_lr1", "sr64", mlp)
    if t["nn_features"].calls != 2 or t["mlp_ratios_f64"].calls != 2:
        raise AssertionError("the smooth stencil must run the closure")
    out["raw_sensors"] = t["raw_sensors"].incl_per_call()
    out["plateau_test"] = t["plateau_detected"].excl_per_call()
    out["features"] = t["nn_features"].excl_per_call()
    out["mlp_raw"] = t["mlp_ratios_f64"].incl_per_call()
    out["reflection_average"] = t["sr_ratios"].excl_per_call()
    out["sr_weights"] = t["sr_weights"].incl_per_call()
    out["weighted_sum"] = t["weno5_sr_value"].excl_per_call()
    t, _ = count_hllc(*SUBSONIC_PAIR)
    out["hllc"] = t["hllc_flux"].incl_per_call()
    return out


SMOOTH_STENCIL7 = (1.0, 1.2, 1.5, 1.9, 2.2, 2.3, 2.35)


def block_costs7(mlp7: dict) -> dict[str, np.ndarray]:
    """Operations per call, by kind, of the building blocks of a WENO7 scalar
    reconstruction: candidates_normal / candidates_gauss, beta_poly (JS), beta_sos
    (Z), js_combine / z_combine / sr_combine (weights and weighted sum of one head),
    plateau_test, features, mlp_raw (one 6-24-16-16-4 pass), reflection_average,
    beta_from_ratios, guard (the Z-fallback check: beta_sos and the trip test)."""
    out: dict[str, np.ndarray] = {}
    for gauss, key in ((0, "candidates_normal"), (1, "candidates_gauss")):
        t, tot = count_scalar7(SMOOTH_STENCIL7, gauss, 1, "js")
        out[key] = t["w7_candidates"].incl_per_call()
        out["beta_poly"] = t["w7_beta_poly"].incl_per_call()
        out["js_combine"] = t["w7_scalar_2d"].excl_per_call()
    t, _ = count_scalar7(SMOOTH_STENCIL7, 0, 1, "z")
    out["beta_sos"] = t["w7_beta_sos"].incl_per_call()
    out["z_combine"] = t["w7_scalar_2d"].excl_per_call()
    t, _ = count_scalar7(SMOOTH_STENCIL7, 0, 1, "sr64", mlp7)
    if t["w7_mlp_raw64"].calls != 2:
        raise AssertionError("the smooth stencil must run the closure")
    out["plateau_test"] = t["w7_plateau"].incl_per_call()
    out["features"] = t["w7_features"].incl_per_call()
    out["mlp_raw"] = t["w7_mlp_raw64"].incl_per_call()
    out["reflection_average"] = t["w7_sym_ratios"].excl_per_call()
    out["beta_from_ratios"] = t["w7_provider_beta"].excl_per_call()
    out["sr_combine"] = t["w7_scalar_2d"].excl_per_call()
    t, _ = count_scalar7(SMOOTH_STENCIL7, 0, 1, "sr64", mlp7, eno=2)
    out["guard"] = t["w7_beta_sos"].incl_per_call() + t["w7_guard_trips"].incl_per_call()
    return out
