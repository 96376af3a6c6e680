"""Density line cuts of C.3-C.6 against a 1200-cell reference
(code/snapshot/weno_z_borges_p2_results/plot_riemann.py:44-57,180-235; the table, not the
figures).

The reference field is the 1200 x 1200 density block-averaged 3 x 3 to 400 x 400
(WENO7-JS for C.3, WENO5-JS otherwise); a cut takes the row (``y``) or column (``x``)
of the cell centre nearest the requested coordinate; errors are the mean, RMS and
max of the density difference along the cut.
"""
from __future__ import annotations

import numpy as np

from ..common import primitive

REFERENCE_METHOD = {"c3": "weno7_js", "c4": "weno5_js", "c5": "weno5_js", "c6": "weno5_js"}
CUTS = {
    "c3": (("y", 0.42), ("y", 0.50), ("x", 0.43), ("x", 0.62)),
    "c4": (("y", 0.50), ("y", 0.44), ("y", 0.715), ("y", 0.34)),
    "c5": (("y", 0.35), ("y", 0.665), ("x", 0.665), ("x", 0.745)),
    "c6": (("y", 0.305), ("y", 0.465), ("y", 0.655), ("x", 0.460)),
}
METHODS = ("weno5_js", "weno5_sr_f64", "weno5_sr_f32", "weno7_js", "weno7_z_p3", "weno7_sr_f64")


def reference_density(state1200: np.ndarray) -> np.ndarray:
    if state1200.shape[:2] != (1200, 1200):
        raise ValueError(f"unexpected reference shape: {state1200.shape}")
    return state1200[..., 0].reshape(400, 3, 400, 3).mean(axis=(1, 3))


def profile(field: np.ndarray, direction: str, value: float) -> tuple[np.ndarray, np.ndarray, float]:
    coordinates = (np.arange(400) + 0.5) / 400.0
    index = int(np.argmin(np.abs(coordinates - value)))
    if direction == "y":
        return coordinates, field[index, :], float(coordinates[index])
    return coordinates, field[:, index], float(coordinates[index])


def linecut_rows(case: str, states: dict[str, np.ndarray], reference_state: np.ndarray) -> list[dict]:
    """states: method -> 400 x 400 interior conserved state."""
    reference = reference_density(reference_state)
    rows = []
    for direction, requested in CUTS[case]:
        _, ref_profile, actual = profile(reference, direction, requested)
        for method in METHODS:
            if method not in states:
                continue
            _, values, _ = profile(primitive(states[method])[..., 0], direction, requested)
            error = values - ref_profile
            rows.append({"case": case, "direction": direction, "requested": requested, "actual": actual,
                         "method": method, "status": "complete", "l1": float(np.mean(np.abs(error))),
                         "l2": float(np.sqrt(np.mean(error * error))), "linf": float(np.max(np.abs(error)))})
    return rows
