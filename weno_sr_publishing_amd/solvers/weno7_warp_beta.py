"""``WarpWeno7PointBeta`` of code/snapshot/for_paper_results/solvers/weno7_warp_beta.py:562-682:
the FP64 6-24-16-16-4 provider that run_double_mach.py selects with
``--weno7-beta-backend warp``.  Here both backends are the same HIP closure; the class
keeps the reference's constructor and its shape check.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from ..checkpoints import load_weno7_checkpoint
from .point_rk4_mlp import TorchWeno7PointBeta

INPUTS, HIDDEN1, HIDDEN2, HIDDEN3, OUTPUTS = 6, 24, 16, 16, 4


class WarpWeno7PointBeta(TorchWeno7PointBeta):
    def __init__(self, model_path: Path, device: str, gamma: float = 1.4) -> None:
        data = np.load(Path(model_path), allow_pickle=False)
        expected = {"w1": (1, INPUTS, HIDDEN1), "b1": (1, HIDDEN1), "w2": (1, HIDDEN1, HIDDEN2),
                    "b2": (1, HIDDEN2), "w3": (1, HIDDEN2, HIDDEN3), "b3": (1, HIDDEN3),
                    "w4": (1, HIDDEN3, OUTPUTS), "b4": (1, OUTPUTS)}
        wrong = {name: (data[name].shape if name in data.files else None) for name, shape in expected.items()
                 if name not in data.files or data[name].shape != shape}
        if wrong:
            raise ValueError(f"Warp WENO7 provider requires 6->24->16->16->4, got {wrong}")
        super().__init__(load_weno7_checkpoint(model_path), device, gamma, precision="fp64")
