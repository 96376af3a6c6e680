"""MI355X-native WENO5-SR finite-volume step.

Drop-in for the reference's neural-WENO 2-D Euler path: the solver adapters in
``weno_sr_publishing_amd.solvers`` keep the reference's entry points, the
checkpoints keep its ``.npz`` layout, and the compute runs in hand-written
gfx950 HIP kernels (``libwsr_hip.so``, C ABI in include/wsr.h).
"""
from .params import Params
from .checkpoints import MlpParams, load_mlp_params, load_f32, load_f64

__all__ = ["Params", "MlpParams", "load_mlp_params", "load_f32", "load_f64"]
__version__ = "0.1.0"
