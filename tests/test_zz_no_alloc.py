"""wsr_euler2d_stage_noalloc (the allocation-free stage on the library's file-scope
device scratch) against wsr_euler2d_stage (caller workspace), through the C ABI: the
same kernels, so the results must be bit-identical.  Two calls in a row, the second on
a smaller grid (the scratch keeps the first call's contents), then a grid whose
workspace exceeds the scratch (refused with WSR_E_WORKSPACE, output untouched)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _state(nx, ny, seed):
    from weno_sr_publishing_amd.initial import make_quadrant_initial
    rng = np.random.default_rng(seed)
    u = make_quadrant_initial(nx, ny, "c4")
    # a rough, non-plateau perturbation so that the closure runs on most stencils
    u[..., 0] *= 1.0 + 0.05 * rng.random(u.shape[:2])
    u[..., 3] *= 1.0 + 0.05 * rng.random(u.shape[:2])
    return u


@pytest.mark.parametrize("scheme", ["sr32", "js"])
def test_noalloc_stage_matches_workspace_stage(mlp32, gpu_device, scheme):
    from weno_sr_publishing_amd import _lib
    from weno_sr_publishing_amd.euler2d import Euler2DSolver
    from weno_sr_publishing_amd.params import Params
    for nx, ny, seed in ((400, 400, 1), (96, 72, 2)):   # second call smaller than the first
        print(f"noalloc stage {scheme} {nx}x{ny}", flush=True)
        p = Params(nx=nx, ny=ny, x_length=1.0, y_length=1.0, cfl=0.4, t_end=0.25)
        s = Euler2DSolver(p, scheme, "hllc", "transmissive", mlp32 if scheme == "sr32" else None, False, gpu_device)
        u0 = s.new_state(_state(nx, ny, seed))
        dt = 0.4 * p.dx / 3.0
        ref_prev, got_prev = u0, u0
        for rk in (1, 2, 3):
            ref = u0.clone()
            got = u0.clone()
            s.stage(ref_prev, u0, ref, dt, rk)
            s.stage_noalloc(got_prev, u0, got, dt, rk)
            torch.cuda.synchronize()
            assert torch.equal(ref, got), (nx, ny, rk)
            ref_prev, got_prev = ref, got
        assert torch.isfinite(got).all()


def test_noalloc_stage_refuses_grids_beyond_the_scratch(gpu_device):
    from weno_sr_publishing_amd import _lib
    from weno_sr_publishing_amd.euler2d import Euler2DSolver
    from weno_sr_publishing_amd.params import Params
    L = _lib.lib()
    nx = ny = 1200
    assert L.wsr_euler2d_workspace_bytes(nx, ny, 3) > _lib.NOALLOC_SCRATCH_BYTES
    p = Params(nx=nx, ny=ny, x_length=1.0, y_length=1.0, cfl=0.4, t_end=0.25)
    s = Euler2DSolver(p, "js", "hllc", "transmissive", None, False, gpu_device)
    u0 = s.new_state(_state(nx, ny, 3))
    out = torch.full_like(u0, 7.0)
    with pytest.raises(_lib.WsrError):
        s.stage_noalloc(u0, u0, out, 1e-4, 1)
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())
