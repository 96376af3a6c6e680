"""The benchmark's operation counts against the op-counting build of the oracle.

``oracle/opcount.cpp`` compiles the oracle source unchanged over a counting scalar and
attributes each operation to the oracle function that ran it.  These tests check
(1) every per-block constant of ``bench.OPS`` against the tally of that block run in
isolation, and (2) that the blocks account for every operation of a whole counted
oracle stage: the stage's tally, kind by kind, equals the blocks times the oracle's
own call multiplicities (the loop structure of oracle_euler2d_stage), for WENO5-JS,
WENO5-Z and WENO5-SR with a mix of plateau and closure stencils.
"""
from __future__ import annotations

import numpy as np
import pytest

import bench
from oracle import oracle, opcount
from weno_sr_publishing_amd.checkpoints import load_f64
from weno_sr_publishing_amd.initial import make_quadrant_initial

KIND = {k: i for i, k in enumerate(opcount.KINDS)}


@pytest.fixture(scope="module")
def mlp():
    return load_f64()


@pytest.fixture(scope="module")
def blocks(mlp):
    return opcount.block_costs(mlp)


def test_bench_constants_equal_the_tally(blocks):
    tallied = {k: int(v.sum()) for k, v in blocks.items()}
    assert tallied == bench.OPS


def test_block_costs_do_not_depend_on_values(blocks, mlp):
    rng = np.random.default_rng(3)
    other = {n: rng.normal(size=np.shape(mlp[n])) for n in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4")}
    again = opcount.block_costs(other)
    assert {k: list(v) for k, v in again.items()} == {k: list(v) for k, v in blocks.items()}


def test_mlp_count_by_kind(blocks):
    """One 5-10-6-6-3 pass: 5*10 + 10*6 + 6*6 + 6*3 products and as many sums (the
    bias add included), a swish per hidden unit (exp, add, div), then the head:
    tanh(z/6)*6, exp and the softmax normalisation."""
    mlp = blocks["mlp_raw"]
    macs = 5 * 10 + 10 * 6 + 6 * 6 + 6 * 3
    hidden = 10 + 6 + 6
    assert mlp[KIND["exp"]] == hidden + 3
    assert mlp[KIND["tanh"]] == 3
    assert mlp[KIND["mul"]] >= macs and mlp[KIND["add"]] >= macs


def _stage_expectation(blocks, tallies, nx, ny, scheme):
    """Kind-by-kind operation count of one oracle stage from the block costs and the
    oracle's loop structure (normal_stage, flux_stage, update_rk3)."""
    b = blocks
    mlp_exp = b["mlp_raw"][KIND["exp"]]
    zero = np.zeros(len(opcount.KINDS), dtype=np.int64)

    def scalar(normal: bool, closure: bool):
        cand = b["candidates_normal"] if normal else b["candidates_gauss"]
        if scheme == "js":
            return cand + b["js_beta"] + b["js_combine"]
        if scheme == "z":
            return cand + b["js_beta"] + b["z_combine"]
        v = cand + b["weighted_sum"] + b["raw_sensors"] + b["plateau_test"]
        if closure:
            v = v + 2 * (b["raw_sensors"] + b["features"]) + 2 * b["mlp_raw"] + b["reflection_average"] \
                + b["sr_weights"]
        return v

    # normal_stage: two char_face per stencil position, each one Roe average, five
    # projections and one back-projection (each recomputing the Jacobian values) and
    # four scalar reconstructions
    positions = (ny + 6) * (nx + 2) + (ny + 2) * (nx + 6)
    n_scalar = 2 * 4 * positions
    k_n = int(tallies["normal_stage"].incl[KIND["exp"]] // (2 * mlp_exp)) if scheme == "sr64" else 0
    normal = 2 * positions * (b["roe"] + 5 * (b["project"] + b["jacobian"]) + b["back_project"] + b["jacobian"])
    normal = normal + (n_scalar - k_n) * scalar(True, False) + k_n * scalar(True, True)
    # flux_stage, twice per direction (the two Gauss points): eight scalar
    # reconstructions, one flux (its tally: HLLC's path depends on the state) and the
    # dt/h scaling per face; optimal_weights of a Gauss head computes its constants
    faces = ny * (nx + 1) + (ny + 1) * nx
    g_scalar = 2 * 8 * faces
    k_g = int(tallies["flux_stage"].incl[KIND["exp"]] // (2 * mlp_exp)) if scheme == "sr64" else 0
    flux = (g_scalar - k_g) * scalar(False, False) + k_g * scalar(False, True)
    flux = flux + tallies["hllc_flux"].incl + tallies["optimal_weights"].incl + 2 * faces * 4 * np.eye(
        len(opcount.KINDS), dtype=np.int64)[KIND["mul"]]
    # update_rk3 (rk 1): per component the two Gauss-point fluxes of the four faces
    # summed (4 adds), the two differences, the two half-weighted subtractions
    update = nx * ny * 4 * np.array([8, 2, 0, 0, 0, 0, 0]) + zero
    return normal, flux, update, k_n, k_g, n_scalar, g_scalar


@pytest.mark.parametrize("scheme", ["js", "z", "sr64"])
def test_stage_is_the_sum_of_its_blocks(blocks, mlp, scheme):
    nx, ny = 24, 20
    u = make_quadrant_initial(nx, ny, "c4")
    # a few oracle steps smear the quadrant interfaces: plateau and closure stencils both occur
    u, _, _, _ = oracle.run2d(u, nx, ny, 1.0 / nx, 1.0 / ny, 1.4, 0.4, 1.0, scheme="js", max_steps=4)
    tallies, total = opcount.count_stage(u, nx, ny, 1.0 / nx, 1.0 / ny, 1.4, 1e-3, rk=1, scheme=scheme,
                                         mlp=mlp if scheme == "sr64" else None)
    normal, flux, update, k_n, k_g, n_scalar, g_scalar = _stage_expectation(blocks, tallies, nx, ny, scheme)
    np.testing.assert_array_equal(tallies["normal_stage"].incl, normal)
    np.testing.assert_array_equal(tallies["flux_stage"].incl, flux)
    np.testing.assert_array_equal(tallies["update_rk3"].incl, update)
    # the stage itself: dt/dx and dt/dy
    np.testing.assert_array_equal(total, normal + flux + update + 2 * np.eye(len(opcount.KINDS),
                                                                               dtype=np.int64)[KIND["div"]])
    if scheme == "sr64":
        assert 0 < k_n < n_scalar and 0 < k_g < g_scalar
        assert tallies["mlp_ratios_f64"].calls == 2 * (k_n + k_g)


def test_hllc_paths():
    """The subsonic count is the star-state path in x; supersonic states return the
    upwind physical flux (fewer operations); the y direction computes two extra products."""
    sub = opcount.count_hllc(*opcount.SUBSONIC_PAIR)[0]["hllc_flux"].incl_per_call().sum()
    sup = opcount.count_hllc((1.0, 5.0, 0.0, 15.0), (1.0, 5.0, 0.0, 15.0))[0]["hllc_flux"].incl_per_call().sum()
    y = opcount.count_hllc(*opcount.SUBSONIC_PAIR, direction=2)[0]["hllc_flux"].incl_per_call().sum()
    assert sub == bench.OPS["hllc"] and sup < sub and y == sub + 2


@pytest.fixture(scope="module")
def blocks7():
    from weno_sr_publishing_amd.checkpoints import WENO7_SR_FP64, load_weno7_checkpoint
    return opcount.block_costs7(load_weno7_checkpoint(WENO7_SR_FP64)), load_weno7_checkpoint(WENO7_SR_FP64)


def test_weno7_constants_equal_the_tally(blocks7):
    b, _ = blocks7
    assert {k: int(v.sum()) for k, v in b.items()} == bench.OPS7


@pytest.mark.parametrize("gauss", [0, 1])
@pytest.mark.parametrize("scheme", ["js", "z", "sr64"])
def test_weno7_scalar_is_the_sum_of_its_blocks(blocks7, scheme, gauss):
    """A whole counted WENO7 scalar reconstruction (plateau and closure stencils) equals
    its blocks: candidates, optimal weights (constants of the head), indicators or the
    closure, and the weights of the head."""
    b, mlp7 = blocks7
    cand = b["candidates_gauss"] if gauss else b["candidates_normal"]
    for stencil, closure in ((opcount.SMOOTH_STENCIL7, True), ((2.0,) * 7, False)):
        t, total = opcount.count_scalar7(stencil, gauss, 1, scheme, mlp7 if scheme == "sr64" else None)
        opt = t["w7_optimal"].incl_per_call()
        if scheme == "js":
            want = cand + b["beta_poly"] + b["js_combine"]
        elif scheme == "z":
            want = cand + b["beta_sos"] + b["z_combine"]
        else:
            want = b["plateau_test"] + cand + b["sr_combine"]
            if closure:
                want = want + b["features"] + 2 * b["mlp_raw"] + b["reflection_average"] + b["beta_from_ratios"]
        np.testing.assert_array_equal(total, want + opt)
    assert bench.NORMAL7["sr"][1] == int((b["plateau_test"] + b["candidates_normal"] + b["sr_combine"] + b["features"]
                                          + 2 * b["mlp_raw"] + b["reflection_average"]
                                          + b["beta_from_ratios"]).sum())


def test_blended_peak():
    """The roofline peak of an op mix: FP64 work at 78.6, FP32 work at 157.3 TFLOP/s.
    An all-FP64 launch (JS) is priced at the FP64 rate, an all-FP32 one at the FP32 rate,
    and the time at peak of a mix is the sum of its parts' times."""
    assert bench.blended_peak(1.0e9, 0.0) == bench.PEAK_FP64_TFLOPS
    assert abs(bench.blended_peak(1.0e9, 1.0e9) - bench.PEAK_FP32_TFLOPS) < 1e-9
    tot, f32 = 3.0e9, 2.0e9
    t_peak = (tot - f32) / bench.PEAK_FP64_TFLOPS + f32 / bench.PEAK_FP32_TFLOPS
    assert abs(bench.blended_peak(tot, f32) - tot / t_peak) < 1e-9
    # the SR32 sweep's FP32 work is the closure's: features + direct and reflected MLP passes
    # of every closure evaluation, none at a zero closure fraction
    assert bench.CLOSURE_FP32["sr32"] == bench.OPS["features"] + 2 * bench.OPS["mlp_raw"]
    assert bench.sweep_closures_per_launch(0.0, 0.0) == 0.0
    n = bench.sweep_closures_per_launch(1.0, 1.0)
    faces_lines = (bench.NX + 1) * (bench.NY + 4) + (bench.NY + 1) * (bench.NX + 4)
    faces = (bench.NX + 1) * bench.NY + (bench.NY + 1) * bench.NX
    assert n == 8 * (faces_lines + faces)
    sr = bench.sweep_flops_per_launch("sr32", 1.0, 1.0)
    assert bench.PEAK_FP64_TFLOPS < bench.blended_peak(sr, n * bench.CLOSURE_FP32["sr32"]) < bench.PEAK_FP32_TFLOPS
