"""Pins the oracle's scalar arithmetic: IEEE binary16 / bfloat16 / E4M3
conversions, the deterministic log / exp2, the step schedule
(StepsizeInfo, common.h:93-164) and Morton order (raymarch_common.h:98-110)."""
import math

import numpy as np
import pytest
import torch

from frustumvolumecaching_amd import abi


@pytest.fixture(scope="module")
def L(oracle_mod):
    return oracle_mod.lib()


def test_f16_decode_all_patterns_match_numpy(L):
    bits = np.arange(65536, dtype=np.uint16)
    ref = bits.view(np.float16).astype(np.float32)
    got = np.array([L.orc_f16_to_f32(int(b)) for b in bits], np.float32)
    finite = np.isfinite(ref)
    assert np.array_equal(got[finite].view(np.uint32), ref[finite].view(np.uint32))
    assert np.all(np.isnan(got[np.isnan(ref)]))


def test_f16_encode_round_to_nearest_even_matches_numpy(L):
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 10.0,
        rng.uniform(-1e-5, 1e-5, 5000).astype(np.float32),          # subnormal range
        (np.arange(-2048, 2048, dtype=np.float32) + 0.5) / 1024.0,   # exact ties
        np.array([65504.0, 65520.0, 65519.99, 1e6, -1e6, 0.0, -0.0, 2.0 ** -25, 2.0 ** -24 * 1.5], np.float32),
    ])
    ref = x.astype(np.float16).view(np.uint16)
    got = np.array([L.orc_f32_to_f16(float(v)) for v in x], np.uint16)
    assert np.array_equal(got, ref)


def test_bf16_matches_torch(L):
    x = np.random.default_rng(1).standard_normal(20000).astype(np.float32) * 100
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    got = np.array([L.orc_f32_to_bf16(float(v)) for v in x], np.uint16)
    assert np.array_equal(got, ref)


def test_e4m3_matches_torch_in_range(L):
    x = np.random.default_rng(2).uniform(-440, 440, 20000).astype(np.float32)
    x = np.concatenate([x, np.random.default_rng(3).uniform(-0.05, 0.05, 5000).astype(np.float32)])
    ref = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    got = np.array([L.orc_f32_to_e4m3(float(v)) for v in x], np.uint8)
    assert np.array_equal(got, ref)
    dec = np.array([L.orc_e4m3_to_f32(int(b)) for b in range(256)], np.float32)
    refd = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).float().numpy()
    m = np.isfinite(refd)
    assert np.array_equal(dec[m], refd[m])


def test_det_log_exp2_accuracy(L):
    x = np.exp(np.random.default_rng(4).uniform(-20, 20, 5000)).astype(np.float32)
    got = np.array([L.orc_det_logf(float(v)) for v in x])
    ref = np.log(x.astype(np.float64))
    assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3)) < 5e-7
    y = np.random.default_rng(5).uniform(-30, 30, 5000).astype(np.float32)
    got = np.array([L.orc_det_exp2f(float(v)) for v in y])
    ref = np.exp2(y.astype(np.float64))
    assert np.max(np.abs(got / ref - 1.0)) < 5e-7
    # exact at integers and at 1
    assert L.orc_det_exp2f(10.0) == 1024.0 and L.orc_det_logf(1.0) == 0.0


def _scene(half=1.0, res=128):
    s = abi.SceneInfo()
    s.grid_resolution, s.grid_nlvl = res, 1
    s.aabb_from, s.aabb_to = abi.f3([-half] * 3), abi.f3([half] * 3)
    return s


def _step(c, stepsize, near, far, scene, oracle_mod):
    st = abi.StepsizeInfo()
    st.cone_angle, st.stepsize, st.near_, st.far_ = c, stepsize, near, far
    return oracle_mod.stepsize_setup(st, scene)


def test_step_schedule_linear_closed_form(L, oracle_mod):
    sc = _scene()
    st = _step(0.0, math.sqrt(3) / 1024, 0.1, 8.0, sc, oracle_mod)
    dt = np.float32(st.dt_min)
    for i in [0, 1, 7, 100, 2047, 5000]:
        assert L.orc_t_from_step_f32(st, float(i)) == np.float32(np.float32(i) * dt)
    assert st.near_i == np.float32(np.float32(0.1) / dt)


def test_step_schedule_cone_closed_forms(L, oracle_mod):
    sc = _scene(16.0)
    c = 1.0 / 256.0
    st = _step(c, math.sqrt(3) / 1024, 0.1, 1e10, sc, oracle_mod)
    dt_min, dt_max = st.dt_min, st.dt_max
    assert dt_max == pytest.approx(32 / 128 * 1.732, rel=1e-6)
    assert st.t_min == pytest.approx(dt_min / c, rel=1e-6)
    assert st.i_min == pytest.approx(1 / c, rel=1e-6)
    assert st.i_max - st.i_min == pytest.approx(math.log(dt_max / dt_min) / math.log1p(c), rel=1e-5)
    assert st.t_max == pytest.approx(dt_max / c, rel=1e-5)
    # continuity at both knots and the geometric growth dt = c*t in between
    for knot in (st.i_min, st.i_max):
        lo = L.orc_t_from_step_f64(st, knot - 1e-7)
        hi = L.orc_t_from_step_f64(st, knot + 1e-7)
        assert abs(hi - lo) < 1e-5 * max(1.0, abs(hi))
    for i in np.linspace(st.i_min + 1, st.i_max - 1, 20):
        t0 = L.orc_t_from_step_f64(st, float(i))
        t1 = L.orc_t_from_step_f64(st, float(i) + 1)
        assert (t1 - t0) / t0 == pytest.approx(c, rel=1e-4)
    # float32 schedule vs float64 schedule, and round trip
    for t in np.exp(np.linspace(math.log(0.01), math.log(200), 200)):
        s32 = L.orc_step_from_t_f32(st, float(t))
        s64 = L.orc_step_from_t_f64(st, float(t))
        assert abs(s32 - s64) < 2e-4 * max(1, abs(s64))
        assert L.orc_t_from_step_f32(st, s32) == pytest.approx(t, rel=3e-6)


def test_valid_t_from_t_snaps_to_bin_starts(L, oracle_mod):
    """valid_t_from_t (raymarch_common.h:286-296): the returned t is a bin
    start on the grid offset by near_i, and the nearest one to t."""
    sc = _scene()
    st = _step(0.0, math.sqrt(3) / 1024, 0.1, 8.0, sc, oracle_mod)
    for t in np.random.default_rng(6).uniform(0.2, 6.0, 200):
        v = L.orc_valid_t_from_t_f32(st, float(t))
        k = (v / st.dt_min) - st.near_i
        assert abs(k - round(k)) < 1e-2
        assert abs(v - t) <= 0.5 * st.dt_min * 1.001


def test_morton_matches_bit_interleave(L):
    rng = np.random.default_rng(7)
    for _ in range(300):
        x, y, z = (int(v) for v in rng.integers(0, 1024, 3))
        ref = 0
        for b in range(10):
            ref |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
        assert L.orc_morton3d(x, y, z) == ref
