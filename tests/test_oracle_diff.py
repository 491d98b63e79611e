"""The oracle's differentiable render (FP64 adjoint) against central finite
differences of its own FP64 forward, and its FP32 forward against FP64."""
import numpy as np
import pytest

from tests import diff_util as U


@pytest.fixture(scope="module")
def setup(small_wl, small_model):
    rm = small_wl.rm(small_wl.cam2world, U.DIFF_W, U.DIFF_H)
    pix = np.arange(U.DIFF_W * U.DIFF_H, dtype=np.int32)
    params = U.params32(small_wl)
    target = U.target_image(len(pix))
    rgb, _, n_eval = small_model.diff_render(rm, pix, params)
    d = 2.0 * (rgb - target)
    _, grads, _ = small_model.diff_render(rm, pix, params, d_rgb=d)
    return rm, pix, params, target, rgb, grads, n_eval


def test_forward_f32_matches_f64(setup, small_model):
    rm, pix, params, _, rgb, _, n_eval = setup
    assert n_eval > 1000
    rgb32, _, n32 = small_model.diff_render(rm, pix, params, f64=False)
    assert n32 == n_eval
    assert np.max(np.abs(rgb32 - rgb)) < 1e-5


def test_forward_matches_render_full_without_perturbation(small_wl, small_model, oracle_mod):
    """With the unperturbed fp16 weights the differentiable forward is the
    full render of the same pixels: the same samples as the FP32 render (the
    march is float in both), the network in double instead of float."""
    rm = small_wl.rm(small_wl.cam2world, U.DIFF_W, U.DIFF_H)
    pix = np.arange(U.DIFF_W * U.DIFF_H, dtype=np.int32)
    params = U.params32(small_wl, sigma=0.0)
    rgb, _, n = small_model.diff_render(rm, pix, params)
    ref, ns = small_model.render_full(rm, mode=oracle_mod.F32)
    assert n == ns
    ref = ref.reshape(-1, 4)
    assert np.max(np.abs(rgb - (ref[:, :3] + ref[:, 3:4]))) < 2e-5


def test_adjoint_matches_finite_differences(setup, small_model):
    rm, pix, params, target, _, grads, _ = setup
    picks = U.pick_params(grads)
    g = np.array([grads[b][i] for b, i in picks])
    fd = U.fd_grads(small_model, rm, pix, params, target, picks)
    rel = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    assert rel < 1e-3, rel
    # per parameter: the step of 1e-3 meets ReLU kinks and the curvature of
    # the small hash-table gradients, so only the median is held to 1e-3
    per = np.abs(g - fd) / np.abs(fd)
    assert np.median(per) < 1e-3, per


def test_missed_rays_have_white_colour_and_no_gradient(small_wl, small_model):
    """A pixel whose ray misses the scene box: C + T = 1, no samples."""
    rm = small_wl.rm(small_wl.cam2world, U.DIFF_W, U.DIFF_H)
    params = U.params32(small_wl)
    pix = np.array([0], np.int32)  # corner pixel: outside the blobs' box at this distance
    rgb, grads, n = small_model.diff_render(rm, pix, params, d_rgb=np.ones((1, 3)))
    if n == 0:
        assert np.allclose(rgb, 1.0)
        assert all(np.count_nonzero(x) == 0 for x in grads)
