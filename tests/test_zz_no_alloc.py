"""fvc_render_full_noalloc (per-frame buffers in the library's static
scratch) against fvc_render_full: two calls in a row, the second on a
smaller view, images and float results bit-equal; oversized views refused."""
import numpy as np
import pytest
import torch

from frustumvolumecaching_amd import lib, synthetic as S

pytestmark = pytest.mark.gpu


def _both(r, rm):
    W, H = rm.cam_info.res_x, rm.cam_info.res_y
    n = W * H
    a_img = torch.zeros(n * 4, dtype=torch.uint8, device="cuda")
    a_res = torch.empty(n * 4, dtype=torch.float32, device="cuda")
    st_a = r.render_full(rm, a_img)
    r.last_result(a_res, n * 4)
    b_img = torch.zeros_like(a_img)
    b_res = torch.full_like(a_res, float("nan"))
    st_b = r.render_full_noalloc(rm, b_img, b_res)
    torch.cuda.synchronize()
    return (a_img.cpu().numpy(), a_res.cpu().numpy(), st_a), (b_img.cpu().numpy(), b_res.cpu().numpy(), st_b)


@pytest.mark.parametrize("which", ["small", "unbounded"])
def test_noalloc_render_equals_render_full(which):
    if which == "small":
        wl = S.small_plain(stepsize_setup=lib.stepsize_setup)
        sizes = [(256, 256), (96, 80)]
    else:
        wl = S.unbounded_1080p(width=240, height=136, stepsize_setup=lib.stepsize_setup)
        sizes = [(240, 136), (120, 68)]
    model = lib.NerfNetwork.from_workload(wl)
    r = lib.Renderer(model)
    for w, h in sizes:
        rm = wl.rm(width=w, height=h)
        (ai, ar, sa), (bi, br, sb) = _both(r, rm)
        print(f"[noalloc] {which} {w}x{h}: {sb.n_resamples} samples, {sb.n_subiters} sub-iterations "
              f"(render_full {sa.n_resamples}, {sa.n_subiters})", flush=True)
        assert np.array_equal(ai, bi)
        assert np.array_equal(ar.view(np.uint32), br.view(np.uint32))
        assert sb.n_resamples > 0
        if which == "small":
            assert sb.n_resamples == sa.n_resamples   # no ray ends early in this scene
    big = wl.rm(width=257, height=256)
    with pytest.raises(lib.FvcError) as e:
        r.render_full_noalloc(big)
    assert e.value.code == lib.E_INVALID
    r.close()
    model.close()
