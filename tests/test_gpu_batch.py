"""Multi-view batches, MLP rows per wavefront and the autotuner: launch
shapes only, so every image must equal the single-view render bit for bit
(a batch or a chunk size only moves sub-iteration boundaries; a ray
composites the same samples in the same order)."""
import numpy as np
import pytest
import torch

from frustumvolumecaching_amd import lib, synthetic as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    wl = S.small_plain(stepsize_setup=lib.stepsize_setup)
    model = lib.NerfNetwork.from_workload(wl)
    r = lib.Renderer(model)
    views = [wl.rm(S.yaw(wl.cam2world, a), w, h) for a, (w, h) in ((0.0, (128, 96)), (3.0, (64, 64)),
                                                                     (-5.0, (200, 120)), (8.0, (33, 17)))]
    yield wl, model, r, views
    r.close()
    model.close()


def _img(rm):
    return torch.zeros(rm.cam_info.res_y, rm.cam_info.res_x, 4, dtype=torch.uint8, device="cuda")


def _results(r, n_px):
    out = torch.empty(n_px * 4, dtype=torch.float32, device="cuda")
    r.last_result(out, n_px * 4)
    return out.cpu().numpy()


def test_full_batch_equals_single_views(setup):
    wl, model, r, views = setup
    single, res_single, n_single = [], [], 0
    for rm in views:
        img = _img(rm)
        st = r.render_full(rm, img)
        n_single += st.n_resamples
        single.append(img.cpu().numpy())
        res_single.append(_results(r, rm.cam_info.res_x * rm.cam_info.res_y))
    imgs = [_img(rm) for rm in views]
    st = r.render_full_batch(views, imgs)
    print(f"[batch] {len(views)} views, {st.n_resamples} samples, {st.n_subiters} sub-iterations", flush=True)
    assert st.n_resamples == n_single
    for a, b in zip(imgs, single):
        assert np.array_equal(a.cpu().numpy(), b)
    n_px = sum(rm.cam_info.res_x * rm.cam_info.res_y for rm in views)
    assert np.array_equal(_results(r, n_px).view(np.uint32), np.concatenate(res_single).view(np.uint32))


def test_cached_batch_equals_single_views(setup):
    wl, model, r, views = setup
    r.initCacheBlockwise(wl.rm(wl.cam2world, 128, 96))
    single = []
    for rm in views:
        img = _img(rm)
        r.render_cached(rm, img)
        single.append(img.cpu().numpy())
    imgs = [_img(rm) for rm in views]
    st = r.render_cached_batch(views, imgs)
    assert st.n_cache_hits > 0
    for a, b in zip(imgs, single):
        assert np.array_equal(a.cpu().numpy(), b)


def test_batch_rejects_mixed_stepsize(setup):
    wl, model, r, views = setup
    other = S.small_plain(stepsize_setup=lib.stepsize_setup)
    rm2 = other.rm(width=32, height=32)
    rm2.stepsize_info.stepsize *= 2.0
    lib.stepsize_setup(rm2.stepsize_info, other.scene)
    with pytest.raises(lib.FvcError):
        r.render_full_batch([views[0], rm2])


def test_mlp_rows_32_and_64_equal_16(setup):
    wl, model, r, views = setup
    rm = views[2]
    out = {}
    r.initCacheBlockwise(views[0])
    for rows in (16, 32, 64):
        r.set_mlp_rows(rows)
        a, b = _img(rm), _img(rm)
        r.render_full(rm, a)
        r.render_cached(rm, b)
        out[rows] = (a.cpu().numpy(), b.cpu().numpy())
    r.set_mlp_rows(16)
    for rows in (32, 64):
        assert np.array_equal(out[16][0], out[rows][0]), rows
        assert np.array_equal(out[16][1], out[rows][1]), rows


def test_autotune_keeps_images(setup):
    wl, model, r, views = setup
    rm = views[0]
    before = _img(rm)
    r.render_full(rm, before)
    res = r.autotune(rm, cached=False)
    print(f"[tune] rows {res.mlp_rows} target 2^{int(np.log2(res.target_batch))} steps {res.max_steps}: "
          f"{res.ms_best:.3f} ms (default {res.ms_default:.3f} ms, {res.n_tried} tried)", flush=True)
    assert res.n_tried == 27 and res.mlp_rows in (16, 32, 64) and res.max_steps in (16, 32, 64)
    assert res.ms_best <= min(res.ms[:res.n_tried]) + 1e-6
    after = _img(rm)
    r.render_full(rm, after)
    assert np.array_equal(before.cpu().numpy(), after.cpu().numpy())


def test_graph_mode_replays_equal_eager(setup):
    """Frames recorded and replayed as hipGraphs (updated in place from frame
    to frame, eager work in between) give the eager images; a frame that
    needs more sub-iterations than recorded finishes with ordinary launches."""
    wl, model, r, views = setup
    path = [wl.rm(S.yaw(wl.cam2world, 0.5 * k), 160, 120) for k in range(6)]
    r.initCacheBlockwise(path[0])
    eager = []
    for rm in path:
        a, b = _img(rm), _img(rm)
        r.render_full(rm, a)
        r.render_cached(rm, b)
        eager.append((a.cpu().numpy(), b.cpu().numpy()))
    r.set_graph(True)
    try:
        small = wl.rm(wl.cam2world, 16, 12)
        r.render_full(small, _img(small))     # few sub-iterations recorded ...
        for k, rm in enumerate(path):         # ... then a larger frame continues eagerly
            a, b = _img(rm), _img(rm)
            r.render_full(rm, a)
            torch.cuda.synchronize()
            _ = torch.ones(1024, device="cuda").sum().item()  # eager work between replays
            r.render_cached(rm, b)
            torch.cuda.synchronize()
            assert np.array_equal(a.cpu().numpy(), eager[k][0]), k
            assert np.array_equal(b.cpu().numpy(), eager[k][1]), k
    finally:
        r.set_graph(False)
