"""Multi-view batches at full size: sixteen views of 320x180, 640x360,
1280x720 and 1920x1080 along the 1080p camera path (hash grid T=2^21, five
cascades, cone steps, L-inf contraction), and scenes of 2 to 64 blobs that
share one network.  Bars:

* full and cached batches equal the single-view renders bit for bit (images
  and float results): a batch only moves sub-iteration boundaries, every ray
  composites the same samples in the same order;
* crops of a 1920x1080 and a 640x360 view of the batch match the
  emulated-FP16 oracle within 2/255 (full render; cached render on the
  GPU's cache contents);
* a mixed-scene batch marches each view through its own scene's occupancy:
  every view equals the single render of its scene, per-ray sample counts
  span the range the blob counts give.
"""
import numpy as np
import pytest
import torch

from frustumvolumecaching_amd import lib, synthetic as S

from .gpu_util import to_8bit

pytestmark = pytest.mark.gpu


def _img(rm):
    return torch.zeros(rm.cam_info.res_y, rm.cam_info.res_x, 4, dtype=torch.uint8, device="cuda")


def _px(rm):
    return rm.cam_info.res_x * rm.cam_info.res_y


def _results(r, n_px):
    out = torch.empty(n_px * 4, dtype=torch.float32, device="cuda")
    r.last_result(out, n_px * 4)
    return out.cpu().numpy()


@pytest.fixture(scope="module")
def setup():
    wl = S.unbounded_1080p(stepsize_setup=lib.stepsize_setup)
    model = lib.NerfNetwork.from_workload(wl)
    r = lib.Renderer(model)
    path = S.camera_path_1080p(wl, 10)
    views = S.batch_views(wl, path)
    yield wl, model, r, path, views
    r.close()
    model.close()


def _singles(r, views, kind):
    imgs, res = [], []
    for rm in views:
        img = _img(rm)
        getattr(r, f"render_{kind}")(rm, img)
        imgs.append(img.cpu().numpy())
        res.append(_results(r, _px(rm)))
    return imgs, res


def _check_batch(r, views, kind, single_imgs, single_res, models=None):
    imgs = [_img(rm) for rm in views]
    if models is None:
        st = getattr(r, f"render_{kind}_batch")(views, imgs)
    else:
        st = r.render_full_batch(views, imgs, models=models)
    for k, (a, b) in enumerate(zip(imgs, single_imgs)):
        assert np.array_equal(a.cpu().numpy(), b), k
    res = _results(r, sum(_px(rm) for rm in views))
    assert np.array_equal(res.view(np.uint32), np.concatenate(single_res).view(np.uint32))
    return st, res


def _view_result(res, views, k):
    off = sum(_px(rm) for rm in views[:k])
    rm = views[k]
    return res[4 * off:4 * (off + _px(rm))].reshape(rm.cam_info.res_y, rm.cam_info.res_x, 4)


CROPS = ((3, (930, 520, 40, 40)), (3, (100, 60, 32, 32)), (1, (300, 160, 40, 40)))


def test_full_batch_equals_single_and_oracle(setup, oracle_mod):
    wl, model, r, path, views = setup
    imgs, res = _singles(r, views, "full")
    st, bres = _check_batch(r, views, "full", imgs, res)
    print(f"[batch16] full: {st.n_resamples} samples over {sum(map(_px, views))} px, "
          f"{st.n_subiters} sub-iterations", flush=True)
    om = oracle_mod.Model.from_workload(wl)
    for k, crop in CROPS:
        x0, y0, w, h = crop
        got = _view_result(bres, views, k)[y0:y0 + h, x0:x0 + w]
        ref, _ = om.render_full(views[k], oracle_mod.EMU16, crop=crop)
        d = np.abs(to_8bit(got) - to_8bit(ref))
        print(f"[batch16] full view {k} crop {crop}: max |d| {d.max()}/255", flush=True)
        assert d.max() <= 2


def test_cached_batch_equals_single_and_oracle(setup, oracle_mod):
    wl, model, r, path, views = setup
    rm0 = wl.rm(path[0])
    r.initCacheBlockwise(rm0)
    imgs, res = _singles(r, views, "cached")
    st, bres = _check_batch(r, views, "cached", imgs, res)
    print(f"[batch16] cached: {st.n_cache_hits} cache / {st.n_resamples} resampled samples", flush=True)
    assert st.n_cache_hits > 0
    om = oracle_mod.Model.from_workload(wl)
    cache = oracle_mod.Cache.from_arrays(om, rm0, r.cache_info(0), r.brick_table(0), r.known_ranges(0),
                                         r.pool_decoded(0), mode=oracle_mod.EMU16)
    for k, crop in CROPS:
        x0, y0, w, h = crop
        got = _view_result(bres, views, k)[y0:y0 + h, x0:x0 + w]
        ref, _, _ = cache.render(views[k], crop=crop)
        d = np.abs(to_8bit(got) - to_8bit(ref))
        print(f"[batch16] cached view {k} crop {crop}: max |d| {d.max()}/255", flush=True)
        assert d.max() <= 2


@pytest.fixture(scope="module")
def scenes(setup):
    wl, model, r, path, views = setup
    out = []
    for nb in S.BATCH_BLOB_COUNTS:
        w = S.unbounded_1080p(stepsize_setup=lib.stepsize_setup, n_blobs=nb)
        out.append((nb, w, lib.NerfNetwork.from_workload(w)))
    yield out
    for _, _, m in out:
        m.close()


def test_mixed_scene_batch_equals_each_scenes_render(setup, scenes, oracle_mod):
    wl, model, r, path, views = setup
    models = [scenes[k % len(scenes)][2] for k in range(len(views))]
    # single renders: a renderer per scene
    imgs, res = [], []
    spp = {}
    for k, rm in enumerate(views):
        rk = lib.Renderer(models[k])
        try:
            img = _img(rm)
            st = rk.render_full(rm, img)
            imgs.append(img.cpu().numpy())
            res.append(_results(rk, _px(rm)))
            nb = scenes[k % len(scenes)][0]
            spp[nb] = spp.get(nb, []) + [st.n_resamples / _px(rm)]
        finally:
            rk.close()
    st, bres = _check_batch(r, views, "full", imgs, res, models=models)
    print(f"[batch16] mixed scenes: samples per pixel by blob count "
          f"{ {nb: round(float(np.mean(v)), 2) for nb, v in spp.items()} }", flush=True)
    # more blobs, more occupied space
    assert np.mean(spp[S.BATCH_BLOB_COUNTS[0]]) < np.mean(spp[S.BATCH_BLOB_COUNTS[-1]])
    # a crop of a 64-blob view (640x360) against that scene's oracle
    k = 5
    om = oracle_mod.Model.from_workload(scenes[k % len(scenes)][1])
    crop = (300, 160, 40, 40)
    x0, y0, w, h = crop
    got = _view_result(bres, views, k)[y0:y0 + h, x0:x0 + w]
    ref, _ = om.render_full(views[k], oracle_mod.EMU16, crop=crop)
    assert np.abs(to_8bit(got) - to_8bit(ref)).max() <= 2


def test_mixed_scene_batch_rejects_other_networks(setup):
    wl, model, r, path, views = setup
    other = S.unbounded_1080p(stepsize_setup=lib.stepsize_setup, log2_hashmap_size=19)
    m2 = lib.NerfNetwork.from_workload(other)
    try:
        with pytest.raises(lib.FvcError):
            r.render_full_batch(views[:2], models=[None, m2])
    finally:
        m2.close()


def test_autotune_batch_over_chunk_sizes(setup):
    """The tuner over a batch of one view per resolution, MLP rows {16, 64}
    and chunk sizes 2^16 .. 2^20 samples per launch: every candidate is
    timed, the choice is the fastest, and the images do not change."""
    wl, model, r, path, views = setup
    vs = views[:4]
    before = [_img(rm) for rm in vs]
    r.render_full_batch(vs, before)
    targets = tuple(1 << k for k in range(16, 21))
    res = r.autotune_batch(vs, rows=(16, 64), targets=targets, steps=(32,))
    print(f"[tune] batch of {len(vs)}: rows {res.mlp_rows} target 2^{int(np.log2(res.target_batch))} "
          f"steps {res.max_steps}: {res.ms_best:.2f} ms (default {res.ms_default:.2f} ms)", flush=True)
    assert res.n_tried == 10
    assert sorted(set(res.targets[:res.n_tried])) == list(targets)
    assert res.ms_best <= min(res.ms[:res.n_tried]) + 1e-6
    assert res.mlp_rows in (16, 64) and res.target_batch in targets and res.max_steps == 32
    after = [_img(rm) for rm in vs]
    r.render_full_batch(vs, after)
    for a, b in zip(before, after):
        assert torch.equal(a, b)
    r.set_batching(1 << 25, 32)
    r.set_mlp_rows(16)
