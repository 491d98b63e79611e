"""Background refill of a double-buffered renderer: the fill of the back
cache runs on its own thread and stream while frames render from the front
cache; published caches give the images of synchronously filled ones."""
import numpy as np
import pytest
import torch

from frustumvolumecaching_amd import abi, lib, synthetic as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    wl = S.small_plain(stepsize_setup=lib.stepsize_setup)
    model = lib.NerfNetwork.from_workload(wl)
    path = [wl.rm(S.translate_forward(S.yaw(wl.cam2world, 1.5 * k), 0.05 * k), 128, 96) for k in range(8)]
    yield wl, model, path
    model.close()


def _img(rm):
    return torch.zeros(rm.cam_info.res_y, rm.cam_info.res_x, 4, dtype=torch.uint8, device="cuda")


def _sync_cached(model, fill_rm, rm):
    r = lib.Renderer(model)
    r.initCacheBlockwise(fill_rm)
    img = _img(rm)
    r.render_cached(rm, img)
    out = img.cpu().numpy()
    r.close()
    return out


def test_async_fill_overlaps_rendering(setup):
    wl, model, path = setup
    r = lib.Renderer(model, double_buffered=True)
    r.initCacheBlockwise(path[0])
    r.swapCache()
    r.initCacheAsync(path[5])
    frames = []
    for rm in path[1:5]:                 # render from the front cache meanwhile
        img = _img(rm)
        r.render_cached(rm, img)
        frames.append(img.cpu().numpy())
    state, stats = r.fill_poll(wait=True)
    assert state == abi.FILL_DONE and stats.n_bricks_set > 0
    assert r.publish()
    assert r.fill_poll()[0] == abi.FILL_IDLE
    img = _img(path[5])
    r.render_cached(path[5], img)
    assert np.array_equal(img.cpu().numpy(), _sync_cached(model, path[5], path[5]))
    for rm, got in zip(path[1:5], frames):
        assert np.array_equal(got, _sync_cached(model, path[0], rm))
    r.close()


def test_refill_policy_triggers_and_publishes(setup):
    wl, model, path = setup
    r = lib.Renderer(model, double_buffered=True)
    r.set_refill_policy(min_hit_ratio=0.0, max_distance=0.12)
    started_at = None
    for k, rm in enumerate(path):
        img = _img(rm)
        st, started = r.render_cached_auto(rm, img)
        assert st.samples_ppx > 0
        if started:
            started_at = k
            r.fill_poll(wait=True)       # make the next frame's source deterministic
            break
    assert started_at is not None and started_at > 0
    nxt = path[started_at + 1]
    img = _img(nxt)
    r.render_cached_auto(nxt, img)       # publishes the fill made at started_at first
    assert np.array_equal(img.cpu().numpy(), _sync_cached(model, path[started_at], nxt))
    r.close()


def test_async_fill_needs_double_buffering(setup):
    wl, model, path = setup
    r = lib.Renderer(model)
    with pytest.raises(lib.FvcError):
        r.initCacheAsync(path[0])
    r.close()


def test_cache_getters_wait_for_a_running_fill(setup):
    """The back cache's contents are read under the fill lock: a getter called
    while a background fill runs returns the finished fill's table, equal to
    a synchronous fill's; the front cache stays readable meanwhile."""
    wl, model, path = setup
    r = lib.Renderer(model, double_buffered=True)
    r.initCacheBlockwise(path[0])
    r.swapCache()
    front = r.brick_table(0)
    r.initCacheAsync(path[3])
    assert np.array_equal(r.brick_table(0), front)
    back = r.brick_table(1)
    state, stats = r.fill_poll(wait=True)
    assert state == abi.FILL_DONE
    assert r.cache_info(1).n_bricks_used == stats.n_bricks_set
    ref = lib.Renderer(model)
    ref.initCacheBlockwise(path[3])
    assert np.array_equal(back, ref.brick_table(0))
    ref.close()
    r.close()
