"""The 1080p workload (hash grid T=2^21, five occupancy cascades, cone-angle
steps, L-inf contraction) at 240x136 through the timed kernels: FP16 full
render, cache fill and reprojected frames along the camera path, against the
oracle's emulated-FP16 path.  Bars: sample counts, brick tables and known
ranges bit-exact; pool within fp16 tolerance; images within 2/255."""
import numpy as np
import pytest
import torch

from frustumvolumecaching_amd import lib, synthetic as S

from .gpu_util import f16_ulps, to_8bit

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(unbounded_wl, unbounded_model, oracle_mod):
    wl = unbounded_wl
    model = lib.NerfNetwork.from_workload(wl)
    r = lib.Renderer(model)
    path = S.camera_path_1080p(wl, 10)
    yield wl, model, r, path, unbounded_model
    r.close()
    model.close()


def _result(r, rm):
    W, H = rm.cam_info.res_x, rm.cam_info.res_y
    res = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
    r.last_result(res, W * H * 4)
    return res.cpu().numpy().reshape(H, W, 4)


def test_hashgrid_encoding_bit_exact(setup, oracle_mod):
    """The T=2^21 grid's encoding at random points and on the faces, edges and
    corners of the unit cube, where the dense levels' corner x + 1 wraps to
    the level's first entry."""
    from .gpu_util import dev, empty, host_f16
    wl, model, r, path, om = setup
    rng = np.random.default_rng(5)
    p = rng.random((8192, 3)).astype(np.float32)
    edge = np.array([0.0, 1e-7, 0.5, 0.9999999, 1.0], np.float32)
    g = np.stack(np.meshgrid(edge, edge, edge, indexing="ij"), -1).reshape(-1, 3)
    p[:len(g)] = g
    out = empty((len(p), 32), torch.int16)
    model.encode(dev(p), len(p), out)
    torch.cuda.synchronize()
    assert np.array_equal(host_f16(out), om.encode(p, oracle_mod.EMU16))


def test_full_render_matches_oracle(setup, oracle_mod):
    wl, model, r, path, om = setup
    for k in (0, 7):
        rm = wl.rm(path[k])
        st = r.render_full(rm)
        got = _result(r, rm)
        ref, n = om.render_full(rm, oracle_mod.EMU16)
        print(f"[unbounded] frame {k}: {st.n_resamples} samples (oracle {n}), "
              f"max |d| {np.abs(to_8bit(got) - to_8bit(ref)).max()}/255", flush=True)
        assert np.abs(to_8bit(got) - to_8bit(ref)).max() <= 2
        assert np.abs(got - ref).max() < 4e-3
        # the full renderer evaluates every sample of its launches, the
        # oracle stops a ray after its terminating sample
        assert st.n_resamples >= n


@pytest.fixture(scope="module")
def filled(setup, oracle_mod):
    wl, model, r, path, om = setup
    rm0 = wl.rm(path[0])
    st = r.initCacheBlockwise(rm0)
    oc = om.cache_fill(rm0, oracle_mod.EMU16)
    return rm0, st, oc


def test_fill_tables_bit_exact(setup, filled):
    wl, model, r, path, om = setup
    rm0, st, oc = filled
    gi, oi = r.cache_info(), oc.info
    for f in ("frustum_x", "frustum_y", "frustum_z", "bricks_x", "bricks_y", "bricks_z", "capacity_bricks",
              "n_bricks_used"):
        assert getattr(gi, f) == getattr(oi, f), f
    assert gi.n_bricks_used > 0
    assert np.array_equal(r.brick_table(), oc.brick_table())
    assert np.array_equal(r.known_ranges().view(np.uint32), oc.known_ranges().view(np.uint32))
    assert st.n_samples_evaluated == oc.stats.n_samples_evaluated
    assert st.n_subiters == oc.stats.n_subiters


def test_frustum_depth_is_the_reference_formula(setup, filled, oracle_mod):
    """The camera is inside the scene AABB: the frustum depth is
    ceil(step_from_t(|AABB diagonal| + near)) padded to 4
    (max_step_in_scene, raymarch_common.h:298-303; cache_renderer.cu:841-842)."""
    import math
    wl, model, r, path, om = setup
    rm0, _, _ = filled
    sc = wl.scene
    e = [np.float32(getattr(sc.aabb_to, a)) - np.float32(getattr(sc.aabb_from, a)) for a in "xyz"]
    diag = np.sqrt(np.float32(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]))
    st = rm0.stepsize_info
    steps = math.ceil(oracle_mod.lib().orc_step_from_t_f32(st, float(np.float32(diag + np.float32(st.near_)))))
    assert r.cache_info().frustum_z == (steps + 3) // 4 * 4


def test_fill_pool_matches_oracle(setup, filled):
    wl, model, r, path, om = setup
    _, _, oc = filled
    gp, op = r.pool(), oc.pool()
    assert np.array_equal(gp["isset"], op[..., 9])
    lat = gp["lat"].view(np.float16).astype(np.float64)
    u = f16_ulps(lat, op[..., :8])
    assert np.quantile(u, 0.999) <= 2
    set_ = op[..., 9] == 1
    rel = np.abs(gp["density"][set_] - op[..., 8][set_]) / op[..., 8][set_]
    assert rel.max() < 5e-3


@pytest.mark.parametrize("k", [1, 5, 9])
def test_cached_frames_match_oracle(setup, filled, oracle_mod, k):
    wl, model, r, path, om = setup
    _, _, oc = filled
    rm = wl.rm(path[k])
    st = r.render_cached(rm)
    got = _result(r, rm)
    ref, nr, nc = oc.render(rm)
    d = np.abs(to_8bit(got) - to_8bit(ref))
    print(f"[unbounded] cached frame {k}: {st.n_cache_hits} cache / {st.n_resamples} resampled "
          f"(oracle {nc} / {nr}), max |d| {d.max()}/255", flush=True)
    assert st.n_cache_hits > 0
    assert d.max() <= 2
