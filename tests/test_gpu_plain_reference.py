"""The reference's plain-model layout (NerfNetwork::init,
nerf_network.h:88-118, 334-342): the density MLP's n_output rows, log sigma
first, are the colour head's latents (latent_offset 0), and the cache keeps
the density in latent 0 (DensityIntermediates, the variant eval.cu:216
picks; cache_renderer.cu:481-486; lookup cache_performance_renderer.cu:
770-779).  Full render, fill, cached render and the nerfacc loader against
the oracle."""
import numpy as np
import pytest
import torch

from frustumvolumecaching_amd import abi, lib, synthetic as S

from .gpu_util import f16_ulps, to_8bit

pytestmark = pytest.mark.gpu

RES = 96


@pytest.fixture(scope="module")
def setup(oracle_mod):
    wl = S.small_plain_reference(stepsize_setup=lib.stepsize_setup, n_output=8)
    model = lib.NerfNetwork.from_workload(wl)
    r = lib.Renderer(model)
    yield wl, model, r, oracle_mod.Model.from_workload(wl)
    r.close()
    model.close()


def _result(r, rm):
    W, H = rm.cam_info.res_x, rm.cam_info.res_y
    res = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
    r.last_result(res, W * H * 4)
    return res.cpu().numpy().reshape(H, W, 4)


def test_full_render_matches_oracle(setup, oracle_mod):
    wl, model, r, om = setup
    rm = wl.rm(width=RES, height=RES)
    st = r.render_full(rm)
    ref, n = om.render_full(rm, oracle_mod.EMU16)
    assert st.n_resamples == n
    assert np.abs(to_8bit(_result(r, rm)) - to_8bit(ref)).max() <= 2


def test_fill_and_cached_render_match_oracle(setup, oracle_mod):
    wl, model, r, om = setup
    rm0 = wl.rm(width=RES, height=RES)
    st = r.initCacheBlockwise(rm0)
    oc = om.cache_fill(rm0, oracle_mod.EMU16)
    assert np.array_equal(r.brick_table(), oc.brick_table())
    assert np.array_equal(r.known_ranges().view(np.uint32), oc.known_ranges().view(np.uint32))
    assert st.n_samples_evaluated == oc.stats.n_samples_evaluated
    gp, op = r.pool(), oc.pool()
    assert np.array_equal(gp["isset"], op[..., 9])
    lat = gp["lat"].view(np.float16).astype(np.float64)
    set_ = op[..., 9] == 1
    # DensityIntermediates: latent 0 is the density as half
    assert np.array_equal(lat[..., 0][set_], gp["density"][set_].astype(np.float16).astype(np.float64))
    assert np.quantile(f16_ulps(lat, op[..., :8]), 0.999) <= 2
    rm1 = wl.rm(S.yaw(wl.cam2world, 2.0), RES, RES)
    codes, cls, comp, _ = r.debug_render_cached(rm1, 64)
    got = _result(r, rm1)
    ref, nr, nc = oc.render(rm1)
    assert comp[..., 1].sum() > 0
    assert np.abs(to_8bit(got) - to_8bit(ref)).max() <= 2
    _, o_codes, o_cnt = oc.render_record(rm1, 64)
    m = np.minimum(o_cnt[..., 0], 64)
    idx = np.arange(64)[None, None, :] < m[..., None]
    assert np.array_equal(codes[idx], o_codes[idx])
    assert np.array_equal(comp, o_cnt[..., 1:])


def test_nerfacc_export_loads_the_reference_layout(setup, tmp_path):
    wl, model, r, om = setup
    d = S.write_nerfacc_export(wl, str(tmp_path / "ref"))
    loaded, _, _ = lib.NerfNetwork.from_nerfacc(d)
    r2 = lib.Renderer(loaded)
    rm = wl.rm(width=RES, height=RES)
    a = torch.zeros(RES * RES * 4, dtype=torch.uint8, device="cuda")
    b = torch.zeros_like(a)
    r.render_full(rm, a)
    r2.render_full(rm, b)
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    r2.close()
    loaded.close()


def test_sixteen_outputs_render_full_and_refuse_the_cache(oracle_mod):
    wl = S.small_plain_reference(stepsize_setup=lib.stepsize_setup, n_output=16)
    model = lib.NerfNetwork.from_workload(wl)
    r = lib.Renderer(model)
    rm = wl.rm(width=64, height=64)
    r.render_full(rm)
    ref, _ = oracle_mod.Model.from_workload(wl).render_full(rm, oracle_mod.EMU16)
    assert np.abs(to_8bit(_result(r, rm)) - to_8bit(ref)).max() <= 2
    # CachePerformanceRenderer refuses latent widths above 8 (cache_performance_renderer.cu:964-967)
    with pytest.raises(lib.FvcError) as e:
        r.initCacheBlockwise(rm)
    assert e.value.code == lib.E_UNSUPPORTED
    r.close()
    model.close()
