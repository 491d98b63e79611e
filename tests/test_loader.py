"""Model file layout of the reference's nerfacc export (load_nerfacc_export,
loader.cu:232-285; parse_helper.h:25-63): config.json + raw .dat blocks."""
import json
import os

import numpy as np
import pytest

from frustumvolumecaching_amd import abi, lib, synthetic as S


def _fields(st):
    return {k: getattr(st, k) for k, _ in st._fields_ if not k.startswith("_")}


@pytest.fixture(scope="module")
def wl():
    return S.unbounded_1080p(width=64, height=48, log2_hashmap_size=19)


@pytest.mark.parametrize("pos_separate", [True, False])
def test_export_roundtrip(tmp_path, wl, pos_separate):
    d = S.write_nerfacc_export(wl, str(tmp_path / "m"), pos_separate=pos_separate)
    scene, step, net, params, occ = lib.nerfacc_read(d)
    ref_scene = _fields(wl.scene)
    got_scene = _fields(scene)
    for k in ("grid_resolution", "grid_nlvl", "vals_per_lvl", "normalized", "contraction_type", "is_open_gl",
              "alpha_thre"):
        assert got_scene[k] == ref_scene[k], k
    for k in ("aabb_from", "aabb_to", "contraction_aabb_from", "contraction_aabb_to"):
        a, b = getattr(scene, k), getattr(wl.scene, k)
        assert (a.x, a.y, a.z) == (b.x, b.y, b.z), k
    # StepsizeInfo::setAdditionalInfo applied, bit-identical to the direct setup
    assert bytes(step) == bytes(wl.step)
    for k in ("n_levels", "n_features_per_level", "log2_hashmap_size", "base_resolution", "per_level_scale",
              "base_width", "base_n_hidden_layers", "base_n_output", "head_width", "head_n_hidden_layers",
              "n_latents", "latent_offset", "sh_degree"):
        assert getattr(net, k) == getattr(wl.net, k), k
    for got, ref in zip(params, wl.params):
        assert np.array_equal(got, np.asarray(ref, np.uint16))
    assert np.array_equal(occ, np.asarray(wl.occupancy, np.uint8).reshape(-1))


def test_bounded_scene_has_aabb_contraction(tmp_path):
    small = S.small_plain(width=32, height=32)
    d = S.write_nerfacc_export(small, str(tmp_path / "s"))
    scene, step, net, _, _ = lib.nerfacc_read(d)
    assert scene.contraction_type == abi.CONTRACTION_AABB
    assert bytes(step) == bytes(small.step)


def _expect_error(d, code):
    with pytest.raises(lib.FvcError) as e:
        lib.nerfacc_read(d)
    assert e.value.code == code


def test_export_errors(tmp_path, wl):
    _expect_error(str(tmp_path / "absent"), lib.E_INVALID)
    d = S.write_nerfacc_export(wl, str(tmp_path / "trunc"))
    with open(os.path.join(d, "mlp_head.dat"), "r+b") as f:
        f.truncate(100)
    _expect_error(d, lib.E_INVALID)
    d = S.write_nerfacc_export(wl, str(tmp_path / "split"))
    cfg = json.load(open(os.path.join(d, "config.json")))
    cfg["mlp_first_head"] = {"n_output": 8}  # no "network": malformed
    json.dump(cfg, open(os.path.join(d, "config.json"), "w"))
    _expect_error(d, lib.E_INVALID)
    # a split head outside the built instance (first head 128 wide)
    cfg["mlp_first_head"] = {"n_output": 8, "n_input": 32, "encoding": {"otype": "SphericalHarmonics", "degree": 4},
                             "network": {"n_neurons": 128, "n_hidden_layers": 1}}
    cfg["mlp_head"]["freq_encoding_degree"] = 4
    json.dump(cfg, open(os.path.join(d, "config.json"), "w"))
    _expect_error(d, lib.E_UNSUPPORTED)
    with open(os.path.join(d, "config.json"), "w") as f:
        f.write('{"scene": {"grid_resolution": 128,')
    _expect_error(d, lib.E_INVALID)
    d = S.write_nerfacc_export(wl, str(tmp_path / "wide"))
    cfg = json.load(open(os.path.join(d, "config.json")))
    cfg["mlp_base"]["network"]["n_neurons"] = 128
    json.dump(cfg, open(os.path.join(d, "config.json"), "w"))
    _expect_error(d, lib.E_UNSUPPORTED)


@pytest.mark.gpu
def test_loaded_model_renders_like_the_direct_one(tmp_path, wl):
    import torch
    d = S.write_nerfacc_export(wl, str(tmp_path / "g"), pos_separate=False)
    direct = lib.NerfNetwork.from_workload(wl)
    loaded, scene, step = lib.NerfNetwork.from_nerfacc(d)
    imgs = []
    for m in (direct, loaded):
        r = lib.Renderer(m)
        img = torch.zeros(wl.height * wl.width * 4, dtype=torch.uint8, device="cuda")
        r.render_full(wl.rm(), img)
        torch.cuda.synchronize()
        imgs.append(img.cpu().numpy())
        r.close()
    assert np.array_equal(imgs[0], imgs[1])
    assert np.array_equal(direct.bitfield(), loaded.bitfield())


def test_reference_plain_export_layout(tmp_path):
    """A plain export as the reference writes it (no "latent_offset" key):
    NerfNetwork::init makes the n_output rows, log sigma first, the head's
    latents (nerf_network.h:93, 334-342) -- latent_offset 0,
    first_latent_is_density 1, n_latents = n_output."""
    ref = S.small_plain_reference(width=32, height=32, n_output=8)
    d = S.write_nerfacc_export(ref, str(tmp_path / "r"))
    cfg = json.load(open(os.path.join(d, "config.json")))
    assert "latent_offset" not in cfg["mlp_head"] and cfg["mlp_head"]["n_input"] == 24
    _, _, net, params, _ = lib.nerfacc_read(d)
    assert (net.latent_offset, net.n_latents, net.first_latent_is_density, net.base_n_output) == (0, 8, 1, 8)
    for got, exp in zip(params, ref.params):
        assert np.array_equal(got, np.asarray(exp, np.uint16))
    # the same file without n_input (its default): still the reference layout
    del cfg["mlp_head"]["n_input"]
    cfg["mlp_base"]["n_output"] = 16
    json.dump(cfg, open(os.path.join(d, "config.json"), "w"))
    _, _, net, _, _ = lib.nerfacc_read(d)
    assert (net.latent_offset, net.n_latents, net.first_latent_is_density) == (0, 16, 1)
    # the synthetic separate-density layout carries its own key
    sep = S.small_plain(width=32, height=32)
    d2 = S.write_nerfacc_export(sep, str(tmp_path / "s"))
    _, _, net2, _, _ = lib.nerfacc_read(d2)
    assert (net2.latent_offset, net2.n_latents, net2.first_latent_is_density, net2.base_n_output) == (1, 8, 0, 9)
    # a head n_input that is not latents + SH is refused
    cfg2 = json.load(open(os.path.join(d2, "config.json")))
    cfg2["mlp_head"]["n_input"] = 40
    json.dump(cfg2, open(os.path.join(d2, "config.json"), "w"))
    _expect_error(d2, lib.E_UNSUPPORTED)


def test_param_counts_of_reference_layouts(oracle_mod):
    for n_out in (1, 8, 16):
        net = S.net_desc(layout="reference", n_output=n_out)
        assert lib.net_param_counts(net) == S.param_counts(net) == oracle_mod.net_param_counts(net)
    bad = S.net_desc(layout="reference", n_output=8)
    bad.n_latents = 7
    with pytest.raises(lib.FvcError):
        lib.net_param_counts(bad)
