"""Pins of the oracle's view-dependent split model (NerfNetwork::init_viewdep,
nerf_network.h:120-158; custom colour head, my_fully_fused_mlp.cu:677-895;
DensityIntermediates cache variant, cache_renderer.cu:481-486 and
cache_performance_renderer.cu:770-787), on the CPU.

* The head's input assembly: with first-layer rows +e_k, -e_k and an output
  row (1, -1), the colour output is input k itself, for every k, and is
  compared with closed forms: the given latents, SH4 of the view direction
  (orc_sh4, pinned in test_oracle_pins), SH4 differences, sin / cos(pi 2^f p).
* The first head against float64 matmuls of the weights (a transposed
  operand or a shifted latent row fails).
* The DensityIntermediates lookup at a voxel centre returns log(sigma) as
  latent 0 and as the density row.
* Cached render of a split model at the fill pose against its full render.
* The nerfacc export of a split model reads back through the product's
  host-side loader.
"""
import ctypes as C
import math

import numpy as np
import pytest

from frustumvolumecaching_amd import synthetic as S


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def _wl(O, di=True, fl=1, freq=4, log2_hashmap_size=12, w=64, h=64):
    wl = S.small_viewdep(w, h, first_latent_is_density=di, first_head_hidden_layers=fl, freq_encoding_degree=freq,
                         stepsize_setup=O.stepsize_setup)
    if log2_hashmap_size != wl.net.log2_hashmap_size:
        wl.net.log2_hashmap_size = log2_hashmap_size
        wl.params = S.make_params(wl.net, S.SEEDS["weights"], S.SEEDS["table"])
    return wl


def _f16(bits):
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float64)


def _sh4(O, d01):
    d64 = np.ascontiguousarray(d01, np.float64)
    out = np.zeros((len(d64), 16))
    O.lib().orc_sh4(C.c_void_p(d64.ctypes.data), len(d64), C.c_void_p(out.ctypes.data))
    return out


def _unit01(rng, n):
    d = rng.standard_normal((n, 3))
    return ((d / np.linalg.norm(d, axis=1, keepdims=True) + 1) / 2).astype(np.float32)


@pytest.mark.parametrize("freq", [4, 0])
def test_head_input_assembly(O, freq):
    """Custom head input k (my_fully_fused_mlp.cu:738-790): latents 0..7,
    SH4(dir) 8..23, SH4(dir) - SH4(init) 24..39, then per axis x, y, z
    sin(pi 2^f p) for f = 0..3 and cos likewise (freqEncChunk4), or 8 padding
    ones without frequencies."""
    wl = _wl(O, freq=freq)
    pos, base, first, head = wl.params
    hin_w = S.head_input_width(wl.net)
    rng = np.random.default_rng(21)
    n = 16
    lat = rng.uniform(-2, 2, (n, 16)).astype(np.float32)
    p = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    d01, i01 = _unit01(rng, n), _unit01(rng, n)
    sh_d, sh_i = _sh4(O, d01), _sh4(O, i01)
    expect = np.zeros((n, hin_w))
    expect[:, :8] = lat[:, :8]
    expect[:, 8:24] = sh_d
    expect[:, 24:40] = sh_d - sh_i
    if freq == 4:
        for a in range(3):
            for c in range(2):
                for f in range(4):
                    v = np.pi * 2.0 ** f * p[:, a].astype(np.float64)
                    expect[:, 40 + 8 * a + 4 * c + f] = np.cos(v) if c else np.sin(v)
    else:
        expect[:, 40:48] = 1.0
    HW = wl.net.head_width
    for k0 in range(0, hin_w, 3):
        H1 = np.zeros((HW, hin_w))
        H2 = np.zeros((16, HW))
        ks = list(range(k0, min(k0 + 3, hin_w)))
        for c, k in enumerate(ks):
            H1[2 * c, k], H1[2 * c + 1, k] = 1.0, -1.0
            H2[c, 2 * c], H2[c, 2 * c + 1] = 1.0, -1.0
        head_k = S.f16_bits(np.concatenate([H1.ravel(), H2.ravel()]))
        model = O.Model(wl.scene, wl.net, wl.occupancy, (pos, base, first, head_k))
        got = model.head_viewdep(lat, p, d01, i01, mode=O.F64)
        for c, k in enumerate(ks):
            assert np.allclose(got[:, c], expect[:, k], rtol=0, atol=1e-12), k


@pytest.mark.parametrize("di,fl", [(True, 1), (False, 1), (True, 2), (False, 2)])
def test_first_head_matches_float64_matmul(O, di, fl):
    """latents = the reference buffer rows latent_offset.. after
    inferenceLatents: row 0 = log sigma (density-MLP row 0), row r >= 1 =
    first-head output r - 1; first head F3 relu((F2) relu(F1 [base, SH4(init)]))."""
    wl = _wl(O, di=di, fl=fl)
    model = O.Model.from_workload(wl)
    first = _f16(wl.params[2])
    FW = wl.net.first_head_width
    F1 = first[:32 * FW].reshape(FW, 32)
    off = 32 * FW
    F2 = None
    if fl == 2:
        F2 = first[off:off + FW * FW].reshape(FW, FW)
        off += FW * FW
    F3 = first[off:].reshape(16, FW)
    rng = np.random.default_rng(22)
    pts = rng.random((64, 3)).astype(np.float32)
    i01 = _unit01(rng, 64)
    base16 = model.latents(pts, mode=O.F64)
    x = np.concatenate([base16, _sh4(O, i01)], 1)
    h = np.maximum(F1 @ x.T, 0)
    if F2 is not None:
        h = np.maximum(F2 @ h, 0)
    fo = (F3 @ h).T
    rows = np.concatenate([base16[:, :1], fo], 1)  # reference buffer rows 0..16
    lo = wl.net.latent_offset
    got = model.latents_viewdep(pts, i01, mode=O.F64)
    scale = max(1.0, np.abs(rows).max())
    assert np.abs(got[:, :8] - rows[:, lo:lo + 8]).max() < 1e-12 * scale
    assert np.array_equal(got[:, 8], base16[:, 0])
    assert not np.any(got[:, 9:])


def test_inference_uses_the_view_direction_as_initial_direction(O):
    """Full render (performance_renderer.cu:269): the SH differences are 0 and
    the density is the density-MLP row 0."""
    wl = _wl(O)
    model = O.Model.from_workload(wl)
    rng = np.random.default_rng(23)
    pts = rng.random((32, 3)).astype(np.float32)
    d01 = _unit01(rng, 32)
    for mode in (O.F64, O.EMU16):
        lat = model.latents_viewdep(pts, d01, mode=mode)
        ref = model.head_viewdep(lat.astype(np.float32), pts, d01, d01, mode=mode)
        got = model.inference(pts, d01, mode=mode)
        assert np.abs(got - ref).max() < (1e-6 if mode == O.F64 else 1e-3)
        assert np.array_equal(got[:, 3], model.latents(pts, mode=mode)[:, 0])


@pytest.fixture(scope="module")
def filled_vd(O):
    out = {}
    for di in (True, False):
        wl = _wl(O, di=di, log2_hashmap_size=19, w=128, h=128)
        model = O.Model.from_workload(wl)
        rm = wl.rm()
        crop = (32, 32, 64, 64)
        out[di] = (wl, model, rm, crop, model.cache_fill(rm, O.F32, crop=crop))
    return out


def test_density_intermediates_fill_and_lookup(O, filled_vd):
    """DensityIntermediates: a set voxel's latent 0 is its density; the
    lookup at its centre returns log(density) as latent 0 and as the density
    row, and the other latents as stored."""
    wl, model, rm, crop, cache = filled_vd[True]
    tab = cache.brick_table()
    pool = cache.pool()
    set_ = pool[..., 9] == 1
    assert set_.sum() > 100
    assert np.array_equal(pool[..., 0][set_], pool[..., 8][set_])
    bz, by, bx = np.nonzero(tab >= 0)
    rng = np.random.default_rng(24)
    pts, exp = [], []
    for k in rng.choice(len(bz), size=min(40, len(bz)), replace=False):
        slot = tab[bz[k], by[k], bx[k]]
        zz, yy, xx = np.nonzero(pool[slot, ..., 9] == 1)
        j = rng.integers(len(zz))
        pts.append(np.array((bx[k] * 8 + xx[j], by[k] * 8 + yy[j], bz[k] * 8 + zz[j]), np.float64) + 0.5)
        rec = pool[slot, zz[j], yy[j], xx[j]]
        exp.append(np.concatenate([[math.log(rec[8])], [math.log(rec[8])], rec[1:8]]))
    got = cache.lookup(np.array(pts, np.float32))
    assert np.allclose(got, np.array(exp), rtol=1e-6, atol=1e-6)
    # without the variant, row 0 is log(max(sum w sigma, 1e-6)) and latent 0 the stored value
    plain = cache.lookup(np.array(pts, np.float32), di=False)
    assert np.allclose(plain[:, 0], np.array(exp)[:, 0], rtol=1e-6, atol=1e-6)
    assert np.allclose(plain[:, 1], [math.exp(e[0]) for e in exp], rtol=1e-6)


@pytest.mark.parametrize("di", [True, False])
def test_split_cached_render_at_fill_pose(O, filled_vd, di):
    """The reference's sanity gate (35 dB at the fill pose, eval.cu:315-326)
    holds with margin for the split model, and a 2 deg yaw stays above it."""
    wl, model, rm, crop, cache = filled_vd[di]
    inner = (crop[0] + 8, crop[1] + 8, crop[2] - 16, crop[3] - 16)
    full, _ = model.render_full(rm, O.F32, crop=inner)
    cached, nr, nc = cache.render(rm, crop=inner)
    assert nc > 10 * max(nr, 1)
    assert O.psnr_8bit(O.to_8bit(cached), O.to_8bit(full)) > 40.0


def test_split_export_reads_back(tmp_path):
    """config.json with "mlp_first_head" and mlp_first_head.dat (loader.cu:251-262)
    through the product's host-side reader."""
    from frustumvolumecaching_amd import lib
    try:
        lib.load()
    except FileNotFoundError:
        pytest.skip("libfvc.so not built")
    for di, fl, freq in ((True, 1, 4), (False, 2, 0)):
        wl = S.small_viewdep(first_latent_is_density=di, first_head_hidden_layers=fl, freq_encoding_degree=freq,
                             stepsize_setup=lambda st, sc: lib.stepsize_setup(st, sc))
        d = tmp_path / f"vd_{int(di)}_{fl}_{freq}"
        S.write_nerfacc_export(wl, str(d))
        scene, step, net, params, occ = lib.nerfacc_read(str(d))
        assert net.split_head == 1 and net.first_latent_is_density == int(di)
        assert net.latent_offset == (0 if di else 1) and net.n_latents == 8
        assert net.first_head_n_hidden_layers == fl and net.freq_encoding_degree == freq
        assert len(params) == 4
        for a, b in zip(params, wl.params):
            assert np.array_equal(a, b)
        assert lib.net_param_counts_viewdep(net) == S.param_counts(wl.net)
