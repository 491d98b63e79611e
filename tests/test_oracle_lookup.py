"""Pins the oracle's trilinear cache lookup off the voxel centres.

Reference: generateNetworkInput_cache_shfl_perf_kernel_cg
(cache_performance_renderer.cu:678-787) with applyInterpolationFunction
(raymarch_common.h:393-402) and, for use_inter_brick_interpolation = false,
the brick clamp of froxel2data (raymarch_common.h:341-357).

The expected values never reuse the lookup's own weight expression
(1 - off)(1 - 2t) + t: they come from closed forms of the field that was
written into a hand-built cache (an affine field interpolates exactly; a
constant field is returned exactly by the Sum w*l / Sum w*isset
re-weighting, whatever corners are unset) and from the textbook trilinear
weights (1 - t, t).  Each pin was checked to fail under a plausible mistake
(marked "fails under" below).
"""
import math

import numpy as np
import pytest

from frustumvolumecaching_amd import abi, synthetic as S

FX, FY, FZ = 40, 24, 32          # frustum dims (voxels); 5 x 3 x 4 bricks


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


@pytest.fixture(scope="module")
def model(O):
    return O.Model.from_workload(S.small_plain(width=32, height=32))


@pytest.fixture(scope="module")
def rm():
    wl = S.small_plain(width=32, height=32)
    st = S.stepsize_info(0.0, math.sqrt(3.0) / 1024.0, 0.1, 8.0)
    return S.raymarch_info(S.camera_info(wl.cam2world, FX, FY), st)


def _info(n_slots):
    i = abi.CacheInfo()
    i.frustum_x, i.frustum_y, i.frustum_z = FX, FY, FZ
    i.bricks_x, i.bricks_y, i.bricks_z = FX // 8, FY // 8, FZ // 8
    i.capacity_bricks = i.n_bricks_used = n_slots
    i.initialized, i.brick_size = 1, 8
    return i


def _cache(O, model, rm, field, unset_brick=None, unset_voxels=None, seed=0, mode=None):
    """A cache whose voxel (x, y, z) (global voxel indices) holds field(x, y, z)
    = (8 latents, density) with isset 1, except voxels in unset bricks or
    marked in unset_voxels (all zero).  Bricks get shuffled pool slots, so
    neighbouring bricks are never neighbouring slots."""
    bx, by, bz = FX // 8, FY // 8, FZ // 8
    rng = np.random.default_rng(seed)
    order = rng.permutation(bx * by * bz)
    table = np.full((bz, by, bx), -1, np.int32)
    n = 0
    for b in order:
        z, y, x = np.unravel_index(b, (bz, by, bx))
        if unset_brick is not None and unset_brick(x, y, z):
            continue
        table[z, y, x] = n
        n += 1
    pool = np.zeros((n, 8, 8, 8, 10), np.float32)
    zz, yy, xx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    for z in range(bz):
        for y in range(by):
            for x in range(bx):
                s = table[z, y, x]
                if s < 0:
                    continue
                gx, gy, gz = x * 8 + xx, y * 8 + yy, z * 8 + zz
                vals = field(gx, gy, gz)                    # [9, 8, 8, 8]
                pool[s, ..., :9] = np.moveaxis(vals, 0, -1)
                pool[s, ..., 9] = 1.0
                if unset_voxels is not None:
                    m = unset_voxels(gx, gy, gz)
                    pool[s][m] = 0.0
    known = np.full((FY, FX), 10.0, np.float32)
    return O.Cache.from_arrays(model, rm, _info(n), table, known, pool, mode=O.F64 if mode is None else mode)


# affine latents and density with dyadic coefficients (exact in float32 at
# these coordinates)
_LA = np.array([[0.5, 1 / 64, -1 / 32, 1 / 128],
                [-0.25, 1 / 16, 1 / 64, -1 / 64],
                [1.0, -1 / 32, 1 / 32, 1 / 32],
                [0.0, 1 / 128, 1 / 128, 1 / 16],
                [0.75, -1 / 64, -1 / 128, 0.0],
                [-1.0, 0.0, 1 / 16, -1 / 128],
                [0.125, 1 / 32, 0.0, 1 / 64],
                [0.25, -1 / 128, 1 / 64, 1 / 32]])
_DA = np.array([0.5, 1 / 64, 1 / 32, 1 / 128])          # density > 0 on the frustum


def _affine(x, y, z):
    c = np.stack([np.ones_like(x, dtype=np.float64), x, y, z]).astype(np.float64)
    lat = np.tensordot(_LA, c, axes=(1, 0))
    den = np.tensordot(_DA, c, axes=(0, 0))[None]
    return np.concatenate([lat, den])


def _aff_at(p):
    """The affine field at continuous voxel-index coordinates p [n, 3]."""
    c = np.concatenate([np.ones((len(p), 1)), p], axis=1)
    return c @ _LA.T, c @ _DA


def _interior_points(rng, n, faces=True):
    """Froxels whose 8 lookup corners all lie inside the frustum; with faces,
    half of them straddle a brick face on one to three axes."""
    lo, hi = 0.5, np.array([FX, FY, FZ]) - 0.5 - 1e-3
    p = lo + rng.random((n, 3)) * (hi - lo)
    if faces:
        k = n // 2
        for a in range(3):
            pick = rng.random(k) < 0.6
            b = rng.integers(0, (np.array([FX, FY, FZ])[a] // 8) - 1, k)
            p[:k, a] = np.where(pick, 8 * b + 7.5 + rng.random(k), p[:k, a])   # corners 8b+7, 8b+8
    return p.astype(np.float32)


def _textbook_weights(f):
    """Trilinear weights of the 8 corners (corner c = 4x + 2y + z) of a
    froxel f [3] with the voxel centres at integer + 1/2."""
    q = np.asarray(f, np.float64) - 0.5
    i0 = np.floor(q)
    t = q - i0
    out = []
    for c in range(8):
        off = np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
        w = np.prod(np.where(off == 1, t, 1.0 - t))
        out.append((i0 + off, w))
    return out


def test_affine_field_is_reproduced_off_centre(O, model, rm):
    """Every corner set: an affine field comes back exactly at the continuous
    point f - 1/2, across brick faces (fails under a swapped corner weight,
    a dropped -1/2 offset, a wrong corner-to-voxel address or slot)."""
    cache = _cache(O, model, rm, _affine)
    p = _interior_points(np.random.default_rng(1), 400)
    got = cache.lookup(p, di=False)
    lat, den = _aff_at(p.astype(np.float64) - 0.5)
    assert np.allclose(got[:, 1:], lat, rtol=0, atol=1e-12)
    assert np.allclose(got[:, 0], np.log(den), rtol=0, atol=1e-12)


def test_affine_field_density_intermediates(O, model, rm):
    """DensityIntermediates: latent 0 holds the density, and the lookup's
    first head input is log(Sum w l0) (cache_performance_renderer.cu:770-779)."""
    def field(x, y, z):
        v = _affine(x, y, z)
        v[0] = v[8]                      # latent 0 = density
        return v
    cache = _cache(O, model, rm, field)
    p = _interior_points(np.random.default_rng(2), 300)
    got = cache.lookup(p, di=True)
    lat, den = _aff_at(p.astype(np.float64) - 0.5)
    assert np.allclose(got[:, 0], np.log(den), rtol=0, atol=1e-12)
    assert np.allclose(got[:, 1], np.log(den), rtol=0, atol=1e-12)
    assert np.allclose(got[:, 2:], lat[:, 1:], rtol=0, atol=1e-12)


def _constant(x, y, z):
    shp = np.shape(x)
    vals = np.array([0.3, -0.7, 0.11, 0.5, -0.05, 0.9, -0.4, 0.2, 2.5])
    return np.broadcast_to(vals.reshape(9, *([1] * len(shp))), (9,) + shp).copy()


def _unset_voxels(x, y, z):
    h = (x * 73856093) ^ (y * 19349663) ^ (z * 83492791)
    return (h % 5) == 0                   # about one voxel in five unset


def _unset_brick(x, y, z):
    return (x + 2 * y + 3 * z) % 4 == 1   # about a quarter of the bricks unset


def test_constant_field_with_unset_corners(O, model, rm):
    """Unset voxels, unset bricks and corners outside the frustum: latents
    come back as the constant (Sum w*l / Sum w*isset), while the density row
    is log of the un-normalised Sum w*sigma (fails under dividing by Sum w,
    under normalising the density, under reading an unset brick)."""
    cache = _cache(O, model, rm, _constant, unset_brick=_unset_brick, unset_voxels=_unset_voxels)
    rng = np.random.default_rng(3)
    p = np.concatenate([_interior_points(rng, 300),
                        # corners outside the frustum on the low and high faces
                        np.stack([rng.random(60) * 0.5, rng.random(60) * FY, rng.random(60) * FZ], 1),
                        np.stack([FX - 0.5 + rng.random(60) * 0.5, rng.random(60) * FY,
                                  rng.random(60) * FZ], 1)]).astype(np.float32)
    tab = cache.brick_table()
    exp_w, keep = [], []
    for f in p:
        wsum = 0.0
        for corner, w in _textbook_weights(f):
            x, y, z = (int(v) for v in corner)
            inside = 0 <= x < FX and 0 <= y < FY and 0 <= z < FZ
            if inside and tab[z // 8, y // 8, x // 8] >= 0 and not _unset_voxels(x, y, z):
                wsum += w
        exp_w.append(wsum)
        keep.append(wsum > 1e-9)
    exp_w, keep = np.array(exp_w), np.array(keep)
    assert keep.sum() > 300 and np.any(exp_w[keep] < 0.999)
    got = cache.lookup(p[keep], di=False)
    c = _constant(np.zeros(1), np.zeros(1), np.zeros(1))[:, 0].astype(np.float32).astype(np.float64)
    assert np.allclose(got[:, 1:], c[:8], rtol=0, atol=1e-12)
    assert np.allclose(got[:, 0], np.log(np.maximum(c[8] * exp_w[keep], 1e-6)), rtol=0, atol=1e-12)


def test_density_floor_of_the_log(O, model, rm):
    """log(max(Sum w sigma, 1e-6)): a density far below 1e-6 reads 1e-6."""
    def tiny(x, y, z):
        v = _constant(x, y, z)
        v[8] = 1e-9
        return v
    cache = _cache(O, model, rm, tiny)
    got = cache.lookup(_interior_points(np.random.default_rng(4), 20), di=False)
    assert np.allclose(got[:, 0], math.log(1e-6), rtol=0, atol=1e-9)


def _smooth(t):
    return 3 * t ** 2 - 2 * t ** 3


def _smoother(t):
    return 6 * t ** 5 - 15 * t ** 4 + 10 * t ** 3


@pytest.mark.parametrize("fn", [1, 2, 3])
def test_interpolation_functions(O, model, rm, fn):
    """Smoothstep / smootherstep / nearest weights (applyInterpolationFunction):
    an affine field comes back at i0 + s(t) per axis, s from the textbook
    polynomials or round-half-to-even (fails under the linear weight, an
    unapplied function or swapped polynomial terms)."""
    cache = _cache(O, model, rm, _affine).set_options(interp_fn=fn)
    p = _interior_points(np.random.default_rng(5 + fn), 300)
    q = p.astype(np.float64) - 0.5
    i0 = np.floor(q)
    t = q - i0
    s = {1: _smooth, 2: _smoother, 3: np.round}[fn](t)        # np.round: half to even
    lat, den = _aff_at(i0 + s)
    got = cache.lookup(p, di=False)
    assert np.allclose(got[:, 1:], lat, rtol=0, atol=1e-9)
    assert np.allclose(got[:, 0], np.log(den), rtol=0, atol=1e-9)
    assert not np.allclose(got[:, 1:], _aff_at(q)[0], rtol=0, atol=1e-6)


def _brick_constant(x, y, z):
    """Each brick holds its own constant: latent k = brick id + k / 16."""
    bid = (x // 8) + 5 * (y // 8) + 15 * (z // 8)
    lat = np.stack([bid + k / 16.0 for k in range(8)]).astype(np.float64)
    return np.concatenate([lat, (1.0 + bid)[None] * np.ones_like(lat[:1])])


def test_clamp_to_brick(O, model, rm):
    """use_inter_brick_interpolation = false: the point is clamped into its
    brick, so a brick-wise constant field returns the brick's own constant
    even right at a face, and an affine field returns its value at the
    clamped point (fails without the clamp: neighbouring bricks mix in)."""
    rng = np.random.default_rng(9)
    p = _interior_points(rng, 400)
    bid = ((p[:, 0] // 8) + 5 * (p[:, 1] // 8) + 15 * (p[:, 2] // 8)).astype(np.float64)
    cache = _cache(O, model, rm, _brick_constant).set_options(inter_brick=False)
    got = cache.lookup(p, di=False)
    assert np.allclose(got[:, 1:], bid[:, None] + np.arange(8)[None] / 16.0, rtol=0, atol=1e-12)
    assert np.allclose(got[:, 0], np.log(1.0 + bid), rtol=0, atol=1e-12)
    free = _cache(O, model, rm, _brick_constant).lookup(p, di=False)
    assert not np.allclose(free[:, 1:], got[:, 1:], rtol=0, atol=1e-6)
    aff = _cache(O, model, rm, _affine).set_options(inter_brick=False)
    pd = p.astype(np.float64)
    b = np.floor(pd / 8) * 8
    pc = b + np.clip(pd - b, 0.5, 7.5)
    lat, den = _aff_at(pc - 0.5)
    got = aff.lookup(p, di=False)
    assert np.allclose(got[:, 1:], lat, rtol=0, atol=1e-9)
