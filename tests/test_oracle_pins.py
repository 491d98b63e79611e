"""Pins the CPU oracle of the render path to mathematics.

The reference ships no golden vectors, known-answer tests or fixtures, so
every oracle function is checked here against something it does not share
code with: closed forms, brute force on small inputs, published definitions
(iNGP hash grid, real spherical harmonics) and invariants chosen so that a
plausible mistake (dropped term, wrong sign, transposed operand, swapped
axis, wrong corner weight, midpoint vs bin start) fails one of them.

Rows of the reference path covered (file:line of the reference):
  rays / AABB        raymarch_common.h:191-251, 482-532
  occupancy          raymarch_common.h:660-710, loader.cu:20-55
  contraction        raymarch_common.h:610-653
  hash grid          tcnn GridEncoding via nerf_network.h:173-202
  MLPs, SH           nerf_network.h:199-257, my_fully_fused_mlp.cu:206-234
  compositing        performance_renderer.cu:93-120
  marcher            performance_renderer.cu:39-62
  cache fill         cache_renderer.cu:358-528
  cache lookup       cache_performance_renderer.cu:678-787
"""
import ctypes as C
import math

import numpy as np
import pytest

from frustumvolumecaching_amd import abi, synthetic as S


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def _stepsize(O, scene, cone=0.0, step=math.sqrt(3.0) / 1024.0, near=0.1, far=8.0):
    st = S.stepsize_info(cone, step, near, far)
    O.stepsize_setup(st, scene)
    return st


def _init_rays(O, scene, rm, w, h):
    n = w * h
    o6 = np.zeros((n, 6), np.float32)
    fs = np.zeros((n, 2), np.float32)
    alive = np.zeros(n, np.uint8)
    O.lib().orc_init_rays(C.byref(scene), C.byref(rm), 0, 0, w, h, C.c_void_p(o6.ctypes.data),
                          C.c_void_p(fs.ctypes.data), C.c_void_p(alive.ctypes.data))
    return o6, fs, alive.astype(bool)


# ------------------------------------------------------------------ rays
def test_pinhole_rays_closed_form(O):
    """Pixel (x, y) looks along R @ ((x+0.5-cx)/fx, (y+0.5-cy)/fy, 1) from the
    camera centre; the principal pixel looks along the camera's third axis."""
    scene = S.scene_info(1.0, 1)
    st = _stepsize(O, scene)
    c2w = S.look_at(S.eye_on_sphere(4.0, 30.0, 20.0), (0.1, -0.2, 0.05))
    w, h = 21, 15
    cam = S.camera_info(c2w, w, h)
    cam.principal = abi.Float2(10.5, 7.5)
    cam.focal = abi.Float2(17.0, 13.0)
    rm = S.raymarch_info(cam, st)
    o6, _, _ = _init_rays(O, scene, rm, w, h)
    R = c2w[:, :3].astype(np.float64)
    ys, xs = np.mgrid[0:h, 0:w]
    v = np.stack([(xs + 0.5 - 10.5) / 17.0, (ys + 0.5 - 7.5) / 13.0, np.ones_like(xs, np.float64)], -1).reshape(-1, 3)
    d = v @ R.T
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    assert np.abs(o6[:, 3:] - d).max() < 2e-6
    assert np.abs(o6[:, :3] - c2w[:, 3]).max() == 0.0
    principal = o6[7 * w + 10, 3:]
    assert np.abs(principal - R[:, 2] / np.linalg.norm(R[:, 2])).max() < 1e-6


def test_ray_aabb_against_dense_sampling(O):
    """AABB entry/exit of every ray against brute-force dense sampling."""
    scene = S.scene_info(1.0, 1)
    st = _stepsize(O, scene, far=50.0)
    c2w = S.look_at(S.eye_on_sphere(2.5, 10.0, 35.0), (0.3, 0.0, 0.0))
    w, h = 24, 18
    rm = S.raymarch_info(S.camera_info(c2w, w, h, fov_y_deg=70.0), st)
    o6, fs, alive = _init_rays(O, scene, rm, w, h)
    ts = np.arange(0.0, 8.0, 1e-3)
    hits = 0
    for i in range(w * h):
        o, d = o6[i, :3].astype(np.float64), o6[i, 3:].astype(np.float64)
        p = o[None, :] + ts[:, None] * d[None, :]
        excess = np.abs(p).max(axis=1) - 1.0
        inside = excess <= 0.0
        if alive[i] != inside.any():
            # only a ray grazing the box within the sampling step may disagree
            assert excess.min() < 2e-3, i
            continue
        if inside.any():
            hits += 1
            t_last = ts[inside][-1]
            assert abs(fs[i, 0] - t_last) <= 1.5e-3
            t_first = max(ts[inside][0], st.near_)
            # start snaps to the bin grid around the entry
            assert abs(fs[i, 1] - t_first) <= st.dt_min + 1.5e-3
    assert 0 < hits < w * h


# ------------------------------------------------------------------ occupancy
def _occ_brute(scene, occ, pts):
    """Cascade and cell of each point from the definition, in float64: unit
    coordinate u with the innermost cascade = [0,1]; cascade l >= 1 holds
    max|u-1/2| in [2^(l-2), 2^(l-1)), cascade 0 below 1/2 (mip_from_unit,
    raymarch_common.h:669-677)."""
    nl = scene.grid_nlvl
    lo = np.array([scene.aabb_from.x, scene.aabb_from.y, scene.aabb_from.z])
    hi = np.array([scene.aabb_to.x, scene.aabb_to.y, scene.aabb_to.z])
    Sx = 2.0 ** (nl - 1)
    u = (pts - lo / Sx) / ((hi - lo) / Sx)
    m = np.abs(u - 0.5).max(axis=1)
    mip = np.where(m < 0.5, 0, np.floor(np.log2(np.maximum(m, 1e-300))) + 2).astype(int)
    q = ((u - 0.5) / (2.0 ** mip)[:, None] + 0.5) * 128.0
    ok = mip < nl
    idx = np.clip(np.floor(q).astype(int), 0, 127)
    res = np.zeros(len(pts), bool)
    res[ok] = occ[mip[ok], idx[ok, 0], idx[ok, 1], idx[ok, 2]] > 0
    # points too close to a cell face or a cascade boundary are ambiguous
    frac = np.abs(q - np.round(q)).min(axis=1)
    edge = np.abs(np.log2(np.maximum(m, 1e-300)) - np.round(np.log2(np.maximum(m, 1e-300))))
    clear = (frac > 1e-3) & (edge > 1e-6)
    return res, clear


def test_occupancy_matches_cell_brute_force(O):
    scene = S.scene_info(4.0, 3)
    rng = np.random.default_rng(11)
    occ = (rng.random((3, 128, 128, 128)) < 0.3).astype(np.uint8)
    wl = S.small_plain()
    model = O.Model(scene, wl.net, occ, wl.params)
    pts = rng.uniform(-4.5, 4.5, (20000, 3))
    ref, clear = _occ_brute(scene, occ, pts)
    got = model.occupied(pts.astype(np.float32))
    assert clear.sum() > 19000
    assert np.array_equal(got[clear], ref[clear])
    # Morton bitfield agrees with the linear grid bit by bit (loader.cu:20-55)
    bf = model.bitfield()
    lin = np.nonzero(occ.reshape(-1))[0][:5000]
    l, rem = np.divmod(lin, 128 ** 3)
    x, rem = np.divmod(rem, 128 * 128)
    y, z = np.divmod(rem, 128)
    def spread(v):
        r = np.zeros_like(v)
        for b in range(7):
            r |= ((v >> b) & 1) << (3 * b)
        return r
    bit = l * 128 ** 3 + (spread(x) | (spread(y) << 1) | (spread(z) << 2))
    assert np.all((bf[bit >> 3] >> (bit & 7)) & 1)
    assert int(np.unpackbits(bf).sum()) == int(occ.sum())


def test_occupancy_special_cases(O):
    """The exact centre has max|u-1/2| = 0 and frexp exponent 0, so it is
    looked up in cascade 1 (a quirk of mip_from_unit kept for parity); points
    beyond the outer cascade are empty."""
    scene = S.scene_info(2.0, 2)
    wl = S.small_plain()
    occ = np.zeros((2, 128, 128, 128), np.uint8)
    occ[1, 64, 64, 64] = 1
    model = O.Model(scene, wl.net, occ, wl.params)
    assert model.occupied(np.zeros((1, 3), np.float32))[0]
    occ2 = np.zeros_like(occ)
    occ2[0, 64, 64, 64] = 1
    model2 = O.Model(scene, wl.net, occ2, wl.params)
    assert not model2.occupied(np.zeros((1, 3), np.float32))[0]
    # a point just off centre falls in cascade 0
    assert model2.occupied(np.array([[1e-3, 1e-3, 1e-3]], np.float32))[0]
    full = O.Model(scene, wl.net, np.ones_like(occ), wl.params)
    far = np.array([[2.5, 0, 0], [0, -3.0, 0], [0, 0, 10.0]], np.float32)
    assert not full.occupied(far).any()
    assert full.occupied(np.array([[1.9, 0, 0], [0, -1.9, 0.5]], np.float32)).all()


# ------------------------------------------------------------------ contraction
def test_contraction_closed_forms(O):
    L = O.lib()
    def contract(scene, pts):
        pts = np.ascontiguousarray(pts, np.float32)
        out = np.zeros((len(pts), 3), np.float64)
        L.orc_contract(C.byref(scene), C.c_void_p(pts.ctypes.data), len(pts), 1, C.c_void_p(out.ctypes.data))
        return out
    # AABB: corners map to 0 / 1, centre to 1/2
    sc = S.scene_info(3.0, 1)
    got = contract(sc, [[-3, -3, -3], [3, 3, 3], [0, 0, 0], [1.5, -1.5, 0]])
    assert np.allclose(got, [[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [0.75, 0.25, 0.5]], atol=1e-12)
    # L-inf warp of the [-1,1]^3 contraction box
    sc = S.scene_info(16.0, 5, abi.CONTRACTION_WARP_AABB_LINF, 1.0)
    rng = np.random.default_rng(5)
    inside = rng.uniform(-1, 1, (200, 3))
    assert np.allclose(contract(sc, inside), inside / 4 + 0.5, atol=1e-12)
    out = rng.uniform(-30, 30, (2000, 3))
    out = out[np.abs(out).max(axis=1) > 1.01]
    c = contract(sc, out) - 0.5
    n = np.abs(out).max(axis=1)
    # direction kept, L-inf radius (2 - 1/n) / 4 in (1/4, 1/2)
    cos = np.sum(c * out, axis=1) / np.linalg.norm(c, axis=1) / np.linalg.norm(out, axis=1)
    assert np.allclose(cos, 1.0, atol=1e-12)
    assert np.allclose(np.abs(c).max(axis=1), (2 - 1 / n) / 4, atol=1e-12)
    # continuity on the box face and the limit at infinity
    assert np.allclose(contract(sc, [[1.0, 0.3, -0.2]]), [[0.75, 0.575, 0.45]], atol=1e-12)
    assert np.allclose(contract(sc, [[1e9, 0, 0], [0, -1e9, 0]]), [[1.0, 0.5, 0.5], [0.5, 0.0, 0.5]], atol=1e-8)


# ------------------------------------------------------------------ hash grid
PRIMES = (1, 2654435761, 805459861)


def _encode_brute(net, table, p):
    """iNGP multiresolution encoding (Mueller et al. 2022, eq. 3 and the
    spatial hash), tcnn conventions for level scale and dense levels."""
    T = 1 << net.log2_hashmap_size
    out = []
    off = 0
    pls = float(np.float32(net.per_level_scale))
    for l in range(net.n_levels):
        s = 2.0 ** (l * math.log2(pls)) * net.base_resolution - 1.0
        r = int(math.ceil(float(np.float32(s)))) + 1
        size = min((r ** 3 + 7) // 8 * 8, T)
        pos = np.clip(p, 0, 1) * s + 0.5
        g = np.floor(pos).astype(np.int64)
        f = pos - g
        acc = np.zeros(2)
        for c in range(8):
            q = [g[k] + ((c >> k) & 1) for k in range(3)]
            w = np.prod([f[k] if (c >> k) & 1 else 1 - f[k] for k in range(3)])
            if r ** 3 <= size:
                idx = q[0] + q[1] * r + q[2] * r * r
            else:
                idx = ((q[0] * PRIMES[0]) ^ (q[1] * PRIMES[1]) ^ (q[2] * PRIMES[2])) & 0xFFFFFFFF
            idx %= size
            acc += w * table[off + idx]
        out.append(acc)
        off += size
    return np.concatenate(out)


def test_hashgrid_matches_published_definition(O):
    wl = S.small_plain()
    net = wl.net
    table = np.asarray(wl.params[0], np.uint16).view(np.float16).astype(np.float64).reshape(-1, 2)
    model = O.Model.from_workload(wl)
    rng = np.random.default_rng(6)
    pts = rng.random((64, 3))
    got = model.encode(pts.astype(np.float32), mode=O.F64)
    ref = np.stack([_encode_brute(net, table, p.astype(np.float32).astype(np.float64)) for p in pts])
    assert np.abs(got - ref).max() < 1e-12


def test_hashgrid_dense_levels_reproduce_affine_fields(O):
    """Trilinear interpolation reproduces an affine function of the vertex
    index exactly (up to the fp16 rounding of the stored values)."""
    wl = S.small_plain()
    net = wl.net
    lay = S.grid_layout(net)
    table = np.zeros((sum(size for _, _, size, _ in lay), 2), np.float64)
    a0, a1 = np.array([0.25, -0.5, 0.125]), np.array([-0.0625, 0.375, 0.5])
    dense = []
    for l, (s, r, size, off) in enumerate(lay):
        if r ** 3 > size:
            continue
        dense.append(l)
        z, y, x = np.meshgrid(np.arange(r), np.arange(r), np.arange(r), indexing="ij")
        idx = (x + y * r + z * r * r).reshape(-1)
        g = np.stack([x.reshape(-1), y.reshape(-1), z.reshape(-1)], 1) / r
        table[off + idx, 0] = g @ a0 + 0.1
        table[off + idx, 1] = g @ a1 - 0.2
    assert len(dense) >= 4
    t16 = table.astype(np.float16)
    params = (t16.view(np.uint16).reshape(-1), wl.params[1], wl.params[2])
    model = O.Model(wl.scene, net, wl.occupancy, params)
    pts = np.random.default_rng(7).random((200, 3)).astype(np.float32)
    got = model.encode(pts, mode=O.F64)
    for l in dense:
        s, r, _, _ = lay[l]
        s64 = 2.0 ** (l * math.log2(float(np.float32(net.per_level_scale)))) * net.base_resolution - 1.0
        pos = pts.astype(np.float64) * s64 + 0.5
        # with an integer scale the upper corner of the last cell is vertex
        # r, which wraps to the next row (tcnn's dense indexing); skip those
        ok = np.all(np.floor(pos) + 1 < r, axis=1)
        assert ok.sum() > 150
        exp0 = pos / r @ a0 + 0.1
        exp1 = pos / r @ a1 - 0.2
        tol = 2.0 ** -10 * 1.5
        assert np.abs(got[ok, 2 * l] - exp0[ok]).max() < tol, l
        assert np.abs(got[ok, 2 * l + 1] - exp1[ok]).max() < tol, l


# ------------------------------------------------------------------ MLPs
def _weights(wl):
    W, HW = wl.net.base_width, wl.net.head_width
    base = np.asarray(wl.params[1], np.uint16).view(np.float16).astype(np.float64)
    head = np.asarray(wl.params[2], np.uint16).view(np.float16).astype(np.float64)
    W1 = base[:W * 32].reshape(W, 32)
    W2 = base[W * 32:W * 32 + W * W].reshape(W, W)
    W3 = base[W * 32 + W * W:].reshape(16, W)
    H1 = head[:HW * 32].reshape(HW, 32)
    H2 = head[HW * 32:].reshape(16, HW)
    return W1, W2, W3, H1, H2


def test_density_mlp_matches_float64_matmul(O):
    """mlp_base = W3 relu(W2 relu(W1 enc)), weights row-major [out][in]."""
    wl = S.small_plain()
    model = O.Model.from_workload(wl)
    W1, W2, W3, _, _ = _weights(wl)
    pts = np.random.default_rng(8).random((128, 3)).astype(np.float32)
    enc = model.encode(pts, mode=O.F64)
    ref = (W3 @ np.maximum(W2 @ np.maximum(W1 @ enc.T, 0), 0)).T
    got = model.latents(pts, mode=O.F64)
    assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_head_mlp_matches_float64_matmul(O):
    """Head input = [8 latents, SH4(dir) 16, 1.0 x 8] -> relu(H1 .) -> H2."""
    wl = S.small_plain()
    model = O.Model.from_workload(wl)
    _, _, _, H1, H2 = _weights(wl)
    rng = np.random.default_rng(9)
    hin = rng.uniform(-2, 2, (64, 16)).astype(np.float32)
    d = rng.standard_normal((64, 3))
    d01 = ((d / np.linalg.norm(d, axis=1, keepdims=True) + 1) / 2).astype(np.float32)
    sh = np.zeros((64, 16))
    d64 = np.ascontiguousarray(d01, np.float64)
    O.lib().orc_sh4(C.c_void_p(d64.ctypes.data), 64, C.c_void_p(sh.ctypes.data))
    x = np.concatenate([hin[:, 1:9].astype(np.float64), sh, np.ones((64, 8))], 1)
    ref = (H2 @ np.maximum(H1 @ x.T, 0)).T[:, :3]
    got = model.head(hin, d01, mode=O.F64)
    assert np.abs(got[:, :3] - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
    assert np.array_equal(got[:, 3], hin[:, 0].astype(np.float64))


def test_sh4_known_answers_and_orthonormality(O):
    L = O.lib()
    def sh(dirs):
        d01 = np.ascontiguousarray((dirs + 1) / 2, np.float64)
        out = np.zeros((len(d01), 16))
        L.orc_sh4(C.c_void_p(d01.ctypes.data), len(d01), C.c_void_p(out.ctypes.data))
        return out
    z = sh(np.array([[0.0, 0.0, 1.0]]))[0]
    # real SH (Sloan, Stupid SH Tricks, App. A1) on the +z axis
    kat = np.zeros(16)
    kat[0] = 0.5 * math.sqrt(1 / math.pi)
    kat[2] = 0.5 * math.sqrt(3 / math.pi)
    kat[6] = 0.25 * math.sqrt(5 / math.pi) * 2.0
    kat[12] = 0.25 * math.sqrt(7 / math.pi) * 2.0
    assert np.allclose(z, kat, atol=1e-12)
    # orthonormality by quadrature on a Fibonacci sphere
    n = 40000
    k = np.arange(n) + 0.5
    th = np.arccos(1 - 2 * k / n)
    ph = math.pi * (1 + 5 ** 0.5) * k
    dirs = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    Y = sh(dirs)
    G = Y.T @ Y * (4 * math.pi / n)
    assert np.abs(G - np.eye(16)).max() < 2e-3
    # signs: Y_1^-1 ~ -y, Y_1^1 ~ -x (Condon-Shortley phase)
    e = sh(np.array([[1.0, 0, 0], [0, 1.0, 0]]))
    assert e[0, 3] < 0 and e[1, 1] < 0


# ------------------------------------------------------------------ compositing
def test_compositing_partition_of_unity(O):
    """With a colour MLP whose output is zero every sample has colour
    sigmoid(0) = 1/2, so rgb = (1 - T) / 2 exactly when the weights
    alpha_k T_k and the final T partition one."""
    wl = S.small_plain(width=64, height=64)
    st = _stepsize(O, wl.scene)
    head = np.asarray(wl.params[2], np.uint16).copy()
    head[64 * 32:] = 0
    model = O.Model(wl.scene, wl.net, wl.occupancy, (wl.params[0], wl.params[1], head))
    rm = S.raymarch_info(S.camera_info(wl.cam2world, 64, 64), st)
    img, n = model.render_full(rm, O.F64)
    assert n > 1000
    rgb, T = img[..., :3].astype(np.float64), img[..., 3].astype(np.float64)
    assert np.abs(rgb - 0.5 * (1 - T)[..., None]).max() < 2e-7
    assert T.min() >= 0 and T.max() <= 1 and T.min() < 0.5


# ------------------------------------------------------------------ marcher
def _march_setup(O, occ):
    wl = S.small_plain(width=24, height=24)
    st = _stepsize(O, wl.scene)
    model = O.Model(wl.scene, wl.net, occ, wl.params)
    rm = S.raymarch_info(S.camera_info(wl.cam2world, 24, 24, fov_y_deg=30.0), st)
    return wl, st, model, rm


def test_march_bins_follow_the_constant_step(O):
    """All cells occupied, cone angle 0: every bin between the snapped start
    and far is a sample and consecutive starts differ by dt_min (up to the
    rounding of the running sum)."""
    occ = np.ones((1, 128, 128, 128), np.uint8)
    wl, st, model, rm = _march_setup(O, occ)
    counts, t0 = model.march(rm, max_per_ray=4096)
    o6, fs, alive = _init_rays(O, wl.scene, rm, 24, 24)
    assert np.array_equal(counts > 0, alive)
    for i in np.nonzero(alive)[0]:
        t = t0[i, :counts[i]].astype(np.float64)
        assert t[0] == fs[i, 1]
        dt = np.diff(t)
        assert np.abs(dt - st.dt_min).max() <= 4 * np.spacing(np.float32(t.max()))
        # the last sample's midpoint is inside far, the next one beyond
        assert t[-1] + st.dt_min / 2 <= fs[i, 0] + 1e-6
        assert t[-1] + 1.5 * st.dt_min > fs[i, 0] - 1e-6
    model0 = O.Model(wl.scene, wl.net, np.zeros_like(occ), wl.params)
    c0, _ = model0.march(rm, max_per_ray=4)
    assert c0.sum() == 0


def test_march_samples_are_bins_with_occupied_midpoints(O):
    """One occupied block: the samples are exactly the bins (of the
    all-occupied march) whose midpoints fall inside its world box."""
    full = np.ones((1, 128, 128, 128), np.uint8)
    wl, st, model_full, rm = _march_setup(O, full)
    counts_f, t0_f = model_full.march(rm, max_per_ray=4096)
    occ = np.zeros_like(full)
    lo_c, hi_c = np.array([60, 58, 62]), np.array([70, 72, 69])
    occ[0, lo_c[0]:hi_c[0], lo_c[1]:hi_c[1], lo_c[2]:hi_c[2]] = 1
    model = O.Model(wl.scene, wl.net, occ, wl.params)
    counts, t0 = model.march(rm, max_per_ray=4096)
    o6, fs, alive = _init_rays(O, wl.scene, rm, 24, 24)
    box_lo, box_hi = -1 + lo_c * 2 / 128, -1 + hi_c * 2 / 128
    checked = 0
    for i in np.nonzero(alive)[0]:
        o, d = o6[i, :3].astype(np.float64), o6[i, 3:].astype(np.float64)
        tb = t0_f[i, :counts_f[i]].astype(np.float64)
        mid = tb + st.dt_min / 2
        p = o[None] + mid[:, None] * d[None]
        dist = np.minimum(p - box_lo, box_hi - p).min(axis=1)
        inside = dist > 0
        clear = np.abs(dist) > 1e-5
        got = set(t0[i, :counts[i]].tolist())
        exp_in = set(tb[inside & clear].astype(np.float32).tolist())
        exp_out = set(tb[~inside & clear].astype(np.float32).tolist())
        assert exp_in <= got and not (exp_out & got)
        checked += len(exp_in)
    assert checked > 100


# ------------------------------------------------------------------ cache
@pytest.fixture(scope="module")
def filled(O):
    wl = S.small_plain()
    st = _stepsize(O, wl.scene)
    model = O.Model.from_workload(wl)
    rm = S.raymarch_info(S.camera_info(wl.cam2world, wl.width, wl.height), st)
    crop = (96, 96, 64, 64)
    return wl, model, rm, crop, model.cache_fill(rm, O.F32, crop=crop)


def test_fill_invariants(O, filled):
    wl, model, rm, crop, cache = filled
    tab = cache.brick_table()
    used = tab[tab >= 0]
    n = cache.info.n_bricks_used
    assert n == cache.stats.n_bricks_set > 0
    assert np.array_equal(np.sort(used), np.arange(n))          # slots form a bijection
    pool = cache.pool()
    isset, dens, lat = pool[..., 9], pool[..., 8], pool[..., :8]
    assert set(np.unique(isset).tolist()) <= {0.0, 1.0}
    assert np.all(dens[isset == 1] > 0)
    assert np.all(dens[isset == 0] == 0) and np.all(lat[isset == 0] == 0)
    assert np.all(isset.reshape(n, -1).max(axis=1) == 1)         # no empty brick allocated
    known = cache.known_ranges()[crop[1]:crop[1] + crop[3], crop[0]:crop[0] + crop[2]]
    assert np.all(known >= rm.stepsize_info.near_)
    # deterministic: a second fill gives the same table
    assert np.array_equal(model.cache_fill(rm, O.F32, crop=crop).brick_table(), tab)
    # capacity overflow stops columns instead of failing
    small = model.cache_fill(rm, O.F32, capacity=16, crop=crop)
    assert small.stats.overflow and small.stats.n_bricks_set <= 16


def test_lookup_at_voxel_centres_returns_the_voxel(O, filled):
    """Froxel (v + 1/2) puts the whole trilinear weight on voxel v (the
    -1/2 offset of the lookup), so the lookup returns its log density and
    latents."""
    wl, model, rm, crop, cache = filled
    tab = cache.brick_table()
    pool = cache.pool()
    bz, by, bx = np.nonzero(tab >= 0)
    rng = np.random.default_rng(10)
    pts, exp = [], []
    for k in rng.choice(len(bz), size=min(40, len(bz)), replace=False):
        slot = tab[bz[k], by[k], bx[k]]
        zz, yy, xx = np.nonzero(pool[slot, ..., 9] == 1)
        j = rng.integers(len(zz))
        v = (bx[k] * 8 + xx[j], by[k] * 8 + yy[j], bz[k] * 8 + zz[j])
        pts.append(np.array(v, np.float64) + 0.5)
        rec = pool[slot, zz[j], yy[j], xx[j]]
        exp.append(np.concatenate([[math.log(max(rec[8], 1e-6))], rec[:8]]))
    got = cache.lookup(np.array(pts, np.float32))
    assert np.allclose(got, np.array(exp), rtol=1e-6, atol=1e-6)


def test_cached_render_at_fill_pose_matches_full_render(O, filled):
    """Reprojection at the fill pose samples the cache at (almost) the fill's
    own sample points; the reference's gate is 35 dB (eval.cu:315-326)."""
    wl, model, rm, crop, cache = filled
    inner = (crop[0] + 8, crop[1] + 8, crop[2] - 16, crop[3] - 16)
    full, _ = model.render_full(rm, O.F32, crop=inner)
    cached, nr, nc = cache.render(rm, crop=inner)
    assert nc > 10 * max(nr, 1)
    assert O.psnr_8bit(O.to_8bit(cached), O.to_8bit(full)) > 40.0


def _bf16(a):
    """Round to bfloat16 (nearest even) from the float32 value."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) & 0xffff0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _split_layer(W, X):
    """Two-term bfloat16 product of W [out, in] and X [in, n]: hi*hi + hi*lo +
    lo*hi with hi = bf16(v), lo = bf16(v - hi) (float64 sums)."""
    wh = _bf16(W)
    wl = _bf16(W - wh)
    xh = _bf16(X)
    xl = _bf16(X.astype(np.float32).astype(np.float64) - xh)
    return wh @ xh + wh @ xl + wl @ xh


def test_bf16x3_mode_is_the_two_term_split(O):
    """EMU_BF16X3: each MLP product is the two-term bfloat16 split (lo*lo
    dropped), hidden activations unrounded, outputs stored as fp16.  The
    oracle's float accumulation agrees with the float64 split to one fp16 ulp
    of the output, and its error against F64 is far below plain BF16's (a
    split reduced to hi*hi is plain BF16)."""
    wl = S.small_plain()
    model = O.Model.from_workload(wl)
    W1, W2, W3, _, _ = _weights(wl)
    pts = np.random.default_rng(18).random((256, 3)).astype(np.float32)
    enc = model.encode(pts, mode=O.EMU16)
    h1 = np.maximum(_split_layer(W1, enc.T), 0)
    h2 = np.maximum(_split_layer(W2, h1), 0)
    ref = _split_layer(W3, h2).T.astype(np.float16).astype(np.float64)
    got = model.latents(pts, mode=O.EMU_BF16X3)
    ulp = 2.0 ** (np.floor(np.log2(np.maximum(np.abs(ref), 2.0 ** -14))) - 10)
    assert (np.abs(got - ref) / ulp).max() <= 1.0
    f64 = model.latents(pts, mode=O.F64)
    e_x3 = np.abs(got - f64).mean()
    e_bf16 = np.abs(model.latents(pts, mode=O.EMU_BF16) - f64).mean()
    assert e_x3 < 0.1 * e_bf16, (e_x3, e_bf16)
