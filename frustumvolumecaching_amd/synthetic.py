"""Deterministic synthetic scenes and models for the render path.

Everything here is seeded with a SplitMix64 stream (Steele et al., "Fast
splittable pseudorandom number generators", OOPSLA 2014), so the GPU renderer
and the CPU oracle always receive bit-identical inputs:

* hash-grid table: fp16 U[-0.1, 0.1];
* MLP weights: Xavier-uniform U[-a, a], a = sqrt(6 / (fan_in + fan_out)),
  stored fp16 in the tcnn FullyFusedMLP layout (row-major [n_out][n_in] per
  layer, output rows padded to 16 with zeros; nerf_network.h:269-308);
* occupancy: Gaussian blobs evaluated at the cell centres of every cascade,
  thresholded, as the linear uint8 grid [lvl][x][y][z] that
  loader.cu:232-285 reads from occupancy_grid.dat.

Cameras use the reference's non-OpenGL convention (x right, y down, z
forward; raymarch_common.h:190-201) and CameraMatrix::inverse
(common.h:44-56) for world2cam.
"""
import math

import numpy as np

from . import abi

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, n):
    """n consecutive SplitMix64 outputs of stream `seed` (uint64 array)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * np.uint64(0xD1B54A32D192ED03) + np.arange(1, n + 1, dtype=np.uint64) * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, lo=0.0, hi=1.0):
    """n doubles U[lo, hi) from stream `seed`."""
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def f16_bits(a):
    return np.asarray(a, dtype=np.float32).astype(np.float16).view(np.uint16)


# ---------------------------------------------------------------------------
# network
# ---------------------------------------------------------------------------
def net_desc(log2_hashmap_size=19, n_levels=16, base_resolution=16, max_resolution=2048, precision=abi.PREC_FP16,
             layout="separate", n_output=None):
    """layout "separate": log sigma in output row 0 and 8 latents after it
    (the density kept apart in the cache); "reference": the reference's plain
    models, the head reads output rows 0..n_output-1 with log sigma first
    (latent_offset 0, DensityIntermediates cache; nerf_network.h:88-118,
    334-342)."""
    pls = math.exp(math.log(max_resolution / base_resolution) / (n_levels - 1))
    if layout == "separate":
        n_out, n_lat, off, fld = n_output or 9, 8, 1, 0
    elif layout == "reference":
        n_out = n_output or 8
        n_lat, off, fld = n_out, 0, 1
    else:
        raise ValueError(layout)
    return abi.NetDesc(
        n_levels=n_levels, n_features_per_level=2, log2_hashmap_size=log2_hashmap_size,
        base_resolution=base_resolution, per_level_scale=float(np.float32(pls)),
        base_width=64, base_n_hidden_layers=2, base_n_output=n_out,
        head_width=64, head_n_hidden_layers=1, n_latents=n_lat, latent_offset=off, sh_degree=4,
        precision=precision, first_latent_is_density=fld)


def net_desc_viewdep(log2_hashmap_size=19, first_latent_is_density=True, first_head_hidden_layers=1,
                     freq_encoding_degree=4, **kw):
    """Split model (NerfNetwork::init_viewdep): 16 density-MLP outputs,
    mlp_first_head 32 -> 64 (x hidden layers) -> 8 latents, custom colour head
    [8 latents, SH4, SH4 difference, frequency encoding] -> 64 -> 3."""
    net = net_desc(log2_hashmap_size=log2_hashmap_size, **kw)
    net.base_n_output = 16
    net.latent_offset = 0 if first_latent_is_density else 1
    net.split_head = 1
    net.first_head_width = 64
    net.first_head_n_hidden_layers = first_head_hidden_layers
    net.first_head_n_input = 32
    net.first_latent_is_density = int(bool(first_latent_is_density))
    net.first_sh_degree = 4
    net.freq_encoding_degree = freq_encoding_degree
    return net


def head_input_width(net):
    """Input width of the colour head: 32, or for split models 8 latents +
    16 SH + 16 SH differences + 6 * freq degree rounded up to 16."""
    if not net.split_head:
        return 32
    return (8 + 16 + 16 + 6 * net.freq_encoding_degree + 15) // 16 * 16


def grid_layout(net):
    """Per-level (scale, resolution, entries, entry offset) of the hash grid.

    tcnn GridEncoding conventions: scale = exp2(l*log2(b))*N_min - 1,
    resolution = ceil(scale)+1, entries = min(next_multiple(res^3, 8), T).
    """
    T = 1 << net.log2_hashmap_size
    out, off = [], 0
    for l in range(net.n_levels):
        s = np.float32(2.0 ** (l * math.log2(float(np.float32(net.per_level_scale)))) * net.base_resolution - 1.0)
        r = int(math.ceil(float(s))) + 1
        p = min((r ** 3 + 7) // 8 * 8, T)
        out.append((float(s), r, p, off))
        off += p
    return out


def param_counts(net):
    """(pos_enc, mlp_base, mlp_head) counts, with mlp_first_head before
    mlp_head for split models."""
    lay = grid_layout(net)
    n_pos = (lay[-1][3] + lay[-1][2]) * net.n_features_per_level
    W, HW = net.base_width, net.head_width
    n_head = head_input_width(net) * HW + HW * 16
    if net.split_head:
        FW = net.first_head_width
        n_first = 32 * FW + (net.first_head_n_hidden_layers - 1) * FW * FW + FW * 16
        return n_pos, 32 * W + W * W + W * 16, n_first, n_head
    return n_pos, 32 * W + W * W + W * 16, n_head


def xavier(seed, n_out_real, n_in, n_out_pad):
    a = math.sqrt(6.0 / (n_in + n_out_real))
    w = np.zeros((n_out_pad, n_in), dtype=np.float32)
    w[:n_out_real] = uniform(seed, n_out_real * n_in, -a, a).reshape(n_out_real, n_in)
    return w


def make_params(net, seed_weights=1, seed_table=2, table_range=0.1):
    """(pos_enc, mlp_base, mlp_head) fp16 bit arrays in the nerfacc export
    layout; split models (pos_enc, mlp_base, mlp_first_head, mlp_head)."""
    n_pos = param_counts(net)[0]
    pos = f16_bits(uniform(seed_table, n_pos, -table_range, table_range))
    W, HW = net.base_width, net.head_width
    base = np.concatenate([
        xavier(seed_weights * 16 + 0, W, 32, W).ravel(),
        xavier(seed_weights * 16 + 1, W, W, W).ravel(),
        xavier(seed_weights * 16 + 2, net.base_n_output, W, 16).ravel()])
    hin = head_input_width(net)
    head = np.concatenate([
        xavier(seed_weights * 16 + 3, HW, hin, HW).ravel(),
        xavier(seed_weights * 16 + 4, 3, HW, 16).ravel()])
    if not net.split_head:
        return pos, f16_bits(base), f16_bits(head)
    FW = net.first_head_width
    first = [xavier(seed_weights * 16 + 5, FW, 32, FW).ravel()]
    for k in range(net.first_head_n_hidden_layers - 1):
        first.append(xavier(seed_weights * 16 + 6 + k, FW, FW, FW).ravel())
    first.append(xavier(seed_weights * 16 + 8, net.n_latents, FW, 16).ravel())
    return pos, f16_bits(base), f16_bits(np.concatenate(first)), f16_bits(head)


# ---------------------------------------------------------------------------
# scene
# ---------------------------------------------------------------------------
def scene_info(aabb_half, nlvl, contraction=abi.CONTRACTION_AABB, contraction_half=1.0, alpha_thre=0.0):
    s = abi.SceneInfo()
    s.grid_resolution = 128
    s.grid_nlvl = nlvl
    s.vals_per_lvl = 128 ** 3
    s.normalized = 0
    s.aabb_from = abi.f3([-aabb_half] * 3)
    s.aabb_to = abi.f3([aabb_half] * 3)
    s.contraction_aabb_from = abi.f3([-contraction_half] * 3)
    s.contraction_aabb_to = abi.f3([contraction_half] * 3)
    s.contraction_type = contraction
    s.is_open_gl = 0
    s.alpha_thre = alpha_thre
    return s


def stepsize_info(cone_angle, stepsize, near, far):
    s = abi.StepsizeInfo()
    s.cone_angle, s.stepsize, s.near_, s.far_ = cone_angle, stepsize, near, far
    return s


def blobs(seed, n, center_radius, sigma_range, amp=1.0):
    u = uniform(seed, n * 5).reshape(n, 5)
    c = (u[:, :3] * 2.0 - 1.0) * center_radius
    s = sigma_range[0] + u[:, 3] * (sigma_range[1] - sigma_range[0])
    return c, s, np.full(n, amp)


def blobs_spread(seed, n, radius, sigma_frac=(0.04, 0.1), min_sigma=0.08):
    """Blobs with centres uniform in direction and in distance from the origin
    (so every cascade holds some) and widths growing with that distance."""
    u = uniform(seed, n * 6).reshape(n, 6)
    d = u[:, :3] * 2.0 - 1.0
    d /= np.linalg.norm(d, axis=1, keepdims=True) + 1e-12
    r = radius * u[:, 3]
    c = d * r[:, None]
    s = np.maximum(min_sigma, (sigma_frac[0] + u[:, 4] * (sigma_frac[1] - sigma_frac[0])) * r)
    return c, s, np.full(n, 1.0)


def occupancy_from_blobs(scene, centers, sigmas, amps, threshold=0.01):
    """Linear uint8 grid [nlvl][x][y][z]: 1 where the blob density at the cell
    centre exceeds `threshold` (cascade l covers the inner AABB scaled by 2^l)."""
    res = 128
    nlvl = scene.grid_nlvl
    lo = np.array([scene.aabb_from.x, scene.aabb_from.y, scene.aabb_from.z], np.float64)
    hi = np.array([scene.aabb_to.x, scene.aabb_to.y, scene.aabb_to.z], np.float64)
    S = float(1 << (nlvl - 1))
    inner_lo, inner_ext = lo / S, (hi - lo) / S
    occ = np.zeros((nlvl, res, res, res), dtype=np.uint8)
    cell = (np.arange(res) + 0.5) / res
    for l in range(nlvl):
        # unit coordinate of cell centres in cascade l, then world
        u = (cell - 0.5) * (2.0 ** l) + 0.5
        coords = [inner_lo[k] + u * inner_ext[k] for k in range(3)]
        step = coords[0][1] - coords[0][0]
        dens = np.zeros((res, res, res), dtype=np.float64)
        for c, s, a in zip(centers, sigmas, amps):
            reach = s * math.sqrt(2.0 * math.log(max(a / threshold, 1.0 + 1e-9))) + step
            idx = []
            for k in range(3):
                i0 = int(np.searchsorted(coords[k], c[k] - reach))
                i1 = int(np.searchsorted(coords[k], c[k] + reach))
                idx.append((i0, i1))
            if any(i1 <= i0 for i0, i1 in idx):
                continue
            (x0, x1), (y0, y1), (z0, z1) = idx
            dx = (coords[0][x0:x1] - c[0]) ** 2
            dy = (coords[1][y0:y1] - c[1]) ** 2
            dz = (coords[2][z0:z1] - c[2]) ** 2
            r2 = dx[:, None, None] + dy[None, :, None] + dz[None, None, :]
            dens[x0:x1, y0:y1, z0:z1] += a * np.exp(-r2 / (2.0 * s * s))
        occ[l] = dens > threshold
    return occ


# ---------------------------------------------------------------------------
# camera
# ---------------------------------------------------------------------------
def camera_matrix_inverse(m):
    """CameraMatrix::inverse (common.h:44-56) in float32."""
    m = np.asarray(m, np.float32).reshape(3, 4)
    r = m[:, :3].T.copy()
    t = m[:, 3]
    out = np.zeros((3, 4), np.float32)
    for i in range(3):
        out[i, :3] = r[i]
        acc = np.float32(r[i, 0] * t[0])
        acc = np.float32(acc + np.float32(r[i, 1] * t[1]))
        acc = np.float32(acc + np.float32(r[i, 2] * t[2]))
        out[i, 3] = -acc
    return out


def look_at(eye, target, up=(0.0, 0.0, 1.0)):
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    m = np.zeros((3, 4))
    m[:, 0], m[:, 1], m[:, 2], m[:, 3] = r, d, f, eye
    return m.astype(np.float32)


def camera_info(cam2world, width, height, fov_y_deg=50.0):
    c = abi.CameraInfo()
    c.res_x, c.res_y, c.is_open_gl = width, height, 0
    fy = 0.5 * height / math.tan(math.radians(fov_y_deg) * 0.5)
    c.focal = abi.Float2(float(np.float32(fy)), float(np.float32(fy)))
    c.principal = abi.Float2(width * 0.5, height * 0.5)
    c.cam2world = abi.CameraMatrix.from_array(cam2world)
    c.world2cam = abi.CameraMatrix.from_array(camera_matrix_inverse(cam2world))
    return c


def yaw(cam2world, deg, up=(0.0, 0.0, 1.0)):
    """Rotate the camera orientation about the world up axis through its centre."""
    m = np.asarray(cam2world, np.float64).reshape(3, 4).copy()
    a = math.radians(deg)
    u = np.asarray(up, np.float64)
    u /= np.linalg.norm(u)
    K = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])
    R = np.eye(3) + math.sin(a) * K + (1 - math.cos(a)) * K @ K
    m[:, :3] = R @ m[:, :3]
    return m.astype(np.float32)


def translate_forward(cam2world, dist):
    m = np.asarray(cam2world, np.float64).reshape(3, 4).copy()
    m[:, 3] += dist * m[:, 2]
    return m.astype(np.float32)


def raymarch_info(cam, step, sample_index=0):
    r = abi.RaymarchInfo()
    r.cam_info = cam
    r.stepsize_info = step
    r.sample_index = sample_index
    r.motion_blur = 0
    r.deterministic = 1
    return r


# ---------------------------------------------------------------------------
# the workloads (seeds fixed here)
# ---------------------------------------------------------------------------
SEEDS = {"weights": 1, "table": 2, "blobs": 3}


def eye_on_sphere(radius, azimuth_deg, elevation_deg):
    az, el = math.radians(azimuth_deg), math.radians(elevation_deg)
    return np.array([radius * math.cos(el) * math.cos(az), radius * math.cos(el) * math.sin(az),
                     radius * math.sin(el)])


class Workload:
    """Everything one render configuration needs, as host arrays/structs."""

    def __init__(self, name, scene, step, net, occupancy, params, cam2world, width, height, fov=50.0):
        self.name, self.scene, self.step, self.net = name, scene, step, net
        self.occupancy = occupancy
        self.params = params
        self.cam2world = cam2world
        self.width, self.height, self.fov = width, height, fov

    def camera(self, cam2world=None, width=None, height=None):
        return camera_info(self.cam2world if cam2world is None else cam2world,
                           width or self.width, height or self.height, self.fov)

    def rm(self, cam2world=None, width=None, height=None, sample_index=0):
        return raymarch_info(self.camera(cam2world, width, height), self.step, sample_index)


def small_plain(width=256, height=256, log2_hashmap_size=19, stepsize_setup=None, n_blobs=8):
    """256x256 plain workload: L=16 F=2 T=2^19 hash grid, one 128^3 cascade over
    [-1,1]^3 from 8 Gaussian blobs, camera at radius 4 looking at the origin,
    50 deg FOV, near 0.1, far 8, step sqrt(3)/1024, cone angle 0."""
    scene = scene_info(1.0, 1)
    step = stepsize_info(0.0, math.sqrt(3.0) / 1024.0, 0.1, 8.0)
    (stepsize_setup or _default_setup)(step, scene)
    net = net_desc(log2_hashmap_size=log2_hashmap_size)
    c, s, a = blobs(SEEDS["blobs"], n_blobs, 0.55, (0.06, 0.16))
    occ = occupancy_from_blobs(scene, c, s, a)
    params = make_params(net, SEEDS["weights"], SEEDS["table"])
    eye = eye_on_sphere(4.0, 30.0, 20.0)
    return Workload("small_plain_256", scene, step, net, occ, params, look_at(eye, (0, 0, 0)), width, height)


def small_plain_reference(width=256, height=256, n_output=8, stepsize_setup=None):
    """The small scene with the reference's plain-model layout: the density
    MLP's n_output rows (log sigma first) are the head's latents, and the
    cache stores the density in latent 0 (DensityIntermediates)."""
    wl = small_plain(width, height, stepsize_setup=stepsize_setup)
    net = net_desc(layout="reference", n_output=n_output)
    params = make_params(net, SEEDS["weights"], SEEDS["table"])
    return Workload("small_plain_reference_256", wl.scene, wl.step, net, wl.occupancy, params, wl.cam2world, width,
                    height)


def small_viewdep(width=256, height=256, first_latent_is_density=True, first_head_hidden_layers=1,
                  freq_encoding_degree=4, stepsize_setup=None):
    """small_plain's scene and camera with a split (view-dependent) model."""
    wl = small_plain(width, height, stepsize_setup=stepsize_setup)
    net = net_desc_viewdep(first_latent_is_density=first_latent_is_density,
                           first_head_hidden_layers=first_head_hidden_layers,
                           freq_encoding_degree=freq_encoding_degree)
    params = make_params(net, SEEDS["weights"], SEEDS["table"])
    return Workload("small_viewdep_256", wl.scene, wl.step, net, wl.occupancy, params, wl.cam2world, width, height)


def unbounded_1080p_viewdep(width=1920, height=1080, stepsize_setup=None):
    """unbounded_1080p's scene and camera with a split model (DensityIntermediates,
    one hidden first-head layer, frequency degree 4)."""
    wl = unbounded_1080p(width, height, stepsize_setup=stepsize_setup)
    net = net_desc_viewdep(log2_hashmap_size=wl.net.log2_hashmap_size)
    params = make_params(net, SEEDS["weights"], SEEDS["table"])
    return Workload("unbounded_1080p_viewdep", wl.scene, wl.step, net, wl.occupancy, params, wl.cam2world,
                    width, height)


def unbounded_1080p(width=1920, height=1080, log2_hashmap_size=21, stepsize_setup=None, n_blobs=64):
    """1920x1080 workload: T=2^21, 5 cascades of 128^3 over [-16,16]^3 with
    L-inf contraction of [-1,1]^3, 64 blobs out to radius 16, cone angle 1/256."""
    scene = scene_info(16.0, 5, abi.CONTRACTION_WARP_AABB_LINF, 1.0)
    step = stepsize_info(1.0 / 256.0, math.sqrt(3.0) / 1024.0, 0.1, 1e10)
    (stepsize_setup or _default_setup)(step, scene)
    net = net_desc(log2_hashmap_size=log2_hashmap_size)
    c, s, a = blobs_spread(SEEDS["blobs"], n_blobs, 16.0)
    occ = occupancy_from_blobs(scene, c, s, a)
    params = make_params(net, SEEDS["weights"], SEEDS["table"])
    eye = eye_on_sphere(4.0, 30.0, 20.0)
    return Workload("unbounded_1080p", scene, step, net, occ, params, look_at(eye, (0, 0, 0)), width, height)


def camera_path_1080p(wl, n_frames=30, yaw_deg=1.0, forward=0.02):
    """Frame k: yaw k deg about world up, then k*forward along the view axis."""
    mats = []
    m = wl.cam2world
    for k in range(n_frames):
        if k > 0:
            m = translate_forward(yaw(m, yaw_deg), forward)
        mats.append(m)
    return mats


BATCH_RESOLUTIONS = ((320, 180), (640, 360), (1280, 720), (1920, 1080))
BATCH_BLOB_COUNTS = (2, 4, 8, 16, 32, 64)


def batch_views(wl, path, n_views=16, resolutions=BATCH_RESOLUTIONS):
    """Multi-view batch: view k at pose 1 + k % 8 of the camera path with
    resolution resolutions[k % len(resolutions)]."""
    return [wl.rm(path[1 + k % 8], *resolutions[k % len(resolutions)]) for k in range(n_views)]


def _default_setup(step, scene):
    """StepsizeInfo::setAdditionalInfo through the product library (host code)."""
    from . import lib
    lib.stepsize_setup(step, scene)


# ---------------------------------------------------------------------------
# model export (the reference's nerfacc layout, loader.cu:232-285)
# ---------------------------------------------------------------------------
def nerfacc_config(wl):
    """config.json content of a workload in the nerfacc export layout."""
    sc, st, net = wl.scene, wl.step, wl.net
    scene = {
        "grid_resolution": int(sc.grid_resolution), "grid_nlvl": int(sc.grid_nlvl),
        "aabb": [sc.aabb_from.x, sc.aabb_from.y, sc.aabb_from.z, sc.aabb_to.x, sc.aabb_to.y, sc.aabb_to.z],
        "is_open_gl": bool(sc.is_open_gl), "alpha_thre": float(sc.alpha_thre),
        "cone_angle": float(st.cone_angle), "stepsize": float(st.stepsize), "near": float(st.near_),
        "far": float(st.far_),
    }
    if sc.contraction_type == abi.CONTRACTION_WARP_AABB_LINF:
        c0, c1 = sc.contraction_aabb_from, sc.contraction_aabb_to
        scene["contraction"] = {"aabb": [c0.x, c0.y, c0.z, c1.x, c1.y, c1.z]}
    mlp = {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "None"}
    cfg = {
        "scene": scene,
        "mlp_base": {
            "encoding": {"otype": "HashGrid", "n_levels": net.n_levels, "n_features_per_level": net.n_features_per_level,
                         "log2_hashmap_size": net.log2_hashmap_size, "base_resolution": net.base_resolution,
                         "per_level_scale": float(np.float32(net.per_level_scale))},
            "network": dict(mlp, n_neurons=net.base_width, n_hidden_layers=net.base_n_hidden_layers),
            "n_output": net.base_n_output,
        },
        "mlp_head": {
            "encoding": {"otype": "Composite", "nested": [
                {"n_dims_to_encode": 3, "otype": "SphericalHarmonics", "degree": net.sh_degree},
                {"otype": "Identity"}]},
            "network": dict(mlp, n_neurons=net.head_width, n_hidden_layers=net.head_n_hidden_layers),
        },
    }
    if not net.split_head:
        cfg["mlp_head"]["n_input"] = net.n_latents + 16
        if net.latent_offset == 1:
            # this library's key for the separate-density layout (reference
            # plain exports read the first n_output rows as latents)
            cfg["mlp_head"]["latent_offset"] = 1
    if net.split_head:
        # init_viewdep (nerf_network.h:120-158) keys
        cfg["mlp_head"]["n_input"] = head_input_width(net)
        cfg["mlp_head"]["freq_encoding_degree"] = net.freq_encoding_degree
        cfg["mlp_first_head"] = {
            "encoding": {"otype": "Composite", "nested": [
                {"n_dims_to_encode": 3, "otype": "SphericalHarmonics", "degree": net.first_sh_degree},
                {"otype": "Identity"}]},
            "network": dict(mlp, n_neurons=net.first_head_width, n_hidden_layers=net.first_head_n_hidden_layers),
            "n_input": net.first_head_n_input, "n_output": net.n_latents,
            "first_latent_is_density": bool(net.first_latent_is_density),
        }
    return cfg


def write_nerfacc_export(wl, out_dir, pos_separate=True):
    """Write a workload as config.json + raw fp16 / uint8 .dat files; with
    pos_separate=False the hash table is appended to mlp_base.dat."""
    import json
    import os
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "config.json"), "w") as f:
        json.dump(nerfacc_config(wl), f, indent=1)
    blocks = [np.ascontiguousarray(p, np.uint16) for p in wl.params]
    pos, base, head = blocks[0], blocks[1], blocks[-1]
    if len(blocks) == 4:
        blocks[2].tofile(os.path.join(out_dir, "mlp_first_head.dat"))
    if pos_separate:
        pos.tofile(os.path.join(out_dir, "pos_enc.dat"))
        base.tofile(os.path.join(out_dir, "mlp_base.dat"))
    else:
        np.concatenate([base, pos]).tofile(os.path.join(out_dir, "mlp_base.dat"))
    head.tofile(os.path.join(out_dir, "mlp_head.dat"))
    np.ascontiguousarray(wl.occupancy, np.uint8).tofile(os.path.join(out_dir, "occupancy_grid.dat"))
    return out_dir
