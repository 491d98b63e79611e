"""ctypes mirrors of the POD structs declared in include/fvc.h.

The layouts follow the reference types field by field (include/common.h:22-218,
include/cache_renderer.h:24-32, 104-109, include/cache_performance_renderer.h:123-131),
with C++ bools widened to int32.
"""
import ctypes as C

import numpy as np


class Float2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Float4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class CameraMatrix(C.Structure):
    """CameraMatrix (common.h:22-62): rows m0, m1, m2 of a 3x4 transform."""
    _fields_ = [("m0", Float4), ("m1", Float4), ("m2", Float4)]

    @classmethod
    def from_array(cls, a):
        a = np.asarray(a, dtype=np.float32).reshape(3, 4)
        return cls(*[Float4(*map(float, a[i])) for i in range(3)])

    def to_array(self):
        return np.array([[r.x, r.y, r.z, r.w] for r in (self.m0, self.m1, self.m2)], dtype=np.float32)


CONTRACTION_NONE, CONTRACTION_AABB, CONTRACTION_WARP_AABB_L2, CONTRACTION_WARP_AABB_LINF = 0, 1, 2, 3


class SceneInfo(C.Structure):
    """SceneInfo (common.h:74-91)."""
    _fields_ = [
        ("grid_resolution", C.c_int32), ("grid_nlvl", C.c_int32), ("vals_per_lvl", C.c_int32),
        ("normalized", C.c_int32),
        ("aabb_from", Float3), ("aabb_to", Float3),
        ("contraction_aabb_from", Float3), ("contraction_aabb_to", Float3),
        ("contraction_type", C.c_int32), ("is_open_gl", C.c_int32), ("alpha_thre", C.c_float),
    ]


class StepsizeInfo(C.Structure):
    """StepsizeInfo (common.h:93-164)."""
    _fields_ = [(n, C.c_float) for n in (
        "cone_angle", "log1p_c", "stepsize", "near_", "near_i", "far_",
        "dt_min", "dt_max", "t_min", "t_max", "i_min", "i_max")]


class CameraInfo(C.Structure):
    """CameraInfo (common.h:166-195)."""
    _fields_ = [
        ("res_x", C.c_int32), ("res_y", C.c_int32), ("is_open_gl", C.c_int32),
        ("focal", Float2), ("principal", Float2),
        ("cam2world", CameraMatrix), ("world2cam", CameraMatrix),
    ]


class RaymarchInfo(C.Structure):
    """RaymarchInfo (common.h:210-218)."""
    _fields_ = [
        ("cam_info", CameraInfo), ("stepsize_info", StepsizeInfo),
        ("sample_index", C.c_int32), ("motion_blur", C.c_int32), ("deterministic", C.c_int32),
    ]


PREC_FP16, PREC_BF16, PREC_FP8, PREC_BF16X3 = 0, 1, 2, 3


class NetDesc(C.Structure):
    """Network description: the tcnn config subset of nerf_network.h:88-118."""
    _fields_ = [
        ("n_levels", C.c_int32), ("n_features_per_level", C.c_int32), ("log2_hashmap_size", C.c_int32),
        ("base_resolution", C.c_int32), ("per_level_scale", C.c_float),
        ("base_width", C.c_int32), ("base_n_hidden_layers", C.c_int32), ("base_n_output", C.c_int32),
        ("head_width", C.c_int32), ("head_n_hidden_layers", C.c_int32),
        ("n_latents", C.c_int32), ("latent_offset", C.c_int32), ("sh_degree", C.c_int32),
        ("precision", C.c_int32),
        ("split_head", C.c_int32), ("first_head_width", C.c_int32), ("first_head_n_hidden_layers", C.c_int32),
        ("first_head_n_input", C.c_int32), ("first_latent_is_density", C.c_int32), ("first_sh_degree", C.c_int32),
        ("freq_encoding_degree", C.c_int32), ("pad_", C.c_int32),
    ]


class CacheSettings(C.Structure):
    """CacheSettings (cache_renderer.h:24-32); defaults = default_cache_settings (55-63)."""
    _fields_ = [(n, C.c_int32) for n in (
        "interpol_variant", "interpol_function", "use_hw_interpol", "reweight_intermediates",
        "disable_reweighting_first_intermediate", "use_inter_brick_interpolation")]

    @classmethod
    def default(cls):
        return cls(0, 0, 1, 1, 1, 1)


class InitStats(C.Structure):
    """CacheRenderer::InitStats (cache_renderer.h:104-109) + diagnostics."""
    _fields_ = [("samples_ppx", C.c_float), ("n_bricks_set", C.c_int32), ("n_bricks_reserved", C.c_int32),
                ("overflow", C.c_int32), ("n_subiters", C.c_int32), ("n_samples_evaluated", C.c_int64)]


FILL_IDLE, FILL_RUNNING, FILL_DONE = 0, 1, 2

# DensityInterpolVariant (cache_renderer.h:15-22), InterpolFunction
# (common.h:221-227), CachePerformanceRenderer::RenderMode
# (cache_performance_renderer.h:23-28)
INTERPOL_DENSITY, INTERPOL_DENSITY_INTERMEDIATES = 0, 3
FN_LINEAR, FN_SMOOTHSTEP, FN_SMOOTHERSTEP, FN_NEAREST = 0, 1, 2, 3
RENDER_BOTH, RENDER_ONLY_CACHE, RENDER_ONLY_RESAMPLED = 0, 1, 2


class RefillPolicy(C.Structure):
    """fvc_refill_policy (the viewer's m_cache_miss_limit / m_cache_update_distance)."""
    _fields_ = [("min_hit_ratio", C.c_float), ("max_distance", C.c_float), ("enabled", C.c_int32),
                ("pad_", C.c_int32)]


TUNE_MAX = 64


class TuneResult(C.Structure):
    """fvc_tune_result (autotuner)."""
    _fields_ = [("mlp_rows", C.c_int32), ("max_steps", C.c_int32), ("target_batch", C.c_int64),
                ("ms_best", C.c_float), ("ms_default", C.c_float), ("n_tried", C.c_int32), ("pad_", C.c_int32),
                ("ms", C.c_float * TUNE_MAX), ("rows", C.c_int32 * TUNE_MAX), ("steps", C.c_int32 * TUNE_MAX),
                ("targets", C.c_int64 * TUNE_MAX)]


class TuneSpace(C.Structure):
    """fvc_tune_space (autotuner candidate sets)."""
    _fields_ = [("n_rows", C.c_int32), ("rows", C.c_int32 * 3), ("n_steps", C.c_int32), ("steps", C.c_int32 * 3),
                ("n_targets", C.c_int32), ("pad_", C.c_int32), ("targets", C.c_int64 * 8)]

    @classmethod
    def of(cls, rows, targets, steps):
        sp = cls()
        sp.n_rows, sp.n_steps, sp.n_targets = len(rows), len(steps), len(targets)
        for i, v in enumerate(rows):
            sp.rows[i] = v
        for i, v in enumerate(steps):
            sp.steps[i] = v
        for i, v in enumerate(targets):
            sp.targets[i] = v
        return sp


class DiffStats(C.Structure):
    """fvc_diff_stats (differentiable render)."""
    _fields_ = [("n_rays", C.c_int64), ("n_samples", C.c_int64), ("n_cache_samples", C.c_int64)]


class RenderStats(C.Structure):
    """RenderStats (cache_performance_renderer.h:123-131)."""
    _fields_ = [("samples_ppx", C.c_float), ("resample_samples_evaluated", C.c_float),
                ("cache_samples_evaluated", C.c_float), ("n_resamples", C.c_uint64),
                ("n_cache_hits", C.c_uint64), ("n_subiters", C.c_int32), ("pad_", C.c_int32)]


class CacheInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "frustum_x", "frustum_y", "frustum_z", "bricks_x", "bricks_y", "bricks_z",
        "capacity_bricks", "n_bricks_used", "initialized", "brick_size", "latent_format")]


class KernelStats(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("units", C.c_double), ("alg_bytes", C.c_double), ("alg_flops", C.c_double)]


# voxel record of the brick pool export: 8 fp16 latents, density, isset, density and isset of the +z neighbour
VOXEL_DTYPE = np.dtype([("lat", np.uint16, 8), ("density", np.float32), ("isset", np.float32),
                        ("znext", np.float32, 2)])
assert VOXEL_DTYPE.itemsize == 32


def f3(v):
    return Float3(*map(float, v))


def ptr(a, ctype=C.c_void_p):
    """Pointer to a numpy array's data (contiguous)."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctype) if ctype is not C.c_void_p else C.c_void_p(a.ctypes.data)
