"""ctypes binding of libfvc.so and a thin mirror of the reference's renderer
classes (PerformanceRenderer, CachePerformanceRenderer / CacheRenderer,
NerfNetwork) over the C ABI of include/fvc.h.

Device buffers are passed as raw device addresses: anything with a
``data_ptr()`` (a torch tensor on the GPU) or a plain int.  The library is
the only compute path: if libfvc.so is missing this module raises.
"""
import ctypes as C
import os

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfvc.so")

_lib = None


class FvcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fvc error {code}: {msg}")
        self.code = code


OK, E_INVALID, E_UNSUPPORTED, E_NOT_INITIALIZED, E_HIP, E_OUT_OF_MEMORY, E_NO_DEVICE = 0, -1, -2, -3, -4, -5, -6

_P = C.c_void_p
_SIGS = {
    "fvc_last_error_message": (C.c_char_p, []),
    "fvc_version": (C.c_int, []),
    "fvc_stepsize_setup": (C.c_int, [C.POINTER(abi.StepsizeInfo), C.POINTER(abi.SceneInfo)]),
    "fvc_net_param_counts": (C.c_int, [C.POINTER(abi.NetDesc), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64)]),
    "fvc_model_create": (C.c_int, [C.POINTER(abi.SceneInfo), C.POINTER(abi.NetDesc), _P, _P, C.c_int64, _P,
                                   C.c_int64, _P, C.c_int64, C.POINTER(_P)]),
    "fvc_net_param_counts_viewdep": (C.c_int, [C.POINTER(abi.NetDesc), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                               C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "fvc_model_create_viewdep": (C.c_int, [C.POINTER(abi.SceneInfo), C.POINTER(abi.NetDesc), _P, _P, C.c_int64, _P,
                                           C.c_int64, _P, C.c_int64, _P, C.c_int64, C.POINTER(_P)]),
    "fvc_nerfacc_read_data_viewdep": (C.c_int, [C.c_char_p, _P, _P, _P, _P, _P]),
    "fvc_network_latents_viewdep": (C.c_int, [_P, _P, _P, C.c_int64, _P, _P]),
    "fvc_network_head_viewdep": (C.c_int, [_P, _P, _P, _P, _P, C.c_int64, _P, _P]),
    "fvc_model_destroy": (None, [_P]),
    "fvc_nerfacc_read_config": (C.c_int, [C.c_char_p, C.POINTER(abi.SceneInfo), C.POINTER(abi.StepsizeInfo),
                                          C.POINTER(abi.NetDesc), C.POINTER(C.c_int64)]),
    "fvc_nerfacc_read_data": (C.c_int, [C.c_char_p, _P, _P, _P, _P]),
    "fvc_model_load_nerfacc": (C.c_int, [C.c_char_p, C.POINTER(abi.SceneInfo), C.POINTER(abi.StepsizeInfo),
                                         C.POINTER(abi.NetDesc), C.POINTER(_P)]),
    "fvc_model_get_bitfield": (C.c_int, [_P, _P, C.c_int64]),
    "fvc_network_latents": (C.c_int, [_P, _P, C.c_int64, _P, _P]),
    "fvc_network_inference": (C.c_int, [_P, _P, _P, C.c_int64, _P, _P]),
    "fvc_network_head": (C.c_int, [_P, _P, _P, C.c_int64, _P, _P]),
    "fvc_network_encode": (C.c_int, [_P, _P, C.c_int64, _P, _P]),
    "fvc_renderer_create": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "fvc_renderer_destroy": (None, [_P]),
    "fvc_renderer_set_stream": (C.c_int, [_P, _P]),
    "fvc_resize_renderbuffers": (C.c_int, [_P, C.c_int, C.c_int]),
    "fvc_renderer_autotune": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo), C.c_int, C.POINTER(abi.TuneResult)]),
    "fvc_renderer_autotune_batch": (C.c_int, [_P, C.POINTER(_P), C.POINTER(abi.RaymarchInfo), C.c_int32, C.c_int,
                                              C.POINTER(abi.TuneSpace), C.POINTER(abi.TuneResult)]),
    "fvc_renderer_set_mlp_rows": (C.c_int, [_P, C.c_int32]),
    "fvc_renderer_set_graph": (C.c_int, [_P, C.c_int]),
    "fvc_renderer_set_batching": (C.c_int, [_P, C.c_int64, C.c_int32]),
    "fvc_render_full": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo), _P, C.POINTER(abi.RenderStats)]),
    "fvc_cache_fill": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo), C.POINTER(abi.InitStats)]),
    "fvc_render_cached": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo), _P, C.POINTER(abi.RenderStats)]),
    "fvc_cache_swap": (C.c_int, [_P]),
    "fvc_cache_fill_async": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo)]),
    "fvc_cache_fill_poll": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int32), C.POINTER(abi.InitStats)]),
    "fvc_cache_publish": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "fvc_renderer_set_refill_policy": (C.c_int, [_P, C.POINTER(abi.RefillPolicy)]),
    "fvc_render_cached_auto": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo), _P, C.POINTER(abi.RenderStats),
                                         C.POINTER(C.c_int32)]),
    "fvc_render_full_batch": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo), C.c_int32, C.POINTER(_P),
                                        C.POINTER(abi.RenderStats)]),
    "fvc_render_cached_batch": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo), C.c_int32, C.POINTER(_P),
                                          C.POINTER(abi.RenderStats)]),
    "fvc_render_full_batch_scenes": (C.c_int, [_P, C.POINTER(_P), C.POINTER(abi.RaymarchInfo), C.c_int32,
                                               C.POINTER(_P), C.POINTER(abi.RenderStats)]),
    "fvc_cache_set_settings": (C.c_int, [_P, C.POINTER(abi.CacheSettings)]),
    "fvc_cache_set_capacity": (C.c_int, [_P, C.c_int32]),
    "fvc_cache_set_latent_format": (C.c_int, [_P, C.c_int]),
    "fvc_cache_get_pool_decoded": (C.c_int, [_P, C.c_int, _P, C.c_int64]),
    "fvc_get_last_result": (C.c_int, [_P, _P, C.c_int64]),
    "fvc_cache_get_info": (C.c_int, [_P, C.c_int, C.POINTER(abi.CacheInfo)]),
    "fvc_cache_get_brick_table": (C.c_int, [_P, C.c_int, _P, C.c_int64]),
    "fvc_cache_get_known_ranges": (C.c_int, [_P, C.c_int, _P, C.c_int64]),
    "fvc_cache_get_pool": (C.c_int, [_P, C.c_int, _P, C.c_int64]),
    "fvc_debug_march": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo), C.c_int32, _P, _P]),
    "fvc_cache_set_render_mode": (C.c_int, [_P, C.c_int]),
    "fvc_render_full_noalloc": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo), _P, _P, C.POINTER(abi.RenderStats)]),
    "fvc_debug_cache_lookup": (C.c_int, [_P, C.c_int, _P, C.c_int64, _P]),
    "fvc_debug_render_cached": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo), C.c_int32, _P, _P, _P,
                                          C.POINTER(abi.RenderStats)]),
    "fvc_diff_forward": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo), _P, C.c_int32, _P, _P, _P, _P,
                                   C.POINTER(abi.DiffStats)]),
    "fvc_diff_backward": (C.c_int, [_P, _P, _P, _P, _P]),
    "fvc_diff_cached_forward": (C.c_int, [_P, C.POINTER(abi.RaymarchInfo), _P, C.c_int32, _P, _P, _P, _P,
                                          C.POINTER(abi.DiffStats)]),
    "fvc_diff_cached_backward": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "fvc_profile_enable": (C.c_int, [_P, C.c_int]),
    "fvc_profile_get": (C.c_int, [_P, _P, C.c_int32, C.POINTER(C.c_int32)]),
}


def exported_symbols():
    return list(_SIGS)


def load():
    """Load libfvc.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing: run `python -m frustumvolumecaching_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        raise FvcError(rc, load().fvc_last_error_message().decode(errors="replace"))


def _addr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return C.c_void_p(x)
    if hasattr(x, "data_ptr"):
        return C.c_void_p(x.data_ptr())
    if hasattr(x, "ctypes"):
        return C.c_void_p(x.ctypes.data)
    raise TypeError(f"cannot take the address of {type(x)}")


def _after_caller(*tensors):
    """Finish the work the caller queued on torch's current stream before the
    renderer reads or writes these tensors.  The renderer launches on its
    own non-blocking stream, which does not wait for torch's stream: without
    this a zero fill of a gradient block, a parameter update or the loss
    gradient queued there could still be running when the library's kernels
    read or accumulate into the same memory."""
    for t in tensors:
        if getattr(t, "is_cuda", False):
            import torch
            torch.cuda.current_stream(t.device).synchronize()
            return


def _stream(s):
    if s is None:
        return None
    if hasattr(s, "cuda_stream"):
        return C.c_void_p(s.cuda_stream)
    return C.c_void_p(int(s))


def stepsize_setup(step, scene):
    """StepsizeInfo::setAdditionalInfo (common.h:109-128)."""
    check(load().fvc_stepsize_setup(C.byref(step), C.byref(scene)))
    return step


def nerfacc_read(data_dir):
    """load_nerfacc_export without the GPU part (loader.cu:232-285): returns
    (scene, stepsize, net, (pos_enc, mlp_base, mlp_head) fp16 bits, occupancy);
    split models (pos_enc, mlp_base, mlp_first_head, mlp_head)."""
    import numpy as np
    scene, step, net = abi.SceneInfo(), abi.StepsizeInfo(), abi.NetDesc()
    counts = (C.c_int64 * 4)()
    d = os.fsencode(str(data_dir))
    check(load().fvc_nerfacc_read_config(d, C.byref(scene), C.byref(step), C.byref(net), counts))
    pos, base, head = (np.zeros(int(counts[k]), np.uint16) for k in range(3))
    occ = np.zeros(int(counts[3]), np.uint8)
    if net.split_head:
        first = np.zeros(net_param_counts_viewdep(net)[2], np.uint16)
        check(load().fvc_nerfacc_read_data_viewdep(d, _addr(pos), _addr(base), _addr(first), _addr(head),
                                                   _addr(occ)))
        return scene, step, net, (pos, base, first, head), occ
    check(load().fvc_nerfacc_read_data(d, _addr(pos), _addr(base), _addr(head), _addr(occ)))
    return scene, step, net, (pos, base, head), occ


def net_param_counts_viewdep(net):
    """(pos_enc, mlp_base, mlp_first_head, mlp_head) counts of a split model."""
    a, b, f, c = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
    check(load().fvc_net_param_counts_viewdep(C.byref(net), C.byref(a), C.byref(b), C.byref(f), C.byref(c)))
    return a.value, b.value, f.value, c.value


def net_param_counts(net):
    a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
    check(load().fvc_net_param_counts(C.byref(net), C.byref(a), C.byref(b), C.byref(c)))
    return a.value, b.value, c.value


class NerfNetwork:
    """Model handle: NerfNetwork<half> + Morton occupancy bitfield
    (nerf_network.h:84-369, loader.cu:20-55)."""

    def __init__(self, scene, net, occupancy_linear, params):
        """params: (pos_enc, mlp_base, mlp_head) fp16 bits, or (pos_enc,
        mlp_base, mlp_first_head, mlp_head) for a split model (set_params,
        nerf_network.h:269-293)."""
        import numpy as np
        blocks = [np.ascontiguousarray(p, dtype=np.uint16) for p in params]
        occ = np.ascontiguousarray(occupancy_linear, dtype=np.uint8)
        self.scene, self.net = scene, net
        h = C.c_void_p()
        if len(blocks) == 4:
            pos, base, first, head = blocks
            check(load().fvc_model_create_viewdep(C.byref(scene), C.byref(net), _addr(occ), _addr(pos), pos.size,
                                                  _addr(base), base.size, _addr(first), first.size, _addr(head),
                                                  head.size, C.byref(h)))
        else:
            pos, base, head = blocks
            check(load().fvc_model_create(C.byref(scene), C.byref(net), _addr(occ), _addr(pos), pos.size,
                                          _addr(base), base.size, _addr(head), head.size, C.byref(h)))
        self.h = h

    @classmethod
    def from_workload(cls, wl):
        return cls(wl.scene, wl.net, wl.occupancy, wl.params)

    @classmethod
    def from_nerfacc(cls, data_dir):
        """load_data for a nerfacc export directory (loader.cu:232-301);
        returns (model, scene, stepsize)."""
        self = cls.__new__(cls)
        scene, step, net = abi.SceneInfo(), abi.StepsizeInfo(), abi.NetDesc()
        h = C.c_void_p()
        check(load().fvc_model_load_nerfacc(os.fsencode(str(data_dir)), C.byref(scene), C.byref(step),
                                            C.byref(net), C.byref(h)))
        self.scene, self.net, self.h = scene, net, h
        return self, scene, step

    def close(self):
        if getattr(self, "h", None):
            load().fvc_model_destroy(self.h)
            self.h = None

    __del__ = close

    def bitfield(self):
        import numpy as np
        out = np.zeros(self.scene.grid_nlvl * 128 ** 3 // 8, np.uint8)
        check(load().fvc_model_get_bitfield(self.h, _addr(out), out.size))
        return out

    # NerfNetwork::inferenceLatents / inference / inferenceHead
    def inferenceLatents(self, pos, n, out16, stream=None):
        check(load().fvc_network_latents(self.h, _addr(pos), n, _addr(out16), _stream(stream)))

    def inference(self, pos, dir01, n, out4, stream=None):
        check(load().fvc_network_inference(self.h, _addr(pos), _addr(dir01), n, _addr(out4), _stream(stream)))

    def inferenceHead(self, head_in16, dir01, n, out4, stream=None):
        check(load().fvc_network_head(self.h, _addr(head_in16), _addr(dir01), n, _addr(out4), _stream(stream)))

    # split models: inferenceLatents with the initial view direction, inferenceHead
    def inferenceLatentsViewdep(self, pos, init_dir01, n, out16, stream=None):
        check(load().fvc_network_latents_viewdep(self.h, _addr(pos), _addr(init_dir01), n, _addr(out16),
                                                 _stream(stream)))

    def inferenceHeadViewdep(self, head_in16, pos, dir01, init_dir01, n, out4, stream=None):
        check(load().fvc_network_head_viewdep(self.h, _addr(head_in16), _addr(pos), _addr(dir01), _addr(init_dir01),
                                              n, _addr(out4), _stream(stream)))

    def encode(self, pos, n, out32, stream=None):
        check(load().fvc_network_encode(self.h, _addr(pos), n, _addr(out32), _stream(stream)))


class Renderer:
    """One handle = PerformanceRenderer + CacheRenderer + CachePerformanceRenderer
    sharing the model (eval.cu:208-227).  double_buffered=False mirrors
    CachePerformanceRenderer(&cache, &cache)."""

    def __init__(self, model, double_buffered=False):
        self.model = model
        h = C.c_void_p()
        check(load().fvc_renderer_create(model.h, int(bool(double_buffered)), C.byref(h)))
        self.h = h
        self.last_render_stats = None
        self.init_stats = None

    def close(self):
        if getattr(self, "h", None):
            load().fvc_renderer_destroy(self.h)
            self.h = None

    __del__ = close

    def set_stream(self, stream):
        check(load().fvc_renderer_set_stream(self.h, _stream(stream)))

    def set_batching(self, target_samples, max_samples_per_ray):
        check(load().fvc_renderer_set_batching(self.h, int(target_samples), int(max_samples_per_ray)))

    def autotune(self, rm, cached=False):
        """Pick MLP rows per wavefront and chunk sizes by timing renders of rm."""
        res = abi.TuneResult()
        check(load().fvc_renderer_autotune(self.h, C.byref(rm), int(bool(cached)), C.byref(res)))
        return res

    def autotune_batch(self, rms, cached=False, models=None, rows=None, targets=None, steps=None):
        """autotune over a batch of views (and per-view models, full only);
        rows / targets / steps: candidate sets (None: the defaults)."""
        res = abi.TuneResult()
        n = len(rms)
        arr = (abi.RaymarchInfo * n)(*rms)
        mp = None if models is None else (_P * n)(*[None if m is None else m.h for m in models])
        sp = None
        if rows or targets or steps:
            sp = abi.TuneSpace.of(rows or (16, 32, 64), targets or (1 << 23, 1 << 24, 1 << 25), steps or (16, 32, 64))
        check(load().fvc_renderer_autotune_batch(self.h, mp, arr, n, int(bool(cached)),
                                                 None if sp is None else C.byref(sp), C.byref(res)))
        return res

    def set_graph(self, enable=True):
        """Record and replay each frame as a hipGraph."""
        check(load().fvc_renderer_set_graph(self.h, int(bool(enable))))

    def set_mlp_rows(self, rows):
        check(load().fvc_renderer_set_mlp_rows(self.h, int(rows)))

    def resizeRenderbuffers(self, width, height):
        check(load().fvc_resize_renderbuffers(self.h, int(width), int(height)))

    # PerformanceRenderer::render
    def render_full(self, rm, rgba=None):
        st = abi.RenderStats()
        _after_caller(rgba)
        check(load().fvc_render_full(self.h, C.byref(rm), _addr(rgba), C.byref(st)))
        self.last_render_stats = st
        return st

    def render_full_noalloc(self, rm, rgba=None, result=None):
        """render_full with the per-frame buffers in the library's static
        scratch (views up to 256 x 256; no device allocation)."""
        st = abi.RenderStats()
        _after_caller(rgba, result)
        check(load().fvc_render_full_noalloc(self.h, C.byref(rm), _addr(rgba), _addr(result), C.byref(st)))
        self.last_render_stats = st
        return st

    # CacheRenderer::initCacheBlockwise
    def initCacheBlockwise(self, rm):
        st = abi.InitStats()
        check(load().fvc_cache_fill(self.h, C.byref(rm), C.byref(st)))
        self.init_stats = st
        return st

    # CachePerformanceRenderer::render
    def render_cached(self, rm, rgba=None):
        st = abi.RenderStats()
        _after_caller(rgba)
        check(load().fvc_render_cached(self.h, C.byref(rm), _addr(rgba), C.byref(st)))
        self.last_render_stats = st
        return st

    def _batch(self, fn, rms, rgbas):
        n = len(rms)
        arr = (abi.RaymarchInfo * n)(*rms)
        ptrs = (_P * n)(*[None if x is None else _addr(x).value for x in (rgbas or [None] * n)])
        st = abi.RenderStats()
        _after_caller(*(rgbas or []))
        check(fn(self.h, arr, n, ptrs, C.byref(st)))
        self.last_render_stats = st
        return st

    def render_full_batch(self, rms, rgbas=None, models=None):
        """Render several views (any resolutions, shared stepsize) in one pass.
        models (optional): the NerfNetwork of each view (None: the renderer's),
        sharing the renderer's network and scene box; their occupancy grids
        may differ."""
        if models is None:
            return self._batch(load().fvc_render_full_batch, rms, rgbas)
        n = len(rms)
        if len(models) != n:
            raise ValueError("one model per view")
        mp = (_P * n)(*[None if m is None else m.h for m in models])
        return self._batch(lambda h, arr, k, ptrs, st: load().fvc_render_full_batch_scenes(h, mp, arr, k, ptrs, st),
                           rms, rgbas)

    def render_cached_batch(self, rms, rgbas=None):
        return self._batch(load().fvc_render_cached_batch, rms, rgbas)

    # background refill (double-buffered renderers)
    def initCacheAsync(self, rm):
        check(load().fvc_cache_fill_async(self.h, C.byref(rm)))

    def fill_poll(self, wait=False):
        st, stats = C.c_int32(), abi.InitStats()
        check(load().fvc_cache_fill_poll(self.h, int(bool(wait)), C.byref(st), C.byref(stats)))
        return st.value, stats

    def publish(self):
        sw = C.c_int32()
        check(load().fvc_cache_publish(self.h, C.byref(sw)))
        return bool(sw.value)

    def set_refill_policy(self, min_hit_ratio=0.9, max_distance=1e30, enabled=True):
        p = abi.RefillPolicy(float(min_hit_ratio), float(max_distance), int(bool(enabled)), 0)
        check(load().fvc_renderer_set_refill_policy(self.h, C.byref(p)))

    def render_cached_auto(self, rm, rgba=None):
        """Cached frame with the refill policy; returns (stats, refill_started)."""
        st, started = abi.RenderStats(), C.c_int32()
        check(load().fvc_render_cached_auto(self.h, C.byref(rm), _addr(rgba), C.byref(st), C.byref(started)))
        self.last_render_stats = st
        return st, bool(started.value)

    def swapCache(self):
        check(load().fvc_cache_swap(self.h))

    def setCacheSettings(self, settings):
        check(load().fvc_cache_set_settings(self.h, C.byref(settings)))

    def cache_settings(self, interpol_function=abi.FN_LINEAR, use_inter_brick_interpolation=True):
        """setCacheSettings with the model's interpolation variant."""
        s = abi.CacheSettings()
        s.interpol_variant = (abi.INTERPOL_DENSITY_INTERMEDIATES if self.model.net.first_latent_is_density
                              else abi.INTERPOL_DENSITY)
        s.interpol_function = int(interpol_function)
        s.use_hw_interpol = 1
        s.reweight_intermediates = 1
        s.disable_reweighting_first_intermediate = 1
        s.use_inter_brick_interpolation = int(bool(use_inter_brick_interpolation))
        self.setCacheSettings(s)

    def set_render_mode(self, mode):
        """CachePerformanceRenderer::_render_mode: abi.RENDER_BOTH,
        RENDER_ONLY_CACHE or RENDER_ONLY_RESAMPLED."""
        check(load().fvc_cache_set_render_mode(self.h, int(mode)))

    def debug_lookup(self, froxel, which=0):
        """The cache head's lookup at froxel points [n, 3]: [n, 9] = (density
        row, latents 0..7) as half values."""
        import numpy as np
        f = np.ascontiguousarray(froxel, np.float32).reshape(-1, 3)
        out = np.zeros((len(f), 9), np.float32)
        check(load().fvc_debug_cache_lookup(self.h, int(which), _addr(f), len(f), _addr(out)))
        return out

    def debug_render_cached(self, rm, k=64):
        """render_cached with the per-pixel records: codes [H, W, k] (0 skip,
        1 resample, 2 cache, -1 past the classified samples), classified
        [H, W], composited [H, W, 2] (resampled, cached)."""
        import numpy as np
        H, W = rm.cam_info.res_y, rm.cam_info.res_x
        codes = np.full((H, W, k), -1, np.int8)
        cls = np.zeros((H, W), np.int32)
        comp = np.zeros((H, W, 2), np.int32)
        st = abi.RenderStats()
        check(load().fvc_debug_render_cached(self.h, C.byref(rm), int(k), _addr(codes), _addr(cls), _addr(comp),
                                             C.byref(st)))
        return codes, cls, comp, st

    def set_latent_format(self, fmt):
        """'fp16' or 'fp8' latents for the next fills."""
        check(load().fvc_cache_set_latent_format(self.h, {"fp16": 0, "fp8": 1}[fmt]))

    def pool_decoded(self, which=0):
        """[n_bricks, 512, 10] float32: 8 latents, density, isset."""
        import numpy as np
        i = self.cache_info(which)
        out = np.zeros((max(i.n_bricks_used, 0), 512, 10), np.float32)
        check(load().fvc_cache_get_pool_decoded(self.h, which, _addr(out), out.size))
        return out

    def set_cache_capacity(self, bricks):
        check(load().fvc_cache_set_capacity(self.h, int(bricks)))

    def last_result(self, out_dev, n_floats):
        _after_caller(out_dev)
        check(load().fvc_get_last_result(self.h, _addr(out_dev), int(n_floats)))

    def cache_info(self, which=0):
        info = abi.CacheInfo()
        check(load().fvc_cache_get_info(self.h, which, C.byref(info)))
        return info

    def brick_table(self, which=0):
        import numpy as np
        i = self.cache_info(which)
        out = np.zeros((i.bricks_z, i.bricks_y, i.bricks_x), np.int32)
        check(load().fvc_cache_get_brick_table(self.h, which, _addr(out), out.size))
        return out

    def known_ranges(self, which=0):
        import numpy as np
        i = self.cache_info(which)
        out = np.zeros((i.frustum_y, i.frustum_x), np.float32)
        check(load().fvc_cache_get_known_ranges(self.h, which, _addr(out), out.size))
        return out

    def pool(self, which=0, n_bricks=None):
        import numpy as np
        i = self.cache_info(which)
        n = i.n_bricks_used if n_bricks is None else n_bricks
        out = np.zeros((n, 8, 8, 8), abi.VOXEL_DTYPE)
        if n:
            check(load().fvc_cache_get_pool(self.h, which, _addr(out), out.nbytes))
        return out

    def profile_enable(self, enable=True):
        check(load().fvc_profile_enable(self.h, int(bool(enable))))

    def profile(self):
        """{kernel: dict(launches, ms, units, bytes, flops)} since profile_enable."""
        n = C.c_int32()
        check(load().fvc_profile_get(self.h, None, 0, C.byref(n)))
        arr = (abi.KernelStats * max(1, n.value))()
        check(load().fvc_profile_get(self.h, C.cast(arr, C.c_void_p), n.value, C.byref(n)))
        return {a.name.decode(): dict(launches=a.launches, ms=a.total_ms, units=a.units, bytes=a.alg_bytes,
                                      flops=a.alg_flops) for a in arr[:n.value]}

    # differentiable full render (training mode)
    def diff_forward(self, rm, pix, params, rgb):
        """rgb [n, 3] (device) = C + T of the pixels pix (device int32, y * W + x)
        with FP32 parameter blocks params = (pos_enc, mlp_base, mlp_head)."""
        st = abi.DiffStats()
        pos, base, head = params
        _after_caller(pix, pos, base, head, rgb)
        check(load().fvc_diff_forward(self.h, C.byref(rm), _addr(pix), int(pix.numel()), _addr(pos), _addr(base),
                                      _addr(head), _addr(rgb), C.byref(st)))
        return st

    def diff_backward(self, d_rgb, grads):
        """Adds d(sum(d_rgb * rgb))/d(params) of the last diff_forward to grads."""
        g_pos, g_base, g_head = grads
        _after_caller(d_rgb, g_pos, g_base, g_head)
        check(load().fvc_diff_backward(self.h, _addr(d_rgb), _addr(g_pos), _addr(g_base), _addr(g_head)))

    def diff_cached_forward(self, rm, pix, params, rgb):
        """diff_forward against the render cache (resample / cache samples)."""
        st = abi.DiffStats()
        pos, base, head = params
        _after_caller(pix, pos, base, head, rgb)
        check(load().fvc_diff_cached_forward(self.h, C.byref(rm), _addr(pix), int(pix.numel()), _addr(pos),
                                             _addr(base), _addr(head), _addr(rgb), C.byref(st)))
        return st

    def diff_cached_backward(self, d_rgb, grads, g_pool):
        """diff_backward of the last cached forward; g_pool [bricks, 512, 10]
        receives the voxel latent / density gradients."""
        g_pos, g_base, g_head = grads
        _after_caller(d_rgb, g_pos, g_base, g_head, g_pool)
        check(load().fvc_diff_cached_backward(self.h, _addr(d_rgb), _addr(g_pos), _addr(g_base), _addr(g_head),
                                              _addr(g_pool)))

    def debug_march(self, rm, max_per_ray=64):
        import numpy as np
        n = rm.cam_info.res_x * rm.cam_info.res_y
        counts = np.zeros(n, np.int32)
        t0s = np.zeros((n, max_per_ray), np.float32)
        check(load().fvc_debug_march(self.h, C.byref(rm), max_per_ray, _addr(counts), _addr(t0s)))
        return counts, t0s


# reference class names
PerformanceRenderer = Renderer
CachePerformanceRenderer = Renderer
