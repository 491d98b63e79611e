// capi.hip -- extern "C" surface of libfvc.so (include/fvc.h).  Every entry
// point catches the C++ exceptions of the implementation and turns them into
// a status code + thread-local message; nothing throws across the ABI.
#include <cmath>
#include <cstring>
#include <string>
#include <thread>

#include "fvc_internal.h"
#include "march.h"

namespace fvc {
static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }
void pack_fragments(const uint16_t* base, const uint16_t* head, int precision, int lat_lo, int n_lat,
                    std::vector<uint8_t>& out, MlpScales& sc);
}  // namespace fvc

using namespace fvc;

struct fvc_model { Model m; };
struct fvc_renderer { Renderer r; };

template <class F> static int guard(F&& f) {
  try {
    f();
    return FVC_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return FVC_E_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return FVC_E_INVALID;
  }
}

#define REQUIRE(cond, msg) \
  do {                     \
    if (!(cond)) throw Error(FVC_E_INVALID, msg); \
  } while (0)

extern "C" {

const char* fvc_last_error_message(void) { return g_last_error.c_str(); }
int fvc_version(void) { return 1; }

// StepsizeInfo::setAdditionalInfo (common.h:109-128)
int fvc_stepsize_setup(fvc_stepsize_info* s, const fvc_scene_info* sc) {
  return guard([&] {
    REQUIRE(s && sc, "null argument");
    s->dt_min = s->stepsize;
    const float res = (float)sc->grid_resolution;
    s->dt_max = std::fmax(std::fmax((sc->aabb_to.x - sc->aabb_from.x) / res * 1.732f,
                                    (sc->aabb_to.y - sc->aabb_from.y) / res * 1.732f),
                          (sc->aabb_to.z - sc->aabb_from.z) / res * 1.732f);
    s->t_min = (float)(((double)s->dt_min) / s->cone_angle);
    const double d_i_min = 1.0 / s->cone_angle;
    s->i_min = (float)d_i_min;
    const double d_log1p_c = std::log(1.0 + s->cone_angle);
    s->log1p_c = (float)d_log1p_c;
    const double d_i_max = std::log(s->dt_max / (double)s->dt_min) / d_log1p_c + d_i_min;
    s->i_max = (float)d_i_max;
    s->t_max = (float)(s->t_min * std::pow(1.0 + s->cone_angle, d_i_max - d_i_min));
    s->near_i = step_from_t(*s, s->near_);
  });
}

int fvc_net_param_counts(const fvc_net_desc* net, int64_t* a, int64_t* b, int64_t* c) {
  return guard([&] {
    REQUIRE(net && a && b && c, "null argument");
    GridParams g;
    if (!net_layout(*net, g, *a, *b, *c)) throw Error(FVC_E_UNSUPPORTED, "unsupported network description");
  });
}

int fvc_net_param_counts_viewdep(const fvc_net_desc* net, int64_t* a, int64_t* b, int64_t* f, int64_t* c) {
  return guard([&] {
    REQUIRE(net && a && b && f && c, "null argument");
    REQUIRE(net->split_head, "not a split model: use fvc_net_param_counts");
    GridParams g;
    if (!net_layout(*net, g, *a, *b, *c, f)) throw Error(FVC_E_UNSUPPORTED, "unsupported network description");
  });
}

// NerfNetwork::init + set_params (nerf_network.h:88-118, 269-308) and the
// occupancy bitfield; p_first (split models only) is mlp_first_head
// 64-bit hash of a byte range (8-byte words mixed with the splitmix64
// finaliser, then the tail bytes), chained through h
static uint64_t hash_bytes(const void* p, size_t n, uint64_t h) {
  auto mix = [](uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  const uint8_t* b = static_cast<const uint8_t*>(p);
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    std::memcpy(&w, b + i, 8);
    h = mix(h ^ w) + 0x9e3779b97f4a7c15ull;
  }
  for (; i < n; i++) h = mix(h ^ b[i]);
  return mix(h ^ n);
}

static void create_model(const fvc_scene_info* scene, const fvc_net_desc* net, const uint8_t* occ,
                         const uint16_t* p_pos, int64_t n_pos, const uint16_t* p_base, int64_t n_base,
                         const uint16_t* p_first, int64_t n_first, const uint16_t* p_head, int64_t n_head,
                         fvc_model** out) {
  {
    REQUIRE(scene && net && occ && p_pos && p_base && p_head && out, "null argument");
    *out = nullptr;
    if (net->split_head && !p_first) throw Error(FVC_E_INVALID, "split models need mlp_first_head: use fvc_model_create_viewdep");
    if (!net->split_head && p_first) throw Error(FVC_E_INVALID, "mlp_first_head given for a model without split head");
    // PerformanceRenderer only supports the fixed 128^3 grid (performance_renderer.cu:162-163)
    if (scene->grid_resolution != 128) throw Error(FVC_E_UNSUPPORTED, "only grid_resolution 128 is supported");
    REQUIRE(scene->grid_nlvl >= 1 && scene->grid_nlvl <= 16, "grid_nlvl out of range");
    if (net->precision != FVC_PREC_FP16 && net->precision != FVC_PREC_BF16 && net->precision != FVC_PREC_FP8 &&
        net->precision != FVC_PREC_BF16X3)
      throw Error(FVC_E_UNSUPPORTED, "unknown MLP precision");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) throw Error(FVC_E_NO_DEVICE, "no HIP device");
    auto* h = new fvc_model();
    Model& m = h->m;
    try {
      FVC_HIP(hipGetDevice(&m.device));
      m.scene = *scene;
      m.net = *net;
      if (!net_layout(*net, m.grid, m.n_pos_enc, m.n_base, m.n_head, &m.n_first))
        throw Error(FVC_E_UNSUPPORTED, "unsupported network description");
      if (n_pos != m.n_pos_enc || n_base != m.n_base || n_head != m.n_head || (p_first && n_first != m.n_first))
        throw Error(FVC_E_INVALID, "parameter counts do not match the network description");
      // Morton bitfield (occupancy_grid_to_morton_bitfield, loader.cu:20-55, 120-144)
      const int64_t vals = 128 * 128 * 128;
      m.bitfield_host.assign((size_t)(vals * scene->grid_nlvl / 8), 0);
      for (int l = 0; l < scene->grid_nlvl; l++)
        for (int x = 0; x < 128; x++)
          for (int y = 0; y < 128; y++)
            for (int z = 0; z < 128; z++) {
              if (!occ[l * vals + ((int64_t)x * 128 + y) * 128 + z]) continue;
              const int64_t bit = l * vals + morton3d(x, y, z);
              m.bitfield_host[bit >> 3] |= (uint8_t)(1u << (bit & 7));
            }
      m.bitfield.ensure(m.bitfield_host.size());
      FVC_HIP(hipMemcpy(m.bitfield.p, m.bitfield_host.data(), m.bitfield_host.size(), hipMemcpyHostToDevice));
      // coarse empty-space table of the marchers
      const std::vector<uint32_t> coarse = build_coarse(occ, scene->grid_nlvl);
      m.coarse.ensure(coarse.size() * sizeof(uint32_t));
      FVC_HIP(hipMemcpy(m.coarse.p, coarse.data(), coarse.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      m.table.ensure(n_pos * sizeof(uint16_t));
      FVC_HIP(hipMemcpy(m.table.p, p_pos, n_pos * sizeof(uint16_t), hipMemcpyHostToDevice));
      std::vector<uint8_t> frags;
      if (net->split_head) pack_fragments_vd(p_base, p_first, p_head, *net, frags);
      else pack_fragments(p_base, p_head, net->precision, net->latent_offset, net->n_latents, frags, m.mlp_scales);
      m.frags.ensure(frags.size());
      FVC_HIP(hipMemcpy(m.frags.p, frags.data(), frags.size(), hipMemcpyHostToDevice));
      uint64_t hsh = hash_bytes(net, sizeof(*net), 0x9e3779b97f4a7c15ull);
      hsh = hash_bytes(p_pos, n_pos * sizeof(uint16_t), hsh);
      hsh = hash_bytes(frags.data(), frags.size(), hsh);
      m.net_hash = hsh;
      m.occ = occ_params(*scene);
      const V3 from = cvt(scene->aabb_from), to = cvt(scene->aabb_to);
      // apply_contraction constants (raymarch_common.h:610-653)
      m.con.type = scene->contraction_type;
      if (scene->contraction_type == FVC_CONTRACTION_AABB) {
        m.con.from = from;
        m.con.ext = to - from;
      } else {
        m.con.from = cvt(scene->contraction_aabb_from);
        m.con.ext = cvt(scene->contraction_aabb_to) - m.con.from;
      }
      m.con.inv = V3{1.0f / m.con.ext.x, 1.0f / m.con.ext.y, 1.0f / m.con.ext.z};
      m.con.pow2 = is_pow2f(m.con.ext.x) && is_pow2f(m.con.ext.y) && is_pow2f(m.con.ext.z);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  }
}

int fvc_model_create(const fvc_scene_info* scene, const fvc_net_desc* net, const uint8_t* occ,
                     const uint16_t* p_pos, int64_t n_pos, const uint16_t* p_base, int64_t n_base,
                     const uint16_t* p_head, int64_t n_head, fvc_model** out) {
  return guard([&] { create_model(scene, net, occ, p_pos, n_pos, p_base, n_base, nullptr, 0, p_head, n_head, out); });
}

int fvc_model_create_viewdep(const fvc_scene_info* scene, const fvc_net_desc* net, const uint8_t* occ,
                             const uint16_t* p_pos, int64_t n_pos, const uint16_t* p_base, int64_t n_base,
                             const uint16_t* p_first, int64_t n_first, const uint16_t* p_head, int64_t n_head,
                             fvc_model** out) {
  return guard([&] {
    REQUIRE(p_first, "null argument");
    create_model(scene, net, occ, p_pos, n_pos, p_base, n_base, p_first, n_first, p_head, n_head, out);
  });
}

void fvc_model_destroy(fvc_model* m) { delete m; }

int fvc_nerfacc_read_config(const char* dir, fvc_scene_info* scene, fvc_stepsize_info* stepsize, fvc_net_desc* net,
                            int64_t* counts) {
  return guard([&] {
    REQUIRE(dir && scene && stepsize && net && counts, "null argument");
    NerfaccExport ex;
    nerfacc_read_config(dir, ex);
    *scene = ex.scene;
    *stepsize = ex.step;
    *net = ex.net;
    counts[0] = ex.n_pos;
    counts[1] = ex.n_base;
    counts[2] = ex.n_head;
    counts[3] = ex.n_occ;
  });
}

int fvc_nerfacc_read_data(const char* dir, uint16_t* pos, uint16_t* base, uint16_t* head, uint8_t* occ) {
  return guard([&] {
    REQUIRE(dir && pos && base && head && occ, "null argument");
    NerfaccExport ex;
    nerfacc_read_config(dir, ex);
    if (ex.net.split_head) throw Error(FVC_E_INVALID, "split model: use fvc_nerfacc_read_data_viewdep");
    nerfacc_read_data(dir, ex, pos, base, head, occ);
  });
}

int fvc_nerfacc_read_data_viewdep(const char* dir, uint16_t* pos, uint16_t* base, uint16_t* first, uint16_t* head,
                                  uint8_t* occ) {
  return guard([&] {
    REQUIRE(dir && pos && base && first && head && occ, "null argument");
    NerfaccExport ex;
    nerfacc_read_config(dir, ex);
    if (!ex.net.split_head) throw Error(FVC_E_INVALID, "config.json has no mlp_first_head: use fvc_nerfacc_read_data");
    nerfacc_read_data(dir, ex, pos, base, head, occ, first);
  });
}

int fvc_model_load_nerfacc(const char* dir, fvc_scene_info* scene, fvc_stepsize_info* stepsize, fvc_net_desc* net,
                           fvc_model** out) {
  return guard([&] {
    REQUIRE(dir && out, "null argument");
    *out = nullptr;
    NerfaccExport ex;
    nerfacc_read_config(dir, ex);
    std::vector<uint16_t> pos((size_t)ex.n_pos), base((size_t)ex.n_base), head((size_t)ex.n_head);
    std::vector<uint16_t> first((size_t)std::max<int64_t>(1, ex.n_first));
    std::vector<uint8_t> occ((size_t)ex.n_occ);
    nerfacc_read_data(dir, ex, pos.data(), base.data(), head.data(), occ.data(), first.data());
    const int rc = ex.net.split_head
                       ? fvc_model_create_viewdep(&ex.scene, &ex.net, occ.data(), pos.data(), ex.n_pos, base.data(),
                                                  ex.n_base, first.data(), ex.n_first, head.data(), ex.n_head, out)
                       : fvc_model_create(&ex.scene, &ex.net, occ.data(), pos.data(), ex.n_pos, base.data(),
                                          ex.n_base, head.data(), ex.n_head, out);
    if (rc != FVC_OK) throw Error(rc, fvc_last_error_message());
    if (scene) *scene = ex.scene;
    if (stepsize) *stepsize = ex.step;
    if (net) *net = ex.net;
  });
}

int fvc_model_get_bitfield(fvc_model* h, uint8_t* out, int64_t n) {
  return guard([&] {
    REQUIRE(h && out, "null argument");
    REQUIRE(n == (int64_t)h->m.bitfield_host.size(), "size mismatch");
    std::memcpy(out, h->m.bitfield_host.data(), (size_t)n);
  });
}

int fvc_network_latents(fvc_model* h, const float* pos, int64_t n, uint16_t* out, void* stream) {
  return guard([&] {
    REQUIRE(h && pos && out && n >= 0, "invalid argument");
    launch_network_latents(h->m, pos, n, out, (hipStream_t)stream);
  });
}
int fvc_network_inference(fvc_model* h, const float* pos, const float* dir, int64_t n, uint16_t* out, void* stream) {
  return guard([&] {
    REQUIRE(h && pos && dir && out && n >= 0, "invalid argument");
    launch_network_full(h->m, pos, dir, n, (uint32_t*)out, (hipStream_t)stream);
  });
}
int fvc_network_head(fvc_model* h, const uint16_t* hin, const float* dir, int64_t n, uint16_t* out, void* stream) {
  return guard([&] {
    REQUIRE(h && hin && dir && out && n >= 0, "invalid argument");
    launch_network_head(h->m, hin, dir, n, (uint32_t*)out, (hipStream_t)stream);
  });
}
int fvc_network_latents_viewdep(fvc_model* h, const float* pos, const float* init_dir, int64_t n, uint16_t* out,
                                void* stream) {
  return guard([&] {
    REQUIRE(h && pos && init_dir && out && n >= 0, "invalid argument");
    if (!h->m.net.split_head) throw Error(FVC_E_UNSUPPORTED, "not a split model: use fvc_network_latents");
    launch_network_latents_vd(h->m, pos, init_dir, n, out, (hipStream_t)stream, 16);
  });
}
int fvc_network_head_viewdep(fvc_model* h, const uint16_t* hin, const float* pos, const float* dir,
                             const float* init_dir, int64_t n, uint16_t* out, void* stream) {
  return guard([&] {
    REQUIRE(h && hin && pos && dir && init_dir && out && n >= 0, "invalid argument");
    if (!h->m.net.split_head) throw Error(FVC_E_UNSUPPORTED, "not a split model: use fvc_network_head");
    launch_network_head_vd(h->m, hin, pos, dir, init_dir, n, (uint32_t*)out, (hipStream_t)stream);
  });
}
int fvc_network_encode(fvc_model* h, const float* pos, int64_t n, uint16_t* out, void* stream) {
  return guard([&] {
    REQUIRE(h && pos && out && n >= 0, "invalid argument");
    launch_network_encode(h->m, pos, n, out, (hipStream_t)stream);
  });
}

int fvc_renderer_create(fvc_model* model, int double_buffered, fvc_renderer** out) {
  return guard([&] {
    REQUIRE(model && out, "null argument");
    *out = nullptr;
    auto* h = new fvc_renderer();
    h->r.model = &model->m;
    // the interpolation variant eval.cu:216 picks for the model
    if (model->m.net.first_latent_is_density)
      h->r.settings.interpol_variant = FVC_INTERPOL_DENSITY_INTERMEDIATES;
    h->r.double_buffered = double_buffered != 0;
    h->r.fill_idx = h->r.double_buffered ? 1 : 0;
    // default priority on purpose: non-default stream priorities are avoided
    hipError_t e = hipStreamCreateWithFlags(&h->r.own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete h;
      throw Error(FVC_E_HIP, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
    }
    *out = h;
  });
}
void fvc_renderer_destroy(fvc_renderer* r) { delete r; }

int fvc_renderer_set_stream(fvc_renderer* h, void* stream) {
  return guard([&] {
    REQUIRE(h, "null argument");
    std::scoped_lock lk(h->r.read_mutex);
    h->r.user_stream = (hipStream_t)stream;
  });
}

int fvc_renderer_set_batching(fvc_renderer* h, int64_t target_samples, int32_t max_samples_per_ray) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    REQUIRE(target_samples >= 4096 && target_samples <= (int64_t)1 << 28, "target_samples out of range");
    REQUIRE(max_samples_per_ray >= 1 && max_samples_per_ray <= MAX_STEPS_PER_LAUNCH,
            "max_samples_per_ray must be in [1, 64]");
    // the render buffers may grow: not while a render uses them
    std::scoped_lock lk(h->r.read_mutex);
    h->r.target_batch = target_samples;
    h->r.max_steps = max_samples_per_ray;
    if (h->r.width > 0) resize_renderbuffers(h->r, h->r.width, h->r.height);
  });
}

static void check_rm(const fvc_raymarch_info* rm);

int fvc_renderer_autotune(fvc_renderer* h, const fvc_raymarch_info* rm, int cached, fvc_tune_result* result) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    check_rm(rm);
    std::scoped_lock lk(h->r.read_mutex);
    autotune(h->r, *rm, cached != 0, result);
  });
}

int fvc_renderer_autotune_batch(fvc_renderer* h, fvc_model* const* models, const fvc_raymarch_info* rms,
                                int32_t n_views, int cached, const fvc_tune_space* space, fvc_tune_result* result) {
  return guard([&] {
    REQUIRE(h && rms && n_views > 0, "invalid argument");
    for (int v = 0; v < n_views; v++) check_rm(&rms[v]);
    std::vector<const Model*> ms((size_t)n_views, nullptr);
    for (int v = 0; models && v < n_views; v++) ms[v] = models[v] ? &models[v]->m : nullptr;
    std::scoped_lock lk(h->r.read_mutex);
    autotune(h->r, Views(rms, n_views, nullptr, h->r.model, ms.data()), cached != 0, space, result);
  });
}

int fvc_renderer_set_graph(fvc_renderer* h, int enable) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    std::scoped_lock lk(h->r.read_mutex);
    h->r.use_graph = enable != 0;
  });
}

int fvc_renderer_set_mlp_rows(fvc_renderer* h, int32_t rows) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    REQUIRE(rows == 16 || rows == 32 || rows == 64, "rows must be 16, 32 or 64");
    std::scoped_lock lk(h->r.read_mutex);
    h->r.mlp_rows = rows;
  });
}

int fvc_resize_renderbuffers(fvc_renderer* h, int w, int hh) {
  return guard([&] {
    REQUIRE(h, "null argument");
    std::scoped_lock lk(h->r.read_mutex);
    resize_renderbuffers(h->r, w, hh);
  });
}

static void check_rm(const fvc_raymarch_info* rm) {
  REQUIRE(rm, "null raymarch info");
  REQUIRE(rm->cam_info.res_x > 0 && rm->cam_info.res_y > 0, "resolution must be positive");
  if (!rm->deterministic || rm->motion_blur)
    throw Error(FVC_E_UNSUPPORTED, "only deterministic rendering without motion blur is supported");
}

int fvc_render_full(fvc_renderer* h, const fvc_raymarch_info* rm, uint8_t* rgba, fvc_render_stats* stats) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    check_rm(rm);
    // the full render shares the ray and sample buffers of cached renders
    std::scoped_lock lk(h->r.read_mutex);
    render_full(h->r, *rm, rgba, stats);
  });
}

int fvc_render_full_noalloc(fvc_renderer* h, const fvc_raymarch_info* rm, uint8_t* rgba, float* result,
                            fvc_render_stats* stats) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    check_rm(rm);
    std::scoped_lock lk(h->r.read_mutex);
    render_full_noalloc(h->r, *rm, rgba, result, stats);
  });
}

int fvc_cache_fill(fvc_renderer* h, const fvc_raymarch_info* rm, fvc_init_stats* stats) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    check_rm(rm);
    std::scoped_lock lk(h->r.write_mutex);
    cache_fill(h->r, *rm, stats);
  });
}

int fvc_render_cached(fvc_renderer* h, const fvc_raymarch_info* rm, uint8_t* rgba, fvc_render_stats* stats) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    check_rm(rm);
    std::scoped_lock lk(h->r.read_mutex);
    render_cached(h->r, *rm, rgba, stats);
  });
}

int fvc_render_full_batch(fvc_renderer* h, const fvc_raymarch_info* rms, int32_t n_views, uint8_t* const* rgba,
                          fvc_render_stats* stats) {
  return guard([&] {
    REQUIRE(h && rms && n_views > 0, "invalid argument");
    for (int v = 0; v < n_views; v++) check_rm(&rms[v]);
    std::scoped_lock lk(h->r.read_mutex);
    render_full(h->r, Views(rms, n_views, rgba, h->r.model), stats);
  });
}

int fvc_render_full_batch_scenes(fvc_renderer* h, fvc_model* const* models, const fvc_raymarch_info* rms,
                                 int32_t n_views, uint8_t* const* rgba, fvc_render_stats* stats) {
  return guard([&] {
    REQUIRE(h && models && rms && n_views > 0, "invalid argument");
    for (int v = 0; v < n_views; v++) check_rm(&rms[v]);
    std::vector<const Model*> ms((size_t)n_views);
    for (int v = 0; v < n_views; v++) ms[v] = models[v] ? &models[v]->m : nullptr;
    std::scoped_lock lk(h->r.read_mutex);
    render_full(h->r, Views(rms, n_views, rgba, h->r.model, ms.data()), stats);
  });
}

int fvc_render_cached_batch(fvc_renderer* h, const fvc_raymarch_info* rms, int32_t n_views, uint8_t* const* rgba,
                            fvc_render_stats* stats) {
  return guard([&] {
    REQUIRE(h && rms && n_views > 0, "invalid argument");
    for (int v = 0; v < n_views; v++) check_rm(&rms[v]);
    std::scoped_lock lk(h->r.read_mutex);
    render_cached(h->r, Views(rms, n_views, rgba, h->r.model), stats);
  });
}

// The differentiable render covers the separate-density layout (log sigma
// in row 0, 8 latents after it, FP16 model parameters) with the default
// cache lookup (linear, inter-brick, both sample classes).
static void check_diff_model(const Renderer& r, bool cached) {
  const fvc_net_desc& n = r.model->net;
  if (n.split_head || n.latent_offset != 1)
    throw Error(FVC_E_UNSUPPORTED,
                "the differentiable render covers models with a separate density row (latent_offset 1), no split head");
  if (cached && (r.settings.interpol_function != FVC_FN_LINEAR || !r.settings.use_inter_brick_interpolation ||
                 r.render_mode != FVC_RENDER_BOTH))
    throw Error(FVC_E_UNSUPPORTED, "the differentiable cached render uses the default cache settings and render mode");
}

int fvc_diff_forward(fvc_renderer* h, const fvc_raymarch_info* rm, const int32_t* pix, int32_t n_rays,
                     const float* pos_enc, const float* mlp_base, const float* mlp_head, float* rgb,
                     fvc_diff_stats* stats) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    check_diff_model(h->r, false);
    check_rm(rm);
    REQUIRE(n_rays >= 0, "negative ray count");
    REQUIRE(n_rays == 0 || (pix && pos_enc && mlp_base && mlp_head && rgb), "null argument");
    std::scoped_lock lk(h->r.read_mutex);
    diff_forward(h->r, *rm, pix, n_rays, pos_enc, mlp_base, mlp_head, rgb, stats, false);
  });
}

int fvc_diff_backward(fvc_renderer* h, const float* d_rgb, float* g_pos_enc, float* g_mlp_base, float* g_mlp_head) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    REQUIRE(!h->r.diff.cached, "the last forward was a cached one: use fvc_diff_cached_backward");
    REQUIRE(h->r.diff.n == 0 || (d_rgb && g_pos_enc && g_mlp_base && g_mlp_head), "null argument");
    std::scoped_lock lk(h->r.read_mutex);
    diff_backward(h->r, d_rgb, g_pos_enc, g_mlp_base, g_mlp_head, nullptr);
  });
}

int fvc_diff_cached_forward(fvc_renderer* h, const fvc_raymarch_info* rm, const int32_t* pix, int32_t n_rays,
                            const float* pos_enc, const float* mlp_base, const float* mlp_head, float* rgb,
                            fvc_diff_stats* stats) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    check_diff_model(h->r, true);
    check_rm(rm);
    REQUIRE(n_rays >= 0, "negative ray count");
    REQUIRE(n_rays == 0 || (pix && pos_enc && mlp_base && mlp_head && rgb), "null argument");
    std::scoped_lock lk(h->r.read_mutex);
    diff_forward(h->r, *rm, pix, n_rays, pos_enc, mlp_base, mlp_head, rgb, stats, true);
  });
}

int fvc_diff_cached_backward(fvc_renderer* h, const float* d_rgb, float* g_pos_enc, float* g_mlp_base,
                             float* g_mlp_head, float* g_pool) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    REQUIRE(h->r.diff.cached, "the last forward was not a cached one: use fvc_diff_backward");
    REQUIRE(h->r.diff.n == 0 || (d_rgb && g_pos_enc && g_mlp_base && g_mlp_head && g_pool), "null argument");
    std::scoped_lock lk(h->r.read_mutex);
    diff_backward(h->r, d_rgb, g_pos_enc, g_mlp_base, g_mlp_head, g_pool);
  });
}

int fvc_cache_swap(fvc_renderer* h) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    // a background fill (running or finished) is joined and published: that
    // swap is the one swapCache asks for (a second swap here would bring the
    // stale cache back to the front)
    if (h->r.fill_state.load() == FILL_RUNNING) fill_poll(h->r, true, nullptr);
    if (fill_publish(h->r)) return;
    std::scoped_lock lk(h->r.read_mutex, h->r.write_mutex);
    if (h->r.double_buffered) std::swap(h->r.render_idx, h->r.fill_idx);
  });
}

int fvc_cache_fill_async(fvc_renderer* h, const fvc_raymarch_info* rm) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    check_rm(rm);
    fill_async(h->r, *rm);
  });
}

int fvc_cache_fill_poll(fvc_renderer* h, int wait, int32_t* state, fvc_init_stats* stats) {
  return guard([&] {
    REQUIRE(h && state, "null argument");
    *state = fill_poll(h->r, wait != 0, stats);
  });
}

int fvc_cache_publish(fvc_renderer* h, int32_t* swapped) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    const bool sw = fill_publish(h->r);
    if (swapped) *swapped = sw ? 1 : 0;
  });
}

int fvc_renderer_set_refill_policy(fvc_renderer* h, const fvc_refill_policy* p) {
  return guard([&] {
    REQUIRE(h && p, "null argument");
    REQUIRE(p->min_hit_ratio >= 0.0f && p->min_hit_ratio <= 1.0f && p->max_distance > 0.0f, "invalid policy");
    h->r.policy = *p;
  });
}

int fvc_render_cached_auto(fvc_renderer* h, const fvc_raymarch_info* rm, uint8_t* rgba, fvc_render_stats* stats,
                           int32_t* refill_started) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    check_rm(rm);
    int started = 0;
    render_cached_auto(h->r, *rm, rgba, stats, &started);
    if (refill_started) *refill_started = started;
  });
}

int fvc_cache_set_settings(fvc_renderer* h, const fvc_cache_settings* s) {
  return guard([&] {
    REQUIRE(h && s, "null argument");
    const fvc_net_desc& net = h->r.model->net;
    // the variant must match what the model's latents hold (eval.cu:216)
    const int variant = net.first_latent_is_density ? FVC_INTERPOL_DENSITY_INTERMEDIATES : FVC_INTERPOL_DENSITY;
    if (s->interpol_variant != variant)
      throw Error(FVC_E_UNSUPPORTED, variant == FVC_INTERPOL_DENSITY
                                         ? "this model's cache needs the Density interpolation variant"
                                         : "this model's cache needs the DensityIntermediates interpolation variant");
    REQUIRE(s->interpol_function >= FVC_FN_LINEAR && s->interpol_function <= FVC_FN_NEAREST,
            "unknown interpolation function");
    std::scoped_lock lk(h->r.read_mutex);
    h->r.settings = *s;
    h->r.settings.use_inter_brick_interpolation = s->use_inter_brick_interpolation != 0;
  });
}

int fvc_cache_set_render_mode(fvc_renderer* h, int mode) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    REQUIRE(mode == FVC_RENDER_BOTH || mode == FVC_RENDER_ONLY_CACHE || mode == FVC_RENDER_ONLY_RESAMPLED,
            "unknown render mode");
    std::scoped_lock lk(h->r.read_mutex);
    h->r.render_mode = mode;
  });
}

int fvc_debug_cache_lookup(fvc_renderer* h, int which, const float* froxel_host, int64_t n, float* out_host) {
  return guard([&] {
    REQUIRE(h && (n == 0 || (froxel_host && out_host)) && n >= 0, "invalid argument");
    REQUIRE(which == 0 || which == 1, "which must be 0 or 1");
    std::scoped_lock lk(h->r.read_mutex);
    debug_lookup(h->r, which, froxel_host, n, out_host);
  });
}

int fvc_debug_render_cached(fvc_renderer* h, const fvc_raymarch_info* rm, int32_t max_codes, int8_t* codes_host,
                            int32_t* classified_host, int32_t* composited_host, fvc_render_stats* stats) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    check_rm(rm);
    REQUIRE(max_codes >= 0 && (max_codes == 0 || codes_host), "invalid argument");
    std::scoped_lock lk(h->r.read_mutex);
    debug_render_cached(h->r, *rm, max_codes, codes_host, classified_host, composited_host, stats);
  });
}

static CacheStore& pick(fvc_renderer* h, int which);
static std::unique_lock<std::mutex> cache_lock(fvc_renderer* h, int which);

int fvc_cache_set_latent_format(fvc_renderer* h, int format) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    REQUIRE(format == FVC_LATENT_FP16 || format == FVC_LATENT_FP8, "unknown latent format");
    if (format == FVC_LATENT_FP8 && h->r.model->net.split_head)
      throw Error(FVC_E_UNSUPPORTED, "FP8 cache latents are not built for split models");
    std::scoped_lock lk(h->r.write_mutex);
    h->r.latent_format = format;
  });
}

// OCP E4M3 code -> float (bias 7, no infinities, 0x7f / 0xff NaN)
static float e4m3_value(uint8_t c) {
  const int sgn = c >> 7, e = (c >> 3) & 15, m = c & 7;
  float v;
  if (e == 15 && m == 7) v = NAN;
  else if (e == 0) v = std::ldexp((float)m, -9);
  else v = std::ldexp((float)(8 + m), e - 10);
  return sgn ? -v : v;
}

int fvc_cache_get_pool_decoded(fvc_renderer* h, int which, float* out, int64_t n) {
  return guard([&] {
    REQUIRE(h && out, "null argument");
    const auto lk = cache_lock(h, which);
    const CacheStore& C = pick(h, which);
    if (!C.initialized) throw Error(FVC_E_NOT_INITIALIZED, "Cache is not initialized!");
    const int64_t nv = (int64_t)C.n_used * BRICK_VOXELS;
    REQUIRE(n >= nv * 10, "buffer too small");
    // voxel i of the export (reference order) is record src(i) of the pool
    auto src = [](int64_t i) {
      const int64_t b = i / BRICK_VOXELS, c = i % BRICK_VOXELS;
      return b * BRICK_VOXELS + brick_voxel((int)(c & 7), (int)((c >> 3) & 7), (int)(c >> 6));
    };
    if (C.lat8) {
      std::vector<Voxel8> v((size_t)nv);
      std::vector<float> sc((size_t)std::max(C.n_used, 1));
      if (nv) FVC_HIP(hipMemcpy(v.data(), C.pool.p, nv * sizeof(Voxel8), hipMemcpyDeviceToHost));
      if (C.n_used) FVC_HIP(hipMemcpy(sc.data(), C.scales.p, C.n_used * sizeof(float), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < nv; i++) {
        const Voxel8& x = v[src(i)];
        for (int k = 0; k < 8; k++)
          out[i * 10 + k] = e4m3_value((uint8_t)(x.lat[k >> 2] >> (8 * (k & 3)))) * sc[i / BRICK_VOXELS];
        out[i * 10 + 8] = x.density;
        out[i * 10 + 9] = x.isset;
      }
    } else {
      std::vector<Voxel> v((size_t)nv);
      if (nv) FVC_HIP(hipMemcpy(v.data(), C.pool.p, nv * sizeof(Voxel), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < nv; i++) {
        const Voxel& x = v[src(i)];
        for (int k = 0; k < 8; k++) {
          _Float16 hv;
          std::memcpy(&hv, &x.lat[k], 2);
          out[i * 10 + k] = (float)hv;
        }
        out[i * 10 + 8] = x.density;
        out[i * 10 + 9] = x.isset;
      }
    }
  });
}

int fvc_cache_set_capacity(fvc_renderer* h, int32_t cap) {
  return guard([&] {
    REQUIRE(h && cap >= 0, "invalid argument");
    REQUIRE(cap <= MAX_BRICKS, "capacity above 2^23 - 1 bricks");
    std::scoped_lock lk(h->r.write_mutex);
    h->r.capacity_override = cap;
  });
}

int fvc_get_last_result(fvc_renderer* h, float* out, int64_t n) {
  return guard([&] {
    REQUIRE(h && out, "null argument");
    std::scoped_lock lk(h->r.read_mutex);
    const int64_t need = h->r.last_n_px * 4;
    REQUIRE(n >= need && need > 0, "buffer too small or nothing rendered");
    FVC_HIP(hipMemcpyAsync(out, h->r.results.p, need * sizeof(float), hipMemcpyDeviceToDevice, h->r.stream()));
    FVC_HIP(hipStreamSynchronize(h->r.stream()));
  });
}

static CacheStore& pick(fvc_renderer* h, int which) {
  REQUIRE(which == 0 || which == 1, "which must be 0 or 1");
  return h->r.caches[which == 0 ? h->r.render_idx : h->r.fill_idx];
}

// The lock that keeps cache `which` still while it is read: renders hold
// read_mutex and never change the front cache; fills write the fill cache
// under write_mutex (a background fill on its own thread); a publish swaps
// the two under both.  A background fill is marked running before its
// thread takes write_mutex, so a reader of the fill cache that gets the lock
// first lets it go and waits for the fill: the contents it reads are the
// finished fill's, as for a reader that comes later.
static std::unique_lock<std::mutex> cache_lock(fvc_renderer* h, int which) {
  REQUIRE(which == 0 || which == 1, "which must be 0 or 1");
  if (which == 0) return std::unique_lock<std::mutex>(h->r.read_mutex);
  for (;;) {
    std::unique_lock<std::mutex> lk(h->r.write_mutex);
    if (h->r.fill_state.load() != FILL_RUNNING) return lk;
    lk.unlock();
    std::this_thread::yield();
  }
}

int fvc_cache_get_info(fvc_renderer* h, int which, fvc_cache_info* info) {
  return guard([&] {
    REQUIRE(h && info, "null argument");
    const auto lk = cache_lock(h, which);
    const CacheStore& C = pick(h, which);
    info->frustum_x = C.fx; info->frustum_y = C.fy; info->frustum_z = C.fz;
    info->bricks_x = C.bx; info->bricks_y = C.by; info->bricks_z = C.bz;
    info->capacity_bricks = C.capacity;
    info->n_bricks_used = C.n_used;
    info->initialized = C.initialized;
    info->brick_size = BRICK;
    info->latent_format = C.lat8 ? FVC_LATENT_FP8 : FVC_LATENT_FP16;
  });
}

int fvc_cache_get_brick_table(fvc_renderer* h, int which, int32_t* out, int64_t n) {
  return guard([&] {
    REQUIRE(h && out, "null argument");
    const auto lk = cache_lock(h, which);
    const CacheStore& C = pick(h, which);
    if (!C.initialized) throw Error(FVC_E_NOT_INITIALIZED, "Cache is not initialized!");
    const int64_t need = (int64_t)C.bx * C.by * C.bz;
    REQUIRE(n >= need, "buffer too small");
    FVC_HIP(hipMemcpy(out, C.table.p, need * sizeof(int32_t), hipMemcpyDeviceToHost));
  });
}

int fvc_cache_get_known_ranges(fvc_renderer* h, int which, float* out, int64_t n) {
  return guard([&] {
    REQUIRE(h && out, "null argument");
    const auto lk = cache_lock(h, which);
    const CacheStore& C = pick(h, which);
    if (!C.initialized) throw Error(FVC_E_NOT_INITIALIZED, "Cache is not initialized!");
    const int64_t need = (int64_t)C.fx * C.fy;
    REQUIRE(n >= need, "buffer too small");
    FVC_HIP(hipMemcpy(out, C.known.p, need * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int fvc_cache_get_pool(fvc_renderer* h, int which, void* out, int64_t n_bytes) {
  return guard([&] {
    REQUIRE(h && out, "null argument");
    const auto lk = cache_lock(h, which);
    const CacheStore& C = pick(h, which);
    if (!C.initialized) throw Error(FVC_E_NOT_INITIALIZED, "Cache is not initialized!");
    const size_t rec = C.lat8 ? sizeof(Voxel8) : sizeof(Voxel);
    const int64_t nv = (int64_t)C.n_used * BRICK_VOXELS;
    REQUIRE(n_bytes >= 0 && n_bytes <= nv * (int64_t)rec, "request exceeds the used part of the pool");
    // records in the reference's voxel order (x fastest inside a brick)
    std::vector<uint8_t> raw((size_t)(nv * rec));
    if (nv) FVC_HIP(hipMemcpy(raw.data(), C.pool.p, raw.size(), hipMemcpyDeviceToHost));
    uint8_t* o = static_cast<uint8_t*>(out);
    for (int64_t i = 0; i * (int64_t)rec < n_bytes; i++) {
      const int64_t b = i / BRICK_VOXELS, c = i % BRICK_VOXELS;
      const int64_t j = b * BRICK_VOXELS + brick_voxel((int)(c & 7), (int)((c >> 3) & 7), (int)(c >> 6));
      std::memcpy(o + i * rec, raw.data() + j * rec, (size_t)std::min<int64_t>((int64_t)rec, n_bytes - i * (int64_t)rec));
    }
  });
}

int fvc_profile_enable(fvc_renderer* h, int enable) {
  return guard([&] {
    REQUIRE(h, "null renderer");
    std::scoped_lock lk(h->r.read_mutex);
    h->r.prof.table.clear();
    h->r.prof.enabled = enable != 0;
  });
}

int fvc_profile_get(fvc_renderer* h, fvc_kernel_stats* out, int32_t max_entries, int32_t* n_entries) {
  return guard([&] {
    REQUIRE(h && n_entries && (out || max_entries == 0), "null argument");
    std::scoped_lock lk(h->r.read_mutex);
    int i = 0;
    for (const auto& kv : h->r.prof.table) {
      if (i < max_entries) {
        fvc_kernel_stats& o = out[i];
        std::memset(&o, 0, sizeof(o));
        std::strncpy(o.name, kv.first.c_str(), sizeof(o.name) - 1);
        o.launches = kv.second.launches;
        o.total_ms = kv.second.ms;
        o.units = kv.second.units;
        o.alg_bytes = kv.second.bytes;
        o.alg_flops = kv.second.flops;
      }
      i++;
    }
    *n_entries = i;
  });
}

int fvc_debug_march(fvc_renderer* h, const fvc_raymarch_info* rm, int32_t max_per_ray, int32_t* counts, float* t0s) {
  return guard([&] {
    REQUIRE(h && counts && max_per_ray >= 0, "invalid argument");
    check_rm(rm);
    std::scoped_lock lk(h->r.read_mutex);
    debug_march(h->r, *rm, max_per_ray, counts, t0s);
  });
}

}  // extern "C"
