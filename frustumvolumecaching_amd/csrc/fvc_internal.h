// fvc_internal.h -- host-side objects behind the C ABI and the kernel
// parameter blocks they hand to the device code.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <functional>
#include <vector>
#include <atomic>
#include <thread>

#include "fvc_math.h"

namespace fvc {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void set_last_error(const std::string& m);

#define FVC_HIP(expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      throw ::fvc::Error(e_ == hipErrorOutOfMemory ? FVC_E_OUT_OF_MEMORY : FVC_E_HIP,          \
                         std::string(#expr) + ": " + hipGetErrorString(e_));                   \
  } while (0)

// ---------------------------------------------------------------- memory
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  // grow-only allocation; contents are not preserved
  void ensure(size_t n) {
    if (n <= bytes) return;
    release();
    if (n == 0) return;
    FVC_HIP(hipMalloc(&p, n));
    bytes = n;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// ---------------------------------------------------------------- network
static const int MAX_LEVELS = 16;
// Hash-grid level table (tcnn GridEncoding conventions).
struct GridLevel {
  float scale;
  uint32_t res;
  uint32_t res2;     // res * res
  uint32_t size;     // entries in the level
  uint32_t offset;   // first entry of the level
  uint32_t first;    // the value of that entry (set by the kernels' level copy)
};
struct GridParams {
  GridLevel lv[MAX_LEVELS];
  uint32_t dense_mask;  // bit l: level l indexed densely
  uint32_t quad_dense;  // bit j: levels 4j..4j+3 all dense
  uint32_t quad_hash;   // bit j: levels 4j..4j+3 all hashed
  int n_levels;
};

// Weight fragments: 20 MFMA A-operands of 64 lanes x 8 halves, in the order
// W1[nb=0..3], W2[nb=0..3][kb=0..1], W3[kb=0..1], H1[nb=0..3], H2[kb=0..1].
static const int N_FRAGS = 20;
static const int FRAG_W1 = 0, FRAG_W2 = 4, FRAG_W3 = 12, FRAG_H1 = 14, FRAG_H2 = 18;
static const int N_NET_FRAGS = 20;
// Split (view-dependent) models, FP16 only: the density MLP as above, then
// mlp_first_head VD_F1[nb] (32 inputs), VD_F2[nb][kb] (second hidden layer,
// if any), VD_F3[kb] (output rows = the 8 head latents in order); the
// custom colour head's first layer over two 32-input blocks, VD_HA[nb]
// (latents / position frequencies + SH of the view direction) and
// VD_HB[nb] (SH differences + frequencies), and its output layer VD_HO[kb].
// The network kernels stage all of them, the head kernels VD_HA..VD_HO+1.
static const int VD_F1 = 14, VD_F2 = 18, VD_F3 = 26, VD_HA = 28, VD_HB = 32, VD_HO = 36;
static const int N_VD_FRAGS = 38, N_VD_NET_FRAGS = 38;

// per-layer accumulator scales of the MLPs (FP8: s_w[l] * s_a; otherwise 1)
// and 1 / s_a (mlp_mfma.h), and the density-MLP output rows [lat_lo, lat_hi)
// that the colour head reads as latents (the other rows of its first 16
// inputs are the constant 1.0 padding): rows 1..8 for models with a separate
// density row, rows 0..n_output-1 for the reference's plain models
struct MlpScales {
  float layer[5];   // W1, W2, W3, H1, H2
  float inv_act;
  int lat_lo, lat_hi;
};

struct Model {
  int device = 0;
  fvc_scene_info scene{};
  fvc_net_desc net{};
  GridParams grid{};
  OccParams occ{};
  ContractParams con{};
  int64_t n_pos_enc = 0, n_base = 0, n_head = 0, n_first = 0;
  DevBuf bitfield;   // Morton occupancy bitfield
  DevBuf coarse;     // per cascade 16^3 "any of 8^3 cells occupied" bits
  DevBuf table;      // hash grid, uint32 per entry (2 x fp16)
  DevBuf frags;      // packed weight fragments in the MLP precision
  MlpScales mlp_scales{};
  std::vector<uint8_t> bitfield_host;
  // hash of the network description and parameters: models with equal
  // hashes share the network (views of a mixed-scene batch must)
  uint64_t net_hash = 0;
};

// The occupancy of one scene of a mixed-scene batch, as the marcher reads it
// (Morton bitfield, coarse empty-space table); the scenes of one batch share
// the scene box, so their OccParams are equal.
struct SceneRef {
  const uint64_t* bf64;
  const uint32_t* coarse;
};

// input width of a split model's custom colour head: 8 latents + 16 SH +
// 16 SH differences + 6 * frequency degree, padded to 16
inline int vd_head_inputs(const fvc_net_desc& d) { return (8 + 16 + 16 + 6 * d.freq_encoding_degree + 15) / 16 * 16; }
// layout of the parameter blocks (nerfacc export) for a description; n_first
// (split models: mlp_first_head) is 0 for plain models
bool net_layout(const fvc_net_desc& d, GridParams& g, int64_t& n_pos, int64_t& n_base, int64_t& n_head,
                int64_t* n_first = nullptr);

// model export of the reference (loader.hip)
struct NerfaccExport {
  fvc_scene_info scene{};
  fvc_stepsize_info step{};
  fvc_net_desc net{};
  bool pos_separate = true;
  int64_t n_pos = 0, n_base = 0, n_head = 0, n_occ = 0, n_first = 0;
};
void nerfacc_read_config(const std::string& dir, NerfaccExport& ex);
// first: mlp_first_head buffer (split models; else unused)
void nerfacc_read_data(const std::string& dir, const NerfaccExport& ex, uint16_t* pos, uint16_t* base, uint16_t* head,
                       uint8_t* occ, uint16_t* first = nullptr);

// network launches (network.hip)
// n_dev (optional): the sample count in device memory, n its upper bound
void launch_network_full(const Model& m, const float* pos, const float* dir, int64_t n, uint32_t* out_h4,
                         hipStream_t s, const int* n_dev = nullptr, int rows = 16);
// the renderer's sample stream: pos4 = (x, y, z, pixel index bits) per
// sample, ray_sh the per-pixel SH table (sh4_all)
void launch_network_full_pix(const Model& m, const float4* pos4, const uint4* ray_sh, int64_t n, uint32_t* out_h4,
                             hipStream_t s, const int* n_dev, int rows);
void launch_network_full_ex(const Model& m, const float* pos, const float* dir, bool pix, int64_t n,
                            uint32_t* out_h4, hipStream_t s, const int* n_dev, int rows);
// log_sigma (optional): row 0 of each sample once more, as a compact half array
void launch_network_latents(const Model& m, const float* pos, int64_t n, uint16_t* out16, hipStream_t s,
                            int rows = 16, const int* n_dev = nullptr, uint16_t* log_sigma = nullptr);
void launch_network_head(const Model& m, const uint16_t* head_in16, const float* dir, int64_t n, uint32_t* out_h4,
                         hipStream_t s);
void launch_network_encode(const Model& m, const float* pos, int64_t n, uint16_t* out32, hipStream_t s);
// split models (viewdep.hip): full inference (the initial view direction is
// dir), latents with the initial view direction init_dir ([8 head latents,
// log sigma] per sample), the custom head alone
void launch_network_full_vd(const Model& m, const float* pos, const float* dir, bool pix, int64_t n,
                            uint32_t* out_h4, hipStream_t s, const int* n_dev, int rows);
void launch_network_latents_vd(const Model& m, const float* pos, const float* init_dir, int64_t n, uint16_t* out16,
                               hipStream_t s, int rows, const int* n_dev = nullptr);
void launch_network_head_vd(const Model& m, const uint16_t* head_in16, const float* pos, const float* dir,
                            const float* init_dir, int64_t n, uint32_t* out_h4, hipStream_t s);
// weight fragments of a split model (FP16)
void pack_fragments_vd(const uint16_t* base, const uint16_t* first, const uint16_t* head, const fvc_net_desc& d,
                       std::vector<uint8_t>& out);

// ---------------------------------------------------------------- cache
static const int BRICK = 8;               // Cache::BRICK_SIZE (cache.h:17)
static const int BRICK_VOXELS = 512;
// pool capacity limit: voxel indices (slot * 512 + voxel) stay below 2^32
static const int MAX_BRICKS = (1 << 23) - 1;

// Voxel order inside a brick of the pool: z fastest, then y, then x.  A
// ray's consecutive samples are one froxel apart in z (the froxel z is the
// step index), so the trilinear corners of neighbouring samples -- the lanes
// of a wave in the cache head -- share cache lines.  Exports and voxel
// gradients use the reference's order (x fastest, z slowest: canon_voxel).
FVC_HD int brick_voxel(int x, int y, int z) { return (x & 7) * 64 + (y & 7) * 8 + (z & 7); }
FVC_HD int canon_voxel(int x, int y, int z) { return (z & 7) * 64 + (y & 7) * 8 + (x & 7); }

// voxel record of the brick pool (32 bytes).  znext repeats density and
// isset of the +z neighbour in the same brick column (zeros at z = 7), so
// that one 16-byte load reads both corners of a z pair of a trilinear lookup.
struct alignas(16) Voxel {
  uint16_t lat[8];   // fp16 latents
  float density;
  float isset;
  float znext[2];
};

// voxel record with FP8-E4M3 latents (16 bytes): codes of l / s with the
// brick's scale s = absmax / 448 (CacheStore::scales)
struct alignas(16) Voxel8 {
  uint32_t lat[2];   // 8 E4M3 codes, latent i in byte i
  float density;
  float isset;
};

struct CacheStore {
  bool initialized = false;
  bool lat8 = false;  // latents as FP8-E4M3 (Voxel8) with per-brick scales
  fvc_raymarch_info rm{};
  int fx = 0, fy = 0, fz = 0;
  int bx = 0, by = 0, bz = 0;
  int capacity = 0;
  int n_used = 0;
  DevBuf table;     // int32 [bz][by][bx]
  DevBuf known;     // float [fy][fx]
  DevBuf pool;      // Voxel [capacity][512], or Voxel8 with lat8
  DevBuf scales;    // float [capacity] (lat8)
};

// ---------------------------------------------------------------- profiling
// Per-kernel HIP-event timing (the reference's CudaTimer per TimingState,
// renderer.h:79-105, cuda_timer.h:59-161) plus the algorithmic work each
// launch did, for roofline reporting.
struct ProfEntry {
  int64_t launches = 0;
  double ms = 0, units = 0, bytes = 0, flops = 0;
};
struct Profiler {
  bool enabled = false;
  std::map<std::string, ProfEntry> table;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
    double units, bytes, flops;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> free_events;
  hipEvent_t get_event() {
    if (!free_events.empty()) {
      hipEvent_t e = free_events.back();
      free_events.pop_back();
      return e;
    }
    hipEvent_t e;
    FVC_HIP(hipEventCreate(&e));
    return e;
  }
  // returns an opaque start token (nullptr when disabled)
  hipEvent_t begin(hipStream_t s) {
    if (!enabled) return nullptr;
    hipEvent_t a = get_event();
    FVC_HIP(hipEventRecord(a, s));
    return a;
  }
  // records the stop event; returns a handle for work() (-1 when disabled)
  int end(hipEvent_t a, hipStream_t s, const char* name, double units = 0, double bytes = 0, double flops = 0) {
    if (!enabled || !a) return -1;
    hipEvent_t b = get_event();
    FVC_HIP(hipEventRecord(b, s));
    pending.push_back({name, a, b, units, bytes, flops});
    return (int)pending.size() - 1;
  }
  // attach work known only after the launch (counters read back later)
  void work(int h, double units, double bytes, double flops) {
    if (h < 0) return;
    pending[h].units = units;
    pending[h].bytes = bytes;
    pending[h].flops = flops;
  }
  // after the stream is synchronised
  void collect() {
    for (auto& p : pending) {
      float ms = 0;
      FVC_HIP(hipEventElapsedTime(&ms, p.a, p.b));
      ProfEntry& e = table[p.name];
      e.launches++;
      e.ms += ms;
      e.units += p.units;
      e.bytes += p.bytes;
      e.flops += p.flops;
      free_events.push_back(p.a);
      free_events.push_back(p.b);
    }
    pending.clear();
  }
  ~Profiler() {
    for (auto& p : pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : free_events) (void)hipEventDestroy(e);
  }
};

// Algorithmic work per unit (one evaluated sample unless noted):
//   hash-grid gather: 16 levels x 8 corners x 2 features x 2 B = 512 B
//   density MLP: 2*(32*64 + 64*64 + 64*9) = 13440 FLOP
//   colour MLP:  2*(24*64 + 64*3)        =  3456 FLOP
//   cache trilinear lookup: 8 corners x (4 B index + 4 B isset + 4 B density
//                           + 16 B latents) = 224 B
static const double ALG_GATHER_B = 512, ALG_LOOKUP_B = 224;
static const double ALG_DENSITY_FLOP = 13440, ALG_COLOUR_FLOP = 3456;
// split models: first head 2 (32*64 + 64*8) (+ 2 * 64 * 64 per extra hidden
// layer), custom colour head 2 (64*64 + 64*3)
static const double ALG_FIRST_HEAD_FLOP = 5120, ALG_VD_COLOUR_FLOP = 8576;
// algorithmic FLOP of one full network evaluation
inline double net_flop(const fvc_net_desc& d) {
  if (!d.split_head) return ALG_DENSITY_FLOP + ALG_COLOUR_FLOP;
  return ALG_DENSITY_FLOP + ALG_FIRST_HEAD_FLOP + 8192.0 * (d.first_head_n_hidden_layers - 1) + ALG_VD_COLOUR_FLOP;
}

// ---------------------------------------------------------------- renderer
static const int MAX_STEPS_PER_LAUNCH = 64;   // cap of fvc_renderer_set_batching (64-bit sample masks)
static const int MAX_SCENES = 64;             // distinct models in one mixed-scene batch

// the render cache as the kernels see it
struct CacheView {
  fvc_camera_info cam;
  fvc_stepsize_info st;
  int fx, fy, fz, bx, by;
  const int* table;
  const float* known;
  const Voxel* pool;     // FP16 latents
  const Voxel8* pool8;   // FP8 latents (lat8)
  const float* scales;   // per brick (lat8)
  int lat8;
  // CacheSettings (cache_renderer.h:24-32): interpol_function (FVC_FN_*) and
  // use_inter_brick_interpolation
  int interp_fn;
  int inter_brick;
};
__device__ __forceinline__ float vox_density(const CacheView& cv, int64_t v) {
  return cv.lat8 ? cv.pool8[v].density : cv.pool[v].density;
}

// state of the differentiable render between forward and backward (diff.hip)
struct DiffState {
  DevBuf frag_index, frags;        // FP32 MFMA fragments of the caller's weights
  DevBuf rays, counts, offsets, ray_out, n_walk, scan_tmp;
  DevBuf s_pos, s_dt, s_ray, out4, d_out4;
  DevBuf counts_c, off_r, off_c, counts_all, list_r, list_c;  // cached mode
  int n = 0;
  bool cached = false;
  int64_t n_samples = 0, n_r = 0, n_c = 0;
  CacheView cv{};
  const float* pos = nullptr;  // parameters of the last completed forward
  const float* base = nullptr;
  const float* head = nullptr;
};

struct Renderer {
  Model* model = nullptr;
  DiffState diff;
  hipStream_t own_stream = nullptr;
  hipStream_t fill_stream = nullptr;
  hipStream_t user_stream = nullptr;
  bool double_buffered = false;
  CacheStore caches[2];
  int render_idx = 0, fill_idx = 0;
  fvc_cache_settings settings{0, 0, 1, 1, 1, 1};
  int render_mode = FVC_RENDER_BOTH;  // CachePerformanceRenderer::_render_mode
  // inspection records of a cached render (device; set only by
  // fvc_debug_render_cached): per-pixel sample classes, their count, and
  // the composited resample / cache samples
  int8_t* dbg_codes = nullptr;
  int dbg_k = 0;
  int* dbg_n = nullptr;
  int2* dbg_comp = nullptr;
  int capacity_override = 0;
  std::mutex read_mutex, write_mutex;
  Profiler prof;

  int width = 0, height = 0;
  int64_t sample_capacity = 0;
  int64_t ray_capacity = 0;    // pixels the ray buffers hold
  int64_t last_n_px = 0;       // pixels of the last render (all views)
  // sub-iteration batching: samples per launch targeted, and the cap on
  // samples per ray per launch (the reference's TARGET_BATCH_SIZE and
  // MAX_N_STEPS_PER_SUBITER, performance_renderer.h:54-58)
  int64_t target_batch = 1 << 25;
  int max_steps = 32;
  int mlp_rows = 16;     // MLP rows (samples) per wavefront and iteration: 16, 32 or 64
  int latent_format = FVC_LATENT_FP16;  // latent storage of the next fills
  int last_iters = 0;    // sub-iterations of the previous frame (first burst length)
  int last_fill_iters = 0;  // ... of the previous fill
  // step table of the marchers (render.hip: step_table)
  DevBuf tab_t, tab_off;
  fvc_stepsize_info tab_key{};
  float tab_limit = 0.0f;
  int tab_rows = -1;   // -1: not built; 0: disabled for this stepsize
  // ray state (double buffered for compaction)
  DevBuf rays[2];        // RayDyn of the alive list
  DevBuf ray_const;      // RayState per pixel (frame start)
  DevBuf ray_pix[2];     // int32 pixel index
  DevBuf ray_res[2];     // float4 rgb + T
  DevBuf results;        // float4 [H*W] final
  DevBuf ray_entry;      // int4 per ray: offset/count per stream
  DevBuf ray_order;      // uint64 per ray: stream of each sample, in march order
  DevBuf t_scratch;      // float [2 * MAX_STEPS_PER_LAUNCH][n]: sample t0 per launch
  DevBuf ray_alive;      // int32 per ray
  DevBuf counters;       // int32 device counters
  int* pinned = nullptr; // pinned host mirror of the counters
  // sample streams
  DevBuf s_pos, s_t0, s_out;  // resample / full stream: float4 (position, pixel index bits), t0, half4 output
  DevBuf c_rec, c_t0, c_out;  // cache-hit stream: c_rec = (froxel x, y, z, pixel index bits) per sample
  DevBuf ray_od;              // split models: ray origin, direction per pixel (2 float4)
  DevBuf ray_sh;                                // 16 fp16 SH coefficients per pixel (first launch of its ray)
  // mixed-scene batches: occupancy of each scene (SceneRef per scene)
  DevBuf scene_refs;
  std::vector<SceneRef> scene_refs_host;
  // fill scratch (own counters and pinned mirror: the fill may run on the
  // fill stream from another thread)
  DevBuf f_rays, f_tiles, f_list, f_pos, f_t0, f_meta, f_out16, f_counters;
  DevBuf f_alive8, f_alloc8;  // fill: alive byte per tile, brick-request byte per list entry
  DevBuf f_lsig;              // fill: log sigma per sample (half), plain models
  DevBuf f_dir;  // fill sample directions (split models: the initial view direction)
  int* f_pinned = nullptr;
  // background refill (double-buffered renderers)
  std::thread fill_thread;
  std::atomic<int> fill_state{0};  // FILL_*
  std::string fill_error;
  fvc_init_stats fill_stats{};
  Profiler fill_prof;
  V3 fill_pose{0, 0, 0}, front_pose{0, 0, 0};
  bool have_front_pose = false;
  fvc_refill_policy policy{0.9f, 1.0e30f, 0, 0};
  int fill_tiles = 0;
  int64_t fill_sample_capacity = 0;
  // graph mode: executable graphs of a full (0) and a cached (1) frame and
  // of a fill (2)
  bool use_graph = false;
  hipGraphExec_t graph_exec[3] = {nullptr, nullptr, nullptr};
  int graph_iters[2] = {0, 0};  // sub-iterations of the last frame of each kind
  ~Renderer() {
    if (fill_thread.joinable()) fill_thread.join();
    if (f_pinned) (void)hipHostFree(f_pinned);
    if (fill_stream) (void)hipStreamDestroy(fill_stream);
    for (hipGraphExec_t& g : graph_exec)
      if (g) (void)hipGraphExecDestroy(g);
    if (pinned) (void)hipHostFree(pinned);
    if (own_stream) (void)hipStreamDestroy(own_stream);
  }

  hipStream_t stream() const { return user_stream ? user_stream : own_stream; }
};

// render entry points (render.hip / cache.hip)
void render_full(Renderer& r, const fvc_raymarch_info& rm, uint8_t* rgba, fvc_render_stats* stats);
void cache_fill(Renderer& r, const fvc_raymarch_info& rm, fvc_init_stats* stats);
void render_cached(Renderer& r, const fvc_raymarch_info& rm, uint8_t* rgba, fvc_render_stats* stats);
enum { FILL_IDLE = 0, FILL_RUNNING = 1, FILL_DONE = 2, FILL_FAILED = 3 };
void cache_fill(Renderer& r, const fvc_raymarch_info& rm, fvc_init_stats* stats, hipStream_t s, Profiler& P);
void fill_async(Renderer& r, const fvc_raymarch_info& rm);
int fill_poll(Renderer& r, bool wait, fvc_init_stats* stats);
bool fill_publish(Renderer& r);
void render_cached_auto(Renderer& r, const fvc_raymarch_info& rm, uint8_t* rgba, fvc_render_stats* stats,
                        int* refill_started);
// autotuner of launch shapes (tune.hip)
void autotune(Renderer& r, const fvc_raymarch_info& rm, bool cached, fvc_tune_result* out);
// differentiable full / cached render (diff.hip)
void diff_forward(Renderer& r, const fvc_raymarch_info& rm, const int* pix, int n, const float* pos,
                  const float* base, const float* head, float* rgb, fvc_diff_stats* stats, bool cached);
void diff_backward(Renderer& r, const float* d_rgb, float* g_pos, float* g_base, float* g_head, float* g_pool);
void debug_march(Renderer& r, const fvc_raymarch_info& rm, int max_per_ray, int32_t* counts, float* t0s);
void debug_lookup(Renderer& r, int which, const float* fr_host, int64_t n, float* out_host);
// full render with its per-frame buffers in the library's static scratch
void render_full_noalloc(Renderer& r, const fvc_raymarch_info& rm, uint8_t* rgba, float* result,
                         fvc_render_stats* stats);
size_t noalloc_scratch_bytes();
void debug_render_cached(Renderer& r, const fvc_raymarch_info& rm, int k, int8_t* codes_host, int32_t* classified_host,
                         int32_t* composited_host, fvc_render_stats* stats);
void resize_renderbuffers(Renderer& r, int w, int h);
void ensure_ray_capacity(Renderer& r, int64_t n_pixels, int64_t target);
void graph_run(Renderer& r, int which, hipStream_t s, const std::function<void()>& record);

// a batch of views rendered together: pixel base of view v in the packed
// pixel space, images (device, may be null)
struct Views {
  const fvc_raymarch_info* rms;
  int n;
  uint8_t* const* rgba;
  std::vector<int64_t> base;
  int64_t n_px;
  int64_t max_px;  // pixels of the largest view
  // scenes (mixed-scene batches): the distinct models of the views, the
  // renderer's own model first when no per-view models are given; scene[v]
  // indexes them
  std::vector<const Model*> scenes;
  std::vector<int> scene;
  // models (optional): the model of each view (null: the renderer's); they
  // must share the renderer's network and scene box (checked)
  Views(const fvc_raymarch_info* rms, int n, uint8_t* const* rgba, const Model* own,
        const Model* const* models = nullptr);
  // samples per launch for a target set per view: a batch scales it by its
  // pixels over its largest view's, so every view keeps the per-ray budget
  // of a single render (at most 2^28)
  int64_t target(int64_t per_view) const {
    if (n_px <= max_px) return per_view;
    return (int64_t)std::min(268435456.0, (double)per_view * (double)n_px / (double)max_px);
  }
};
void render_full(Renderer& r, const Views& vs, fvc_render_stats* stats);
void render_cached(Renderer& r, const Views& vs, fvc_render_stats* stats);
// space: candidate sets (null: the default sweep)
void autotune(Renderer& r, const Views& vs, bool cached, const fvc_tune_space* space, fvc_tune_result* out);

}  // namespace fvc
