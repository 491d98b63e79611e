// render.hip -- full render (PerformanceRenderer::render,
// src/performance_renderer.cu:214-286) as a sub-iteration loop of
//   march (compacted sample stream) -> fused network -> composite + ray compaction.
//
// Rays live in a compacted SoA list that shrinks every sub-iteration; dead
// rays scatter their result to the per-pixel result buffer (as
// compactPayloads_split_kernel does, renderer.cu:158-189).  Samples of one ray
// are contiguous in the stream (offsets from a wave prefix sum + one atomic
// per wave), so compositing walks them in order.
#include <algorithm>
#include <cstring>
#include <functional>
#include <mutex>

#include "march.h"
#include "mlp_mfma.h"
#include "render_common.h"

namespace fvc {

// initRayPayloads_split_kernel (renderer.cu:86-134), deterministic branch:
// per pixel ray, AABB clip, start t snapped by valid_t_from_t; rays that hit
// the AABB are appended to the alive list, every pixel result starts as
// (0,0,0, T=1).
// Pixels of a view are numbered pix_base + y * W + x in the packed pixel
// space of a batch of views.
__global__ __launch_bounds__(256) void k_init_rays(fvc_raymarch_info rm, int pix_base, int scene, V3 aabb_from,
                                                   V3 aabb_to,
                                                   RayDyn* __restrict__ rays, RayState* __restrict__ rconst,
                                                   int* __restrict__ pix, float4* __restrict__ rres,
                                                   float4* __restrict__ results, int* __restrict__ counter) {
  const int W = rm.cam_info.res_x, H = rm.cam_info.res_y;
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  const bool valid = x < W && y < H;
  RayState st{};
  const bool alive = valid && init_ray_state(rm, aabb_from, aabb_to, x, y, st);
  if (valid) results[pix_base + y * W + x] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  int slot[1];
  block_reserve<1>({alive ? 1 : 0}, {counter}, slot);
  if (alive) {
    st.scene = scene;
    rconst[pix_base + y * W + x] = st;
    rays[slot[0]] = RayDyn{st.t, st.j};
    pix[slot[0]] = pix_base + y * W + x;
    rres[slot[0]] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  }
}

// sample_performance_kernel (performance_renderer.cu:15-69): march until
// n_steps occupied bin midpoints or t_mid > far; emit the contracted
// position with the pixel index (float4) and t0 of every occupied sample
// into the compacted stream.  The view direction is per ray: at its first
// launch a ray writes the SH coefficients of unit_to_01(dir) to its pixel's
// entry of ray_sh, which the network reads instead of a direction per
// sample.  The t0 of a ray's k-th sample of the launch goes to
// t_scr[k * n_alive + i] (coalesced across the wave when read back for the
// stream).
static constexpr int EMIT = 16;  // samples whose t0 the full marcher's emission loads together

// MULTI (mixed-scene batches): each ray reads the occupancy of its view's
// scene (scenes[scene]) with the coarse table from global memory (L2
// resident) instead of the block's LDS copy.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_march(int* __restrict__ ctl, int par, MarchGrid mg, ContractParams con,
                                               const uint32_t* __restrict__ coarse_g,
                                               const SceneRef* __restrict__ scenes, RayDyn* __restrict__ rays,
                                               const RayState* __restrict__ rconst, const int* __restrict__ pix_in,
                                               int4* __restrict__ entry, int* __restrict__ alive_flag,
                                               float* __restrict__ t_scr, float4* __restrict__ s_pos,
                                               uint4* __restrict__ ray_sh, float* __restrict__ s_t0) {
  const int n_alive = ctl[CTL_ALIVE + par];
  if ((int)(blockIdx.x * blockDim.x) >= n_alive) return;  // whole block past the alive rays
  const int n_steps = ctl[CTL_NSTEPS];
  int* s_counter = ctl + CTL_R;
  int* step_counter = ctl + CTL_BINS;
  __shared__ uint32_t coarse[16 * COARSE_WORDS];
  if constexpr (!MULTI) load_coarse(coarse_g, mg.occ.nlvl, coarse);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n_alive;
  int cnt = 0, bins = 0;
  bool alive = false;
  RayState rs{};
  int px = 0;
  if (valid) {
    const RayDyn dy = rays[i];
    px = pix_in[i];
    rs = rconst[px];
    rs.t = dy.t;
    rs.j = dy.j;
    if (rs.j == 0) {
      // first launch of the ray: the colour MLP's SH input of its direction
      uint4 lo, hi;
      const V3 d01 = unit_to_01(V3{rs.dx, rs.dy, rs.dz});
      sh4_all(d01.x, d01.y, d01.z, lo, hi);
      ray_sh[2 * (int64_t)px] = lo;
      ray_sh[2 * (int64_t)px + 1] = hi;
    }
    alive = true;
    MarchGrid g = mg;
    const uint32_t* cz;
    if constexpr (MULTI) {
      const SceneRef sr = scenes[rs.scene];
      g.bf64 = sr.bf64;
      cz = sr.coarse;
    } else {
      cz = coarse;
    }
    MarchRay r = march_ray(rs, g.tab);
    while (cnt < n_steps) {
      float t0;
      const int res = march_trip(g, cz, r, t0, bins);
      if (res == TRIP_END) { alive = false; break; }
      if (res == TRIP_SAMPLE) {
        t_scr[(int64_t)cnt * n_alive + i] = t0;
        cnt++;
      }
    }
    rs.t = r.t1;
    rs.j = r.j;
  }
  int slot[2];
  block_reserve<2>({cnt, bins}, {s_counter, step_counter}, slot);
  const int off = slot[0];
  if (valid) {
    const V3 o = {rs.ox, rs.oy, rs.oz}, d = {rs.dx, rs.dy, rs.dz};
    const float pxf = __builtin_bit_cast(float, px);
    // the t0 of EMIT samples are loaded together before they are used
    for (int k0 = 0; k0 < cnt; k0 += EMIT) {
      float t0v[EMIT];
#pragma unroll
      for (int u = 0; u < EMIT; u++) t0v[u] = t_scr[(int64_t)min(k0 + u, cnt - 1) * n_alive + i];
#pragma unroll
      for (int u = 0; u < EMIT; u++) {
        const int k = k0 + u;
        if (k >= cnt) break;
        const float t0 = t0v[u];
        const float t1 = t0 + step_dt(mg.st, t0);
        const float tm = (t0 + t1) * 0.5f;
        const V3 p = contract(o + tm * d, con);
        const int64_t q = (int64_t)off + k;
        s_pos[q] = make_float4(p.x, p.y, p.z, pxf);
        s_t0[q] = t0;
      }
    }
    rays[i] = RayDyn{rs.t, rs.j};
    entry[i] = make_int4(off, cnt, 0, 0);
    alive_flag[i] = alive;
  }
}

// accumulate_performance_kernel (performance_renderer.cu:71-125) + the ray
// compaction of the next sub-iteration (renderer.cu:158-189).  The samples
// of a ray are read in batches of COMPOSITE_BATCH independent loads ahead of
// the sequential compositing.
__global__ __launch_bounds__(256) void k_composite(int* __restrict__ ctl, int par, fvc_stepsize_info st,
                                                   float alpha_thre, const RayDyn* __restrict__ rays_in,
                                                   const int* __restrict__ pix_in, const float4* __restrict__ res_in,
                                                   const int4* __restrict__ entry, const int* __restrict__ alive_flag,
                                                   const uint2* __restrict__ s_out, const float* __restrict__ s_t0,
                                                   RayDyn* __restrict__ rays_out, int* __restrict__ pix_out,
                                                   float4* __restrict__ res_out, float4* __restrict__ results) {
  const int n_alive = ctl[CTL_ALIVE + par];
  if ((int)(blockIdx.x * blockDim.x) >= n_alive) return;
  int* counter = ctl + CTL_ALIVE + (par ^ 1);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n_alive;
  bool keep = false;
  int px = 0;
  float4 res = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  if (valid) {
    res = res_in[i];
    const int4 e = entry[i];
    // loaded with the entry: they are only needed after the samples
    const int ray_alive = alive_flag[i];
    px = pix_in[i];
    bool visible = true;  // transmittance still above the early-stop threshold
    for (int k0 = 0; k0 < e.y && visible; k0 += COMPOSITE_BATCH) {
      uint2 v[COMPOSITE_BATCH];
      float t[COMPOSITE_BATCH];
#pragma unroll
      for (int u = 0; u < COMPOSITE_BATCH; u++) {
        if (k0 + u < e.y) {
          v[u] = s_out[e.x + k0 + u];
          t[u] = s_t0[e.x + k0 + u];
        }
      }
#pragma unroll
      for (int u = 0; u < COMPOSITE_BATCH; u++)
        if (visible && k0 + u < e.y) visible = composite_sample(res, v[u], t[u], st, alpha_thre);
    }
    keep = visible && ray_alive != 0;
    if (!keep) results[px] = res;
  }
  int slot[1];
  block_reserve<1>({keep ? 1 : 0}, {counter}, slot);
  if (keep) {
    rays_out[slot[0]] = rays_in[i];
    pix_out[slot[0]] = px;
    res_out[slot[0]] = res;
  }
}

// writeImageBuffer_performance_kernel (performance_renderer.cu:127-149)
__global__ void k_write_image(int W, int H, int pix_base, const float4* __restrict__ results,
                              uchar4* __restrict__ rgba, int sample_index) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= W || y >= H) return;
  const int64_t p = (int64_t)y * W + x;
  const float4 r = results[pix_base + p];
  const uchar4 prev = rgba[p];
  const float si = (float)sample_index, den = (float)(sample_index + 1);
  const float cr = (r.x + r.w + (float)prev.x / 255.0f * si) / den;
  const float cg = (r.y + r.w + (float)prev.y / 255.0f * si) / den;
  const float cb = (r.z + r.w + (float)prev.z / 255.0f * si) / den;
  rgba[p] = make_uchar4((unsigned char)to_8bit(cr), (unsigned char)to_8bit(cg), (unsigned char)to_8bit(cb), 255);
}

// march only: counts of occupied samples + first max_per_ray t0 per pixel
__global__ void k_debug_march(fvc_raymarch_info rm, V3 aabb_from, V3 aabb_to, MarchGrid mg,
                              const uint32_t* __restrict__ coarse_g, int max_per_ray, int* __restrict__ counts,
                              float* __restrict__ t0s) {
  __shared__ uint32_t coarse[16 * COARSE_WORDS];
  load_coarse(coarse_g, mg.occ.nlvl, coarse);
  const int W = rm.cam_info.res_x, H = rm.cam_info.res_y;
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= W || y >= H) return;
  const int64_t p = (int64_t)y * W + x;
  RayState rs;
  int cnt = 0;
  if (init_ray_state(rm, aabb_from, aabb_to, x, y, rs)) {
    MarchRay r = march_ray(rs, mg.tab);
    int bins = 0;
    while (true) {
      float t0;
      const int res = march_trip(mg, coarse, r, t0, bins);
      if (res == TRIP_END) break;
      if (res != TRIP_SAMPLE) continue;
      if (cnt < max_per_ray) t0s[p * max_per_ray + cnt] = t0;
      cnt++;
    }
  }
  counts[p] = cnt;
}

// ---------------------------------------------------------------- host

void resize_renderbuffers(Renderer& r, int w, int h) {
  if (w <= 0 || h <= 0) throw Error(FVC_E_INVALID, "resolution must be positive");
  ensure_ray_capacity(r, (int64_t)w * h, r.target_batch);
  r.width = w;
  r.height = h;
}

// ray and sample buffers for n pixels (all views of a batch) and a launch
// target of `target` samples
void ensure_ray_capacity(Renderer& r, int64_t n, int64_t target) {
  if (n >= ((int64_t)1 << 31) / 4) throw Error(FVC_E_INVALID, "too many pixels in one render call");
  n = std::max(n, r.ray_capacity);  // buffers only grow
  // a launch emits at most n_alive * n_steps <= max(target, n) samples per
  // stream, and never more than n * max_steps
  const int64_t cap = std::min<int64_t>(std::max<int64_t>(target, n), n * r.max_steps) + 64;
  for (int b = 0; b < 2; b++) {
    r.rays[b].ensure(n * sizeof(RayDyn));
    r.ray_pix[b].ensure(n * sizeof(int));
    r.ray_res[b].ensure(n * sizeof(float4));
  }
  r.results.ensure(n * sizeof(float4));
  r.ray_const.ensure(n * sizeof(RayState));
  r.ray_entry.ensure(n * sizeof(int4));
  r.ray_order.ensure(n * sizeof(uint64_t));
  r.t_scratch.ensure(n * 2 * MAX_STEPS_PER_LAUNCH * sizeof(float));
  r.ray_alive.ensure(n * sizeof(int));
  r.counters.ensure(CTL_WORDS * sizeof(int));
  r.s_pos.ensure(cap * sizeof(float4));
  r.s_t0.ensure(cap * sizeof(float));
  r.ray_sh.ensure(n * 32);
  r.s_out.ensure(cap * sizeof(uint2));
  r.ray_capacity = n;
  r.sample_capacity = cap;
  if (!r.pinned) FVC_HIP(hipHostMalloc((void**)&r.pinned, 64 * sizeof(int), hipHostMallocDefault));
}

__global__ void k_plan(int* __restrict__ ctl, int it, int64_t target, int max_steps, int64_t cap, int log_only) {
  if (threadIdx.x != 0) return;
  if (it > 0 && it - 1 <= MAX_SUBITER) {
    int* lg = ctl + CTL_LOG + CTL_LOG_FIELDS * (it - 1);
    lg[0] = ctl[CTL_ALIVE + ((it - 1) & 1)];
    lg[1] = ctl[CTL_NSTEPS];
    lg[2] = ctl[CTL_R];
    lg[3] = ctl[CTL_C];
    lg[4] = ctl[CTL_BINS];
  }
  if (log_only) return;
  const int64_t n = ctl[CTL_ALIVE + (it & 1)];
  int64_t k = target / (n > 1 ? n : 1);
  k = k < 1 ? 1 : (k > max_steps ? max_steps : k);
  if (n > 0 && n * k > cap) k = cap / n > 1 ? cap / n : 1;
  ctl[CTL_NSTEPS] = (int)k;
  ctl[CTL_R] = 0;
  ctl[CTL_C] = 0;
  ctl[CTL_BINS] = 0;
  ctl[CTL_ALIVE + ((it + 1) & 1)] = 0;
}

std::vector<IterLog> read_iter_log(Renderer& r, int n_it, hipStream_t s) {
  std::vector<IterLog> out((size_t)n_it);
  if (n_it > 0) {
    FVC_HIP(hipMemcpyAsync(out.data(), r.counters.as<int>() + CTL_LOG, sizeof(IterLog) * n_it, hipMemcpyDeviceToHost,
                           s));
    FVC_HIP(hipStreamSynchronize(s));
  }
  return out;
}

int read_counter(Renderer& r, int idx, hipStream_t s) {
  FVC_HIP(hipMemcpyAsync(r.pinned + idx, r.counters.as<int>() + idx, sizeof(int), hipMemcpyDeviceToHost, s));
  FVC_HIP(hipStreamSynchronize(s));
  return r.pinned[idx];
}

__global__ void k_zero_ints(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// a kernel, not a memset: the frame may be recorded into a graph
void zero_counters(Renderer& r, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(CTL_LOG), 0, s, r.counters.as<int>(), CTL_LOG);
  FVC_HIP(hipGetLastError());
}

// Record the launches of `record` on stream s into a graph and run it: the
// executable graph of slot `which` is updated in place when the topology is
// unchanged (same sub-iteration count), else instantiated again.
void graph_run(Renderer& r, int which, hipStream_t s, const std::function<void()>& record) {
  FVC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    record();
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hipGraph_t g = nullptr;
  FVC_HIP(hipStreamEndCapture(s, &g));
  bool updated = false;
  if (r.graph_exec[which]) {
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res;
    updated = hipGraphExecUpdate(r.graph_exec[which], g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
    if (!updated) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(r.graph_exec[which]);
      r.graph_exec[which] = nullptr;
    }
  }
  if (!updated) {
    const hipError_t e = hipGraphInstantiate(&r.graph_exec[which], g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      FVC_HIP(e);
    }
  }
  (void)hipGraphDestroy(g);
  FVC_HIP(hipGraphLaunch(r.graph_exec[which], s));
}

Views::Views(const fvc_raymarch_info* rms_, int n_, uint8_t* const* rgba_, const Model* own,
             const Model* const* models)
    : rms(rms_), n(n_), rgba(rgba_) {
  if (n <= 0) throw Error(FVC_E_INVALID, "no views");
  base.resize(n);
  scenes.assign(1, own);
  scene.assign(n, 0);
  for (int v = 0; models && v < n; v++) {
    const Model* mv = models[v] ? models[v] : own;
    if (mv == own) continue;
    if (mv->device != own->device) throw Error(FVC_E_INVALID, "the models of a batch must live on one device");
    if (mv->net_hash != own->net_hash || std::memcmp(&mv->net, &own->net, sizeof(fvc_net_desc)) != 0)
      throw Error(FVC_E_UNSUPPORTED, "the models of a batch must share the renderer's network");
    if (std::memcmp(&mv->scene, &own->scene, sizeof(fvc_scene_info)) != 0)
      throw Error(FVC_E_UNSUPPORTED, "the models of a batch must share the scene box and cascades");
    const auto it = std::find(scenes.begin(), scenes.end(), mv);
    scene[v] = (int)(it - scenes.begin());
    if (it == scenes.end()) scenes.push_back(mv);
  }
  if ((int)scenes.size() > MAX_SCENES) throw Error(FVC_E_UNSUPPORTED, "too many scenes in one batch");
  n_px = 0;
  max_px = 0;
  for (int v = 0; v < n; v++) {
    const int W = rms[v].cam_info.res_x, H = rms[v].cam_info.res_y;
    if (W <= 0 || H <= 0) throw Error(FVC_E_INVALID, "resolution must be positive");
    if (std::memcmp(&rms[v].stepsize_info, &rms[0].stepsize_info, sizeof(fvc_stepsize_info)) != 0)
      throw Error(FVC_E_INVALID, "the views of a batch must share the stepsize info");
    base[v] = n_px;
    n_px += (int64_t)W * H;
    max_px = std::max<int64_t>(max_px, (int64_t)W * H);
  }
}

// Rays of every view appended to one alive list (one launch per view, one
// counter); returns the number of alive rays.
int launch_init_rays(Renderer& r, const Views& vs, hipStream_t s) {
  launch_init_rays_async(r, vs, s);
  return read_counter(r, CTL_ALIVE, s);
}

void launch_init_rays_async(Renderer& r, const Views& vs, hipStream_t s) {
  const Model& m = *r.model;
  zero_counters(r, s);
  for (int v = 0; v < vs.n; v++) {
    const fvc_raymarch_info& rm = vs.rms[v];
    const int W = rm.cam_info.res_x, H = rm.cam_info.res_y;
    hipLaunchKernelGGL(k_init_rays, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, s, rm, (int)vs.base[v],
                       vs.scene[v], cvt(m.scene.aabb_from), cvt(m.scene.aabb_to), r.rays[0].as<RayDyn>(),
                       r.ray_const.as<RayState>(), r.ray_pix[0].as<int>(), r.ray_res[0].as<float4>(),
                       r.results.as<float4>(), r.counters.as<int>() + CTL_ALIVE);
    FVC_HIP(hipGetLastError());
  }
}

void launch_write_image(Renderer& r, const Views& vs, hipStream_t s) {
  for (int v = 0; v < vs.n; v++) {
    const fvc_raymarch_info& rm = vs.rms[v];
    const int W = rm.cam_info.res_x, H = rm.cam_info.res_y;
    if (!vs.rgba || !vs.rgba[v]) continue;
    hipLaunchKernelGGL(k_write_image, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, s, W, H, (int)vs.base[v],
                       r.results.as<float4>(), (uchar4*)vs.rgba[v], rm.sample_index);
    FVC_HIP(hipGetLastError());
  }
}

// pixels, renderbuffers and step table for a batch of views
void prepare_views(Renderer& r, const Views& vs) {
  r.last_n_px = vs.n_px;
  if (vs.n == 1) {
    const int W = vs.rms[0].cam_info.res_x, H = vs.rms[0].cam_info.res_y;
    if (W != r.width || H != r.height) resize_renderbuffers(r, W, H);
  } else {
    ensure_ray_capacity(r, vs.n_px, vs.target(r.target_batch));
  }
}

// Largest t any ray of this view can march to: the AABB exit is at most the
// distance to the farthest AABB corner, and never beyond `far`.
static float max_ray_t(const fvc_raymarch_info& rm, const fvc_scene_info& sc) {
  const V3 o = translation(rm.cam_info.cam2world);
  const V3 from = cvt(sc.aabb_from), to = cvt(sc.aabb_to);
  float tf = 0.0f;
  for (int c = 0; c < 8; c++) {
    const V3 corner = {(c & 1) ? to.x : from.x, (c & 2) ? to.y : from.y, (c & 4) ? to.z : from.z};
    const V3 v = corner - o;
    tf = std::max(tf, std::sqrt(dot3(v, v)));
  }
  return std::min(tf, rm.stepsize_info.far_);
}

// Build (or reuse) the step table for this stepsize and view (see
// build_step_rows); when it is disabled the marchers simply step.
StepTableView step_table(Renderer& r, const fvc_raymarch_info* rms, int n_views) {
  const fvc_stepsize_info& st = rms[0].stepsize_info;
  float need = 0.0f;
  for (int v = 0; v < n_views; v++) need = std::max(need, max_ray_t(rms[v], r.model->scene));
  StepTableView v = step_view(st);
  if (!(r.tab_rows >= 0 && std::memcmp(&r.tab_key, &st, sizeof(st)) == 0 && need <= r.tab_limit)) {
    const float lim = need * 1.02f + 4.0f * st.dt_max;
    std::vector<float> t;
    std::vector<int> off;
    const int rows = build_step_rows(st, lim, t, off);
    if (rows > 0) {
      r.tab_t.ensure(t.size() * sizeof(float));
      r.tab_off.ensure(off.size() * sizeof(int));
      FVC_HIP(hipMemcpy(r.tab_t.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
      FVC_HIP(hipMemcpy(r.tab_off.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    r.tab_rows = rows;
    r.tab_key = st;
    r.tab_limit = rows > 0 ? lim : 3.0e38f;
  }
  v.t = r.tab_rows > 0 ? r.tab_t.as<float>() : nullptr;
  v.off = r.tab_rows > 0 ? r.tab_off.as<int>() : nullptr;
  v.n_rows = r.tab_rows;
  return v;
}

MarchGrid march_grid(Renderer& r, const fvc_raymarch_info* rms, int n_views) {
  MarchGrid g;
  g.occ = r.model->occ;
  g.st = rms[0].stepsize_info;
  g.tab = step_table(r, rms, n_views);
  g.bf64 = r.model->bitfield.as<uint64_t>();
  return g;
}

// the occupancy table of a mixed-scene batch's scenes (uploaded when it changes)
void upload_scene_refs(Renderer& r, const Views& vs) {
  std::vector<SceneRef> refs;
  for (const Model* m : vs.scenes) refs.push_back(SceneRef{m->bitfield.as<uint64_t>(), m->coarse.as<uint32_t>()});
  if (refs.size() == r.scene_refs_host.size() &&
      std::memcmp(refs.data(), r.scene_refs_host.data(), refs.size() * sizeof(SceneRef)) == 0)
    return;
  r.scene_refs.ensure(MAX_SCENES * sizeof(SceneRef));
  FVC_HIP(hipStreamSynchronize(r.stream()));
  FVC_HIP(hipMemcpy(r.scene_refs.p, refs.data(), refs.size() * sizeof(SceneRef), hipMemcpyHostToDevice));
  r.scene_refs_host = refs;
}

int steps_for(int n_alive, int64_t target, int max_steps) {
  int64_t k = target / std::max(n_alive, 1);
  if (k < 1) k = 1;
  if (k > max_steps) k = max_steps;
  return (int)k;
}

// Sub-iterations are launched in bursts with their counts kept on the
// device (k_plan); the host reads the alive count once per burst.  The
// first burst is as long as the previous frame's loop.
int next_burst(const Renderer& r, int it) { return it == 0 ? std::max(2, r.last_iters) : 2; }

void render_full(Renderer& r, const Views& vs, fvc_render_stats* stats) {
  const Model& m = *r.model;
  prepare_views(r, vs);
  const fvc_raymarch_info& rm = vs.rms[0];
  const double n_px = (double)vs.n_px;
  hipStream_t s = r.stream();
  const MarchGrid mg = march_grid(r, vs.rms, vs.n);
  const bool multi = vs.scenes.size() > 1;
  if (multi) upload_scene_refs(r, vs);
  Profiler& P = r.prof;
  // graph mode: the frame's launches up to the previous frame's sub-iteration
  // count + 1 are recorded and replayed as one graph; their grids are sized
  // for every pixel (the kernels read the live counts on the device)
  const bool graph = r.use_graph && r.graph_iters[0] > 0 && !P.enabled;
  hipEvent_t ev = P.begin(s);
  const int n0 = graph ? (int)vs.n_px : launch_init_rays(r, vs, s);
  P.work(P.end(ev, s, "init_rays"), n_px, n_px * 16 + n0 * 60.0, 0);
  int* ctl = r.counters.as<int>();
  const int64_t cap = r.sample_capacity - 64;
  const int64_t target = std::min<int64_t>(vs.target(r.target_batch), cap);
  const int64_t n_bound = std::min<int64_t>(cap, (int64_t)n0 * r.max_steps);  // samples per launch
  const dim3 grid((n0 + 255) / 256);
  struct Handles { int march, net, comp; };
  std::vector<Handles> hs;
  auto iteration = [&](int it) {
      const int cur = it & 1;
      hipLaunchKernelGGL(k_plan, dim3(1), dim3(64), 0, s, ctl, it, target, r.max_steps, cap, 0);
      Handles h;
      ev = P.begin(s);
      hipLaunchKernelGGL(multi ? k_march<true> : k_march<false>, grid, dim3(256), 0, s, ctl, cur, mg, m.con,
                         m.coarse.as<uint32_t>(), r.scene_refs.as<SceneRef>(), r.rays[cur].as<RayDyn>(), r.ray_const.as<RayState>(), r.ray_pix[cur].as<int>(),
                         r.ray_entry.as<int4>(), r.ray_alive.as<int>(), r.t_scratch.as<float>(),
                         r.s_pos.as<float4>(), r.ray_sh.as<uint4>(), r.s_t0.as<float>());
      FVC_HIP(hipGetLastError());
      h.march = P.end(ev, s, "march");
      ev = P.begin(s);
      launch_network_full_pix(m, r.s_pos.as<float4>(), r.ray_sh.as<uint4>(), n_bound, r.s_out.as<uint32_t>(), s,
                              ctl + CTL_R, r.mlp_rows);
      h.net = P.end(ev, s, "network_full");
      ev = P.begin(s);
      hipLaunchKernelGGL(k_composite, grid, dim3(256), 0, s, ctl, cur, rm.stepsize_info, m.scene.alpha_thre,
                         r.rays[cur].as<RayDyn>(), r.ray_pix[cur].as<int>(), r.ray_res[cur].as<float4>(),
                         r.ray_entry.as<int4>(), r.ray_alive.as<int>(), r.s_out.as<uint2>(), r.s_t0.as<float>(),
                         r.rays[cur ^ 1].as<RayDyn>(), r.ray_pix[cur ^ 1].as<int>(),
                         r.ray_res[cur ^ 1].as<float4>(), r.results.as<float4>());
      FVC_HIP(hipGetLastError());
      h.comp = P.end(ev, s, "composite");
      hs.push_back(h);
  };
  int it = 0;
  bool done = n0 == 0;
  if (graph) {
    const int k = std::min(r.graph_iters[0] + 1, MAX_SUBITER);
    graph_run(r, 0, s, [&] {
      launch_init_rays_async(r, vs, s);
      for (int i = 0; i < k; i++) iteration(i);
    });
    it = k;
    done = read_counter(r, CTL_ALIVE + (it & 1), s) == 0;
  }
  while (!done && it < MAX_SUBITER) {
    const int burst = next_burst(r, it);
    for (int b = 0; b < burst && it < MAX_SUBITER; b++, it++) iteration(it);
    done = read_counter(r, CTL_ALIVE + (it & 1), s) == 0;
  }
  hipLaunchKernelGGL(k_plan, dim3(1), dim3(64), 0, s, ctl, it, target, r.max_steps, cap, 1);
  ev = P.begin(s);
  launch_write_image(r, vs, s);
  P.end(ev, s, "write_image", n_px, n_px * 24, 0);
  const std::vector<IterLog> lg = read_iter_log(r, it, s);
  int64_t n_eval = 0;
  int n_real = 0;
  for (int i = 0; i < it; i++) {
    const IterLog& l = lg[i];
    if (l.alive > 0) n_real = i + 1;
    const double bins = (double)(uint32_t)l.bins;
    // 1 B occupancy per bin, 20 B per emitted sample (position + pixel, t0),
    // ray state + entry per ray
    P.work(hs[i].march, bins, bins + 20.0 * l.n_r + 84.0 * l.alive, 0);
    P.work(hs[i].net, l.n_r, l.n_r * (ALG_GATHER_B + 16.0 + 8.0), l.n_r * net_flop(m.net));
    P.work(hs[i].comp, l.n_r, 12.0 * l.n_r + 124.0 * l.alive, 0);
    n_eval += l.n_r;
  }
  r.last_iters = n_real;
  r.graph_iters[0] = n_real;
  P.collect();
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->samples_ppx = (float)((double)n_eval / n_px);
    stats->resample_samples_evaluated = stats->samples_ppx;
    stats->n_resamples = (uint64_t)n_eval;
    stats->n_subiters = n_real;
  }
}

void render_full(Renderer& r, const fvc_raymarch_info& rm, uint8_t* rgba, fvc_render_stats* stats) {
  render_full(r, Views(&rm, 1, &rgba, r.model), stats);
}

// ---------------------------------------------------------------- no-allocation render
// The full render of one view with every per-frame device buffer (ray state,
// sample streams, counters) inside `scratch`, a file-scope device array of
// NOALLOC_SCRATCH_BYTES: for callers that cannot allocate device memory.
// Views of at most NOALLOC_MAX_PIXELS pixels; at most NOALLOC_STEPS samples
// per ray and NOALLOC_SAMPLES per launch; rays step every bin (no step
// table: the samples are the same, bit for bit).  Each GPU holds its own
// copy of the array; calls that overlap on one GPU race on it.
static constexpr int64_t NOALLOC_MAX_PIXELS = 256 * 256;
static constexpr int NOALLOC_STEPS = 16;
static constexpr int64_t NOALLOC_SAMPLES = (1 << 19) + 64;
static constexpr int NOALLOC_MAX_ITERS = 256;
static constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
static constexpr size_t NOALLOC_CTL = align256(CTL_WORDS * sizeof(int));
static constexpr size_t NOALLOC_PER_PIXEL = 2 * sizeof(RayDyn) + sizeof(RayState) + 2 * sizeof(int) +
                                            2 * sizeof(float4) + sizeof(float4) + sizeof(int4) + sizeof(int) +
                                            NOALLOC_STEPS * sizeof(float) + 2 * sizeof(uint4);
static constexpr size_t NOALLOC_PER_SAMPLE = sizeof(float4) + sizeof(float) + sizeof(uint2);
static constexpr size_t NOALLOC_SCRATCH_BYTES =
    NOALLOC_CTL + align256(NOALLOC_MAX_PIXELS * NOALLOC_PER_PIXEL) + 16 * 256 +
    align256(NOALLOC_SAMPLES * NOALLOC_PER_SAMPLE);
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// the device address of `scratch` on each device, taken once per device
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[64] = {};

static unsigned char* scratch_on(int device) {
  if (device < 0 || device >= 64) throw Error(FVC_E_INVALID, "device index out of range");
  std::scoped_lock lk(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    FVC_HIP(hipGetDevice(&cur));
    if (cur != device) FVC_HIP(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) FVC_HIP(hipSetDevice(cur));
    FVC_HIP(e);
    g_scratch_ptr[device] = ptr;
  }
  return static_cast<unsigned char*>(g_scratch_ptr[device]);
}

size_t noalloc_scratch_bytes() { return NOALLOC_SCRATCH_BYTES; }

void render_full_noalloc(Renderer& r, const fvc_raymarch_info& rm, uint8_t* rgba, float* result, fvc_render_stats* stats) {
  const Model& m = *r.model;
  const int W = rm.cam_info.res_x, H = rm.cam_info.res_y;
  const int64_t n_px = (int64_t)W * H;
  if (n_px > NOALLOC_MAX_PIXELS)
    throw Error(FVC_E_INVALID, "the no-allocation render takes views of at most 65536 pixels");
  hipStream_t s = r.stream();
  unsigned char* base = scratch_on(m.device);
  // carve the scratch: counters, per-pixel ray state, sample streams
  int* ctl = reinterpret_cast<int*>(base);
  unsigned char* p = base + NOALLOC_CTL;
  auto take = [&](size_t bytes) {
    unsigned char* q = p;
    p += align256(bytes);
    return q;
  };
  const int64_t P = n_px;
  RayDyn* rays[2] = {reinterpret_cast<RayDyn*>(take(P * sizeof(RayDyn))),
                     reinterpret_cast<RayDyn*>(take(P * sizeof(RayDyn)))};
  RayState* rconst = reinterpret_cast<RayState*>(take(P * sizeof(RayState)));
  int* pix[2] = {reinterpret_cast<int*>(take(P * sizeof(int))), reinterpret_cast<int*>(take(P * sizeof(int)))};
  float4* rres[2] = {reinterpret_cast<float4*>(take(P * sizeof(float4))),
                     reinterpret_cast<float4*>(take(P * sizeof(float4)))};
  float4* results = reinterpret_cast<float4*>(take(P * sizeof(float4)));
  int4* entry = reinterpret_cast<int4*>(take(P * sizeof(int4)));
  int* alive = reinterpret_cast<int*>(take(P * sizeof(int)));
  float* t_scr = reinterpret_cast<float*>(take(P * NOALLOC_STEPS * sizeof(float)));
  uint4* ray_sh = reinterpret_cast<uint4*>(take(P * 2 * sizeof(uint4)));
  const int64_t S = NOALLOC_SAMPLES;
  float4* s_pos = reinterpret_cast<float4*>(take(S * sizeof(float4)));
  float* s_t0 = reinterpret_cast<float*>(take(S * sizeof(float)));
  uint2* s_out = reinterpret_cast<uint2*>(take(S * sizeof(uint2)));
  if ((size_t)(p - base) > NOALLOC_SCRATCH_BYTES) throw Error(FVC_E_INVALID, "scratch layout exceeds its array");

  MarchGrid mg;
  mg.occ = m.occ;
  mg.st = rm.stepsize_info;
  mg.tab = step_view(rm.stepsize_info);  // no table rows: every bin is stepped
  mg.tab.t = nullptr;
  mg.tab.off = nullptr;
  mg.tab.n_rows = 0;
  mg.bf64 = m.bitfield.as<uint64_t>();

  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(CTL_LOG), 0, s, ctl, CTL_LOG);
  hipLaunchKernelGGL(k_init_rays, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, s, rm, 0, 0,
                     cvt(m.scene.aabb_from), cvt(m.scene.aabb_to), rays[0], rconst, pix[0], rres[0], results, ctl + CTL_ALIVE);
  FVC_HIP(hipGetLastError());
  const int64_t cap = S - 64;
  const int64_t n_bound = std::min<int64_t>(cap, P * NOALLOC_STEPS);
  const dim3 grid((unsigned)((P + 255) / 256));
  auto iteration = [&](int it) {
    const int cur = it & 1;
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(64), 0, s, ctl, it, cap, NOALLOC_STEPS, cap, 0);
    hipLaunchKernelGGL(k_march<false>, grid, dim3(256), 0, s, ctl, cur, mg, m.con, m.coarse.as<uint32_t>(), nullptr,
                       rays[cur], rconst,
                       pix[cur], entry, alive, t_scr, s_pos, ray_sh, s_t0);
    FVC_HIP(hipGetLastError());
    launch_network_full_pix(m, s_pos, ray_sh, n_bound, reinterpret_cast<uint32_t*>(s_out), s, ctl + CTL_R,
                            r.mlp_rows);
    hipLaunchKernelGGL(k_composite, grid, dim3(256), 0, s, ctl, cur, rm.stepsize_info, m.scene.alpha_thre, rays[cur],
                       pix[cur], rres[cur], entry, alive, s_out, s_t0, rays[cur ^ 1], pix[cur ^ 1], rres[cur ^ 1],
                       results);
    FVC_HIP(hipGetLastError());
  };
  // bursts of sub-iterations planned on the device; the host reads the
  // alive count once per burst (a copy to a host variable)
  int it = 0;
  bool done = false;
  while (!done) {
    if (it >= NOALLOC_MAX_ITERS)
      throw Error(FVC_E_INVALID, "the no-allocation render needs more than 256 sub-iterations for this view");
    for (int b = 0; b < 4 && it < NOALLOC_MAX_ITERS; b++, it++) iteration(it);
    int n_alive = 0;
    FVC_HIP(hipMemcpyAsync(&n_alive, ctl + CTL_ALIVE + (it & 1), sizeof(int), hipMemcpyDeviceToHost, s));
    FVC_HIP(hipStreamSynchronize(s));
    done = n_alive == 0;
  }
  hipLaunchKernelGGL(k_plan, dim3(1), dim3(64), 0, s, ctl, it, cap, NOALLOC_STEPS, cap, 1);
  if (rgba) {
    hipLaunchKernelGGL(k_write_image, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, s, W, H, 0, results,
                       (uchar4*)rgba, rm.sample_index);
    FVC_HIP(hipGetLastError());
  }
  if (result) FVC_HIP(hipMemcpyAsync(result, results, P * sizeof(float4), hipMemcpyDeviceToDevice, s));
  IterLog lg[NOALLOC_MAX_ITERS];
  FVC_HIP(hipMemcpyAsync(lg, ctl + CTL_LOG, sizeof(IterLog) * it, hipMemcpyDeviceToHost, s));
  FVC_HIP(hipStreamSynchronize(s));
  if (stats) {
    int64_t n_eval = 0;
    int n_real = 0;
    for (int i = 0; i < it; i++) {
      n_eval += lg[i].n_r;
      if (lg[i].alive > 0) n_real = i + 1;
    }
    std::memset(stats, 0, sizeof(*stats));
    stats->samples_ppx = (float)((double)n_eval / (double)P);
    stats->resample_samples_evaluated = stats->samples_ppx;
    stats->n_resamples = (uint64_t)n_eval;
    stats->n_subiters = n_real;
  }
}

void debug_march(Renderer& r, const fvc_raymarch_info& rm, int max_per_ray, int32_t* counts_host, float* t0_host) {
  const Model& m = *r.model;
  const int W = rm.cam_info.res_x, H = rm.cam_info.res_y;
  const int64_t n = (int64_t)W * H;
  hipStream_t s = r.stream();
  const MarchGrid mg = march_grid(r, &rm, 1);
  DevBuf c, t;
  c.ensure(n * sizeof(int));
  t.ensure(std::max<int64_t>(1, n * max_per_ray) * sizeof(float));
  // slots past a ray's count read as 0
  FVC_HIP(hipMemsetAsync(t.p, 0, std::max<int64_t>(1, n * max_per_ray) * sizeof(float), s));
  hipLaunchKernelGGL(k_debug_march, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, s, rm, cvt(m.scene.aabb_from),
                     cvt(m.scene.aabb_to), mg, m.coarse.as<uint32_t>(), max_per_ray, c.as<int>(), t.as<float>());
  FVC_HIP(hipGetLastError());
  FVC_HIP(hipMemcpyAsync(counts_host, c.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
  if (max_per_ray > 0 && t0_host)
    FVC_HIP(hipMemcpyAsync(t0_host, t.p, n * max_per_ray * sizeof(float), hipMemcpyDeviceToHost, s));
  FVC_HIP(hipStreamSynchronize(s));
}

}  // namespace fvc
