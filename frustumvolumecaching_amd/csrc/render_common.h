// render_common.h -- ray state, compositing and host helpers shared by the
// full renderer (render.hip) and the cache renderers (cache.hip).
#pragma once
#include "fvc_internal.h"
#include "march.h"
#include "wave.h"

namespace fvc {

// BaseRayInitInfo + SegmentedRayPayload of the reference
// (renderer_common.h:65-88) in one 48-byte record: origin, direction, far,
// next bin start t, and the bin position in the step table (start index m,
// bins advanced j, so t == T[m][j]).
struct RayState {
  float ox, oy, oz, dx, dy, dz, far_, t;
  int m, j;
  int scene;  // scene of the ray's view (mixed-scene batches; else 0)
  int pad;
};
// Where a ray stands between sub-iterations: next bin start t == T[m][j].
// The alive list compacts this and the pixel index; the rest of the ray
// (RayState, written once per frame at its pixel) stays in place.
struct RayDyn {
  float t;
  int j;
};

// Ray of pixel (x, y) clipped to the scene AABB, its start snapped by
// valid_t_from_t with the bin index kept: t = t_from_step(m + near_i)
// (initRayPayloads_split_kernel, renderer.cu:86-134, deterministic branch).
// Returns false when the ray misses the AABB.
__device__ __forceinline__ bool init_ray_state(const fvc_raymarch_info& rm, V3 aabb_from, V3 aabb_to, int x, int y,
                                               RayState& st) {
  const Ray r = gen_ray(rm.cam_info, (float)x + 0.5f, (float)y + 0.5f);
  float t0, t1;
  const bool hit = ray_aabb(r, aabb_from, aabb_to, t0, t1);
  const fvc_stepsize_info& s = rm.stepsize_info;
  st.ox = r.o.x; st.oy = r.o.y; st.oz = r.o.z;
  st.dx = r.d.x; st.dy = r.d.y; st.dz = r.d.z;
  st.far_ = hit ? fminf_(t1, s.far_) : s.far_;
  const float mf = hit ? roundf(step_from_t(s, fmaxf_(t0, s.near_)) - s.near_i) : 0.0f;
  st.t = hit ? t_from_step(s, mf + s.near_i) : s.near_;
  st.m = (int)mf;
  st.j = 0;
  st.scene = st.pad = 0;
  return hit;
}

__device__ __forceinline__ MarchRay march_ray(const RayState& s, const StepTableView& tab) {
  MarchRay r;
  r.row = step_row(tab, s.m);
  r.o = V3{s.ox, s.oy, s.oz};
  r.d = V3{s.dx, s.dy, s.dz};
  r.slab = ray_slab(r.o, r.d);
  r.far_ = s.far_;
  r.m = s.m;
  r.j = s.j;
  r.t1 = s.t;
  r.sk = no_skip();
  return r;
}

// samples a compositing lane loads ahead
static const int COMPOSITE_BATCH = 16;

// slots of Renderer::counters
enum { CNT_ALIVE = 0, CNT_SAMPLES = 1, CNT_CACHE = 2, CNT_STEPS = 3, CNT_TILES = 4, CNT_SLOTS = 5, CNT_OVERFLOW = 6 };

// Device control block of a frame's sub-iteration loop (also in
// Renderer::counters): alive counts (ping-pong by iteration parity), the
// samples per ray of the current iteration, the stream and bin counters,
// and a per-iteration log {alive, n_steps, resample, cache, bins}.  The host
// launches iterations in bursts and reads the alive count once per burst.
enum { CTL_ALIVE = 16, CTL_NSTEPS = 18, CTL_R = 19, CTL_C = 20, CTL_BINS = 21, CTL_LOG = 32 };
static const int CTL_LOG_FIELDS = 5;
static const int MAX_SUBITER = 1000;        // PerformanceRenderer::MAX_SUBITER
static const int CTL_WORDS = CTL_LOG + CTL_LOG_FIELDS * (MAX_SUBITER + 1);

// Plan iteration it: log iteration it-1, set the samples per ray
// (TARGET_BATCH_SIZE / n_alive clamped to [1, max_steps] and to the stream
// capacity, performance_renderer.cu:249-250) and zero the counters.
__global__ void k_plan(int* __restrict__ ctl, int it, int64_t target, int max_steps, int64_t cap, int log_only);

inline CacheView cache_view(const CacheStore& C) {
  CacheView cv;
  cv.cam = C.rm.cam_info;
  cv.st = C.rm.stepsize_info;
  cv.fx = C.fx; cv.fy = C.fy; cv.fz = C.fz; cv.bx = C.bx; cv.by = C.by;
  cv.table = C.table.as<int>();
  cv.known = C.known.as<float>();
  cv.pool = C.pool.as<Voxel>();
  cv.pool8 = C.pool.as<Voxel8>();
  cv.scales = C.scales.as<float>();
  cv.lat8 = C.lat8 ? 1 : 0;
  cv.interp_fn = FVC_FN_LINEAR;
  cv.inter_brick = 1;
  return cv;
}

// Class of an occupied sample [t0, t0 + dt) of a ray whose origin / direction
// in the cache camera's frame are oc / dc (sample_cache_performance_kernel,
// cache_performance_renderer.cu:96-160): SAMPLE_RESAMPLE outside the cache
// frustum or outside [near, known_to] of its pixel, SAMPLE_CACHE when the
// nearest voxel's brick is set and its density is > 0, else SAMPLE_SKIP.
enum { SAMPLE_SKIP = 0, SAMPLE_RESAMPLE = 1, SAMPLE_CACHE = 2 };
__device__ __forceinline__ int classify_sample(const CacheView& cv, V3 oc, V3 dc, float t0, float dt) {
  const float t1 = t0 + dt;
  const float tm = (t0 + t1) * 0.5f;
  float tf;
  const V3 f = cam2froxel(oc + tm * dc, cv.cam, cv.st, tf);
  const bool in_fov = f.x >= 0.0f && f.x < (float)cv.fx && f.y >= 0.0f && f.y < (float)cv.fy;
  const bool inside = in_fov && f.z >= 0.0f && f.z < (float)cv.fz;
  bool unsampled = true;
  if (in_fov) {
    const float known = cv.known[(int64_t)(int)f.y * cv.fx + (int)f.x];
    unsampled = tf < cv.st.near_ || tf > known;
  }
  if (!inside || unsampled) return SAMPLE_RESAMPLE;
  const int vx = (int)floorf(f.x), vy = (int)floorf(f.y), vz = (int)floorf(f.z);
  const int slot = cv.table[((int64_t)(vz >> 3) * cv.by + (vy >> 3)) * cv.bx + (vx >> 3)];
  if (slot >= 0 && vox_density(cv, (int64_t)slot * BRICK_VOXELS + brick_voxel(vx, vy, vz)) > 0.0f)
    return SAMPLE_CACHE;
  return SAMPLE_SKIP;
}

__device__ __forceinline__ float h2f(uint32_t bits16) { return (float)__builtin_bit_cast(_Float16, (uint16_t)bits16); }

// One step of front-to-back compositing (accumulate_performance_kernel,
// performance_renderer.cu:93-120): v = half4 (r, g, b raw, log sigma).
// Returns false once T < 1e-4.
__device__ __forceinline__ bool composite_sample(float4& res, uint2 v, float t0, const fvc_stepsize_info& st,
                                                 float alpha_thre) {
  const float cr = h2f(v.x & 0xffffu), cg = h2f(v.x >> 16), cb = h2f(v.y & 0xffffu), ls = h2f(v.y >> 16);
  const float density = expf(ls);
  const float dt = step_dt(st, t0);
  float alpha = 1.0f - expf(-density * dt);
  alpha = fminf_(fmaxf_(alpha, 0.0f), 1.0f);
  if (alpha > alpha_thre) {
    const float w = alpha * res.w;
    res.x = res.x + w * (1.0f / (1.0f + expf(-cr)));
    res.y = res.y + w * (1.0f / (1.0f + expf(-cg)));
    res.z = res.z + w * (1.0f / (1.0f + expf(-cb)));
    res.w = res.w * (1.0f - alpha);
  }
  return !(res.w < 1e-4f);
}

// to_8bit (raymarch_common.h:554-557)
FVC_HD int to_8bit(float f) {
  const int v = (int)(f * 256.0f);
  return v < 0 ? 0 : (v > 255 ? 255 : v);
}

// host helpers (render.hip)
StepTableView step_table(Renderer& r, const fvc_raymarch_info* rms, int n_views);
MarchGrid march_grid(Renderer& r, const fvc_raymarch_info* rms, int n_views);
void upload_scene_refs(Renderer& r, const Views& vs);
int read_counter(Renderer& r, int idx, hipStream_t s);
void zero_counters(Renderer& r, hipStream_t s);
int launch_init_rays(Renderer& r, const Views& vs, hipStream_t s);
void launch_init_rays_async(Renderer& r, const Views& vs, hipStream_t s);
void prepare_views(Renderer& r, const Views& vs);
// per-iteration log entries {alive, n_steps, resample, cache, bins} of a frame
struct IterLog { int alive, n_steps, n_r, n_c, bins; };
std::vector<IterLog> read_iter_log(Renderer& r, int n_it, hipStream_t s);
void launch_write_image(Renderer& r, const Views& vs, hipStream_t s);
int steps_for(int n_alive, int64_t target, int max_steps);
int next_burst(const Renderer& r, int it);

}  // namespace fvc
