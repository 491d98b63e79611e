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
  int m, j, pad0, pad1;
};

// slots of Renderer::counters
enum { CNT_ALIVE = 0, CNT_SAMPLES = 1, CNT_CACHE = 2, CNT_STEPS = 3, CNT_TILES = 4, CNT_SLOTS = 5, CNT_OVERFLOW = 6 };

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
StepTableView step_table(Renderer& r, const fvc_raymarch_info& rm);
int read_counter(Renderer& r, int idx, hipStream_t s);
void zero_counters(Renderer& r, hipStream_t s);
int launch_init_rays(Renderer& r, const fvc_raymarch_info& rm, hipStream_t s);
void launch_write_image(Renderer& r, const fvc_raymarch_info& rm, uint8_t* rgba, hipStream_t s);
int steps_for(int n_alive, int64_t target, int max_steps);

}  // namespace fvc
