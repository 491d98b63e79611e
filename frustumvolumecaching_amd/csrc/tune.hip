// tune.hip -- autotuner of the renderer's launch shapes.
//
// Candidates: MLP rows per wavefront and loop iteration (16, 32 or 64
// samples: one, two or four 16-sample MFMA tiles per wave, whose gathers are
// issued together) for the network kernels, and the sub-iteration chunk sizes
// (target samples per launch, by default 2^23 / 2^24 / 2^25, and samples per
// ray per launch, by default 16 / 32 / 64: the reference's TARGET_BATCH_SIZE
// and MAX_N_STEPS_PER_SUBITER, performance_renderer.h:54-58); a caller may
// give its own candidate sets (fvc_tune_space).  Each candidate renders the
// calibration view (or batch of views) once untimed and five times timed
// (each render synchronises its stream); the fastest render decides.
// Results do not depend on the choice: only launch shapes change.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "fvc_internal.h"

namespace fvc {

// The default candidate space.
static fvc_tune_space default_space() {
  fvc_tune_space sp{};
  sp.n_rows = 3;
  sp.rows[0] = 16; sp.rows[1] = 32; sp.rows[2] = 64;
  sp.n_steps = 3;
  sp.steps[0] = 16; sp.steps[1] = 32; sp.steps[2] = 64;
  sp.n_targets = 3;
  sp.targets[0] = (int64_t)1 << 23; sp.targets[1] = (int64_t)1 << 24; sp.targets[2] = (int64_t)1 << 25;
  return sp;
}

static void check_space(const fvc_tune_space& sp) {
  if (sp.n_rows < 1 || sp.n_rows > 3 || sp.n_steps < 1 || sp.n_steps > 3 || sp.n_targets < 1 || sp.n_targets > 8)
    throw Error(FVC_E_INVALID, "tune space: 1..3 rows, 1..3 steps and 1..8 targets");
  if (sp.n_rows * sp.n_steps * sp.n_targets > FVC_TUNE_MAX)
    throw Error(FVC_E_INVALID, "tune space: too many candidates");
  for (int i = 0; i < sp.n_rows; i++)
    if (sp.rows[i] != 16 && sp.rows[i] != 32 && sp.rows[i] != 64)
      throw Error(FVC_E_INVALID, "tune space: MLP rows must be 16, 32 or 64");
  for (int i = 0; i < sp.n_steps; i++)
    if (sp.steps[i] < 1 || sp.steps[i] > MAX_STEPS_PER_LAUNCH)
      throw Error(FVC_E_INVALID, "tune space: samples per ray per launch must be in [1, 64]");
  for (int i = 0; i < sp.n_targets; i++)
    if (sp.targets[i] < ((int64_t)1 << 10) || sp.targets[i] > ((int64_t)1 << 28))
      throw Error(FVC_E_INVALID, "tune space: target samples per launch must be in [2^10, 2^28]");
}

void autotune(Renderer& r, const Views& vs, bool cached, const fvc_tune_space* space, fvc_tune_result* out) {
  if (cached && !r.caches[r.render_idx].initialized) throw Error(FVC_E_NOT_INITIALIZED, "Cache is not initialized!");
  if (cached && vs.scenes.size() > 1) throw Error(FVC_E_UNSUPPORTED, "a cached batch renders the scene of the renderer's cache");
  const fvc_tune_space sp = space ? *space : default_space();
  check_space(sp);
  const int rows0 = r.mlp_rows, steps0 = r.max_steps;
  const int64_t target0 = r.target_batch;
  fvc_tune_result res;
  std::memset(&res, 0, sizeof(res));
  // buffers for the chunk sizes of a candidate (the sample capacity
  // follows the target and the samples per ray per launch)
  auto shape = [&](int rows, int64_t target, int steps) {
    r.mlp_rows = rows;
    r.target_batch = target;
    r.max_steps = steps;
    ensure_ray_capacity(r, vs.n_px, vs.target(target));
    if (vs.n == 1) {
      r.width = vs.rms[0].cam_info.res_x;
      r.height = vs.rms[0].cam_info.res_y;
    }
  };
  auto time_one = [&](int rows, int64_t target, int steps) {
    shape(rows, target, steps);
    auto run = [&] {
      if (cached) render_cached(r, vs, nullptr);
      else render_full(r, vs, nullptr);
    };
    run();
    // the fastest of five renders: the candidates lie within a few per cent
    // of each other, and the minimum is the least noisy ranking
    double best_ms = 1e30;
    for (int k = 0; k < 5; k++) {
      const auto t0 = std::chrono::steady_clock::now();
      run();
      best_ms = std::min(best_ms, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    return (float)best_ms;
  };
  float best = 3.0e38f;
  int k = 0;
  try {
    res.ms_default = time_one(rows0, target0, steps0);
    for (int a = 0; a < sp.n_rows; a++)
      for (int b = 0; b < sp.n_targets; b++)
        for (int c = 0; c < sp.n_steps; c++) {
          const int rows = sp.rows[a], steps = sp.steps[c];
          const int64_t target = sp.targets[b];
          const float ms = time_one(rows, target, steps);
          res.ms[k] = ms;
          res.rows[k] = rows;
          res.steps[k] = steps;
          res.targets[k] = target;
          k++;
          if (ms < best) {
            best = ms;
            res.mlp_rows = rows;
            res.target_batch = target;
            res.max_steps = steps;
          }
        }
  } catch (...) {
    // the renderer keeps the shapes it had before the sweep
    r.mlp_rows = rows0;
    r.target_batch = target0;
    r.max_steps = steps0;
    throw;
  }
  res.n_tried = k;
  res.ms_best = best;
  shape(res.mlp_rows, res.target_batch, res.max_steps);
  if (out) *out = res;
}

void autotune(Renderer& r, const fvc_raymarch_info& rm, bool cached, fvc_tune_result* out) {
  autotune(r, Views(&rm, 1, nullptr, r.model), cached, nullptr, out);
}

}  // namespace fvc
