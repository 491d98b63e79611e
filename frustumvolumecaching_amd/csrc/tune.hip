// tune.hip -- autotuner of the renderer's launch shapes.
//
// Candidates: MLP rows per wavefront and loop iteration (16, 32 or 64
// samples: one, two or four 16-sample MFMA tiles per wave, whose gathers are
// issued together) for the network kernels, and the sub-iteration chunk sizes
// (target samples per launch 2^23 / 2^24 / 2^25, samples per ray per launch
// 16 / 32 / 64: the reference's TARGET_BATCH_SIZE and
// MAX_N_STEPS_PER_SUBITER, performance_renderer.h:54-58).  Each candidate
// renders the calibration view once untimed and five times timed (each
// render synchronises its stream); the fastest render decides.  Results do
// not depend on the choice: only launch shapes change.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "fvc_internal.h"

namespace fvc {

void autotune(Renderer& r, const fvc_raymarch_info& rm, bool cached, fvc_tune_result* out) {
  if (cached && !r.caches[r.render_idx].initialized) throw Error(FVC_E_NOT_INITIALIZED, "Cache is not initialized!");
  const int rows_c[3] = {16, 32, 64};
  const int64_t target_c[3] = {(int64_t)1 << 23, (int64_t)1 << 24, (int64_t)1 << 25};
  const int steps_c[3] = {16, 32, 64};
  const int W = rm.cam_info.res_x, H = rm.cam_info.res_y;
  const int rows0 = r.mlp_rows, steps0 = r.max_steps;
  const int64_t target0 = r.target_batch;
  fvc_tune_result res;
  std::memset(&res, 0, sizeof(res));
  auto time_one = [&](int rows, int64_t target, int steps) {
    r.mlp_rows = rows;
    r.target_batch = target;
    r.max_steps = steps;
    resize_renderbuffers(r, W, H);  // sample capacity follows the chunk sizes
    auto run = [&] {
      if (cached) render_cached(r, rm, nullptr, nullptr);
      else render_full(r, rm, nullptr, nullptr);
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
  res.ms_default = time_one(rows0, target0, steps0);
  float best = 3.0e38f;
  int k = 0;
  for (int rows : rows_c)
    for (int64_t target : target_c)
      for (int steps : steps_c) {
        const float ms = time_one(rows, target, steps);
        if (k < FVC_TUNE_MAX) {
          res.ms[k] = ms;
          res.rows[k] = rows;
          res.steps[k] = steps;
          res.targets[k] = target;
        }
        k++;
        if (ms < best) {
          best = ms;
          res.mlp_rows = rows;
          res.target_batch = target;
          res.max_steps = steps;
        }
      }
  res.n_tried = std::min(k, FVC_TUNE_MAX);
  res.ms_best = best;
  r.mlp_rows = res.mlp_rows;
  r.target_batch = res.target_batch;
  r.max_steps = res.max_steps;
  resize_renderbuffers(r, W, H);
  if (out) *out = res;
}

}  // namespace fvc
