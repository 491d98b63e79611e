// Host check of the empty-space skipping marcher (march.h) against plain
// stepping: every ray of a pixel subsample is marched to its far end twice,
// once with march_trip (hierarchical occupancy, empty boxes, step-table
// jumps) and once bin by bin with the reference's loop (t1 = t0 + dt(t0),
// sample when the midpoint is occupied: performance_renderer.cu:39-62,
// grid_occupied_at raymarch_common.h:680-710); the sample start times must
// be bit-identical.  Reads <dir>/scene.bin, rm.bin (the fvc_scene_info and
// fvc_raymarch_info records) and occ.bin (uint8 [nlvl][128][128][128]).
// Usage: march_skip_check <dir> <pixel stride>.  Prints counts; exits 1 on
// any mismatch.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../frustumvolumecaching_amd/csrc/march.h"

using namespace fvc;

static std::vector<uint8_t> read_file(const std::string& p) {
  std::vector<uint8_t> v;
  FILE* f = std::fopen(p.c_str(), "rb");
  if (!f) { std::fprintf(stderr, "cannot open %s\n", p.c_str()); std::exit(2); }
  std::fseek(f, 0, SEEK_END);
  v.resize((size_t)std::ftell(f));
  std::fseek(f, 0, SEEK_SET);
  if (std::fread(v.data(), 1, v.size(), f) != v.size()) std::exit(2);
  std::fclose(f);
  return v;
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const std::string dir = argv[1];
  const int stride = std::atoi(argv[2]);
  const std::vector<uint8_t> scb = read_file(dir + "/scene.bin"), rmb = read_file(dir + "/rm.bin");
  const std::vector<uint8_t> occ = read_file(dir + "/occ.bin");
  fvc_scene_info sc;
  fvc_raymarch_info rm;
  if (scb.size() != sizeof(sc) || rmb.size() != sizeof(rm)) return 2;
  std::memcpy(&sc, scb.data(), sizeof(sc));
  std::memcpy(&rm, rmb.data(), sizeof(rm));
  const int nlvl = sc.grid_nlvl, V = 128 * 128 * 128;
  if (occ.size() != (size_t)nlvl * V) return 2;
  // Morton bitfield (loader.cu:20-55) and the coarse tables
  std::vector<uint64_t> bf((size_t)nlvl * V / 64, 0);
  for (int l = 0; l < nlvl; l++)
    for (int x = 0; x < 128; x++)
      for (int y = 0; y < 128; y++)
        for (int z = 0; z < 128; z++)
          if (occ[(size_t)l * V + ((size_t)x * 128 + y) * 128 + z]) {
            const uint64_t bit = (uint64_t)l * V + morton3d(x, y, z);
            bf[bit >> 6] |= 1ull << (bit & 63);
          }
  const std::vector<uint32_t> coarse = build_coarse(occ.data(), nlvl);
  const uint8_t* bf8 = reinterpret_cast<const uint8_t*>(bf.data());
  MarchGrid g;
  g.occ = occ_params(sc);
  g.st = rm.stepsize_info;
  const fvc_stepsize_info& s = g.st;
  // step table rows up to the farthest AABB corner (as the renderer builds it)
  const V3 from = cvt(sc.aabb_from), to = cvt(sc.aabb_to), eye = translation(rm.cam_info.cam2world);
  float need = 0.0f;
  for (int c = 0; c < 8; c++) {
    const V3 q = {(c & 1) ? to.x : from.x, (c & 2) ? to.y : from.y, (c & 4) ? to.z : from.z};
    const V3 v = q - eye;
    need = std::max(need, std::sqrt(dot3(v, v)));
  }
  need = std::min(need, s.far_);
  std::vector<float> tt;
  std::vector<int> off;
  const int rows = build_step_rows(s, need * 1.02f + 4.0f * s.dt_max, tt, off);
  g.tab = step_view(s);
  g.tab.t = tt.data();
  g.tab.off = off.data();
  g.tab.n_rows = rows;
  g.bf64 = bf.data();
  const int W = rm.cam_info.res_x, H = rm.cam_info.res_y;
  long long n_rays = 0, n_samples = 0, n_trips = 0, n_bins = 0, bad = 0;
  std::vector<float> a, b;
  for (int y = 0; y < H; y += stride)
    for (int x = 0; x < W; x += stride) {
      const Ray r = gen_ray(rm.cam_info, (float)x + 0.5f, (float)y + 0.5f);
      float t0, t1;
      if (!ray_aabb(r, from, to, t0, t1)) continue;
      const float far_ = std::min(t1, s.far_);
      const float mf = std::round(step_from_t(s, std::max(t0, s.near_)) - s.near_i);
      const float tstart = t_from_step(s, mf + s.near_i);
      // skipping marcher
      MarchRay mr;
      mr.o = r.o;
      mr.d = r.d;
      mr.slab = ray_slab(r.o, r.d);
      mr.far_ = far_;
      mr.m = (int)mf;
      mr.j = 0;
      mr.t1 = tstart;
      mr.sk = no_skip();
      mr.row = step_row(g.tab, mr.m);
      a.clear();
      int bins = 0;
      for (;;) {
        float ts;
        const int res = march_trip(g, coarse.data(), mr, ts, bins);
        if (res == TRIP_END) break;
        n_trips++;
        if (res == TRIP_SAMPLE) a.push_back(ts);
      }
      // plain stepping
      b.clear();
      float tc = tstart;
      for (;;) {
        const float tn = tc + step_dt(s, tc);
        const float tm = (tc + tn) * 0.5f;
        if (tm > far_) break;
        if (occupied(r.o + tm * r.d, g.occ, bf8)) b.push_back(tc);
        tc = tn;
        n_bins++;
      }
      n_rays++;
      n_samples += (long long)b.size();
      if (a.size() != b.size() || std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) != 0) {
        if (bad < 5)
          std::printf("mismatch at pixel (%d, %d): %zu vs %zu samples\n", x, y, a.size(), b.size());
        bad++;
      }
    }
  std::printf("%lld rays, %lld samples, %lld bins stepped, %lld marcher trips (%.2f per ray), %lld mismatching rays\n",
              n_rays, n_samples, n_bins, n_trips, n_rays ? (double)n_trips / n_rays : 0.0, bad);
  return bad != 0 || n_samples == 0;
}
