// march.h -- occupancy queries with empty-space skipping for the marchers.
//
// The reference steps through every bin and tests the occupancy bitfield at
// each bin midpoint (performance_renderer.cu:39-62).  The bin sequence
// t_{k+1} = t_k + dt(t_k) must stay exactly that float recurrence, so the
// marchers here still run it step by step, but skip the occupancy lookup of
// every step whose midpoint lies inside a box already known to be empty:
//
//  * a coarse table per cascade (16^3 blocks of 8^3 cells, one bit each,
//    "any cell occupied") held in LDS answers most queries in empty space;
//  * on an empty block or an empty 128^3 cell the ray's parameter interval
//    through that box, shrunk by 1/64 of the box size on every face, becomes
//    the skip interval.
//
// The shrink margin is orders of magnitude above the float error of the
// lookup's position arithmetic (a few ulps of |p|), so every skipped midpoint
// would have been classified into the same empty box by grid_occupied_at
// (raymarch_common.h:680-710): the emitted samples stay bit-identical.
// Cascade boundaries are aligned to the 8-cell blocks (level l-1 covers cells
// 32..95 of level l), so a box never straddles two cascades.
#pragma once
#include "fvc_internal.h"

namespace fvc {

static const int COARSE_WORDS = 16 * 16 * 16 / 32;  // uint32 words per cascade

struct Skip { float lo, hi; };
__device__ __forceinline__ Skip no_skip() { return Skip{1.0f, 0.0f}; }
__device__ __forceinline__ bool in_skip(const Skip& s, float tm) { return tm >= s.lo && tm <= s.hi; }

__device__ __forceinline__ void load_coarse(const uint32_t* __restrict__ g, int nlvl, uint32_t* lds) {
  for (int i = threadIdx.x; i < nlvl * COARSE_WORDS; i += blockDim.x) lds[i] = g[i];
  __syncthreads();
}

// unit-cube coordinate of a world point (world_to_aabb_unit, raymarch_common.h:660-666)
__device__ __forceinline__ V3 occ_unit(V3 p, const OccParams& o) {
  if (o.normalized) return p;
  if (o.pow2) return (p - o.unit_off) * o.unit_inv;  // exact: the extent is a power of two
  return (p - o.unit_off) / o.unit_ext;
}

// Parameter interval of the ray inside the world box of cells [c0, c0+g) of
// cascade mip (per axis), each face pulled in by 1/64 of the box size.
__device__ __forceinline__ Skip box_skip(int cx, int cy, int cz, int g, int mip, const OccParams& o, V3 ro, V3 inv_d) {
  const float sc = ldexpf(1.0f, mip) * (1.0f / 128.0f);
  const V3 boff = o.normalized ? V3{0.0f, 0.0f, 0.0f} : o.unit_off;
  const V3 bext = o.normalized ? V3{1.0f, 1.0f, 1.0f} : o.unit_ext;
  // unit coordinate of cell index i: (i/128 - 0.5) * 2^mip + 0.5
  const float half = 0.5f - 0.5f * ldexpf(1.0f, mip);
  float tin = -3.0e38f, tout = 3.0e38f;
  const int ci[3] = {cx, cy, cz};
  const float off[3] = {boff.x, boff.y, boff.z}, ext[3] = {bext.x, bext.y, bext.z};
  const float oo[3] = {ro.x, ro.y, ro.z}, id[3] = {inv_d.x, inv_d.y, inv_d.z};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float lo = off[a] + (half + (float)ci[a] * sc) * ext[a];
    const float hi = off[a] + (half + (float)(ci[a] + g) * sc) * ext[a];
    const float m = (hi - lo) * (1.0f / 64.0f);
    const float l2 = lo + m, h2 = hi - m;
    if (id[a] == 0.0f || fabsf(id[a]) > 1e30f) {  // axis-parallel ray: coordinate is constant
      if (!(oo[a] >= l2 && oo[a] <= h2)) return no_skip();
      continue;
    }
    float ta = (l2 - oo[a]) * id[a], tb = (h2 - oo[a]) * id[a];
    if (ta > tb) { const float t = ta; ta = tb; tb = t; }
    // the margin must dwarf the rounding of o + t*d at this distance
    if (m < 1e-5f * (fabsf(oo[a]) + fabsf(tb))) return no_skip();
    tin = ta > tin ? ta : tin;
    tout = tb < tout ? tb : tout;
  }
  return Skip{tin, tout};
}

// grid_occupied_at<128, true> (raymarch_common.h:680-710) for p = o + tm*d,
// updating the skip interval when the answer is "empty".
__device__ __forceinline__ bool occ_query(V3 p, const OccParams& o, const uint8_t* __restrict__ bf,
                                          const uint32_t* coarse, V3 ro, V3 inv_d, Skip& sk) {
  const V3 u = occ_unit(p, o);
  const float ax = fabsf(u.x - 0.5f), ay = fabsf(u.y - 0.5f), az = fabsf(u.z - 0.5f);
  const float mx = fmaxf_(fmaxf_(ax, ay), az);
  int e;
  frexpf(mx, &e);
  const int mip = e + 1 > 0 ? e + 1 : 0;
  if (mip >= o.nlvl) return false;
  const float scm = ldexpf(1.0f, -mip);
  const float px = (u.x - 0.5f) * scm + 0.5f, py = (u.y - 0.5f) * scm + 0.5f, pz = (u.z - 0.5f) * scm + 0.5f;
  int ix = (int)(px * 128.0f), iy = (int)(py * 128.0f), iz = (int)(pz * 128.0f);
  ix = ix < 0 ? 0 : (ix > 127 ? 127 : ix);
  iy = iy < 0 ? 0 : (iy > 127 ? 127 : iy);
  iz = iz < 0 ? 0 : (iz > 127 ? 127 : iz);
  const int cb = (((ix >> 3) * 16) + (iy >> 3)) * 16 + (iz >> 3);
  if (!((coarse[mip * COARSE_WORDS + (cb >> 5)] >> (cb & 31)) & 1u)) {
    sk = box_skip(ix & ~7, iy & ~7, iz & ~7, 8, mip, o, ro, inv_d);
    return false;
  }
  const uint32_t idx = morton3d((uint32_t)ix, (uint32_t)iy, (uint32_t)iz);
  if ((bf[(idx >> 3) + mip * (128 * 128 * 128 / 8)] >> (idx & 7)) & 1) return true;
  sk = box_skip(ix, iy, iz, 1, mip, o, ro, inv_d);
  return false;
}

// Step table: row m holds the bin starts of a ray that starts at
// t_from_step(m + near_i) (valid_t_from_t, raymarch_common.h:286-296), i.e.
// T[m][0] = t_from_step(m + near_i), T[m][j+1] = T[m][j] + dt(T[m][j]),
// built on the host with the same float operations as the device marcher.
// A ray's bin sequence depends only on m, so a ray at bin j of row m can
// jump ahead in the table without re-running the recurrence.
struct StepTableView {
  const float* t;    // concatenated rows
  const int* off;    // row m = t[off[m] .. off[m+1])
  int n_rows;        // 0: no table (plain stepping)
};

// The bin before j (midpoint in the skip interval) was skipped; advance to
// the first bin k >= j whose midpoint exceeds hi.  Every bin in between has
// its midpoint in [current midpoint, hi], inside the same empty box.  t1 and
// j keep the invariant t1 == T[m][j].  Returns the number of bins skipped.
__device__ __forceinline__ int table_jump(const StepTableView& tab, const fvc_stepsize_info& st, int m, int& j,
                                          float& t1, float hi) {
  if (m < 0 || m >= tab.n_rows) return 0;
  const int base = tab.off[m];
  const int len = tab.off[m + 1] - base;
  if (j + 2 >= len) return 0;
  const float* row = tab.t + base;
  // estimate from the continuous schedule, two bins early, then verify
  int k = (int)(step_from_t(st, hi) - st.near_i) - m - 2;
  if (k > len - 2) k = len - 2;
  if (k <= j || !((row[k - 1] + row[k]) * 0.5f <= hi)) k = j;
  while (k < len - 2 && (row[k] + row[k + 1]) * 0.5f <= hi) k++;
  const int skipped = k - j;
  j = k;
  t1 = row[k];
  return skipped;
}

__device__ __forceinline__ V3 recip(V3 d) {
  return V3{d.x != 0.0f ? 1.0f / d.x : 0.0f, d.y != 0.0f ? 1.0f / d.y : 0.0f, d.z != 0.0f ? 1.0f / d.z : 0.0f};
}

}  // namespace fvc
