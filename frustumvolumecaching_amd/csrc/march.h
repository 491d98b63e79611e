// march.h -- the occupancy-grid marcher with empty-space skipping, shared by
// the full renderer, the cache renderer and the debug march.
//
// The reference steps through every bin and tests the occupancy bitfield at
// each bin midpoint (performance_renderer.cu:39-62).  The bin sequence
// t_{k+1} = t_k + dt(t_k) must stay exactly that float recurrence, and every
// emitted sample must be a bin whose midpoint grid_occupied_at
// (raymarch_common.h:680-710) calls occupied.  The marcher here keeps both and
// only avoids work that cannot change the answer:
//
//  * Occupancy is tested hierarchically: a coarse table per cascade (one bit
//    per block of 32^3, 16^3 and 8^3 cells) held in LDS, then the 64-bit
//    Morton word of the 4^3 block, then its byte (a 2^3 block), then the
//    cell bit.  The largest level that reads zero names an empty box of 32,
//    16, 8, 4, 2 or 1 cells.
//  * The ray's parameter interval through that box, shrunk on every face by a
//    margin far above the float error of the lookup's position arithmetic
//    (box_skip), becomes the skip interval: every bin midpoint inside it
//    would have been classified into the same empty box, so skipping those
//    bins leaves the emitted samples bit-identical.  Boxes never straddle
//    two cascades (the inner cascade covers cells 32..95, multiples of 32).
//  * Skipped bins are jumped over in the step table (rows of precomputed bin
//    starts, built with the device's own recurrence), so a skip costs one
//    short table lookup instead of one trip per bin.
//
// One marcher trip (march_trip) steps one bin and, when the bin is empty,
// also performs the jump, so a lane spends one loop trip per empty box and
// one per occupied sample.
#pragma once
#include <cmath>
#include <vector>

#include "fvc_internal.h"

namespace fvc {

// Coarse empty-space table, per cascade: one bit per block of 8^3 cells
// (16^3 blocks, 128 words), then per block of 16^3 cells (8^3, 16 words), then
// per block of 32^3 cells (4^3, 2 words); a bit is set when any cell of the
// block is occupied.
static const int COARSE8_WORDS = 128, COARSE16_WORDS = 16, COARSE32_WORDS = 2;
static const int COARSE_WORDS = COARSE8_WORDS + COARSE16_WORDS + COARSE32_WORDS;  // uint32 words per cascade
static const int CASCADE_WORDS64 = 128 * 128 * 128 / 64;

// Coarse table of a linear uint8 occupancy grid [nlvl][x][y][z] of 128^3.
inline std::vector<uint32_t> build_coarse(const uint8_t* occ, int nlvl) {
  std::vector<uint32_t> c((size_t)nlvl * COARSE_WORDS, 0u);
  for (int l = 0; l < nlvl; l++) {
    uint32_t* w = c.data() + (size_t)l * COARSE_WORDS;
    const uint8_t* g = occ + (size_t)l * 128 * 128 * 128;
    for (int x = 0; x < 128; x++)
      for (int y = 0; y < 128; y++)
        for (int z = 0; z < 128; z++) {
          if (!g[((size_t)x * 128 + y) * 128 + z]) continue;
          const int b8 = ((x >> 3) * 16 + (y >> 3)) * 16 + (z >> 3);
          const int b16 = ((x >> 4) * 8 + (y >> 4)) * 8 + (z >> 4);
          const int b32 = ((x >> 5) * 4 + (y >> 5)) * 4 + (z >> 5);
          w[b8 >> 5] |= 1u << (b8 & 31);
          w[COARSE8_WORDS + (b16 >> 5)] |= 1u << (b16 & 31);
          w[COARSE8_WORDS + COARSE16_WORDS + (b32 >> 5)] |= 1u << (b32 & 31);
        }
  }
  return c;
}

__device__ __forceinline__ void load_coarse(const uint32_t* __restrict__ g, int nlvl, uint32_t* lds) {
  for (int i = threadIdx.x; i < nlvl * COARSE_WORDS; i += blockDim.x) lds[i] = g[i];
  __syncthreads();
}

struct Skip { float lo, hi; };
FVC_HD Skip no_skip() { return Skip{1.0f, 0.0f}; }
FVC_HD bool in_skip(const Skip& s, float tm) { return tm >= s.lo && tm <= s.hi; }

// unit-cube coordinate of a world point (world_to_aabb_unit, raymarch_common.h:660-666)
FVC_HD V3 occ_unit(V3 p, const OccParams& o) {
  if (o.normalized) return p;
  if (o.pow2) return (p - o.unit_off) * o.unit_inv;  // exact: the extent is a power of two
  return (p - o.unit_off) / o.unit_ext;
}

// Per-ray terms of box_skip: 1/d with +-1e30 for (near) axis-parallel
// directions (the slab test then needs no special case), o/d, |1/d| and the
// largest |o| component.
struct RaySlab {
  V3 inv, oinv, ainv;
  float omax;
};
FVC_HD RaySlab ray_slab(V3 o, V3 d) {
  RaySlab r;
  const float dd[3] = {d.x, d.y, d.z}, oo[3] = {o.x, o.y, o.z};
  float iv[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    iv[a] = fabsf(dd[a]) < 1e-30f ? (dd[a] < 0.0f ? -1e30f : 1e30f) : 1.0f / dd[a];
  }
  r.inv = V3{iv[0], iv[1], iv[2]};
  r.oinv = V3{oo[0] * iv[0], oo[1] * iv[1], oo[2] * iv[2]};
  r.ainv = V3{fabsf(iv[0]), fabsf(iv[1]), fabsf(iv[2])};
  r.omax = fmaxf(fmaxf(fabsf(oo[0]), fabsf(oo[1])), fabsf(oo[2]));
  return r;
}

// Parameter interval of the ray inside the world box of cells [c, c+g) of
// cascade mip, each face pulled in by a margin.  The lower face of cell i
// lies at box_c + (i - 64) * box_e * 2^(mip-7).  The interval is only ever
// used to skip bins, so it is computed with fused multiply-adds; its margin
// absorbs that rounding.
//
// Margin: the lookup of a point p = o + tm*d rounds at most about ten times
// (the product and sum of p, the unit transform, the cascade rescale), each
// by 2^-24 relative to a magnitude bounded by |o| + |t| + the cascade's
// largest coordinate, so its cell boundaries are uncertain by well under
// 1e-6 of that magnitude.  The margin is 4e-5 of it (over 60x the error).
// A face is pulled in by margin/|d_a| in t; a box thinner than two margins
// along an axis leaves an empty interval (no skip).
FVC_HD Skip box_skip(int cx, int cy, int cz, int g, int mip, const OccParams& o, const RaySlab& rs) {
  const float s = ldexpf(1.0f, mip - 7);
  const float gs = (float)g;
  const int ci[3] = {cx, cy, cz};
  const float bc[3] = {o.box_c.x, o.box_c.y, o.box_c.z}, be[3] = {o.box_e.x, o.box_e.y, o.box_e.z};
  const float iv[3] = {rs.inv.x, rs.inv.y, rs.inv.z}, oi[3] = {rs.oinv.x, rs.oinv.y, rs.oinv.z};
  const float ai[3] = {rs.ainv.x, rs.ainv.y, rs.ainv.z};
  float tn[3], tf[3];
  float tout0 = 3.0e38f;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float w = be[a] * s;
    const float lo = fmaf((float)(ci[a] - 64), w, bc[a]);
    const float hi = fmaf(gs, w, lo);
    const float ta = fmaf(lo, iv[a], -oi[a]), tb = fmaf(hi, iv[a], -oi[a]);
    // entry and exit along the axis (the values are finite: |1/d| <= 1e30)
    tn[a] = fminf(ta, tb);
    tf[a] = fmaxf(ta, tb);
    tout0 = fminf(tout0, tf[a]);
  }
  // the cascade's largest coordinate, bounded by 64 w_max + |box_c|_max
  const float cmag = fmaf(64.0f * s, o.be_max, o.bc_max);
  const float m = 4e-5f * (rs.omax + fabsf(tout0) + cmag);
  float tin = -3.0e38f, tout = 3.0e38f;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    tin = fmaxf(tin, fmaf(m, ai[a], tn[a]));
    tout = fminf(tout, fmaf(-m, ai[a], tf[a]));
  }
  return Skip{tin, tout};
}

// grid_occupied_at<128, true> (raymarch_common.h:680-710) for p = o + tm*d.
// When the answer is "empty" the skip interval of the largest empty box
// around the cell is stored in sk (left unchanged outside every cascade).
FVC_HD bool occ_query(V3 p, const OccParams& o, const uint64_t* __restrict__ bf64, const uint32_t* coarse,
                      const RaySlab& rs, Skip& sk) {
  const V3 u = occ_unit(p, o);
  const float ax = fabsf(u.x - 0.5f), ay = fabsf(u.y - 0.5f), az = fabsf(u.z - 0.5f);
  const float mx = fmaxf(fmaxf(ax, ay), az);
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
  const uint32_t* cw = coarse + mip * COARSE_WORDS;
  const int b8 = (((ix >> 3) * 16) + (iy >> 3)) * 16 + (iz >> 3);
  int g;
  if ((cw[b8 >> 5] >> (b8 & 31)) & 1u) {
    const uint32_t idx = morton3d((uint32_t)ix, (uint32_t)iy, (uint32_t)iz);
    const uint64_t w = bf64[(idx >> 6) + (uint32_t)mip * CASCADE_WORDS64];
    const uint32_t byte = (uint32_t)(w >> (idx & 56u)) & 0xffu;
    if ((byte >> (idx & 7u)) & 1u) return true;
    g = w == 0 ? 4 : (byte == 0 ? 2 : 1);
  } else {
    const int b16 = (((ix >> 4) * 8) + (iy >> 4)) * 8 + (iz >> 4);
    const int b32 = (((ix >> 5) * 4) + (iy >> 5)) * 4 + (iz >> 5);
    const bool e16 = !((cw[COARSE8_WORDS + (b16 >> 5)] >> (b16 & 31)) & 1u);
    const bool e32 = !((cw[COARSE8_WORDS + COARSE16_WORDS + (b32 >> 5)] >> (b32 & 31)) & 1u);
    g = e32 ? 32 : (e16 ? 16 : 8);
  }
  sk = box_skip(ix & -g, iy & -g, iz & -g, g, mip, o, rs);
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
  // coefficients of a cheap estimate of the continuous step index of t
  // (step_from_t - near_i); only used as a starting guess that is verified
  int cone;
  float near_i, lin_inv, t_min, t_max, i_max, log_a, log_b, inv_dt_max;
};

FVC_HD float fast_log2(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_logf(x);
#else
  return log2f(x);
#endif
}

FVC_HD float step_estimate(const StepTableView& v, float t) {
  if (!v.cone || t < v.t_min) return t * v.lin_inv;
  if (t >= v.t_max) return (t - v.t_max) * v.inv_dt_max + v.i_max;
  return fast_log2(t) * v.log_a + v.log_b;
}

// Where row m lies in the table: base = its first entry, last = the largest
// k whose bin end is in the row (base = -1: no row).  A ray reads it once
// per launch (MarchRay), not at every jump.
struct StepRow { int base, last; };
FVC_HD StepRow step_row(const StepTableView& tab, int m) {
  if (m < 0 || m >= tab.n_rows) return StepRow{-1, -1};
  const int base = tab.off[m];
  return StepRow{base, tab.off[m + 1] - base - 2};
}

// The bin before j (midpoint in [lo, hi]) is skipped; advance to the first
// bin k >= j whose midpoint exceeds hi.  Every bin in between has its
// midpoint in [that midpoint, hi], inside the same empty box.  t1 and j keep
// the invariant t1 == T[m][j].  Returns the number of bins skipped.
//
// The continuous schedule puts the answer within a bin of its estimate (the
// float recurrence drifts by a small fraction of a bin over a row), so one
// round of four independent loads around the estimate decides it; a
// bisection or a forward walk covers the rare misses.
FVC_HD int table_jump(const StepTableView& tab, StepRow rw, int m, int& j, float& t1, float hi) {
  if (rw.base < 0) return 0;
  const int last = rw.last;
  if (j >= last) return 0;
  const float* row = tab.t + rw.base;
  // first k with midpoint(k) > hi: k + m + near_i + 1/2 > step(hi)
  int k = (int)floorf(step_estimate(tab, hi) - tab.near_i - (float)m - 0.5f) + 1;
  k = k < j ? j : (k > last ? last : k);
  const float r0 = row[k > 0 ? k - 1 : 0], r1 = row[k], r2 = row[k + 1], r3 = row[k + 1 < last ? k + 2 : last + 1];
  if (k > j && !((r0 + r1) * 0.5f <= hi)) {
    // overshoot: the answer lies in [j, k-1]; bisect
    int a = j, b = k - 1;
    while (a < b) {
      const int c = (a + b) >> 1;
      if ((row[c] + row[c + 1]) * 0.5f <= hi) a = c + 1;
      else b = c;
    }
    k = a;
  } else if ((r1 + r2) * 0.5f <= hi) {
    if (k < last) {
      k++;
      if ((r2 + r3) * 0.5f <= hi)
        while (k < last && (row[k] + row[k + 1]) * 0.5f <= hi) k++;
    }
  }
  const int skipped = k - j;
  j = k;
  t1 = row[k];
  return skipped;
}

// Estimate coefficients of a stepsize (step_from_t, common.h:93-164, with
// log(t / t_min) / log1p(c) = (log2 t - log2 t_min) * ln2 / log1p(c)).
inline StepTableView step_view(const fvc_stepsize_info& st) {
  StepTableView v{};
  v.cone = st.cone_angle > 0.0f;
  v.near_i = st.near_i;
  v.lin_inv = (float)(1.0 / (double)st.dt_min);
  v.t_min = st.t_min;
  v.t_max = st.t_max;
  v.i_max = st.i_max;
  v.inv_dt_max = (float)(1.0 / (double)st.dt_max);
  if (v.cone) {
    const double a = 0.69314718055994530942 / (double)st.log1p_c;
    v.log_a = (float)a;
    v.log_b = (float)((double)st.i_min - std::log2((double)st.t_min) * a);
  }
  return v;
}

// Rows of the step table for rays whose midpoints stay below lim: row m
// starts at t_from_step(m + near_i) and runs the device recurrence until
// the bin midpoint passes lim.  Returns the number of rows, 0 when the
// table would exceed 200000 rows or 64 Mi entries (the marchers then step
// every bin).
inline int build_step_rows(const fvc_stepsize_info& st, float lim, std::vector<float>& t, std::vector<int>& off) {
  t.clear();
  off.clear();
  const double rows_d = std::ceil((double)step_from_t(st, lim) - (double)st.near_i) + 2.0;
  if (!(rows_d > 0 && rows_d < 200000)) return 0;
  const int rows = (int)rows_d;
  off.reserve(rows + 1);
  for (int m = 0; m < rows; m++) {
    off.push_back((int)t.size());
    float t1 = t_from_step(st, (float)m + st.near_i);
    t.push_back(t1);
    for (int k = 0; k < 4 * rows + 4096; k++) {
      const float t0 = t1;
      t1 = t0 + step_dt(st, t0);
      t.push_back(t1);
      if ((t0 + t1) * 0.5f > lim) break;
    }
    if (t.size() > ((size_t)64 << 20)) {
      t.clear();
      off.clear();
      return 0;
    }
  }
  off.push_back((int)t.size());
  return rows;
}

// Read-only inputs of the marcher.
struct MarchGrid {
  OccParams occ;
  fvc_stepsize_info st;
  StepTableView tab;
  const uint64_t* bf64;   // Morton bitfield, 64 cells per word
};

// A ray between marcher trips: the next bin starts at t1 == T[m][j].
struct MarchRay {
  V3 o, d;
  RaySlab slab;
  float far_;
  int m, j;
  float t1;
  Skip sk;
  StepRow row;  // step_row(tab, m)
};

enum { TRIP_EMPTY = 0, TRIP_SAMPLE = 1, TRIP_END = 2 };

// One marcher trip: step one bin (the loop body of
// sample_performance_kernel, performance_renderer.cu:39-62).  TRIP_SAMPLE:
// the bin starting at t0 has an occupied midpoint.  TRIP_END: the midpoint
// passed far.  TRIP_EMPTY: the bin was empty; when it lies in an empty box
// the bins up to the box exit are jumped over as well.  bins counts the bins
// stepped or jumped.
FVC_HD int march_trip(const MarchGrid& g, const uint32_t* coarse, MarchRay& r, float& t0, int& bins) {
  t0 = r.t1;
  r.t1 = t0 + step_dt(g.st, t0);
  r.j++;
  bins++;
  const float tm = (t0 + r.t1) * 0.5f;
  if (tm > r.far_) return TRIP_END;
  if (!in_skip(r.sk, tm)) {
    if (occ_query(r.o + tm * r.d, g.occ, g.bf64, coarse, r.slab, r.sk)) return TRIP_SAMPLE;
    if (!in_skip(r.sk, tm)) return TRIP_EMPTY;
  }
  bins += table_jump(g.tab, r.row, r.m, r.j, r.t1, r.sk.hi);
  return TRIP_EMPTY;
}

}  // namespace fvc
