// fvc_math.h -- geometry of the render path for device and host code.
//
// Everything here is compiled with -ffp-contract=off: each float operation is
// rounded separately, in the order written, so the marcher, the occupancy
// test and the froxel reprojection are bit-reproducible on any IEEE-754
// machine.  Transcendentals of the cone-angle step schedule use fvc_logf /
// fvc_exp2f (range reduction + series from basic operations only) for the
// same reason.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fvc.h"

#define FVC_HD __host__ __device__ __forceinline__

namespace fvc {

struct V3 { float x, y, z; };
FVC_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
FVC_HD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
FVC_HD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
FVC_HD V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
FVC_HD V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
FVC_HD V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
FVC_HD V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
FVC_HD V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
FVC_HD V3 operator+(V3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
FVC_HD V3 operator-(V3 a, float s) { return {a.x - s, a.y - s, a.z - s}; }
FVC_HD float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
FVC_HD V3 cvt(fvc_float3 f) { return {f.x, f.y, f.z}; }
// helper_math normalize (helper_math.h:1326-1330), invLen = 1/sqrt(.)
FVC_HD V3 normalize(V3 v) { const float inv = 1.0f / sqrtf(dot3(v, v)); return v * inv; }
// a / b correctly rounded, given y = 1 / b correctly rounded (one IEEE
// division shared by every numerator over the same b): q = a*y lies within
// an ulp of a/b, r = a - b*q is exact, and q + r*y rounds to a/b
// (Markstein's theorem), for normal b and quotients that neither overflow nor
// underflow; a zero quotient takes the sign of a*y, as a / b does.
FVC_HD float div_recip(float a, float b, float y) {
  const float q = a * y;
  const float r = fmaf(-q, b, a);
  return copysignf(fmaf(r, y, q), q);
}
FVC_HD float fminf_(float a, float b) { return a < b ? a : b; }
FVC_HD float fmaxf_(float a, float b) { return a > b ? a : b; }

// CameraMatrix::transform (common.h:39-42): three float4 dot products
FVC_HD V3 transform(const fvc_camera_matrix& m, float x, float y, float z, float w) {
  return {x * m.m0.x + y * m.m0.y + z * m.m0.z + w * m.m0.w, x * m.m1.x + y * m.m1.y + z * m.m1.z + w * m.m1.w,
          x * m.m2.x + y * m.m2.y + z * m.m2.z + w * m.m2.w};
}
FVC_HD V3 translation(const fvc_camera_matrix& m) { return {m.m0.w, m.m1.w, m.m2.w}; }

// ---- deterministic log / exp2 ----
FVC_HD float fvc_logf(float x) {
  int e;
  float m = frexpf(x, &e);
  if (m < 0.70710678118654752f) { m = m * 2.0f; e = e - 1; }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  float p = 0.111111111111111111f;
  p = p * s2 + 0.142857142857142857f;
  p = p * s2 + 0.2f;
  p = p * s2 + 0.333333333333333333f;
  p = p * s2 + 1.0f;
  const float lm = (2.0f * s) * p;
  const float ef = (float)e;
  return ef * 0.693145751953125f + (ef * 1.428606765330187045e-06f + lm);
}
FVC_HD float fvc_exp2f(float y) {
  const float n = floorf(y + 0.5f);
  const float f = y - n;
  const float u = f * 0.693147180559945309417f;
  float p = 1.98412698412698413e-4f;
  p = p * u + 1.38888888888888889e-3f;
  p = p * u + 8.33333333333333333e-3f;
  p = p * u + 4.16666666666666667e-2f;
  p = p * u + 0.166666666666666667f;
  p = p * u + 0.5f;
  p = p * u + 1.0f;
  p = p * u + 1.0f;
  return ldexpf(p, (int)n);
}
FVC_HD float fvc_powf(float b, float y) { return fvc_exp2f(y * (fvc_logf(b) * 1.44269504088896340736f)); }

// ---- StepsizeInfo (common.h:93-164) ----
FVC_HD float t_from_step(const fvc_stepsize_info& s, float step) {
  if (s.cone_angle <= 0.0f) return step * s.dt_min;
  if (step >= s.i_max) return s.t_max + (step - s.i_max) * s.dt_max;
  else if (step < s.i_min) return step * s.dt_min;
  else return s.t_min * fvc_powf(1.0f + s.cone_angle, step - s.i_min);
}
FVC_HD float step_from_t(const fvc_stepsize_info& s, float t) {
  if (s.cone_angle <= 0.0f) return t / s.dt_min;
  if (t >= s.t_max) return (t - s.t_max) / s.dt_max + s.i_max;
  else if (t < s.t_min) return t / s.dt_min;
  else return fvc_logf(t / s.t_min) / s.log1p_c + s.i_min;
}
// calculate_stepsize (raymarch_common.h:274-284)
FVC_HD float step_dt(const fvc_stepsize_info& s, float t) {
  // the operands are positive and finite: the native minimum / maximum
  return fminf(fmaxf(t * s.cone_angle, s.dt_min), s.dt_max);
}
// valid_t_from_t (raymarch_common.h:286-296)
FVC_HD float valid_t_from_t(const fvc_stepsize_info& s, float t) {
  return t_from_step(s, roundf(step_from_t(s, fmaxf_(t, s.near_)) - s.near_i) + s.near_i);
}

// ---- Morton bitfield occupancy (raymarch_common.h:98-110, 660-710) ----
FVC_HD uint32_t expand_bits_3d(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
FVC_HD uint32_t morton3d(uint32_t x, uint32_t y, uint32_t z) {
  return expand_bits_3d(x) | (expand_bits_3d(y) << 1) | (expand_bits_3d(z) << 2);
}

// Precomputed per-scene constants of grid_occupied_at.
struct OccParams {
  V3 unit_off;    // aabb_from / 2^(nlvl-1)
  V3 unit_ext;    // (aabb_to - aabb_from) / 2^(nlvl-1)
  V3 unit_inv;    // 1 / unit_ext, exact when pow2
  V3 box_c;       // world centre of the cascades (unit coordinate 0.5)
  V3 box_e;       // world extent of cascade 0 (unit_ext, or 1 when normalized)
  float be_max;   // largest box_e component
  float bc_max;   // largest |box_c| component
  int nlvl;
  int normalized;
  int pow2;       // every unit_ext component is a power of two
};

// Byte index (<0: outside every cascade) and bit of the occupancy cell of p.
// The cell index is clamped to [0,127] (the float expression can round to
// exactly 1.0 at the upper face of a cascade).
FVC_HD int occ_cell(V3 p, const OccParams& o, int& bit) {
  const V3 u = o.normalized ? p : (o.pow2 ? (p - o.unit_off) * o.unit_inv : (p - o.unit_off) / o.unit_ext);
  const float ax = fabsf(u.x - 0.5f), ay = fabsf(u.y - 0.5f), az = fabsf(u.z - 0.5f);
  const float mx = fmaxf_(fmaxf_(ax, ay), az);
  int e;
  frexpf(mx, &e);
  const int mip = e + 1 > 0 ? e + 1 : 0;
  if (mip >= o.nlvl) return -1;
  const float scm = ldexpf(1.0f, -mip);
  const float px = (u.x - 0.5f) * scm + 0.5f, py = (u.y - 0.5f) * scm + 0.5f, pz = (u.z - 0.5f) * scm + 0.5f;
  int ix = (int)(px * 128.0f), iy = (int)(py * 128.0f), iz = (int)(pz * 128.0f);
  ix = ix < 0 ? 0 : (ix > 127 ? 127 : ix);
  iy = iy < 0 ? 0 : (iy > 127 ? 127 : iy);
  iz = iz < 0 ? 0 : (iz > 127 ? 127 : iz);
  const uint32_t idx = morton3d((uint32_t)ix, (uint32_t)iy, (uint32_t)iz);
  bit = (int)(idx & 7u);
  return (int)(idx >> 3) + mip * (128 * 128 * 128 / 8);
}
FVC_HD bool occupied(V3 p, const OccParams& o, const uint8_t* bf) {
  int bit;
  const int byte = occ_cell(p, o, bit);
  return byte >= 0 && ((bf[byte] >> bit) & 1);
}

// ---- apply_contraction (raymarch_common.h:610-653) ----
struct ContractParams {
  int type;
  V3 from, ext;    // AABB contraction: scene aabb; warps: contraction aabb
  V3 inv;          // 1 / ext, exact when pow2
  int pow2;
};
FVC_HD V3 contract(V3 p, const ContractParams& c) {
  const V3 q = c.pow2 ? (p - c.from) * c.inv : (p - c.from) / c.ext;
  if (c.type == FVC_CONTRACTION_AABB) return q;
  if (c.type == FVC_CONTRACTION_WARP_AABB_LINF || c.type == FVC_CONTRACTION_WARP_AABB_L2) {
    const V3 u = q * 2.0f - 1.0f;
    float n;
    if (c.type == FVC_CONTRACTION_WARP_AABB_LINF) n = fmaxf_(fmaxf_(fabsf(u.x), fabsf(u.y)), fabsf(u.z));
    else n = sqrtf(dot3(u, u));
    const float f = fmaxf_(n, 1.0f);
    const V3 cp = (2.0f - 1.0f / f) * (u / f);
    return cp / 4.0f + 0.5f;
  }
  return p;
}

// ---- rays (raymarch_common.h:190-251, 482-532; renderer.cu:86-134) ----
struct Ray { V3 o, d; };
FVC_HD Ray gen_ray(const fvc_camera_info& cam, float sx, float sy) {
  const float gl = cam.is_open_gl ? -1.0f : 1.0f;
  const float dx = (sx - cam.principal.x) / cam.focal.x;
  const float dy = (sy - cam.principal.y) / cam.focal.y * gl;
  Ray r;
  r.o = translation(cam.cam2world);
  r.d = normalize(transform(cam.cam2world, dx, dy, gl, 0.0f));
  return r;
}
FVC_HD void swapf(float& a, float& b) { const float c = a; a = b; b = c; }
FVC_HD bool ray_aabb(const Ray& r, V3 from, V3 to, float& nr, float& fr) {
  float tmin = (from.x - r.o.x) / r.d.x, tmax = (to.x - r.o.x) / r.d.x;
  if (tmin > tmax) swapf(tmin, tmax);
  float tymin = (from.y - r.o.y) / r.d.y, tymax = (to.y - r.o.y) / r.d.y;
  if (tymin > tymax) swapf(tymin, tymax);
  if (tmin > tymax || tymin > tmax) { nr = fr = 1e10f; return false; }
  if (tymin > tmin) tmin = tymin;
  if (tymax < tmax) tmax = tymax;
  float tzmin = (from.z - r.o.z) / r.d.z, tzmax = (to.z - r.o.z) / r.d.z;
  if (tzmin > tzmax) swapf(tzmin, tzmax);
  if (tmin > tzmax || tzmin > tmax) { nr = fr = 1e10f; return false; }
  if (tzmin > tmin) tmin = tzmin;
  if (tzmax < tmax) tmax = tzmax;
  nr = tmin; fr = tmax;
  return true;
}

// ---- cam2froxel (raymarch_common.h:316-329) ----
FVC_HD V3 cam2froxel(V3 c, const fvc_camera_info& cam, const fvc_stepsize_info& st, float& t) {
  const float gl = cam.is_open_gl ? -1.0f : 1.0f;
  const float z = c.z * gl;
  t = sqrtf(dot3(c, c));
  return {c.x / z * cam.focal.x + cam.principal.x, c.y * gl / z * cam.focal.y + cam.principal.y,
          (step_from_t(st, t) - st.near_i) * (c.z < 0.0f ? -1.0f : 1.0f)};
}

FVC_HD V3 unit_to_01(V3 d) { return (d + 1.0f) / 2.0f; }

// x is a positive power of two (then a / x == a * (1/x) exactly)
FVC_HD bool is_pow2f(float x) {
  int e;
  return x > 0.0f && frexpf(x, &e) == 0.5f;
}

// grid_occupied_at constants of a scene (raymarch_common.h:660-666)
inline OccParams occ_params(const fvc_scene_info& sc) {
  OccParams o{};
  const float S = (float)(1 << (sc.grid_nlvl - 1));
  const V3 from = cvt(sc.aabb_from), to = cvt(sc.aabb_to);
  o.unit_off = from / S;
  o.unit_ext = (to - from) / S;
  o.unit_inv = V3{1.0f / o.unit_ext.x, 1.0f / o.unit_ext.y, 1.0f / o.unit_ext.z};
  o.pow2 = is_pow2f(o.unit_ext.x) && is_pow2f(o.unit_ext.y) && is_pow2f(o.unit_ext.z);
  o.nlvl = sc.grid_nlvl;
  o.normalized = sc.normalized;
  o.box_e = sc.normalized ? V3{1.0f, 1.0f, 1.0f} : o.unit_ext;
  o.box_c = sc.normalized ? V3{0.5f, 0.5f, 0.5f} : o.unit_off + 0.5f * o.unit_ext;
  o.be_max = std::max(std::max(o.box_e.x, o.box_e.y), o.box_e.z);
  o.bc_max = std::max(std::max(std::fabs(o.box_c.x), std::fabs(o.box_c.y)), std::fabs(o.box_c.z));
  return o;
}

}  // namespace fvc
