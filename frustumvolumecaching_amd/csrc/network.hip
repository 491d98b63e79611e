// network.hip -- multiresolution hash-grid encoding fused with the density /
// latent MLP and the view-dependent colour MLP (NerfNetwork<half>::inference,
// inferenceLatents, inferenceHead; nerf_network.h:160-266).
//
// Hash-grid conventions follow the tcnn GridEncoding the reference calls
// (nerf_network.h:90, 181): per level l, scale = exp2(l*log2(b))*N_min - 1,
// resolution = ceil(scale)+1, entries = min(next_multiple(res^3,8), T);
// dense index x + y*res + z*res^2 while res^3 <= T, otherwise the hash
// x ^ y*2654435761 ^ z*805459861; index mod entries; trilinear weights over
// the 8 corners, accumulated in FP32 and rounded to fp16.
//
// Work split: one wave64 evaluates 16 samples per step.  Lane l computes
// levels g, g+4, g+8, g+12 (g = l/16) of sample l%16 -- the 8 k-values of the
// first layer's B fragment, whose weight columns are permuted to match on the
// host -- so the gathered features feed the MFMA chain in mlp_mfma.h
// directly from registers.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "mlp_mfma.h"

namespace fvc {

// ------------------------------------------------------------------ layout
// Split models (init_viewdep, nerf_network.h:120-158) additionally need:
// mlp_first_head on 32 inputs (16 density-MLP outputs + SH4), 64 wide, 1 or
// 2 hidden layers, 8 latents; the custom head's supported instance
// (my_fully_fused_mlp.cu:1363-1421: width 64, SH degrees 4 and 4, frequency
// degree 4 or 0); latent_offset 0 exactly when the first latent is the
// density; FP16 arithmetic (the custom head is half only, 1271-1281).
bool net_layout(const fvc_net_desc& d, GridParams& g, int64_t& n_pos, int64_t& n_base, int64_t& n_head,
                int64_t* n_first) {
  if (d.n_levels != 16 || d.n_features_per_level != 2 || d.log2_hashmap_size < 10 || d.log2_hashmap_size > 24)
    return false;
  if (d.base_width != 64 || d.base_n_hidden_layers != 2 || d.base_n_output < 1 || d.base_n_output > 16) return false;
  if (d.head_width != 64 || d.head_n_hidden_layers != 1 || d.sh_degree != 4) return false;
  // Plain models (NerfNetwork::init): the reference keeps
  // m_first_latent_is_density = true, so latent_offset() = 0 and the head
  // reads the first n_output rows, log sigma first (nerf_network.h:93, 214,
  // 334-342); the cache stores the density in latent 0 (DensityIntermediates,
  // eval.cu:216).  latent_offset 1 (log sigma in row 0, 8 latents after it,
  // the density in its own cache array) is the synthetic workload layout.
  if (!d.split_head) {
    const bool ref = d.latent_offset == 0 && d.first_latent_is_density && d.n_latents == d.base_n_output;
    const bool sep = d.latent_offset == 1 && !d.first_latent_is_density && d.n_latents == 8 && d.base_n_output >= 9;
    if (!ref && !sep) return false;
  } else if (d.n_latents != 8 || d.base_n_output < 9) {
    return false;
  }
  if (d.split_head &&
      (d.first_head_width != 64 || d.first_head_n_hidden_layers < 1 || d.first_head_n_hidden_layers > 2 ||
       d.first_head_n_input != 32 || d.first_sh_degree != 4 ||
       (d.freq_encoding_degree != 0 && d.freq_encoding_degree != 4) ||
       d.latent_offset != (d.first_latent_is_density ? 0 : 1) || d.precision != FVC_PREC_FP16))
    return false;
  std::memset(&g, 0, sizeof(g));
  g.n_levels = d.n_levels;
  const uint32_t T = 1u << d.log2_hashmap_size;
  uint32_t off = 0;
  for (int l = 0; l < d.n_levels; l++) {
    const double s = std::exp2(l * std::log2((double)d.per_level_scale)) * d.base_resolution - 1.0;
    const float sf = (float)s;
    const uint32_t r = (uint32_t)std::ceil((double)sf) + 1u;
    const uint64_t dense = (uint64_t)r * r * r;
    uint64_t p = (dense + 7) / 8 * 8;
    if (p > T) p = T;
    g.lv[l].scale = sf;
    g.lv[l].res = r;
    g.lv[l].res2 = r * r;
    g.lv[l].size = (uint32_t)p;
    g.lv[l].offset = off;
    if (dense <= p) g.dense_mask |= 1u << l;
    off += (uint32_t)p;
  }
  for (int j = 0; j < 4; j++) {
    const uint32_t q = (g.dense_mask >> (4 * j)) & 15u;
    if (q == 15u) g.quad_dense |= 1u << j;
    if (q == 0u) g.quad_hash |= 1u << j;
  }
  n_pos = (int64_t)off * 2;
  n_base = 32 * 64 + 64 * 64 + 64 * 16;
  n_head = 32 * 64 + 64 * 16;
  if (n_first) *n_first = 0;
  if (d.split_head) {
    n_head = (int64_t)vd_head_inputs(d) * 64 + 64 * 16;
    if (n_first) *n_first = 32 * 64 + (d.first_head_n_hidden_layers - 1) * 64 * 64 + 64 * 16;
  }
  return true;
}

// ------------------------------------------------------------------ encode
// Entries of corners x and x + 1 of one (y, z) row of a level, with one
// 8-byte load (dword aligned) for both where the table allows.  ky, kz: the
// row's y / z index terms (products for dense levels, hash terms otherwise).
//  * dense: entries a and a + 1 (a = the row's x index reduced to the
//    level); only when a is the level's last entry does a + 1 wrap to entry
//    0: then the pair (a - 1, a) is read and corner x + 1 takes L.first.
//  * hashed (the level's size is a power of two): for even x, x and x + 1
//    differ in bit 0 only, so their entries i0 and i0 ^ 1 form the aligned
//    pair holding i0; for odd x corner x + 1 is a load of its own, made by
//    those lanes only.
// No branch waits on a load, so every gather of a level is in flight before
// the first wait.
__device__ __forceinline__ void row_pair(const uint32_t* __restrict__ tl, const GridLevel& L, bool dense, uint32_t gx,
                                         uint32_t ky, uint32_t kz, uint32_t& e0, uint32_t& e1) {
  uint32_t lo, i1 = 0u;
  bool sel, use_first = false, use_extra = false;
  if (dense) {
    const uint32_t t = gx + ky + kz;  // t + 1 < 2 * size
    const uint32_t a = min(t, t - L.size);
    const bool pair = a + 1u < L.size;
    lo = pair ? a : a - 1u;
    sel = !pair;
    use_first = !pair;
  } else {
    const uint32_t k = ky ^ kz;
    const uint32_t i0 = (gx ^ k) & (L.size - 1u);
    lo = i0 & ~1u;
    sel = i0 & 1u;
    use_extra = gx & 1u;
    i1 = ((gx + 1u) ^ k) & (L.size - 1u);
  }
  uint2 v;
  __builtin_memcpy(&v, tl + lo, sizeof(v));
  uint32_t extra = 0u;
  if (use_extra) extra = tl[i1];
  e0 = sel ? v.y : v.x;
  e1 = use_first ? L.first : (use_extra ? extra : (sel ? v.x : v.y));
}

// Features of levels g, g+4, g+8, g+12 at p (clamped to [0,1]^3) -> f[8]
// (f[2j], f[2j+1] = level 4j+g).  With this assignment the four lane groups
// of a wave work on levels 4j..4j+3 together, which are either all dense,
// all hashed, or (one quad at most) mixed: the first two take wave-uniform
// branches.  Per level the y / z index terms of both corners are formed once
// (24-bit products for dense levels, which stay below 2^24); each row of
// corners x, x + 1 then costs one add or xor, the reduction to the level's
// entries and (mostly) one 8-byte load (row_pair).
//
// HASHED_PAIRS: hashed levels read their rows with row_pair as well (the
// render kernels); otherwise they, and the mixed quad, read every corner
// with a load of its own (the fill's latents kernel, whose spatially
// coherent samples run faster that way).
template <bool HASHED_PAIRS>
__device__ __forceinline__ void encode_quads(const GridLevel* __restrict__ lv, const GridParams& gp,
                                            const uint32_t* __restrict__ table, V3 p, int g, float* f) {
  const float px = fminf_(fmaxf_(p.x, 0.0f), 1.0f);
  const float py = fminf_(fmaxf_(p.y, 0.0f), 1.0f);
  const float pz = fminf_(fmaxf_(p.z, 0.0f), 1.0f);
  uint32_t ent[4][8];
  float wx[4], wy[4], wz[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int l = 4 * j + g;
    const GridLevel L = lv[l];
    const float ax = px * L.scale + 0.5f, ay = py * L.scale + 0.5f, az = pz * L.scale + 0.5f;
    const float fx = floorf(ax), fy = floorf(ay), fz = floorf(az);
    const uint32_t gx = (uint32_t)(int)fx, gy = (uint32_t)(int)fy, gz = (uint32_t)(int)fz;
    wx[j] = ax - fx; wy[j] = ay - fy; wz[j] = az - fz;
    const uint32_t* tl = table + L.offset;
    // the four (y, z) rows of the cell, corners x and x + 1 each
    auto rows = [&](bool dense, bool pairs) {
      const uint32_t y0 = dense ? __umul24(gy, L.res) : gy * 2654435761u;
      const uint32_t z0 = dense ? __umul24(gz, L.res2) : gz * 805459861u;
      const uint32_t yy[2] = {y0, y0 + (dense ? L.res : 2654435761u)};
      const uint32_t zz[2] = {z0, z0 + (dense ? L.res2 : 805459861u)};
#pragma unroll
      for (int c = 0; c < 8; c += 2) {
        const uint32_t ky = yy[(c >> 1) & 1], kz = zz[c >> 2];
        if (pairs) {
          row_pair(tl, L, dense, gx, ky, kz, ent[j][c], ent[j][c + 1]);
        } else {
#pragma unroll
          for (int o = 0; o < 2; o++) {
            const uint32_t x = gx + o, ta = x + ky + kz, th = (x ^ ky ^ kz) & (L.size - 1u);
            ent[j][c + o] = tl[dense ? min(ta, ta - L.size) : th];  // ta < 2 * size
          }
        }
      }
    };
    if ((gp.quad_dense >> j) & 1u) rows(true, true);
    else if ((gp.quad_hash >> j) & 1u) rows(false, HASHED_PAIRS);
    else rows((gp.dense_mask >> l) & 1u, HASHED_PAIRS);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const float ux = 1.0f - wx[j], uy = 1.0f - wy[j], uz = 1.0f - wz[j];
    const float wxy[4] = {ux * uy, wx[j] * uy, ux * wy[j], wx[j] * wy[j]};
    float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const float w = wxy[c & 3] * ((c & 4) ? wz[j] : uz);
      const uint32_t e = ent[j][c];
      const float v0 = (float)__builtin_bit_cast(_Float16, (uint16_t)(e & 0xffffu));
      const float v1 = (float)__builtin_bit_cast(_Float16, (uint16_t)(e >> 16));
      // fused: one v_fma_mix_f32 per corner and feature (the half operand
      // is read directly)
      a0 = __builtin_fmaf(v0, w, a0);
      a1 = __builtin_fmaf(v1, w, a1);
    }
    f[2 * j] = a0;
    f[2 * j + 1] = a1;
  }
}

// the block's copy of the level table, with each level's entry 0
__device__ __forceinline__ void load_levels(const GridParams& gp, const uint32_t* __restrict__ table,
                                            GridLevel* lv) {
  if (threadIdx.x < MAX_LEVELS) {
    GridLevel L = gp.lv[threadIdx.x];
    if ((int)threadIdx.x < gp.n_levels) L.first = table[L.offset];
    lv[threadIdx.x] = L;
  }
  __syncthreads();
}

// ------------------------------------------------------------------ kernels
static const int NET_BLOCK = 256;  // 4 waves

__device__ __forceinline__ uint2 pack_h4(_Float16 a, _Float16 b, _Float16 c, _Float16 d) {
  uint2 v;
  v.x = (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
  v.y = (uint32_t)__builtin_bit_cast(uint16_t, c) | ((uint32_t)__builtin_bit_cast(uint16_t, d) << 16);
  return v;
}

// HEAD: full inference -> half4 (r, g, b, log sigma) per sample
// !HEAD: inferenceLatents -> 16 halves per sample
// TILES: 16-sample MFMA tiles per wave and loop iteration (MLP rows per
// wavefront = 16 TILES); the tiles' hash gathers are issued together.
// PIX (renderer sample streams): pos holds float4 (x, y, z, pixel index
// bits) per sample and dir is the per-pixel SH table (sh4_all, 32 bytes per
// pixel, written by the marcher at the ray's first launch) instead of a
// direction per sample.
template <bool HEAD, int P, int TILES, bool PIX = false>
__global__ __launch_bounds__(NET_BLOCK) void k_network(const float* __restrict__ pos, const float* __restrict__ dir,
                                                       int64_t n, const uint32_t* __restrict__ table,
                                                       const typename Mx<P>::T* __restrict__ frags, GridParams gp,
                                                       MlpScales sc, uint2* __restrict__ out_h4,
                                                       uint2* __restrict__ out16, const int* __restrict__ n_dev) {
  // n_dev: sample count written by an earlier kernel (n is then its bound)
  if (n_dev) n = *n_dev;
  if ((int64_t)blockIdx.x * (NET_BLOCK / 64) * 16 * TILES >= n) return;
  __shared__ GridLevel lv[MAX_LEVELS];
  __shared__ typename Mx<P>::T wl[N_NET_FRAGS * 64];
  load_levels(gp, table, lv);
  stage_frags(frags, wl, N_NET_FRAGS);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, g = lane >> 4;
  const int64_t n_groups = (n + 16 * TILES - 1) / (16 * TILES);
  for (int64_t grp = (int64_t)blockIdx.x * (NET_BLOCK / 64) + wave; grp < n_groups;
       grp += (int64_t)gridDim.x * (NET_BLOCK / 64)) {
    const int li = opaque(lane);
    int64_t s[TILES], sl[TILES];
    bool valid[TILES];
    float f[TILES][8];
    V3 pv[TILES], dv[TILES];  // positions; view directions (used after the density MLP)
    uint2 shq[TILES];         // PIX: the lane group's SH coefficients
    int px[TILES];
    // every tile's position (and direction) is loaded before any tile's
    // gathers, which wait for them (the wait counter is in order)
#pragma unroll
    for (int t = 0; t < TILES; t++) {
      s[t] = (grp * TILES + t) * 16 + col;
      valid[t] = s[t] < n;
      sl[t] = valid[t] ? s[t] : n - 1;
      if constexpr (PIX) {
        const float4 q = reinterpret_cast<const float4*>(pos)[sl[t]];
        pv[t] = V3{q.x, q.y, q.z};
        px[t] = __builtin_bit_cast(int, q.w);
      } else {
        pv[t] = V3{pos[3 * sl[t]], pos[3 * sl[t] + 1], pos[3 * sl[t] + 2]};
        if (HEAD) dv[t] = V3{dir[3 * sl[t]], dir[3 * sl[t] + 1], dir[3 * sl[t] + 2]};
      }
    }
#pragma unroll
    for (int t = 0; t < TILES; t++) {
      encode_quads<HEAD>(lv, gp, table, pv[t], g, f[t]);
      // pinned here (after the gathers are issued): otherwise the load sinks
      // to its use after the density MLP and waits there on its own
      if constexpr (HEAD && PIX) {
        shq[t] = reinterpret_cast<const uint2*>(dir)[(int64_t)px[t] * 4 + g];
        asm volatile("" : "+v"(shq[t].x), "+v"(shq[t].y));
      } else if (HEAD) {
        asm volatile("" : "+v"(dv[t].x), "+v"(dv[t].y), "+v"(dv[t].z));
      }
    }
#pragma unroll
    for (int t = 0; t < TILES; t++) {
#pragma unroll
      for (int e = 0; e < 8; e++) f[t][e] = (float)(_Float16)f[t][e];  // the encoding is stored as half
      const typename Mx<P>::T x = Mx<P>::pack(f[t], sc.inv_act);
      const f4 d3 = base_mlp<P>(wl, li, x, sc);
      const _Float16 rows[4] = {(_Float16)d3[0], (_Float16)d3[1], (_Float16)d3[2], (_Float16)d3[3]};
      if (!HEAD) {
        if (valid[t]) out16[s[t] * 4 + g] = pack_h4(rows[0], rows[1], rows[2], rows[3]);
        // out_h4, when given: row 0 (log sigma) of each sample once more, as
        // a compact half array (the fill's transmittance pass reads only it)
        if (out_h4 && g == 0 && valid[t])
          reinterpret_cast<uint16_t*>(out_h4)[s[t]] = __builtin_bit_cast(uint16_t, rows[0]);
      } else {
        typename Mx<P>::T hin;
        if constexpr (PIX) {
          _Float16 sh[4];
          sh4_unpack(shq[t], sh);
          hin = head_input_sh<P>(rows, g, sh, sc.inv_act, sc.lat_lo, sc.lat_hi);
        } else {
          hin = head_input<P>(rows, g, dv[t].x, dv[t].y, dv[t].z, sc.inv_act, sc.lat_lo, sc.lat_hi);
        }
        const f4 o = head_mlp<P>(wl + FRAG_H1 * 64, wl + FRAG_H2 * 64, li, hin, sc);
        if (g == 0 && valid[t]) out_h4[s[t]] = pack_h4((_Float16)o[0], (_Float16)o[1], (_Float16)o[2], rows[0]);
      }
    }
  }
}

// inferenceHead on given head inputs (16 halves per sample in the layout of
// the density-MLP output: row 0 log sigma, latents in rows [lat_lo, lat_hi))
template <int P>
__global__ __launch_bounds__(NET_BLOCK) void k_head(const uint16_t* __restrict__ hin16, const float* __restrict__ dir,
                                                    int64_t n, const typename Mx<P>::T* __restrict__ frags,
                                                    MlpScales sc, uint2* __restrict__ out_h4) {
  __shared__ typename Mx<P>::T wl[N_NET_FRAGS * 64];
  stage_frags(frags, wl, N_NET_FRAGS);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, g = lane >> 4;
  const int64_t n_groups = (n + 15) / 16;
  for (int64_t grp = (int64_t)blockIdx.x * (NET_BLOCK / 64) + wave; grp < n_groups;
       grp += (int64_t)gridDim.x * (NET_BLOCK / 64)) {
    const int li = opaque(lane);
    const int64_t s = grp * 16 + col;
    const bool valid = s < n;
    const int64_t sl = valid ? s : n - 1;
    _Float16 rows[4];
#pragma unroll
    for (int e = 0; e < 4; e++) rows[e] = __builtin_bit_cast(_Float16, hin16[sl * 16 + 4 * g + e]);
    const typename Mx<P>::T x = head_input<P>(rows, g, dir[3 * sl], dir[3 * sl + 1], dir[3 * sl + 2], sc.inv_act,
                                              sc.lat_lo, sc.lat_hi);
    const f4 o = head_mlp<P>(wl + FRAG_H1 * 64, wl + FRAG_H2 * 64, li, x, sc);
    if (g == 0 && valid)
      out_h4[s] = pack_h4((_Float16)o[0], (_Float16)o[1], (_Float16)o[2],
                          __builtin_bit_cast(_Float16, hin16[sl * 16]));
  }
}

// Split models (NerfNetwork::inference / inferenceLatents after
// init_viewdep, nerf_network.h:160-202): hash grid and density MLP as in
// k_network, then mlp_first_head on [the 16 density-MLP outputs, SH4 of the
// initial view direction] and, for FULL, the custom colour head
// (my_fully_fused_mlp.cu:677-895).  FULL: the initial view direction is the
// view direction (performance_renderer.cu:269), so the SH differences are
// exactly 0; output half4 (r, g, b, log sigma).  !FULL: dir is the initial
// view direction; output [8 head latents, log sigma, 0 x 7] per sample.
// di: the first latent is log sigma (the first head's latent row 0 is then
// a zero-weight row replaced here).  FL: first-head hidden layers.
// PIX (FULL only): the renderer's sample stream as in k_network.
template <bool FULL, int TILES, int FL, bool PIX = false>
__global__ __launch_bounds__(NET_BLOCK) void k_network_vd(const float* __restrict__ pos, const float* __restrict__ dir,
                                                          int64_t n, const uint32_t* __restrict__ table,
                                                          const h8* __restrict__ frags, GridParams gp, int di,
                                                          int freq_on, uint2* __restrict__ out,
                                                          const int* __restrict__ n_dev) {
  if (n_dev) n = *n_dev;
  if ((int64_t)blockIdx.x * (NET_BLOCK / 64) * 16 * TILES >= n) return;
  __shared__ GridLevel lv[MAX_LEVELS];
  __shared__ h8 wl[N_VD_NET_FRAGS * 64];
  load_levels(gp, table, lv);
  stage_frags(frags, wl, FULL ? N_VD_NET_FRAGS : VD_HA);
  __syncthreads();
  const MlpScales sc{{1.0f, 1.0f, 1.0f, 1.0f, 1.0f}, 1.0f};
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, g = lane >> 4;
  const int64_t n_groups = (n + 16 * TILES - 1) / (16 * TILES);
  for (int64_t grp = (int64_t)blockIdx.x * (NET_BLOCK / 64) + wave; grp < n_groups;
       grp += (int64_t)gridDim.x * (NET_BLOCK / 64)) {
    const int li = opaque(lane);
    int64_t s[TILES], sl[TILES];
    bool valid[TILES];
    float f[TILES][8];
    V3 pv[TILES], dv[TILES];  // positions and directions, loaded before the gathers
    uint2 shq[TILES];         // PIX: the lane group's SH coefficients
    int px[TILES];
#pragma unroll
    for (int t = 0; t < TILES; t++) {
      s[t] = (grp * TILES + t) * 16 + col;
      valid[t] = s[t] < n;
      sl[t] = valid[t] ? s[t] : n - 1;
      if constexpr (PIX) {
        const float4 q = reinterpret_cast<const float4*>(pos)[sl[t]];
        pv[t] = V3{q.x, q.y, q.z};
        px[t] = __builtin_bit_cast(int, q.w);
      } else {
        pv[t] = V3{pos[3 * sl[t]], pos[3 * sl[t] + 1], pos[3 * sl[t] + 2]};
        dv[t] = V3{dir[3 * sl[t]], dir[3 * sl[t] + 1], dir[3 * sl[t] + 2]};
      }
    }
#pragma unroll
    for (int t = 0; t < TILES; t++) {
      encode_quads<FULL>(lv, gp, table, pv[t], g, f[t]);
      // pinned here (after the gathers are issued), as in k_network
      if constexpr (PIX) {
        shq[t] = reinterpret_cast<const uint2*>(dir)[(int64_t)px[t] * 4 + g];
        asm volatile("" : "+v"(shq[t].x), "+v"(shq[t].y));
      } else {
        asm volatile("" : "+v"(dv[t].x), "+v"(dv[t].y), "+v"(dv[t].z));
      }
    }
#pragma unroll
    for (int t = 0; t < TILES; t++) {
#pragma unroll
      for (int e = 0; e < 8; e++) f[t][e] = (float)(_Float16)f[t][e];
      const f4 d3 = base_mlp<FVC_PREC_FP16>(wl, li, MxH::pack(f[t], 1.0f), sc);
      const _Float16 bh[4] = {(_Float16)d3[0], (_Float16)d3[1], (_Float16)d3[2], (_Float16)d3[3]};
      _Float16 sh[4];
      if constexpr (PIX) sh4_unpack(shq[t], sh);
      else sh4_quarter(dv[t].x, dv[t].y, dv[t].z, g, sh);
      const f4 fo = first_head_mlp<FL>(wl, li, h8_of(bh, sh));
      _Float16 lat[4] = {(_Float16)fo[0], (_Float16)fo[1], (_Float16)fo[2], (_Float16)fo[3]};
      if (di && g == 0) lat[0] = bh[0];
      if constexpr (!FULL) {
        // log sigma (density-MLP row 0, group 0) into group 2
        const int ls = __shfl_xor((int)__builtin_bit_cast(uint16_t, bh[0]), 32);
        const _Float16 z = (_Float16)0.0f;
        uint2 v = pack_h4(lat[0], lat[1], lat[2], lat[3]);
        if (g == 2) v = pack_h4(__builtin_bit_cast(_Float16, (uint16_t)ls), z, z, z);
        if (g == 3) v = make_uint2(0u, 0u);
        if (valid[t]) out[s[t] * 4 + g] = v;
      } else {
        _Float16 fa[4], fb[4];
        freq_quad(pv[t], g + 2, freq_on != 0, fa);
        freq_quad(pv[t], g, freq_on != 0, fb);
        const _Float16 zero4[4] = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
        const h8 a = h8_of(g < 2 ? lat : fa, sh);
        const h8 b = h8_of(zero4, fb);
        const f4 o = head_vd_mlp(wl + VD_HA * 64, wl + VD_HB * 64, wl + VD_HO * 64, li, a, b);
        if (g == 0 && valid[t]) out[s[t]] = pack_h4((_Float16)o[0], (_Float16)o[1], (_Float16)o[2], bh[0]);
      }
    }
  }
}

// the custom colour head alone on given inputs (16 halves per sample in the
// layout of k_network_vd<false>), with the initial view direction init
__global__ __launch_bounds__(NET_BLOCK) void k_head_vd(const uint16_t* __restrict__ hin16, const float* __restrict__ pos,
                                                       const float* __restrict__ dir, const float* __restrict__ init,
                                                       int64_t n, const h8* __restrict__ frags, int di, int freq_on,
                                                       uint2* __restrict__ out) {
  __shared__ h8 wl[10 * 64];  // VD_HA, VD_HB, VD_HO
  stage_frags(frags + VD_HA * 64, wl, 10);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, g = lane >> 4;
  const int64_t n_groups = (n + 15) / 16;
  for (int64_t grp = (int64_t)blockIdx.x * (NET_BLOCK / 64) + wave; grp < n_groups;
       grp += (int64_t)gridDim.x * (NET_BLOCK / 64)) {
    const int li = opaque(lane);
    const int64_t s = grp * 16 + col;
    const bool valid = s < n;
    const int64_t q = valid ? s : n - 1;
    _Float16 lat[4], sh[4], shi[4], dsh[4], fa[4], fb[4];
#pragma unroll
    for (int e = 0; e < 4; e++) lat[e] = __builtin_bit_cast(_Float16, hin16[q * 16 + (4 * g + e) % 8]);
    sh4_quarter(dir[3 * q], dir[3 * q + 1], dir[3 * q + 2], g, sh);
    sh4_quarter(init[3 * q], init[3 * q + 1], init[3 * q + 2], g, shi);
#pragma unroll
    for (int e = 0; e < 4; e++) dsh[e] = (_Float16)((float)sh[e] - (float)shi[e]);
    const V3 p = {pos[3 * q], pos[3 * q + 1], pos[3 * q + 2]};
    freq_quad(p, g + 2, freq_on != 0, fa);
    freq_quad(p, g, freq_on != 0, fb);
    const f4 o = head_vd_mlp(wl, wl + 4 * 64, wl + 8 * 64, li, h8_of(g < 2 ? lat : fa, sh), h8_of(dsh, fb));
    if (g == 0 && valid)
      out[s] = pack_h4((_Float16)o[0], (_Float16)o[1], (_Float16)o[2],
                       __builtin_bit_cast(_Float16, hin16[q * 16 + (di ? 0 : 8)]));
  }
}

// encoding alone: 32 halves per sample (thread per sample)
__global__ void k_encode(const float* __restrict__ pos, int64_t n, const uint32_t* __restrict__ table, GridParams gp,
                         uint16_t* __restrict__ out32) {
  __shared__ GridLevel lv[MAX_LEVELS];
  load_levels(gp, table, lv);
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const V3 p = {pos[3 * s], pos[3 * s + 1], pos[3 * s + 2]};
  for (int q = 0; q < 4; q++) {
    float f[8];
    encode_quads<true>(lv, gp, table, p, q, f);
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 2; k++)
        out32[s * 32 + 2 * (4 * j + q) + k] = __builtin_bit_cast(uint16_t, (_Float16)f[2 * j + k]);
  }
}

static int grid_for(int64_t n_groups) {
  const int64_t waves = n_groups;
  int64_t blocks = (waves + NET_BLOCK / 64 - 1) / (NET_BLOCK / 64);
  // 16384 blocks (64 per CU), grid-stride beyond: measured on the 1080p path
  // against 4096 - 65536 (more blocks help the full render; the resample
  // launches, sized by a bound, pay for blocks that find no work)
  const int64_t cap = 16384;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  return (int)blocks;
}

// one launch per MLP precision of the model
#define FVC_BY_PREC(m, LAUNCH)                                               \
  switch ((m).net.precision) {                                               \
    case FVC_PREC_BF16: LAUNCH(FVC_PREC_BF16); break;                        \
    case FVC_PREC_FP8: LAUNCH(FVC_PREC_FP8); break;                          \
    case FVC_PREC_BF16X3: LAUNCH(FVC_PREC_BF16X3); break;                    \
    default: LAUNCH(FVC_PREC_FP16); break;                                   \
  }

void launch_network_full(const Model& m, const float* pos, const float* dir, int64_t n, uint32_t* out_h4,
                         hipStream_t s, const int* n_dev, int rows) {
  launch_network_full_ex(m, pos, dir, false, n, out_h4, s, n_dev, rows);
}
void launch_network_full_pix(const Model& m, const float4* pos4, const uint4* ray_sh, int64_t n, uint32_t* out_h4,
                             hipStream_t s, const int* n_dev, int rows) {
  launch_network_full_ex(m, reinterpret_cast<const float*>(pos4), reinterpret_cast<const float*>(ray_sh), true, n,
                         out_h4, s, n_dev, rows);
}
void launch_network_full_ex(const Model& m, const float* pos, const float* dir, bool pix, int64_t n, uint32_t* out_h4,
                            hipStream_t s, const int* n_dev, int rows) {
  if (n <= 0) return;
  if (m.net.split_head) {
    launch_network_full_vd(m, pos, dir, pix, n, out_h4, s, n_dev, rows);
    return;
  }
#define L3(P, TL, PX)                                                                                              \
  hipLaunchKernelGGL((k_network<true, P, TL, PX>), dim3(grid_for((n + 16 * TL - 1) / (16 * TL))), dim3(NET_BLOCK), 0, \
                     s, pos, dir, n, m.table.as<uint32_t>(), m.frags.as<typename Mx<P>::T>(), m.grid, m.mlp_scales,  \
                     (uint2*)out_h4, (uint2*)nullptr, n_dev)
#define L2(P, TL)          \
  if (pix) { L3(P, TL, true); } \
  else { L3(P, TL, false); }
#define L(P)                       \
  if (rows == 64) { L2(P, 4); }    \
  else if (rows == 32) { L2(P, 2); } \
  else { L2(P, 1); }
  FVC_BY_PREC(m, L)
#undef L
#undef L2
#undef L3
  FVC_HIP(hipGetLastError());
}
void launch_network_latents(const Model& m, const float* pos, int64_t n, uint16_t* out16, hipStream_t s, int rows,
                            const int* n_dev, uint16_t* log_sigma) {
  if (m.net.split_head) throw Error(FVC_E_UNSUPPORTED, "split models need the initial view direction");
  if (n <= 0) return;
#define L2(P, TL)                                                                                                 \
  hipLaunchKernelGGL((k_network<false, P, TL>), dim3(grid_for((n + 16 * TL - 1) / (16 * TL))), dim3(NET_BLOCK), 0, s, \
                     pos, (const float*)nullptr, n, m.table.as<uint32_t>(), m.frags.as<typename Mx<P>::T>(), m.grid, \
                     m.mlp_scales, (uint2*)log_sigma, (uint2*)out16, n_dev)
#define L(P)                       \
  if (rows == 64) { L2(P, 4); }    \
  else if (rows == 32) { L2(P, 2); } \
  else { L2(P, 1); }
  FVC_BY_PREC(m, L)
#undef L
#undef L2
  FVC_HIP(hipGetLastError());
}
void launch_network_head(const Model& m, const uint16_t* hin16, const float* dir, int64_t n, uint32_t* out_h4,
                         hipStream_t s) {
  if (m.net.split_head) throw Error(FVC_E_UNSUPPORTED, "split models need the position and initial view direction");
  if (n <= 0) return;
#define L(P)                                                                                                    \
  hipLaunchKernelGGL((k_head<P>), dim3(grid_for((n + 15) / 16)), dim3(NET_BLOCK), 0, s, hin16, dir, n,            \
                     m.frags.as<typename Mx<P>::T>(), m.mlp_scales, (uint2*)out_h4)
  FVC_BY_PREC(m, L)
#undef L
  FVC_HIP(hipGetLastError());
}
void launch_network_full_vd(const Model& m, const float* pos, const float* dir, bool pix, int64_t n,
                            uint32_t* out_h4, hipStream_t s, const int* n_dev, int rows) {
  if (n <= 0) return;
  const int di = m.net.first_latent_is_density, fq = m.net.freq_encoding_degree > 0;
#define L2(TL, FL, PX)                                                                                            \
  hipLaunchKernelGGL((k_network_vd<true, TL, FL, PX>), dim3(grid_for((n + 16 * TL - 1) / (16 * TL))),             \
                     dim3(NET_BLOCK), 0, s, pos, dir, n, m.table.as<uint32_t>(), m.frags.as<h8>(), m.grid, di, fq,  \
                     (uint2*)out_h4, n_dev)
#define L(TL, FL)                \
  if (pix) { L2(TL, FL, true); } \
  else { L2(TL, FL, false); }
  if (m.net.first_head_n_hidden_layers == 2) {
    if (rows == 64) { L(4, 2); } else if (rows == 32) { L(2, 2); } else { L(1, 2); }
  } else {
    if (rows == 64) { L(4, 1); } else if (rows == 32) { L(2, 1); } else { L(1, 1); }
  }
#undef L
#undef L2
  FVC_HIP(hipGetLastError());
}
void launch_network_latents_vd(const Model& m, const float* pos, const float* init_dir, int64_t n, uint16_t* out16,
                               hipStream_t s, int rows, const int* n_dev) {
  if (n <= 0) return;
  const int di = m.net.first_latent_is_density, fq = m.net.freq_encoding_degree > 0;
#define L(TL, FL)                                                                                               \
  hipLaunchKernelGGL((k_network_vd<false, TL, FL>), dim3(grid_for((n + 16 * TL - 1) / (16 * TL))), dim3(NET_BLOCK), \
                     0, s, pos, init_dir, n, m.table.as<uint32_t>(), m.frags.as<h8>(), m.grid, di, fq,            \
                     (uint2*)out16, n_dev)
  if (m.net.first_head_n_hidden_layers == 2) {
    if (rows == 64) { L(4, 2); } else if (rows == 32) { L(2, 2); } else { L(1, 2); }
  } else {
    if (rows == 64) { L(4, 1); } else if (rows == 32) { L(2, 1); } else { L(1, 1); }
  }
#undef L
  FVC_HIP(hipGetLastError());
}
void launch_network_head_vd(const Model& m, const uint16_t* head_in16, const float* pos, const float* dir,
                            const float* init_dir, int64_t n, uint32_t* out_h4, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_head_vd, dim3(grid_for((n + 15) / 16)), dim3(NET_BLOCK), 0, s, head_in16, pos, dir, init_dir, n,
                     m.frags.as<h8>(), m.net.first_latent_is_density, m.net.freq_encoding_degree > 0,
                     (uint2*)out_h4);
  FVC_HIP(hipGetLastError());
}

void launch_network_encode(const Model& m, const float* pos, int64_t n, uint16_t* out32, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_encode, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pos, n, m.table.as<uint32_t>(),
                     m.grid, out32);
  FVC_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ packing
// Column of the reference weight matrix that lane-group g, slot e of a
// fragment multiplies (see mlp_mfma.h).
// W1: slot pair j of group g holds level 4j+g (encode_quads)
static int col_first(int g, int e) { return 2 * (4 * (e >> 1) + g) + (e & 1); }
static int col_chain(int kb, int g, int e) { return e < 4 ? 32 * kb + 4 * g + e : 32 * kb + 16 + 4 * g + e - 4; }
// H1: the head's inputs are [n latents, 16 SH, 16 - n padding]; slot e < 4
// of group g carries base-output row r = 4g + e, a latent when r is in
// [lo, lo + n), else one of the padding inputs (in row order)
static int col_head_in(int g, int e, int lo, int n) {
  if (e >= 4) return n + 4 * g + (e - 4);  // SH coefficient 4g + e - 4
  const int r = 4 * g + e;
  if (r >= lo && r < lo + n) return r - lo;  // latent r - lo
  return n + 16 + (r < lo ? r : r - n);      // padding
}

// host conversions for the weight fragments
static float f16_value(uint16_t h) {
  const uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 31u, m = h & 1023u;
  float v;
  if (e == 0) v = std::ldexp((float)m, -24);
  else if (e == 31) v = m ? NAN : INFINITY;
  else v = std::ldexp((float)(m | 1024u), (int)e - 25);
  return s ? -v : v;
}
static uint16_t bf16_bits(float f) {  // round to nearest even
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float bf16_value(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
// OCP E4M3 (bias 7, max 448), round to nearest even, saturating
static uint8_t e4m3_code(float f) {
  const uint8_t sign = std::signbit(f) ? 0x80 : 0x00;
  const float a = std::fabs(f);
  if (!(a < 448.0f)) return (uint8_t)(sign | 0x7e);
  int e;
  std::frexp(a, &e);
  int ue = e - 1;                               // exponent of the leading bit
  if (ue < -6) ue = -6;                         // subnormals share exponent -6
  const float q = std::ldexp(1.0f, ue - 3);     // quantum of 3 fraction bits
  const float r = std::nearbyint(a / q) * q;    // ties to even
  if (r >= 448.0f) return (uint8_t)(sign | 0x7e);
  if (r == 0.0f) return sign;
  int e2;
  std::frexp(r, &e2);
  if (e2 - 1 < -6) return (uint8_t)(sign | (int)std::lround(r / std::ldexp(1.0f, -9)));
  const int ef = e2 - 1 + 7;
  const int mant = (int)std::lround(r / std::ldexp(1.0f, e2 - 1 - 3)) - 8;
  return (uint8_t)(sign | (ef << 3) | mant);
}

// Fragments in the operand format of the precision: 2 bytes per value for
// FP16 / BF16, 1 byte for FP8 (E4M3 codes of w / s_w, per-layer
// s_w = absmax / 448 over the whole fp16 layer).  Fills the scales.
void pack_fragments(const uint16_t* base, const uint16_t* head, int precision, int lat_lo, int n_lat,
                    std::vector<uint8_t>& out, MlpScales& sc) {
  sc.lat_lo = lat_lo;
  sc.lat_hi = lat_lo + n_lat;
  // bytes per lane and fragment: 8 operands, two terms of 8 in BF16X3
  const int lane_bytes = precision == FVC_PREC_FP8 ? 8 : (precision == FVC_PREC_BF16X3 ? 32 : 16);
  out.assign((size_t)N_FRAGS * 64 * lane_bytes, 0);
  const uint16_t* W1 = base;
  const uint16_t* W2 = base + 32 * 64;
  const uint16_t* W3 = base + 32 * 64 + 64 * 64;
  const uint16_t* H1 = head;
  const uint16_t* H2 = head + 32 * 64;
  const uint16_t* Ws[5] = {W1, W2, W3, H1, H2};
  const int counts[5] = {64 * 32, 64 * 64, 16 * 64, 64 * 32, 16 * 64};
  float wscale[5];
  for (int l = 0; l < 5; l++) {
    float amax = 0.0f;
    for (int i = 0; i < counts[l]; i++) amax = std::max(amax, std::fabs(f16_value(Ws[l][i])));
    wscale[l] = amax > 0.0f ? amax / 448.0f : 1.0f;
    sc.layer[l] = precision == FVC_PREC_FP8 ? wscale[l] * FP8_ACT_SCALE : 1.0f;
  }
  sc.inv_act = 1.0f / FP8_ACT_SCALE;
  auto put = [&](int frag, int layer, const uint16_t* W, int n_in, int row0, auto colf) {
    for (int lane = 0; lane < 64; lane++) {
      const int i = lane % 16, g = lane / 16;
      for (int e = 0; e < 8; e++) {
        const uint16_t h = W[(size_t)(row0 + i) * n_in + colf(g, e)];
        const size_t base = ((size_t)frag * 64 + lane) * lane_bytes;
        if (precision == FVC_PREC_FP8) {
          out[base + e] = e4m3_code(f16_value(h) / wscale[layer]);
        } else if (precision == FVC_PREC_BF16X3) {
          // hi = bf16(w), lo = bf16(w - hi): exact for fp16 weights
          const float w = f16_value(h);
          const uint16_t hi = bf16_bits(w);
          const uint16_t lo = bf16_bits(w - bf16_value(hi));
          std::memcpy(&out[base + 2 * e], &hi, 2);
          std::memcpy(&out[base + 16 + 2 * e], &lo, 2);
        } else {
          const uint16_t v = precision == FVC_PREC_BF16 ? bf16_bits(f16_value(h)) : h;
          std::memcpy(&out[base + 2 * e], &v, 2);
        }
      }
    }
  };
  for (int nb = 0; nb < 4; nb++) put(FRAG_W1 + nb, 0, W1, 32, nb * 16, col_first);
  for (int nb = 0; nb < 4; nb++)
    for (int kb = 0; kb < 2; kb++)
      put(FRAG_W2 + nb * 2 + kb, 1, W2, 64, nb * 16, [kb](int g, int e) { return col_chain(kb, g, e); });
  for (int kb = 0; kb < 2; kb++) put(FRAG_W3 + kb, 2, W3, 64, 0, [kb](int g, int e) { return col_chain(kb, g, e); });
  for (int nb = 0; nb < 4; nb++)
    put(FRAG_H1 + nb, 3, H1, 32, nb * 16, [=](int g, int e) { return col_head_in(g, e, lat_lo, n_lat); });
  for (int kb = 0; kb < 2; kb++) put(FRAG_H2 + kb, 4, H2, 64, 0, [kb](int g, int e) { return col_chain(kb, g, e); });
}

// Split models.  Custom head input columns (my_fully_fused_mlp.cu:738-790):
// latents 0..7, SH 8..23, SH differences 24..39, frequencies from 40 in the
// order of freqEncChunk4 (per axis x, y, z: sin for f = 0..3, then cos);
// quad q = 2 axis + cos holds columns 40 + 4q .. 40 + 4q + 3.  Without
// frequencies, quads 0 and 1 carry the 8 padding inputs (value 1.0) and the
// others multiply zero weights.
static int col_freq(const fvc_net_desc& d, int q, int f) {
  if (d.freq_encoding_degree == 4) return 40 + 4 * q + f;
  return q < 2 ? 40 + 4 * q + f : -1;
}
// block A: groups 0, 1 latents 4g..4g+3, groups 2, 3 frequency quads 4, 5;
// slots 4..7 SH coefficients 4g..4g+3
static int col_vd_a(const fvc_net_desc& d, int g, int e) {
  if (e >= 4) return 8 + 4 * g + (e - 4);
  return g < 2 ? 4 * g + e : col_freq(d, g + 2, e);
}
// block B: SH differences 4g..4g+3, frequency quad g
static int col_vd_b(const fvc_net_desc& d, int g, int e) { return e < 4 ? 24 + 4 * g + e : col_freq(d, g, e - 4); }
// mlp_first_head input: density-MLP outputs 4g..4g+3, SH of the initial
// view direction 4g..4g+3 (nerf_network.h:186-194)
static int col_vd_first(int g, int e) { return e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4); }

void pack_fragments_vd(const uint16_t* base, const uint16_t* first, const uint16_t* head, const fvc_net_desc& d,
                       std::vector<uint8_t>& out) {
  out.assign((size_t)N_VD_FRAGS * 64 * 8 * 2, 0);
  // fragment frag, lane (i, g), slot e <- W[rowf(i)][colf(g, e)] (0 where either is -1)
  auto put = [&](int frag, const uint16_t* W, int n_in, auto rowf, auto colf) {
    for (int lane = 0; lane < 64; lane++) {
      const int i = lane % 16, g = lane / 16, r = rowf(i);
      for (int e = 0; e < 8; e++) {
        const int c = colf(g, e);
        const uint16_t h = (r < 0 || c < 0) ? 0 : W[(size_t)r * n_in + c];
        std::memcpy(&out[(((size_t)frag * 64 + lane) * 8 + e) * 2], &h, 2);
      }
    }
  };
  auto rows = [](int nb) { return [nb](int i) { return nb * 16 + i; }; };
  auto chain = [](int kb) { return [kb](int g, int e) { return col_chain(kb, g, e); }; };
  auto ident = [](int i) { return i; };
  const uint16_t* W1 = base;
  const uint16_t* W2 = base + 32 * 64;
  const uint16_t* W3 = base + 32 * 64 + 64 * 64;
  for (int nb = 0; nb < 4; nb++) put(FRAG_W1 + nb, W1, 32, rows(nb), col_first);
  for (int nb = 0; nb < 4; nb++)
    for (int kb = 0; kb < 2; kb++) put(FRAG_W2 + nb * 2 + kb, W2, 64, rows(nb), chain(kb));
  for (int kb = 0; kb < 2; kb++) put(FRAG_W3 + kb, W3, 64, ident, chain(kb));
  // mlp_first_head; output row k (k < 8) = the head's latent k: reference
  // buffer row latent_offset + k, whose row 0 is log sigma (filled in by the
  // kernels: zero weights here) and row r >= 1 first-head output r - 1
  const int FL = d.first_head_n_hidden_layers;
  for (int nb = 0; nb < 4; nb++) put(VD_F1 + nb, first, 32, rows(nb), col_vd_first);
  if (FL == 2)
    for (int nb = 0; nb < 4; nb++)
      for (int kb = 0; kb < 2; kb++) put(VD_F2 + nb * 2 + kb, first + 32 * 64, 64, rows(nb), chain(kb));
  const uint16_t* F3 = first + 32 * 64 + (FL - 1) * 64 * 64;
  const int lo = d.latent_offset;
  auto latent_row = [lo](int i) { return i < 8 && lo + i >= 1 ? lo + i - 1 : -1; };
  for (int kb = 0; kb < 2; kb++) put(VD_F3 + kb, F3, 64, latent_row, chain(kb));
  // custom colour head
  const int HIN = vd_head_inputs(d);
  const uint16_t* HO = head + (size_t)HIN * 64;
  for (int nb = 0; nb < 4; nb++) {
    put(VD_HA + nb, head, HIN, rows(nb), [&d](int g, int e) { return col_vd_a(d, g, e); });
    put(VD_HB + nb, head, HIN, rows(nb), [&d](int g, int e) { return col_vd_b(d, g, e); });
  }
  for (int kb = 0; kb < 2; kb++) put(VD_HO + kb, HO, 64, ident, chain(kb));
}

}  // namespace fvc
