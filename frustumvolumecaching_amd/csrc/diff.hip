// diff.hip -- differentiable full render (training mode) and its adjoint.
//
// Forward, for a list of pixels (y * W + x) of one camera, with FP32
// parameter blocks in the nerfacc layout (hash table [entries][2], density
// MLP W1|W2|W3, colour MLP H1|H2, row-major [out][in]) owned by the caller:
//   k_diff_count   ray setup (initRayPayloads_split_kernel, renderer.cu:86-
//                  134) and the occupancy march of sample_performance_kernel
//                  (performance_renderer.cu:15-69) to the far bound, counting
//                  the occupied samples per ray;
//   exclusive scan of the counts (hipCUB);
//   k_diff_emit    the same march again, writing each sample's contracted
//                  midpoint, step size and ray;
//   k_diff_net_fwd hash-grid encoding + both MLPs in FP32 on the 16x16x4 f32
//                  MFMA: (colour logits, log density) per sample;
//   k_diff_composite  accumulate_performance_kernel (performance_renderer.cu:
//                  71-125) per ray, stopping at T < 1e-4; output C + T (the
//                  white background of applyWhiteBackground).
// Backward, for dL/d(C + T):
//   k_diff_composite_bwd  per ray, dL/d(logits, log density) per sample:
//                  dC/dc_i = w_i and dC/dsigma_i = dt_i (T_{i+1} c_i - S_i),
//                  S_i = (C + T) - sum_{j<=i} w_j c_j; sigmoid' and exp';
//   k_diff_net_bwd per 16 samples: the forward recomputed (no activation is
//                  stored), the chain back through the colour MLP, the
//                  density MLP (ReLU' = [pre-activation > 0]) and the
//                  trilinear interpolation, on MFMA with transposed-weight
//                  fragments; weight gradients dW = sum_s dY_s X_s^T as MFMA
//                  over the samples (operands transposed through LDS) summed
//                  per block in LDS, then added to the caller's gradient;
//                  hash-table gradients scattered with FP32 atomics.
// Samples skipped by alpha_thre and samples after termination carry no
// gradient (the branch decisions are constants of the forward pass).
//
// MFMA operand layout (v_mfma_f32_16x16x4_f32, transposed formulation
// D[neuron][sample] = W[neuron][k] X[k][sample]): lane l = (col = l % 16,
// g = l / 16); A: W[row(i = col)][k(slot, g)], B: X[k(slot, g)][sample col],
// D: rows 4g..4g+3 of a 16-row block for sample col.  A layer's D blocks are
// the next layer's B slots (slot 4 nb + e = block nb element e), so each
// weight matrix is packed once per call into fragments whose k order
// follows the slots (k_diff_pack, from index tables built on the host).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fvc_internal.h"
#include "march.h"
#include "mlp_mfma.h"
#include "render_common.h"
#include "wave.h"

namespace fvc {

// ------------------------------------------------------------ fragments
enum {
  F_W1 = 0,     // 4 blocks x 8 slots (encoding)
  F_W2 = 32,    // 4 x 16
  F_W3 = 96,    // 1 x 16
  F_H1 = 112,   // 4 x 8 (head input)
  F_H2 = 144,   // 1 x 16
  B_H2T = 160,  // 4 blocks (d hidden) x 4 slots (16 outputs)
  B_H1T = 176,  // 1 block (latent / padding inputs) x 16 slots
  B_W3T = 192,  // 4 x 4
  B_W2T = 208,  // 4 x 16
  B_W1T = 272,  // 2 blocks (encoding, quad-level order) x 16
  N_DIFF_FRAGS = 304
};

// k index of slot m in lane group g
static int k_chain(int m, int g) { return 16 * (m / 4) + 4 * g + (m % 4); }
static int k_enc(int m, int g) { return 2 * (g + 4 * (m / 2)) + (m % 2); }  // level g + 4 (m/2), feature m % 2
static int k_head(int m, int g) {                                        // head input column
  if (m >= 4) return 8 + 4 * g + (m - 4);                                // SH coefficient
  const int r = 4 * g + m;                                               // base output row
  if (r >= 1 && r <= 8) return r - 1;                                    // latent r - 1
  return r == 0 ? 24 : r + 16;                                           // 1.0 padding columns
}
// encoding feature of row i of d_enc block nb (the encoder's slot order)
static int row_enc(int nb, int i) { return k_enc(4 * nb + (i % 4), i / 4); }
// head input feature of row i of the d_head_input block
static int row_head(int i) { return k_head(i % 4, i / 4); }

// Index of every fragment value in the concatenated parameters
// [mlp_base | mlp_head] (-1: zero).
std::vector<int> diff_frag_index() {
  const int n_base = 32 * 64 + 64 * 64 + 64 * 16;
  const int W1 = 0, W2 = 32 * 64, W3 = 32 * 64 + 64 * 64, H1 = n_base, H2 = n_base + 32 * 64;
  std::vector<int> t((size_t)N_DIFF_FRAGS * 64, -1);
  auto put = [&](int f0, int n_blocks, int n_slots, auto val) {
    for (int nb = 0; nb < n_blocks; nb++)
      for (int m = 0; m < n_slots; m++)
        for (int lane = 0; lane < 64; lane++) t[((size_t)(f0 + nb * n_slots + m)) * 64 + lane] = val(nb, m, lane % 16, lane / 16);
  };
  // forward: A = W[16 nb + i][k(m, g)]
  put(F_W1, 4, 8, [&](int nb, int m, int i, int g) { return W1 + (16 * nb + i) * 32 + k_enc(m, g); });
  put(F_W2, 4, 16, [&](int nb, int m, int i, int g) { return W2 + (16 * nb + i) * 64 + k_chain(m, g); });
  put(F_W3, 1, 16, [&](int, int m, int i, int g) { return W3 + i * 64 + k_chain(m, g); });
  put(F_H1, 4, 8, [&](int nb, int m, int i, int g) { return H1 + (16 * nb + i) * 32 + k_head(m, g); });
  put(F_H2, 1, 16, [&](int, int m, int i, int g) { return H2 + i * 64 + k_chain(m, g); });
  // backward: A = W^T[row][k(m, g)] = W[k(m, g)][row]
  put(B_H2T, 4, 4, [&](int nb, int m, int i, int g) { return H2 + k_chain(m, g) * 64 + 16 * nb + i; });
  put(B_H1T, 1, 16, [&](int, int m, int i, int g) { return H1 + k_chain(m, g) * 32 + row_head(i); });
  put(B_W3T, 4, 4, [&](int nb, int m, int i, int g) { return W3 + k_chain(m, g) * 64 + 16 * nb + i; });
  put(B_W2T, 4, 16, [&](int nb, int m, int i, int g) { return W2 + k_chain(m, g) * 64 + 16 * nb + i; });
  put(B_W1T, 2, 16, [&](int nb, int m, int i, int g) { return W1 + k_chain(m, g) * 32 + row_enc(nb, i); });
  return t;
}

__global__ void k_diff_pack(const int* __restrict__ idx, const float* __restrict__ base,
                            const float* __restrict__ head, int n_base, float* __restrict__ frags) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N_DIFF_FRAGS * 64) return;
  const int k = idx[s];
  frags[s] = k < 0 ? 0.0f : (k < n_base ? base[k] : head[k - n_base]);
}

// ------------------------------------------------------------ device math
__device__ __forceinline__ f4 mma4(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// NB output blocks from M input slots with fragments fr[(nb * M + m) * 64]
// (the lane index is made opaque per call so the fragment loads of one layer
// are issued together but not hoisted across layers)
template <int NB, int M>
__device__ __forceinline__ void dense4(const float* __restrict__ fr, int lane_, const float* x, f4* y) {
  const int lane = opaque(lane_);
#pragma unroll
  for (int nb = 0; nb < NB; nb++) {
    f4 acc = f4zero();
#pragma unroll
    for (int m = 0; m < M; m++) acc = mma4(fr[(nb * M + m) * 64 + lane], x[m], acc);
    y[nb] = acc;
  }
}
template <int NB> __device__ __forceinline__ void relu_slots(const f4* y, float* x) {
#pragma unroll
  for (int nb = 0; nb < NB; nb++)
#pragma unroll
    for (int e = 0; e < 4; e++) x[4 * nb + e] = relu_bits(y[nb][e]);
}
template <int NB> __device__ __forceinline__ void relu_grad(const f4* pre, f4* d) {
#pragma unroll
  for (int nb = 0; nb < NB; nb++)
#pragma unroll
    for (int e = 0; e < 4; e++) d[nb][e] = pre[nb][e] > 0.0f ? d[nb][e] : 0.0f;
}

// SH degree 4 coefficients 4g..4g+3 of dir01 * 2 - 1 in FP32
// (my_fully_fused_mlp.cu:206-234)
__device__ __forceinline__ void sh4_quarter_f32(V3 d01, int g, float* o) {
  const float x = d01.x * 2.0f - 1.0f, y = d01.y * 2.0f - 1.0f, z = d01.z * 2.0f - 1.0f;
  const float xy = x * y, xz = x * z, yz = y * z, x2 = x * x, y2 = y * y, z2 = z * z;
  if (g == 0) {
    o[0] = 0.28209479177387814f;
    o[1] = -0.48860251190291987f * y;
    o[2] = 0.48860251190291987f * z;
    o[3] = -0.48860251190291987f * x;
  } else if (g == 1) {
    o[0] = 1.0925484305920792f * xy;
    o[1] = -1.0925484305920792f * yz;
    o[2] = 0.94617469575755997f * z2 - 0.31539156525251999f;
    o[3] = -1.0925484305920792f * xz;
  } else if (g == 2) {
    o[0] = 0.54627421529603959f * x2 - 0.54627421529603959f * y2;
    o[1] = 0.59004358992664352f * y * (-3.0f * x2 + y2);
    o[2] = 2.8906114426405538f * xy * z;
    o[3] = 0.45704579946446572f * y * (1.0f - 5.0f * z2);
  } else {
    o[0] = 0.3731763325901154f * z * (5.0f * z2 - 3.0f);
    o[1] = 0.45704579946446572f * x * (1.0f - 5.0f * z2);
    o[2] = 1.4453057213202769f * z * (x2 - y2);
    o[3] = 0.59004358992664352f * x * (-x2 + 3.0f * y2);
  }
}

// Corners of levels g, g+4, g+8, g+12 at p (clamped to [0,1]^3): entry
// index (table entries) and trilinear weight, corner bits x=1, y=2, z=4
// (tcnn GridEncoding, the hash of the iNGP paper).
__device__ __forceinline__ void quad_corners(const GridLevel* __restrict__ lv, const GridParams& gp, V3 p, int g,
                                             uint32_t (&idx)[4][8], float (&w)[4][8]) {
  const float px = fminf_(fmaxf_(p.x, 0.0f), 1.0f);
  const float py = fminf_(fmaxf_(p.y, 0.0f), 1.0f);
  const float pz = fminf_(fmaxf_(p.z, 0.0f), 1.0f);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int l = 4 * j + g;
    const GridLevel L = lv[l];
    const float ax = px * L.scale + 0.5f, ay = py * L.scale + 0.5f, az = pz * L.scale + 0.5f;
    const float fx = floorf(ax), fy = floorf(ay), fz = floorf(az);
    const uint32_t gx = (uint32_t)(int)fx, gy = (uint32_t)(int)fy, gz = (uint32_t)(int)fz;
    const float wx = ax - fx, wy = ay - fy, wz = az - fz;
    const float ux = 1.0f - wx, uy = 1.0f - wy, uz = 1.0f - wz;
    const float wxy[4] = {ux * uy, wx * uy, ux * wy, wx * wy};
    const bool dense = (gp.dense_mask >> l) & 1u;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const uint32_t x = gx + (c & 1), y = gy + ((c >> 1) & 1), z = gz + (c >> 2);
      uint32_t t;
      if (dense) {
        t = x + y * L.res + z * L.res2;
        t = t % L.size;
      } else {
        t = (x ^ (y * 2654435761u) ^ (z * 805459861u)) & (L.size - 1u);
      }
      idx[j][c] = L.offset + t;
      w[j][c] = wxy[c & 3] * ((c & 4) ? wz : uz);
    }
  }
}

// network forward of one 16-sample tile; activations kept for the backward
struct FwdActs {
  float enc[8];
  f4 a1[4], a2[4], out, z1[4], o;
  float xh[8];
};

// colour MLP on the base-output rows 4g..4g+3 in `rows` (rows 1..8 are the
// latents; row 0 and rows 9..15 give way to the 1.0 padding inputs)
__device__ __forceinline__ void head_forward(const float* __restrict__ fr, int lane, int g, f4 rows, V3 d01,
                                             FwdActs& a) {
  float sh[4], h[16];
  sh4_quarter_f32(d01, g, sh);
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int r = 4 * g + m;
    a.xh[m] = (r >= 1 && r <= 8) ? rows[m] : 1.0f;
    a.xh[4 + m] = sh[m];
  }
  dense4<4, 8>(fr + F_H1 * 64, lane, a.xh, a.z1);
  relu_slots<4>(a.z1, h);
  dense4<1, 16>(fr + F_H2 * 64, lane, h, &a.o);
}

__device__ __forceinline__ void net_forward(const float* __restrict__ fr, const float* __restrict__ table,
                                            const GridLevel* __restrict__ lv, const GridParams& gp, V3 p, V3 d01,
                                            int lane, int g, FwdActs& a) {
  uint32_t idx[4][8];
  float w[4][8];
  quad_corners(lv, gp, p, g, idx, w);
  const float2* t2 = reinterpret_cast<const float2*>(table);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    float e0 = 0.0f, e1 = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const float2 v = t2[idx[j][c]];
      e0 = e0 + w[j][c] * v.x;
      e1 = e1 + w[j][c] * v.y;
    }
    a.enc[2 * j] = e0;
    a.enc[2 * j + 1] = e1;
  }
  float h[16];
  dense4<4, 8>(fr + F_W1 * 64, lane, a.enc, a.a1);
  relu_slots<4>(a.a1, h);
  dense4<4, 16>(fr + F_W2 * 64, lane, h, a.a2);
  relu_slots<4>(a.a2, h);
  dense4<1, 16>(fr + F_W3 * 64, lane, h, &a.out);
  head_forward(fr, lane, g, a.out, d01, a);
}

__device__ __forceinline__ V3 ray_dir01(const RayState& r) { return unit_to_01(V3{r.dx, r.dy, r.dz}); }

// ------------------------------------------------------------ march
__global__ __launch_bounds__(256) void k_diff_count(fvc_raymarch_info rm, V3 aabb_from, V3 aabb_to,
                                                    const int* __restrict__ pix, int n, MarchGrid mg,
                                                    const uint32_t* __restrict__ coarse_g,
                                                    RayState* __restrict__ rays, int* __restrict__ counts) {
  __shared__ uint32_t coarse[16 * COARSE_WORDS];
  load_coarse(coarse_g, mg.occ.nlvl, coarse);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int W = rm.cam_info.res_x;
  RayState r;
  const bool hit = init_ray_state(rm, aabb_from, aabb_to, pix[i] % W, pix[i] / W, r);
  rays[i] = r;
  int cnt = 0;
  if (hit) {
    MarchRay mr = march_ray(r, mg.tab);
    int bins = 0;
    while (true) {
      float t0;
      const int res = march_trip(mg, coarse, mr, t0, bins);
      if (res == TRIP_END) break;
      if (res == TRIP_SAMPLE) cnt++;
    }
  }
  counts[i] = cnt;
}

__global__ __launch_bounds__(256) void k_diff_emit(int n, MarchGrid mg, ContractParams con,
                                                   const uint32_t* __restrict__ coarse_g,
                                                   const RayState* __restrict__ rays, const int* __restrict__ counts,
                                                   const int64_t* __restrict__ offsets, float* __restrict__ s_pos,
                                                   float* __restrict__ s_dt, int* __restrict__ s_ray) {
  __shared__ uint32_t coarse[16 * COARSE_WORDS];
  load_coarse(coarse_g, mg.occ.nlvl, coarse);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || counts[i] == 0) return;
  const RayState r = rays[i];
  MarchRay mr = march_ray(r, mg.tab);
  const V3 o = {r.ox, r.oy, r.oz}, d = {r.dx, r.dy, r.dz};
  int q = (int)offsets[i];  // the batch holds fewer than 2^31 samples
  const int end = q + counts[i];
  int bins = 0;
  while (q < end) {
    float t0;
    const int res = march_trip(mg, coarse, mr, t0, bins);
    if (res == TRIP_END) break;
    if (res != TRIP_SAMPLE) continue;
    const float dt = step_dt(mg.st, t0);
    const float tm = (t0 + (t0 + dt)) * 0.5f;
    const V3 p = contract(o + tm * d, con);
    s_pos[3 * q] = p.x; s_pos[3 * q + 1] = p.y; s_pos[3 * q + 2] = p.z;
    s_dt[q] = dt;
    s_ray[q] = i;
    q++;
  }
}

// ------------------------------------------------------------ network
__global__ __launch_bounds__(256) void k_diff_net_fwd(int n, const int* __restrict__ list,
                                                      const float* __restrict__ s_pos,
                                                      const int* __restrict__ s_ray,
                                                      const RayState* __restrict__ rays,
                                                      const float* __restrict__ table, GridParams gp,
                                                      const float* __restrict__ frags, float4* __restrict__ out4) {
  __shared__ GridLevel lv[MAX_LEVELS];
  if (threadIdx.x < MAX_LEVELS) lv[threadIdx.x] = gp.lv[threadIdx.x];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, g = lane >> 4;
  const int n_groups = (n + 15) / 16;
  for (int grp = blockIdx.x * 4 + wave; grp < n_groups; grp += gridDim.x * 4) {
    const int j = grp * 16 + col;
    const bool valid = j < n;
    const int sl = list ? list[valid ? j : n - 1] : (valid ? j : n - 1);
    const V3 p = {s_pos[3 * sl], s_pos[3 * sl + 1], s_pos[3 * sl + 2]};
    const V3 d01 = ray_dir01(rays[s_ray[sl]]);
    FwdActs a;
    net_forward(frags, table, lv, gp, p, d01, lane, g, a);
    if (g == 0 && valid) out4[sl] = make_float4(a.o[0], a.o[1], a.o[2], a.out[0]);
  }
}

// ------------------------------------------------------------ compositing
__device__ __forceinline__ void sample_terms(float4 o4, float dt, float& sigma, float& alpha, float (&c)[3]) {
  sigma = expf(o4.w);
  alpha = fminf_(fmaxf_(1.0f - expf(-sigma * dt), 0.0f), 1.0f);
  c[0] = 1.0f / (1.0f + expf(-o4.x));
  c[1] = 1.0f / (1.0f + expf(-o4.y));
  c[2] = 1.0f / (1.0f + expf(-o4.z));
}

__global__ __launch_bounds__(256) void k_diff_composite(int n, float alpha_thre, const int* __restrict__ counts,
                                                        const int64_t* __restrict__ offsets,
                                                        const float4* __restrict__ out4,
                                                        const float* __restrict__ s_dt, float* __restrict__ rgb,
                                                        float4* __restrict__ ray_out, int* __restrict__ n_used) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int off = (int)offsets[i], cnt = counts[i];
  float C[3] = {0.0f, 0.0f, 0.0f}, T = 1.0f;
  int k = 0;
  for (; k < cnt; k++) {
    float sigma, alpha, c[3];
    sample_terms(out4[off + k], s_dt[off + k], sigma, alpha, c);
    if (alpha > alpha_thre) {
      const float w = alpha * T;
      for (int ch = 0; ch < 3; ch++) C[ch] = C[ch] + w * c[ch];
      T = T * (1.0f - alpha);
    }
    if (T < 1e-4f) { k++; break; }
  }
  const float4 res = make_float4(C[0] + T, C[1] + T, C[2] + T, T);
  ray_out[i] = res;
  n_used[i] = k;  // samples walked (the backward walks the same ones)
  rgb[3 * i] = res.x; rgb[3 * i + 1] = res.y; rgb[3 * i + 2] = res.z;
}

__global__ __launch_bounds__(256) void k_diff_composite_bwd(int n, float alpha_thre, const int* __restrict__ counts,
                                                            const int64_t* __restrict__ offsets,
                                                            const int* __restrict__ n_walk,
                                                            const float4* __restrict__ out4,
                                                            const float* __restrict__ s_dt,
                                                            const float4* __restrict__ ray_out,
                                                            const float* __restrict__ d_rgb,
                                                            float4* __restrict__ d_out4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int off = (int)offsets[i], cnt = counts[i], walk = n_walk[i];
  const float4 ro = ray_out[i];
  const float tot[3] = {ro.x, ro.y, ro.z};
  const float dr[3] = {d_rgb[3 * i], d_rgb[3 * i + 1], d_rgb[3 * i + 2]};
  float T = 1.0f, pre[3] = {0.0f, 0.0f, 0.0f};
  for (int k = 0; k < cnt; k++) {
    float4 d = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (k < walk) {
      const float dt = s_dt[off + k];
      float sigma, alpha, c[3];
      sample_terms(out4[off + k], dt, sigma, alpha, c);
      if (alpha > alpha_thre) {
        const float w = alpha * T;
        const float T1 = T * (1.0f - alpha);
        float dsig = 0.0f, dc[3];
        for (int ch = 0; ch < 3; ch++) {
          pre[ch] = pre[ch] + w * c[ch];
          const float S = tot[ch] - pre[ch];
          dsig = dsig + dr[ch] * (T1 * c[ch] - S);
          dc[ch] = dr[ch] * w * c[ch] * (1.0f - c[ch]);
        }
        d = make_float4(dc[0], dc[1], dc[2], dsig * dt * sigma);
        T = T1;
      }
    }
    d_out4[off + k] = d;
  }
}

// ------------------------------------------------------------ backward
// gradient accumulators of one block in LDS (row-major [out][in])
enum {
  G_W1 = 0, G_W2 = 64 * 32, G_W3 = 64 * 32 + 64 * 64, G_H1 = 64 * 32 + 64 * 64 + 16 * 64,
  G_H2 = G_H1 + 64 * 32, G_WORDS = G_H2 + 16 * 64
};
static const int STG = 20;  // LDS row stride of the transposed staging (floats)


// dW[n_out][n_in] += sum over the tile's 16 samples of dY[row][s] X[col][s]:
// dY as NBY D blocks (rows 16 nb + 4 g + e), X as MX slots with column
// kcol(m, g).  Samples are stored in LDS in the order 4 (s % 4) + s / 4 so
// a lane reads its 4 k-steps (samples g, g+4, g+8, g+12) as one float4.
template <int NBY, int MX, class KCol>
__device__ __forceinline__ void wgrad(float* ys, float* xs, const f4* dY, const float* x, KCol kcol, int col, int g,
                                      float* acc, int n_in) {
  const int ps = 4 * (col % 4) + col / 4;
#pragma unroll
  for (int nb = 0; nb < NBY; nb++)
#pragma unroll
    for (int e = 0; e < 4; e++) ys[(16 * nb + 4 * g + e) * STG + ps] = dY[nb][e];
#pragma unroll
  for (int m = 0; m < MX; m++) xs[kcol(m, g) * STG + ps] = x[m];
  wave_sync();
  const int n_kb = MX * 4 / 16;
#pragma unroll
  for (int jb = 0; jb < NBY; jb++) {
    const float4 a = *reinterpret_cast<const float4*>(&ys[(16 * jb + col) * STG + 4 * g]);
#pragma unroll
    for (int kb = 0; kb < n_kb; kb++) {
      const float4 b = *reinterpret_cast<const float4*>(&xs[(16 * kb + col) * STG + 4 * g]);
      f4 d = mma4(a.x, b.x, f4zero());
      d = mma4(a.y, b.y, d);
      d = mma4(a.z, b.z, d);
      d = mma4(a.w, b.w, d);
#pragma unroll
      for (int e = 0; e < 4; e++) atomicAdd(&acc[(16 * jb + 4 * g + e) * n_in + 16 * kb + col], d[e]);
    }
  }
  wave_sync();
}

// column maps of the B slots (device copies of k_chain, k_enc, k_head)
struct KChain {
  __device__ int operator()(int m, int g) const { return 16 * (m / 4) + 4 * g + (m % 4); }
};
struct KEnc {
  __device__ int operator()(int m, int g) const { return 2 * (g + 4 * (m / 2)) + (m % 2); }
};
struct KHead {
  __device__ int operator()(int m, int g) const {
    if (m >= 4) return 8 + 4 * g + (m - 4);
    const int r = 4 * g + m;
    if (r >= 1 && r <= 8) return r - 1;
    return r == 0 ? 24 : r + 16;
  }
};

// Adds (v0, v1) to entry `key` of the hash-table gradient for the 16
// samples of a lane row (lane col = sample).  Consecutive samples of a tile
// lie along one ray, so at the coarse levels they share corners: a
// segmented inclusive scan over runs of equal keys sums each run, and only
// the run's last lane issues the two FP32 atomics.
__device__ __forceinline__ void scatter_run(float* __restrict__ g_pos, uint32_t key, float v0, float v1, int col,
                                            bool valid) {
  const uint32_t kprev = __shfl_up(key, 1, 16);
  const bool vprev = __shfl_up((int)valid, 1, 16) != 0;
  int f = (col == 0 || kprev != key || !vprev || !valid) ? 1 : 0;  // run head
#pragma unroll
  for (int d = 1; d < 16; d <<= 1) {
    const float u0 = __shfl_up(v0, d, 16), u1 = __shfl_up(v1, d, 16);
    const int fu = __shfl_up(f, d, 16);
    if (col >= d) {
      if (!f) { v0 += u0; v1 += u1; }
      f |= fu;
    }
  }
  const uint32_t knext = __shfl_down(key, 1, 16);
  const bool vnext = __shfl_down((int)valid, 1, 16) != 0;
  const bool tail = col == 15 || knext != key || !vnext;
  if (valid && tail && (v0 != 0.0f || v1 != 0.0f)) {
    unsafeAtomicAdd(&g_pos[2 * (size_t)key], v0);
    unsafeAtomicAdd(&g_pos[2 * (size_t)key + 1], v1);
  }
}

// colour MLP backward of a tile: d_o (output rows 4g..4g+3), d_z1
// (hidden pre-activations) and d_xh (head inputs in the rows of the
// base output, row r = latent r - 1 for r in 1..8)
__device__ __forceinline__ void head_backward(const float* __restrict__ frags, int lane, int g, const FwdActs& a,
                                              float4 dq, f4& d_o, f4* d_z1, f4& d_xh) {
  d_o = g == 0 ? f4{dq.x, dq.y, dq.z, 0.0f} : f4zero();
  float x[16];
#pragma unroll
  for (int e = 0; e < 4; e++) x[e] = d_o[e];
  dense4<4, 4>(frags + B_H2T * 64, lane, x, d_z1);
  relu_grad<4>(a.z1, d_z1);
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
#pragma unroll
    for (int e = 0; e < 4; e++) x[4 * nb + e] = d_z1[nb][e];
  dense4<1, 16>(frags + B_H1T * 64, lane, x, &d_xh);
}

__device__ __forceinline__ void head_wgrad(float* ys, float* xs, const FwdActs& a, const f4& d_o, const f4* d_z1,
                                           int col, int g, float* acc) {
  float h[16];
  relu_slots<4>(a.z1, h);
  wgrad<1, 16>(ys, xs, &d_o, h, KChain(), col, g, acc + G_H2, 64);
  wgrad<4, 8>(ys, xs, d_z1, a.xh, KHead(), col, g, acc + G_H1, 32);
}

// block accumulators into the caller's gradient blocks
__device__ __forceinline__ void flush_grads(const float* acc, float* g_base, float* g_head) {
  for (int k = threadIdx.x; k < G_WORDS; k += blockDim.x) {
    const float v = acc[k];
    if (v == 0.0f) continue;
    if (k < G_H1) unsafeAtomicAdd(&g_base[k], v);
    else unsafeAtomicAdd(&g_head[k - G_H1], v);
  }
}
__device__ __forceinline__ void flush_head_grads(const float* acc, float* g_head) {
  for (int k = G_H1 + threadIdx.x; k < G_WORDS; k += blockDim.x) {
    const float v = acc[k];
    if (v != 0.0f) unsafeAtomicAdd(&g_head[k - G_H1], v);
  }
}

__global__ __launch_bounds__(256) void k_diff_net_bwd(int n, const int* __restrict__ list,
                                                      const float* __restrict__ s_pos,
                                                      const int* __restrict__ s_ray,
                                                      const RayState* __restrict__ rays,
                                                      const float* __restrict__ table, GridParams gp,
                                                      const float* __restrict__ frags,
                                                      const float4* __restrict__ d_out4, float* __restrict__ g_pos,
                                                      float* __restrict__ g_base, float* __restrict__ g_head) {
  __shared__ GridLevel lv[MAX_LEVELS];
  __shared__ float acc[G_WORDS];
  __shared__ __attribute__((aligned(16))) float stage[4][2][64 * STG];
  if (threadIdx.x < MAX_LEVELS) lv[threadIdx.x] = gp.lv[threadIdx.x];
  for (int k = threadIdx.x; k < G_WORDS; k += blockDim.x) acc[k] = 0.0f;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, g = lane >> 4;
  float* ys = stage[wave][0];
  float* xs = stage[wave][1];
  const int n_groups = (n + 15) / 16;
  for (int grp = blockIdx.x * 4 + wave; grp < n_groups; grp += gridDim.x * 4) {
    const int j = grp * 16 + col;
    const bool valid = j < n;
    const int sl = list ? list[valid ? j : n - 1] : (valid ? j : n - 1);
    const float4 dq = valid ? d_out4[sl] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const V3 p = {s_pos[3 * sl], s_pos[3 * sl + 1], s_pos[3 * sl + 2]};
    const V3 d01 = ray_dir01(rays[s_ray[sl]]);
    FwdActs a;
    net_forward(frags, table, lv, gp, p, d01, lane, g, a);
    // colour MLP
    f4 d_o, d_z1[4], d_xh;
    head_backward(frags, lane, g, a, dq, d_o, d_z1, d_xh);
    float x[16];
    // density MLP: output rows 4g + e (0: log density, 1..8: latents)
    f4 d_out;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int r = 4 * g + e;
      d_out[e] = r == 0 ? dq.w : ((r >= 1 && r <= 8) ? d_xh[e] : 0.0f);
    }
    f4 d_a2[4], d_a1[4], d_enc[2];
    for (int e = 0; e < 4; e++) x[e] = d_out[e];
    dense4<4, 4>(frags + B_W3T * 64, lane, x, d_a2);
    relu_grad<4>(a.a2, d_a2);
#pragma unroll
    for (int nb = 0; nb < 4; nb++)
#pragma unroll
      for (int e = 0; e < 4; e++) x[4 * nb + e] = d_a2[nb][e];
    dense4<4, 16>(frags + B_W2T * 64, lane, x, d_a1);
    relu_grad<4>(a.a1, d_a1);
#pragma unroll
    for (int nb = 0; nb < 4; nb++)
#pragma unroll
      for (int e = 0; e < 4; e++) x[4 * nb + e] = d_a1[nb][e];
    dense4<2, 16>(frags + B_W1T * 64, lane, x, d_enc);
    // hash-grid scatter: feature slot 2j + k of level 4j + g is d_enc[j / 2][2 (j % 2) + k]
    {
      uint32_t idx[4][8];
      float w[4][8];
      quad_corners(lv, gp, p, g, idx, w);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const float d0 = valid ? d_enc[j >> 1][2 * (j & 1)] : 0.0f;
        const float d1 = valid ? d_enc[j >> 1][2 * (j & 1) + 1] : 0.0f;
#pragma unroll
        for (int c = 0; c < 8; c++) scatter_run(g_pos, idx[j][c], w[j][c] * d0, w[j][c] * d1, col, valid);
      }
    }
    // weight gradients (activations as the forward's B slots)
    head_wgrad(ys, xs, a, d_o, d_z1, col, g, acc);
    float h[16];
    relu_slots<4>(a.a2, h);
    wgrad<1, 16>(ys, xs, &d_out, h, KChain(), col, g, acc + G_W3, 64);
    relu_slots<4>(a.a1, h);
    wgrad<4, 16>(ys, xs, d_a2, h, KChain(), col, g, acc + G_W2, 64);
    wgrad<4, 8>(ys, xs, d_a1, a.enc, KEnc(), col, g, acc + G_W1, 32);
  }
  __syncthreads();
  flush_grads(acc, g_base, g_head);
}

// ------------------------------------------------------------ cached mode
// Differentiable cached render: classification of every occupied sample as
// in the cache marcher (classify_sample); resample samples take the network
// (k_diff_net_fwd / _bwd over list_r), cache samples the trilinear lookup of
// the brick pool and the colour MLP (k_diffc_head_fwd / _bwd over list_c).
// The adjoint of the lookup (generateNetworkInput_cache_shfl_perf_kernel_cg,
// cache_performance_renderer.cu:678-787): with s = log(max(sum_c w_c
// sigma_c, 1e-6)) and l_i = sum_c w_c l_c,i / sum_c w_c isset_c,
//   dL/dsigma_c = w_c dL/ds / sum (0 when the clamp is active),
//   dL/dl_c,i   = w_c dL/dl_i / sum_c w_c isset_c;
// isset is a mask.  Voxel gradients go to g_pool [slot][512][10] (8 latents,
// density, isset = 0) with FP32 atomics.
__global__ __launch_bounds__(256) void k_diffc_count(fvc_raymarch_info rm, V3 aabb_from, V3 aabb_to,
                                                     const int* __restrict__ pix, int n, MarchGrid mg, CacheView cv,
                                                     const uint32_t* __restrict__ coarse_g,
                                                     RayState* __restrict__ rays, int* __restrict__ counts_r,
                                                     int* __restrict__ counts_c) {
  __shared__ uint32_t coarse[16 * COARSE_WORDS];
  load_coarse(coarse_g, mg.occ.nlvl, coarse);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int W = rm.cam_info.res_x;
  RayState r;
  const bool hit = init_ray_state(rm, aabb_from, aabb_to, pix[i] % W, pix[i] / W, r);
  rays[i] = r;
  int nr = 0, nc = 0;
  if (hit) {
    MarchRay mr = march_ray(r, mg.tab);
    const V3 oc = transform(cv.cam.world2cam, mr.o.x, mr.o.y, mr.o.z, 1.0f);
    const V3 dc = transform(cv.cam.world2cam, mr.d.x, mr.d.y, mr.d.z, 0.0f);
    int bins = 0;
    while (true) {
      float t0;
      const int res = march_trip(mg, coarse, mr, t0, bins);
      if (res == TRIP_END) break;
      if (res != TRIP_SAMPLE) continue;
      const int k = classify_sample(cv, oc, dc, t0, step_dt(mg.st, t0));
      nr += k == SAMPLE_RESAMPLE;
      nc += k == SAMPLE_CACHE;
    }
  }
  counts_r[i] = nr;
  counts_c[i] = nc;
}

__global__ __launch_bounds__(256) void k_diffc_emit(int n, MarchGrid mg, ContractParams con, CacheView cv,
                                                    const uint32_t* __restrict__ coarse_g,
                                                    const RayState* __restrict__ rays,
                                                    const int* __restrict__ counts_r, const int* __restrict__ counts_c,
                                                    const int64_t* __restrict__ off_r,
                                                    const int64_t* __restrict__ off_c, int* __restrict__ counts,
                                                    int64_t* __restrict__ offsets, float* __restrict__ s_pos,
                                                    float* __restrict__ s_dt, int* __restrict__ s_ray,
                                                    int* __restrict__ list_r, int* __restrict__ list_c) {
  __shared__ uint32_t coarse[16 * COARSE_WORDS];
  load_coarse(coarse_g, mg.occ.nlvl, coarse);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int nr = counts_r[i], nc = counts_c[i];
  int q = (int)(off_r[i] + off_c[i]);  // the batch holds fewer than 2^31 samples
  counts[i] = nr + nc;
  offsets[i] = q;
  if (nr + nc == 0) return;
  const RayState r = rays[i];
  MarchRay mr = march_ray(r, mg.tab);
  const V3 o = {r.ox, r.oy, r.oz}, d = {r.dx, r.dy, r.dz};
  const V3 oc = transform(cv.cam.world2cam, o.x, o.y, o.z, 1.0f);
  const V3 dc = transform(cv.cam.world2cam, d.x, d.y, d.z, 0.0f);
  int kr = (int)off_r[i], kc = (int)off_c[i];
  const int end = q + nr + nc;
  int bins = 0;
  while (q < end) {
    float t0;
    const int res = march_trip(mg, coarse, mr, t0, bins);
    if (res == TRIP_END) break;
    if (res != TRIP_SAMPLE) continue;
    const float dt = step_dt(mg.st, t0);
    const int k = classify_sample(cv, oc, dc, t0, dt);
    if (k == SAMPLE_SKIP) continue;
    // network inputs at t0 + dt / 2 (cache_performance_renderer.cu:198-313)
    const float tm2 = t0 + dt * 0.5f;
    V3 p;
    if (k == SAMPLE_RESAMPLE) {
      p = contract(o + tm2 * d, con);
      list_r[kr++] = q;
    } else {
      float tf;
      p = cam2froxel(oc + tm2 * dc, cv.cam, cv.st, tf);
      list_c[kc++] = q;
    }
    s_pos[3 * q] = p.x; s_pos[3 * q + 1] = p.y; s_pos[3 * q + 2] = p.z;
    s_dt[q] = dt;
    s_ray[q] = i;
    q++;
  }
}

// trilinear lookup of one sample by its 4 lanes: lane group g fetches the
// corners (x, y) = (g >> 1, g & 1), z = 0 / 1; every lane ends with the sums
struct LookupLane {
  float w[2];
  int64_t v[2];  // voxel index in the reference's order (g_pool), -1: outside or unset brick
  float wsum, dsum, lat[8];
};
__device__ __forceinline__ void lookup_f32(const CacheView& cv, V3 f, int g, LookupLane& L) {
  const float fcx = f.x - 0.5f, fcy = f.y - 0.5f, fcz = f.z - 0.5f;
  const float ix = floorf(fcx), iy = floorf(fcy), iz = floorf(fcz);
  const float lx = fcx - ix, ly = fcy - iy, lz = fcz - iz;
  const float ox = (float)(g >> 1), oy = (float)(g & 1);
  const float px = ix + ox, py = iy + oy;
  const float wx = (1.0f - ox) * (1.0f - 2.0f * lx) + lx;
  const float wy = (1.0f - oy) * (1.0f - 2.0f * ly) + ly;
  const float wxy = wx * wy;
  const bool in_xy = px >= 0.0f && px < (float)cv.fx && py >= 0.0f && py < (float)cv.fy;
  float wd = 0.0f, dd = 0.0f, ld[8];
#pragma unroll
  for (int cc = 0; cc < 2; cc++) {
    const float pz = iz + (float)cc;
    const float wz = cc == 0 ? (1.0f - 2.0f * lz) + lz : lz;
    const float w = wxy * wz;
    L.w[cc] = w;
    L.v[cc] = -1;
    float cw = 0.0f, cd = 0.0f, cl[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (in_xy && pz >= 0.0f && pz < (float)cv.fz) {
      const int vx = (int)px, vy = (int)py, vz = (int)pz;
      const int slot = cv.table[((int64_t)(vz >> 3) * cv.by + (vy >> 3)) * cv.bx + (vx >> 3)];
      if (slot >= 0) {
        L.v[cc] = (int64_t)slot * BRICK_VOXELS + canon_voxel(vx, vy, vz);
        const Voxel vox = cv.pool[(int64_t)slot * BRICK_VOXELS + brick_voxel(vx, vy, vz)];
        cw = w * vox.isset;
        cd = w * vox.density;
#pragma unroll
        for (int k = 0; k < 8; k++) cl[k] = w * h2f(vox.lat[k]);
      }
    }
    if (cc == 0) {
      wd = cw; dd = cd;
#pragma unroll
      for (int k = 0; k < 8; k++) ld[k] = cl[k];
    } else {
      wd = wd + cw; dd = dd + cd;
#pragma unroll
      for (int k = 0; k < 8; k++) ld[k] = ld[k] + cl[k];
    }
  }
  L.wsum = xor32_sum(xor16_sum(wd));
  L.dsum = xor32_sum(xor16_sum(dd));
#pragma unroll
  for (int k = 0; k < 8; k++) L.lat[k] = xor32_sum(xor16_sum(ld[k])) / L.wsum;
}

// base-output rows 4g..4g+3 of a cache sample: latents in rows 1..8
__device__ __forceinline__ f4 lookup_rows(const LookupLane& L, int g) {
  f4 r;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int k = 4 * g + e - 1;
    float v = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; i++) v = k == i ? L.lat[i] : v;
    r[e] = v;
  }
  return r;
}

__global__ __launch_bounds__(256) void k_diffc_head_fwd(int n, const int* __restrict__ list,
                                                        const float* __restrict__ s_pos,
                                                        const int* __restrict__ s_ray,
                                                        const RayState* __restrict__ rays, CacheView cv,
                                                        const float* __restrict__ frags,
                                                        float4* __restrict__ out4) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, g = lane >> 4;
  const int n_groups = (n + 15) / 16;
  for (int grp = blockIdx.x * 4 + wave; grp < n_groups; grp += gridDim.x * 4) {
    const int j = grp * 16 + col;
    const bool valid = j < n;
    const int sl = list[valid ? j : n - 1];
    const V3 f = {s_pos[3 * sl], s_pos[3 * sl + 1], s_pos[3 * sl + 2]};
    const V3 d01 = ray_dir01(rays[s_ray[sl]]);
    LookupLane L;
    lookup_f32(cv, f, g, L);
    FwdActs a;
    head_forward(frags, lane, g, lookup_rows(L, g), d01, a);
    if (g == 0 && valid) out4[sl] = make_float4(a.o[0], a.o[1], a.o[2], logf(fmaxf_(L.dsum, 1e-6f)));
  }
}

__global__ __launch_bounds__(256) void k_diffc_head_bwd(int n, const int* __restrict__ list,
                                                        const float* __restrict__ s_pos,
                                                        const int* __restrict__ s_ray,
                                                        const RayState* __restrict__ rays, CacheView cv,
                                                        const float* __restrict__ frags,
                                                        const float4* __restrict__ d_out4,
                                                        float* __restrict__ g_head, float* __restrict__ g_pool) {
  __shared__ float acc[G_WORDS];
  __shared__ __attribute__((aligned(16))) float stage[4][2][64 * STG];
  for (int k = G_H1 + threadIdx.x; k < G_WORDS; k += blockDim.x) acc[k] = 0.0f;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, g = lane >> 4;
  float* ys = stage[wave][0];
  float* xs = stage[wave][1];
  const int n_groups = (n + 15) / 16;
  for (int grp = blockIdx.x * 4 + wave; grp < n_groups; grp += gridDim.x * 4) {
    const int j = grp * 16 + col;
    const bool valid = j < n;
    const int sl = list[valid ? j : n - 1];
    const float4 dq = valid ? d_out4[sl] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const V3 f = {s_pos[3 * sl], s_pos[3 * sl + 1], s_pos[3 * sl + 2]};
    const V3 d01 = ray_dir01(rays[s_ray[sl]]);
    LookupLane L;
    lookup_f32(cv, f, g, L);
    FwdActs a;
    head_forward(frags, lane, g, lookup_rows(L, g), d01, a);
    f4 d_o, d_z1[4], d_xh;
    head_backward(frags, lane, g, a, dq, d_o, d_z1, d_xh);
    head_wgrad(ys, xs, a, d_o, d_z1, col, g, acc);
    // dL/dl_i of latent i (row i + 1: lane group (i + 1) / 4, element (i + 1) % 4)
    float d_lat[8];
#pragma unroll
    for (int i = 0; i < 8; i++) d_lat[i] = __shfl(d_xh[(i + 1) & 3], col + 16 * ((i + 1) >> 2), 64);
    const float d_dsum = L.dsum > 1e-6f ? dq.w / L.dsum : 0.0f;
    if (valid) {
#pragma unroll
      for (int cc = 0; cc < 2; cc++) {
        if (L.v[cc] < 0) continue;
        float* gv = g_pool + L.v[cc] * 10;
        const float wl = L.w[cc] / L.wsum;
#pragma unroll
        for (int i = 0; i < 8; i++) unsafeAtomicAdd(&gv[i], wl * d_lat[i]);
        unsafeAtomicAdd(&gv[8], L.w[cc] * d_dsum);
      }
    }
  }
  __syncthreads();
  flush_head_grads(acc, g_head);
}

// ------------------------------------------------------------ host
static int persistent_blocks(int64_t n_groups) {
  const int64_t blocks = (n_groups + 3) / 4;
  return (int)std::max<int64_t>(1, std::min<int64_t>(blocks, 256 * 8));
}

// exclusive scan of counts[0..n) into offs[0..n] (offs[n] = total)
static int64_t scan_counts(DiffState& D, int* counts, int64_t* offs, int n, hipStream_t s) {
  FVC_HIP(hipMemsetAsync(counts + n, 0, sizeof(int), s));
  size_t tmp = 0;
  FVC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, counts, offs, n + 1, s));
  D.scan_tmp.ensure(std::max<size_t>(tmp, 16));
  FVC_HIP(hipcub::DeviceScan::ExclusiveSum(D.scan_tmp.p, tmp, counts, offs, n + 1, s));
  int64_t total = 0;
  FVC_HIP(hipMemcpyAsync(&total, offs + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  FVC_HIP(hipStreamSynchronize(s));
  return total;
}

void diff_forward(Renderer& r, const fvc_raymarch_info& rm, const int* pix, int n, const float* pos,
                  const float* base, const float* head, float* rgb, fvc_diff_stats* stats, bool cached) {
  const Model& m = *r.model;
  DiffState& D = r.diff;
  hipStream_t s = r.stream();
  if (cached && !r.caches[r.render_idx].initialized) throw Error(FVC_E_NOT_INITIALIZED, "Cache is not initialized!");
  if (cached && r.caches[r.render_idx].lat8)
    throw Error(FVC_E_UNSUPPORTED, "the differentiable cached render needs FP16 cache latents");
  if (D.frag_index.p == nullptr) {
    const std::vector<int> t = diff_frag_index();
    D.frag_index.ensure(t.size() * sizeof(int));
    FVC_HIP(hipMemcpy(D.frag_index.p, t.data(), t.size() * sizeof(int), hipMemcpyHostToDevice));
    D.frags.ensure((size_t)N_DIFF_FRAGS * 64 * sizeof(float));
  }
  D.pos = nullptr;  // set once the forward pass has completed
  D.n = n;
  D.cached = cached;
  D.n_samples = D.n_r = D.n_c = 0;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n == 0) {
    D.pos = pos; D.base = base; D.head = head;
    return;
  }
  D.rays.ensure((size_t)n * sizeof(RayState));
  D.counts.ensure(((size_t)n + 1) * sizeof(int));
  D.offsets.ensure(((size_t)n + 1) * sizeof(int64_t));
  D.ray_out.ensure((size_t)n * sizeof(float4));
  D.n_walk.ensure((size_t)n * sizeof(int));
  hipLaunchKernelGGL(k_diff_pack, dim3((N_DIFF_FRAGS * 64 + 255) / 256), dim3(256), 0, s, D.frag_index.as<int>(), base,
                     head, (int)m.n_base, D.frags.as<float>());
  const MarchGrid mg = march_grid(r, &rm, 1);
  const V3 af = {m.scene.aabb_from.x, m.scene.aabb_from.y, m.scene.aabb_from.z};
  const V3 at = {m.scene.aabb_to.x, m.scene.aabb_to.y, m.scene.aabb_to.z};
  const dim3 grid((n + 255) / 256);
  int64_t total64 = 0;
  if (!cached) {
    hipLaunchKernelGGL(k_diff_count, grid, dim3(256), 0, s, rm, af, at, pix, n, mg, m.coarse.as<uint32_t>(),
                       D.rays.as<RayState>(), D.counts.as<int>());
    FVC_HIP(hipGetLastError());
    total64 = scan_counts(D, D.counts.as<int>(), D.offsets.as<int64_t>(), n, s);
    D.n_r = total64;
  } else {
    D.cv = cache_view(r.caches[r.render_idx]);
    D.counts_c.ensure(((size_t)n + 1) * sizeof(int));
    D.off_r.ensure(((size_t)n + 1) * sizeof(int64_t));
    D.off_c.ensure(((size_t)n + 1) * sizeof(int64_t));
    hipLaunchKernelGGL(k_diffc_count, grid, dim3(256), 0, s, rm, af, at, pix, n, mg, D.cv, m.coarse.as<uint32_t>(),
                       D.rays.as<RayState>(), D.counts.as<int>(), D.counts_c.as<int>());
    FVC_HIP(hipGetLastError());
    D.n_r = scan_counts(D, D.counts.as<int>(), D.off_r.as<int64_t>(), n, s);
    D.n_c = scan_counts(D, D.counts_c.as<int>(), D.off_c.as<int64_t>(), n, s);
    total64 = D.n_r + D.n_c;
  }
  if (total64 >= ((int64_t)1 << 31) / 3)
    throw Error(FVC_E_INVALID, "more than 2^31 / 3 samples in one differentiable batch: pass fewer rays");
  const int total = (int)total64;
  D.n_samples = total;
  const size_t ns = (size_t)std::max(total, 1);
  D.s_pos.ensure(ns * 3 * sizeof(float));
  D.s_dt.ensure(ns * sizeof(float));
  D.s_ray.ensure(ns * sizeof(int));
  D.out4.ensure(ns * sizeof(float4));
  D.d_out4.ensure(ns * sizeof(float4));
  if (!cached) {
    hipLaunchKernelGGL(k_diff_emit, grid, dim3(256), 0, s, n, mg, m.con, m.coarse.as<uint32_t>(),
                       D.rays.as<RayState>(), D.counts.as<int>(), D.offsets.as<int64_t>(), D.s_pos.as<float>(),
                       D.s_dt.as<float>(), D.s_ray.as<int>());
  } else {
    D.list_r.ensure((size_t)std::max<int64_t>(D.n_r, 1) * sizeof(int));
    D.list_c.ensure((size_t)std::max<int64_t>(D.n_c, 1) * sizeof(int));
    D.counts_all.ensure((size_t)n * sizeof(int));
    hipLaunchKernelGGL(k_diffc_emit, grid, dim3(256), 0, s, n, mg, m.con, D.cv, m.coarse.as<uint32_t>(),
                       D.rays.as<RayState>(), D.counts.as<int>(), D.counts_c.as<int>(), D.off_r.as<int64_t>(),
                       D.off_c.as<int64_t>(), D.counts_all.as<int>(), D.offsets.as<int64_t>(), D.s_pos.as<float>(),
                       D.s_dt.as<float>(), D.s_ray.as<int>(), D.list_r.as<int>(), D.list_c.as<int>());
  }
  FVC_HIP(hipGetLastError());
  const int* list_r = cached ? D.list_r.as<int>() : nullptr;
  if (D.n_r > 0) {
    hipLaunchKernelGGL(k_diff_net_fwd, dim3(persistent_blocks((D.n_r + 15) / 16)), dim3(256), 0, s, (int)D.n_r,
                       list_r, D.s_pos.as<float>(), D.s_ray.as<int>(), D.rays.as<RayState>(), pos, m.grid,
                       D.frags.as<float>(), D.out4.as<float4>());
    FVC_HIP(hipGetLastError());
  }
  if (D.n_c > 0) {
    hipLaunchKernelGGL(k_diffc_head_fwd, dim3(persistent_blocks((D.n_c + 15) / 16)), dim3(256), 0, s, (int)D.n_c,
                       D.list_c.as<int>(), D.s_pos.as<float>(), D.s_ray.as<int>(), D.rays.as<RayState>(), D.cv,
                       D.frags.as<float>(), D.out4.as<float4>());
    FVC_HIP(hipGetLastError());
  }
  const int* counts = cached ? D.counts_all.as<int>() : D.counts.as<int>();
  hipLaunchKernelGGL(k_diff_composite, grid, dim3(256), 0, s, n, m.scene.alpha_thre, counts, D.offsets.as<int64_t>(),
                     D.out4.as<float4>(), D.s_dt.as<float>(), rgb, D.ray_out.as<float4>(), D.n_walk.as<int>());
  FVC_HIP(hipGetLastError());
  FVC_HIP(hipStreamSynchronize(s));
  D.pos = pos;
  D.base = base;
  D.head = head;
  if (stats) {
    stats->n_rays = n;
    stats->n_samples = total;
    stats->n_cache_samples = D.n_c;
  }
}

void diff_backward(Renderer& r, const float* d_rgb, float* g_pos, float* g_base, float* g_head, float* g_pool) {
  const Model& m = *r.model;
  DiffState& D = r.diff;
  if (D.pos == nullptr) throw Error(FVC_E_NOT_INITIALIZED, "no differentiable forward pass to differentiate");
  if (D.cached && D.n_c > 0 && g_pool == nullptr) throw Error(FVC_E_INVALID, "the cached backward needs g_pool");
  hipStream_t s = r.stream();
  const int n = D.n;
  if (n == 0) return;
  const int* counts = D.cached ? D.counts_all.as<int>() : D.counts.as<int>();
  hipLaunchKernelGGL(k_diff_composite_bwd, dim3((n + 255) / 256), dim3(256), 0, s, n, m.scene.alpha_thre, counts,
                     D.offsets.as<int64_t>(), D.n_walk.as<int>(), D.out4.as<float4>(), D.s_dt.as<float>(),
                     D.ray_out.as<float4>(), d_rgb, D.d_out4.as<float4>());
  FVC_HIP(hipGetLastError());
  if (D.n_r > 0) {
    hipLaunchKernelGGL(k_diff_net_bwd, dim3(persistent_blocks((D.n_r + 15) / 16)), dim3(256), 0, s, (int)D.n_r,
                       D.cached ? D.list_r.as<int>() : nullptr, D.s_pos.as<float>(), D.s_ray.as<int>(),
                       D.rays.as<RayState>(), D.pos, m.grid, D.frags.as<float>(), D.d_out4.as<float4>(), g_pos,
                       g_base, g_head);
    FVC_HIP(hipGetLastError());
  }
  if (D.n_c > 0) {
    hipLaunchKernelGGL(k_diffc_head_bwd, dim3(persistent_blocks((D.n_c + 15) / 16)), dim3(256), 0, s, (int)D.n_c,
                       D.list_c.as<int>(), D.s_pos.as<float>(), D.s_ray.as<int>(), D.rays.as<RayState>(), D.cv,
                       D.frags.as<float>(), D.d_out4.as<float4>(), g_head, g_pool);
    FVC_HIP(hipGetLastError());
  }
  FVC_HIP(hipStreamSynchronize(s));
}

}  // namespace fvc
