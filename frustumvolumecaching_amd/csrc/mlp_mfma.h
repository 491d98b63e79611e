// mlp_mfma.h -- register-resident MLP tiles on v_mfma_f32_16x16x32_f16.
//
// One wave64 evaluates 16 samples per step.  Operands use the transposed
// formulation D[neuron][sample] = W[neuron][k] * X[k][sample]:
//   A (16x32, weights):  lane l -> neuron row l%16, k-group g = l/16 (8 halves)
//   B (32x16, inputs):   lane l -> sample   l%16, k-group g       (8 halves)
//   D (16x16, f32):      lane l -> sample   l%16, neurons 4g..4g+3
// so the D fragments of one layer become the B fragment of the next without
// any data movement: B[kb] = {relu(D[2kb]) , relu(D[2kb+1])} as halves.  The
// k-index order inside a lane is a free permutation, applied to the weights
// once on the host (pack_fragments in network.hip).  Accumulation is FP32;
// activations are rounded to fp16 between layers, as the reference stores them.
#pragma once
#include <hip/hip_runtime.h>

#include "fvc_internal.h"

namespace fvc {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 f4zero() { return f4{0.0f, 0.0f, 0.0f, 0.0f}; }

// B fragment for K block kb from four 16-neuron D blocks (ReLU + fp16)
__device__ __forceinline__ h8 relu_pack(f4 lo, f4 hi) {
  h8 b;
  b[0] = (_Float16)fmaxf(lo[0], 0.0f); b[1] = (_Float16)fmaxf(lo[1], 0.0f);
  b[2] = (_Float16)fmaxf(lo[2], 0.0f); b[3] = (_Float16)fmaxf(lo[3], 0.0f);
  b[4] = (_Float16)fmaxf(hi[0], 0.0f); b[5] = (_Float16)fmaxf(hi[1], 0.0f);
  b[6] = (_Float16)fmaxf(hi[2], 0.0f); b[7] = (_Float16)fmaxf(hi[3], 0.0f);
  return b;
}

struct BaseWeights { h8 w1[4], w2[4][2], w3[2]; };
struct HeadWeights { h8 h1[4], h2[2]; };

__device__ __forceinline__ void load_base(const h8* __restrict__ frags, int lane, BaseWeights& w) {
#pragma unroll
  for (int nb = 0; nb < 4; nb++) w.w1[nb] = frags[(FRAG_W1 + nb) * 64 + lane];
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
#pragma unroll
    for (int kb = 0; kb < 2; kb++) w.w2[nb][kb] = frags[(FRAG_W2 + nb * 2 + kb) * 64 + lane];
#pragma unroll
  for (int kb = 0; kb < 2; kb++) w.w3[kb] = frags[(FRAG_W3 + kb) * 64 + lane];
}
__device__ __forceinline__ void load_head(const h8* __restrict__ frags, int lane, HeadWeights& w) {
#pragma unroll
  for (int nb = 0; nb < 4; nb++) w.h1[nb] = frags[(FRAG_H1 + nb) * 64 + lane];
#pragma unroll
  for (int kb = 0; kb < 2; kb++) w.h2[kb] = frags[(FRAG_H2 + kb) * 64 + lane];
}

// density / latent MLP 32 -> 64 -> 64 -> 16; returns output rows 4g..4g+3 (f32)
__device__ __forceinline__ f4 base_mlp(const BaseWeights& w, h8 x) {
  f4 d1[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) d1[nb] = mfma(w.w1[nb], x, f4zero());
  const h8 b2_0 = relu_pack(d1[0], d1[1]), b2_1 = relu_pack(d1[2], d1[3]);
  f4 d2[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) d2[nb] = mfma(w.w2[nb][1], b2_1, mfma(w.w2[nb][0], b2_0, f4zero()));
  const h8 b3_0 = relu_pack(d2[0], d2[1]), b3_1 = relu_pack(d2[2], d2[3]);
  return mfma(w.w3[1], b3_1, mfma(w.w3[0], b3_0, f4zero()));
}

// head MLP 32 -> 64 -> 16; returns output rows 4g..4g+3 (rows 0..2 = rgb)
__device__ __forceinline__ f4 head_mlp(const HeadWeights& w, h8 x) {
  f4 d1[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) d1[nb] = mfma(w.h1[nb], x, f4zero());
  const h8 b_0 = relu_pack(d1[0], d1[1]), b_1 = relu_pack(d1[2], d1[3]);
  return mfma(w.h2[1], b_1, mfma(w.h2[0], b_0, f4zero()));
}

// The same MLPs with the weight fragments read from LDS at each use (w =
// the N_FRAGS x 64 fragments staged by the block); li is the lane index made
// opaque per loop iteration by the caller so the reads are not hoisted into
// registers.  Frees 80 VGPRs for gathers in flight.
__device__ __forceinline__ f4 base_mlp_lds(const h8* w, int li, h8 x) {
  f4 d1[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) d1[nb] = mfma(w[(FRAG_W1 + nb) * 64 + li], x, f4zero());
  const h8 b2_0 = relu_pack(d1[0], d1[1]), b2_1 = relu_pack(d1[2], d1[3]);
  f4 d2[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
    d2[nb] = mfma(w[(FRAG_W2 + nb * 2 + 1) * 64 + li], b2_1, mfma(w[(FRAG_W2 + nb * 2) * 64 + li], b2_0, f4zero()));
  const h8 b3_0 = relu_pack(d2[0], d2[1]), b3_1 = relu_pack(d2[2], d2[3]);
  return mfma(w[(FRAG_W3 + 1) * 64 + li], b3_1, mfma(w[FRAG_W3 * 64 + li], b3_0, f4zero()));
}
__device__ __forceinline__ f4 head_mlp_lds(const h8* w, int li, h8 x) {
  f4 d1[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) d1[nb] = mfma(w[(FRAG_H1 + nb) * 64 + li], x, f4zero());
  const h8 b_0 = relu_pack(d1[0], d1[1]), b_1 = relu_pack(d1[2], d1[3]);
  return mfma(w[(FRAG_H2 + 1) * 64 + li], b_1, mfma(w[FRAG_H2 * 64 + li], b_0, f4zero()));
}
// stage all weight fragments of the model into LDS (every thread of the block)
__device__ __forceinline__ void stage_frags(const h8* __restrict__ frags, h8* lds) {
  for (int i = threadIdx.x; i < N_FRAGS * 64; i += blockDim.x) lds[i] = frags[i];
  __syncthreads();
}
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// SH degree 4 coefficients 4g..4g+3 of dir01*2-1 (my_fully_fused_mlp.cu:206-234),
// rounded to fp16.
__device__ __forceinline__ void sh4_quarter(float dx01, float dy01, float dz01, int g, _Float16* o) {
  const float x = dx01 * 2.0f - 1.0f, y = dy01 * 2.0f - 1.0f, z = dz01 * 2.0f - 1.0f;
  const float xy = x * y, xz = x * z, yz = y * z, x2 = x * x, y2 = y * y, z2 = z * z;
  float c0, c1, c2, c3;
  if (g == 0) {
    c0 = 0.28209479177387814f;
    c1 = -0.48860251190291987f * y;
    c2 = 0.48860251190291987f * z;
    c3 = -0.48860251190291987f * x;
  } else if (g == 1) {
    c0 = 1.0925484305920792f * xy;
    c1 = -1.0925484305920792f * yz;
    c2 = 0.94617469575755997f * z2 - 0.31539156525251999f;
    c3 = -1.0925484305920792f * xz;
  } else if (g == 2) {
    c0 = 0.54627421529603959f * x2 - 0.54627421529603959f * y2;
    c1 = 0.59004358992664352f * y * (-3.0f * x2 + y2);
    c2 = 2.8906114426405538f * xy * z;
    c3 = 0.45704579946446572f * y * (1.0f - 5.0f * z2);
  } else {
    c0 = 0.3731763325901154f * z * (5.0f * z2 - 3.0f);
    c1 = 0.45704579946446572f * x * (1.0f - 5.0f * z2);
    c2 = 1.4453057213202769f * z * (x2 - y2);
    c3 = 0.59004358992664352f * x * (-x2 + 3.0f * y2);
  }
  o[0] = (_Float16)c0; o[1] = (_Float16)c1; o[2] = (_Float16)c2; o[3] = (_Float16)c3;
}

// Head input fragment.  Slots e=0..3 carry base-output rows 4g+e: rows 1..8
// are the latents, rows 0 and 9..15 are replaced by the constant 1.0 padding
// of the reference's head input (nerf_network.h:248-253); slots e=4..7 carry
// SH coefficients 4g..4g+3.
__device__ __forceinline__ h8 head_input(const _Float16* rows4, int g, float dx01, float dy01, float dz01) {
  h8 b;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int r = 4 * g + e;
    b[e] = (r >= 1 && r <= 8) ? rows4[e] : (_Float16)1.0f;
  }
  _Float16 sh[4];
  sh4_quarter(dx01, dy01, dz01, g, sh);
  b[4] = sh[0]; b[5] = sh[1]; b[6] = sh[2]; b[7] = sh[3];
  return b;
}

}  // namespace fvc
