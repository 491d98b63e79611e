// mlp_mfma.h -- MLP tiles on the 16x16x32 MFMA instructions of gfx950, in
// FP16, BF16, two-term BF16 (BF16X3) or FP8-E4M3 operands with FP32
// accumulation.
//
// One wave64 evaluates 16 samples per step.  Operands use the transposed
// formulation D[neuron][sample] = W[neuron][k] * X[k][sample]:
//   A (16x32, weights):  lane l -> neuron row l%16, k-group g = l/16 (8 values)
//   B (32x16, inputs):   lane l -> sample   l%16, k-group g       (8 values)
//   D (16x16, f32):      lane l -> sample   l%16, neurons 4g..4g+3
// so the D fragments of one layer become the B fragment of the next without
// any data movement: B[kb] = {act(D[2kb]), act(D[2kb+1])}.  The k-index
// order inside a lane is a free permutation, applied to the weights once on
// the host (pack_fragments in network.hip).
//
// Storage rounding follows the reference where it stores half: encodings,
// the density/latent outputs and the colour outputs are fp16.  In BF16 mode
// every MLP operand (weights, inputs, hidden activations) is rounded to
// bfloat16 (RNE).  In FP8 mode weights are E4M3 codes of w / s_w with the
// per-layer scale s_w = absmax(W) / 448, activations codes of a / s_a with
// the fixed s_a = 64 / 448 (saturating at +-448); each layer's accumulator is
// rescaled by s_w * s_a.  In BF16X3 mode every operand is the pair (hi, lo)
// of bfloat16s with hi = bf16(v), lo = bf16(v - hi), and a product is
// hi*hi + hi*lo + lo*hi on three MFMAs (about 16 significant bits per
// operand: the fp16 weights and inputs are exact, hidden activations keep
// more bits than the reference's fp16).  The weight fragments are staged in
// LDS per block.
#pragma once
#include <hip/hip_runtime.h>

#include "fvc_internal.h"

namespace fvc {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 f4zero() { return f4{0.0f, 0.0f, 0.0f, 0.0f}; }

static const float FP8_ACT_SCALE = 64.0f / 448.0f;

// operand traits per precision
template <int P> struct Mx;
template <> struct Mx<FVC_PREC_FP16> {
  typedef h8 T;
  static __device__ __forceinline__ T pack(const float* v, float) {
    typedef float f8v __attribute__((ext_vector_type(8)));
    const f8v w = {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
    return __builtin_convertvector(w, T);
  }
  static __device__ __forceinline__ f4 mma(T a, T b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
template <> struct Mx<FVC_PREC_BF16> {
  typedef b8 T;
  static __device__ __forceinline__ T pack(const float* v, float) {
    typedef float f8v __attribute__((ext_vector_type(8)));
    const f8v w = {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
    return __builtin_convertvector(w, T);
  }
  static __device__ __forceinline__ f4 mma(T a, T b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
// two-term bfloat16 operand: v ~ hi + lo
struct B3 {
  b8 hi, lo;
};
template <> struct Mx<FVC_PREC_BF16X3> {
  typedef B3 T;
  static __device__ __forceinline__ T pack(const float* v, float) {
    typedef float f8v __attribute__((ext_vector_type(8)));
    const f8v w = {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
    const b8 hi = __builtin_convertvector(w, b8);
    const f8v r = w - __builtin_convertvector(hi, f8v);
    return T{hi, __builtin_convertvector(r, b8)};
  }
  // the small terms first: hi*lo and lo*hi, then hi*hi
  static __device__ __forceinline__ f4 mma(T a, T b, f4 c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, b.hi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.lo, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, c, 0, 0, 0);
  }
};
template <> struct Mx<FVC_PREC_FP8> {
  typedef long T;
  // E4M3 codes of v * inv_scale, clamped to the finite range (the hardware
  // conversion rounds to nearest even but does not saturate)
  static __device__ __forceinline__ T pack(const float* v, float inv_scale) {
    float q[8];
#pragma unroll
    for (int e = 0; e < 8; e++) q[e] = fminf(fmaxf(v[e] * inv_scale, -448.0f), 448.0f);
    int lo = __builtin_amdgcn_cvt_pk_fp8_f32(q[0], q[1], 0, false);
    lo = __builtin_amdgcn_cvt_pk_fp8_f32(q[2], q[3], lo, true);
    int hi = __builtin_amdgcn_cvt_pk_fp8_f32(q[4], q[5], 0, false);
    hi = __builtin_amdgcn_cvt_pk_fp8_f32(q[6], q[7], hi, true);
    return (long)(((unsigned long)(unsigned)hi << 32) | (unsigned)lo);
  }
  static __device__ __forceinline__ f4 mma(T a, T b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b, c, 0, 0, 0);
  }
};

template <int P> __device__ __forceinline__ f4 rescale(f4 d, float s) {
  if (P == FVC_PREC_FP8) return d * s;
  return d;
}

// ReLU as a signed-integer max on the bit pattern: every value with the
// sign bit set becomes +0, the others pass unchanged (one instruction, no
// NaN canonicalisation).  On packed 16-bit operands the same max of the
// halves (or bfloat16s) commutes with the rounding, since rounding keeps the
// sign and maps 0 to 0.
__device__ __forceinline__ float relu_bits(float x) {
  const int b = __builtin_bit_cast(int, x);
  return __builtin_bit_cast(float, b > 0 ? b : 0);
}
typedef short s8v __attribute__((ext_vector_type(8)));

// B fragment from two 16-neuron D blocks: ReLU, then operand rounding
template <int P> __device__ __forceinline__ typename Mx<P>::T act_pack(f4 lo, f4 hi, float inv_act) {
  if constexpr (P == FVC_PREC_FP8 || P == FVC_PREC_BF16X3) {
    // ReLU on the float values, then the operand format
    const float v[8] = {relu_bits(lo[0]), relu_bits(lo[1]), relu_bits(lo[2]), relu_bits(lo[3]),
                        relu_bits(hi[0]), relu_bits(hi[1]), relu_bits(hi[2]), relu_bits(hi[3])};
    return Mx<P>::pack(v, inv_act);
  } else {
    const float v[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    const s8v b = __builtin_bit_cast(s8v, Mx<P>::pack(v, inv_act));
    return __builtin_bit_cast(typename Mx<P>::T, __builtin_elementwise_max(b, (s8v)(short)0));
  }
}

// density / latent MLP 32 -> 64 -> 64 -> 16 with the fragments in LDS (w);
// li is the lane index made opaque per loop iteration by the caller so the
// reads are not hoisted into registers.  Returns output rows 4g..4g+3.
template <int P>
__device__ __forceinline__ f4 base_mlp(const typename Mx<P>::T* w, int li, typename Mx<P>::T x, const MlpScales& sc) {
  f4 d1[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) d1[nb] = rescale<P>(Mx<P>::mma(w[(FRAG_W1 + nb) * 64 + li], x, f4zero()), sc.layer[0]);
  const typename Mx<P>::T b2_0 = act_pack<P>(d1[0], d1[1], sc.inv_act), b2_1 = act_pack<P>(d1[2], d1[3], sc.inv_act);
  f4 d2[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
    d2[nb] = rescale<P>(Mx<P>::mma(w[(FRAG_W2 + nb * 2 + 1) * 64 + li], b2_1,
                                   Mx<P>::mma(w[(FRAG_W2 + nb * 2) * 64 + li], b2_0, f4zero())),
                        sc.layer[1]);
  const typename Mx<P>::T b3_0 = act_pack<P>(d2[0], d2[1], sc.inv_act), b3_1 = act_pack<P>(d2[2], d2[3], sc.inv_act);
  return rescale<P>(Mx<P>::mma(w[(FRAG_W3 + 1) * 64 + li], b3_1, Mx<P>::mma(w[FRAG_W3 * 64 + li], b3_0, f4zero())),
                    sc.layer[2]);
}

// colour MLP 32 -> 64 -> 16 with the first-layer fragments at w1 (4) and the
// output-layer fragments at w2 (2); returns output rows 4g..4g+3 (rows 0..2 =
// rgb)
template <int P>
__device__ __forceinline__ f4 head_mlp(const typename Mx<P>::T* w1, const typename Mx<P>::T* w2, int li,
                                       typename Mx<P>::T x, const MlpScales& sc) {
  f4 d1[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) d1[nb] = rescale<P>(Mx<P>::mma(w1[nb * 64 + li], x, f4zero()), sc.layer[3]);
  const typename Mx<P>::T b_0 = act_pack<P>(d1[0], d1[1], sc.inv_act), b_1 = act_pack<P>(d1[2], d1[3], sc.inv_act);
  return rescale<P>(Mx<P>::mma(w2[64 + li], b_1, Mx<P>::mma(w2[li], b_0, f4zero())), sc.layer[4]);
}

// stage n weight fragments into LDS (every thread of the block)
template <class T> __device__ __forceinline__ void stage_frags(const T* __restrict__ frags, T* lds, int n) {
  for (int i = threadIdx.x; i < n * 64; i += blockDim.x) lds[i] = frags[i];
}
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}
__device__ __forceinline__ float opaque_f(float v) {
  asm volatile("" : "+v"(v));
  return v;
}
// float -> half, round to nearest even, of the float value as computed.
// Written as the instruction itself: the compiler would otherwise fuse a
// product or fma that produces the float with the conversion
// (v_fma_mix{lo,hi}_f16, one rounding instead of the reference's two, which
// differs where the float is a half midpoint).
__device__ __forceinline__ _Float16 f16_of(float v) {
  uint32_t h;
  asm volatile("v_cvt_f16_f32 %0, %1" : "=v"(h) : "v"(v));
  return __builtin_bit_cast(_Float16, (uint16_t)h);
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
  // rounded to float, then to half (the reference's __float2half of the
  // float coefficient)
  o[0] = f16_of(c0); o[1] = f16_of(c1); o[2] = f16_of(c2); o[3] = f16_of(c3);
}

// all 16 coefficients as 4 dwords of packed halves per 8 coefficients
__device__ __forceinline__ void sh4_all(float dx01, float dy01, float dz01, uint4& lo, uint4& hi) {
  _Float16 c[16];
#pragma unroll
  for (int g = 0; g < 4; g++) sh4_quarter(dx01, dy01, dz01, g, c + 4 * g);
  uint32_t w[8];
#pragma unroll
  for (int k = 0; k < 8; k++)
    w[k] = (uint32_t)__builtin_bit_cast(uint16_t, c[2 * k]) | ((uint32_t)__builtin_bit_cast(uint16_t, c[2 * k + 1]) << 16);
  lo = make_uint4(w[0], w[1], w[2], w[3]);
  hi = make_uint4(w[4], w[5], w[6], w[7]);
}

// Head input fragment.  Slots e=0..3 carry base-output rows 4g+e: rows
// [lo, hi) are the latents, the other rows are replaced by the constant 1.0
// padding of the reference's head input (nerf_network.h:248-253); slots
// e=4..7 carry SH coefficients 4g..4g+3.  Values are the fp16 ones the
// reference stores, then rounded to the MLP's operand precision.
template <int P>
__device__ __forceinline__ typename Mx<P>::T head_input_sh(const _Float16* rows4, int g, const _Float16* sh,
                                                            float inv_act, int lo, int hi) {
  float v[8];
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int r = 4 * g + e;
    v[e] = (r >= lo && r < hi) ? (float)rows4[e] : 1.0f;
  }
#pragma unroll
  for (int e = 0; e < 4; e++) v[4 + e] = (float)sh[e];
  return Mx<P>::pack(v, inv_act);
}
template <int P>
__device__ __forceinline__ typename Mx<P>::T head_input(const _Float16* rows4, int g, float dx01, float dy01,
                                                         float dz01, float inv_act, int lo, int hi) {
  _Float16 sh[4];
  sh4_quarter(dx01, dy01, dz01, g, sh);
  return head_input_sh<P>(rows4, g, sh, inv_act, lo, hi);
}

// SH coefficients 4g..4g+3 from a pixel's table entry (sh4_all's 16 halves,
// read as four uint2)
__device__ __forceinline__ void sh4_unpack(uint2 q, _Float16* sh) {
  sh[0] = __builtin_bit_cast(_Float16, (uint16_t)(q.x & 0xffffu));
  sh[1] = __builtin_bit_cast(_Float16, (uint16_t)(q.x >> 16));
  sh[2] = __builtin_bit_cast(_Float16, (uint16_t)(q.y & 0xffffu));
  sh[3] = __builtin_bit_cast(_Float16, (uint16_t)(q.y >> 16));
}

// ------------------------------------------------------------ split models
// (FP16 operands, FP32 accumulation; fragment indices VD_* in fvc_internal.h)
typedef Mx<FVC_PREC_FP16> MxH;

// mlp_first_head 32 -> 64 (FL hidden layers) -> 16 with the fragments at w
// (the staged fragment table); returns output rows 4g..4g+3, rows 0..7 being
// the colour head's latents 0..7
template <int FL> __device__ __forceinline__ f4 first_head_mlp(const h8* w, int li, h8 x) {
  f4 d1[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) d1[nb] = MxH::mma(w[(VD_F1 + nb) * 64 + li], x, f4zero());
  h8 b0 = act_pack<FVC_PREC_FP16>(d1[0], d1[1], 1.0f), b1 = act_pack<FVC_PREC_FP16>(d1[2], d1[3], 1.0f);
  if constexpr (FL == 2) {
    f4 d2[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++)
      d2[nb] = MxH::mma(w[(VD_F2 + nb * 2 + 1) * 64 + li], b1, MxH::mma(w[(VD_F2 + nb * 2) * 64 + li], b0, f4zero()));
    b0 = act_pack<FVC_PREC_FP16>(d2[0], d2[1], 1.0f);
    b1 = act_pack<FVC_PREC_FP16>(d2[2], d2[3], 1.0f);
  }
  return MxH::mma(w[(VD_F3 + 1) * 64 + li], b1, MxH::mma(w[VD_F3 * 64 + li], b0, f4zero()));
}

// custom colour head (my_fully_fused_mlp.cu:677-895) on the two input
// blocks a, b: first-layer fragments wa, wb (4 each), output fragments wo
// (2); returns rows 4g..4g+3 (rgb in group 0)
__device__ __forceinline__ f4 head_vd_mlp(const h8* wa, const h8* wb, const h8* wo, int li, h8 a, h8 b) {
  f4 d[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) d[nb] = MxH::mma(wa[nb * 64 + li], a, MxH::mma(wb[nb * 64 + li], b, f4zero()));
  const h8 b0 = act_pack<FVC_PREC_FP16>(d[0], d[1], 1.0f), b1 = act_pack<FVC_PREC_FP16>(d[2], d[3], 1.0f);
  return MxH::mma(wo[64 + li], b1, MxH::mma(wo[li], b0, f4zero()));
}

// Frequency encoding quad q (= 2 axis + cos) of the position p
// (freqEncChunk4, my_fully_fused_mlp.cu:63-110): sin or cos of pi 2^f p_axis
// for f = 0..3, as halves.  v_sin / v_cos take revolutions, so the argument
// p 2^(f-1) is exact.  on = false (no frequency inputs): the padding 1.0.
__device__ __forceinline__ void freq_quad(V3 p, int q, bool on, _Float16* o) {
  const float x = q < 2 ? p.x : (q < 4 ? p.y : p.z);
#pragma unroll
  for (int f = 0; f < 4; f++) {
    const float r = x * (0.5f * (float)(1 << f));
    const float v = (q & 1) ? __builtin_amdgcn_cosf(r) : __builtin_amdgcn_sinf(r);
    o[f] = on ? (_Float16)v : (_Float16)1.0f;
  }
}

__device__ __forceinline__ h8 h8_of(const _Float16* lo, const _Float16* hi) {
  return h8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

}  // namespace fvc
