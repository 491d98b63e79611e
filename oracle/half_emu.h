// half_emu.h -- IEEE-754 binary16 and bfloat16 round-to-nearest-even
// conversions written from the format definitions (binary16: 1 sign,
// 5 exponent bits with bias 15, 10 fraction bits; bfloat16: the upper 16
// bits of binary32).  Test infrastructure of the oracle only.
#pragma once
#include <cstdint>
#include <cstring>
#include <cmath>

namespace orc {

inline uint32_t f32_bits(float f) { uint32_t x; std::memcpy(&x, &f, 4); return x; }
inline float bits_f32(uint32_t x) { float f; std::memcpy(&f, &x, 4); return f; }

// binary32 -> binary16, round to nearest, ties to even; overflow -> inf.
inline uint16_t f32_to_f16(float f) {
  const uint32_t x = f32_bits(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t exp = (x >> 23) & 0xffu;
  uint32_t mant = x & 0x7fffffu;
  if (exp == 0xffu) return (uint16_t)(sign | 0x7c00u | (mant ? (0x200u | (mant >> 13)) : 0u));
  const int e = (int)exp - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    mant |= 0x800000u;
    const int shift = 14 - e;
    uint32_t hm = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t halfway = 1u << (shift - 1);
    if (rem > halfway || (rem == halfway && (hm & 1u))) hm++;
    return (uint16_t)(sign | hm);
  }
  uint32_t h = sign | ((uint32_t)e << 10) | (mant >> 13);
  const uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)h;
}

inline float f16_to_f32(uint16_t h) {
  const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  const uint32_t exp = ((uint32_t)h >> 10) & 0x1fu;
  const uint32_t mant = (uint32_t)h & 0x3ffu;
  if (exp == 0) {
    // zero or subnormal: mant * 2^-24 (exact in binary32)
    float v = std::ldexp((float)mant, -24);
    return sign ? -v : v;
  }
  if (exp == 31) return bits_f32(sign | 0x7f800000u | (mant << 13));
  return bits_f32(sign | ((exp - 15 + 127) << 23) | (mant << 13));
}

inline float round_f16(float f) { return f16_to_f32(f32_to_f16(f)); }
inline double round_f16(double d) { return (double)round_f16((float)d); }

// binary32 -> bfloat16 (round to nearest even), NaN kept quiet.
inline uint16_t f32_to_bf16(float f) {
  uint32_t x = f32_bits(f);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  x += 0x7fffu + ((x >> 16) & 1u);
  return (uint16_t)(x >> 16);
}
inline float bf16_to_f32(uint16_t b) { return bits_f32((uint32_t)b << 16); }
inline float round_bf16(float f) { return bf16_to_f32(f32_to_bf16(f)); }

// OCP FP8 E4M3 (bias 7, no infinities, max finite 448, 0x7f/0xff = NaN),
// round to nearest even, saturating to +-448.
inline uint8_t f32_to_e4m3(float f) {
  if (std::isnan(f)) return 0x7f;
  const uint8_t sign = std::signbit(f) ? 0x80 : 0x00;
  float a = std::fabs(f);
  if (a >= 448.0f) return (uint8_t)(sign | 0x7e);
  if (a < std::ldexp(1.0f, -10)) {
    // below half of the smallest subnormal (2^-9): ties-to-even gives 0
    return sign;
  }
  int e; std::frexp(a, &e);          // a = m * 2^e, m in [0.5, 1)
  int ue = e - 1;                    // unbiased exponent of the leading 1
  if (ue < -6) ue = -6;              // subnormal range uses exponent -6
  // quantum of the 3-bit fraction at this exponent
  const float q = std::ldexp(1.0f, ue - 3);
  float n = a / q;                   // exact: power-of-two scaling
  float r = std::nearbyint(n);       // default rounding mode = ties to even
  float v = r * q;
  if (v >= 448.0f) return (uint8_t)(sign | 0x7e);
  // encode v
  if (v == 0.0f) return sign;
  int e2; float m2 = std::frexp(v, &e2);
  int ue2 = e2 - 1;
  if (ue2 < -6) {
    const int frac = (int)std::lround(v / std::ldexp(1.0f, -9));
    return (uint8_t)(sign | frac);
  }
  const int frac = (int)std::lround((m2 * 2.0f - 1.0f) * 8.0f);
  return (uint8_t)(sign | ((ue2 + 7) << 3) | frac);
}
inline float e4m3_to_f32(uint8_t b) {
  const bool neg = b & 0x80;
  const int exp = (b >> 3) & 0xf;
  const int frac = b & 0x7;
  if (exp == 0xf && frac == 0x7) return NAN;
  float v = exp == 0 ? std::ldexp((float)frac, -9) : std::ldexp(1.0f + frac / 8.0f, exp - 7);
  return neg ? -v : v;
}
inline float round_e4m3(float f) { return e4m3_to_f32(f32_to_e4m3(f)); }

}  // namespace orc
