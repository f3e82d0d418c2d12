// sRGB transfer functions of the reference's legacy-precision build, for the
// colour write and blend of sRGB attachments (device) and their clear value
// (host): linearToSRGB and sRGBtoLinear (ShaderCore.cpp:673-689) over
// Pow<Mediump> = Exp2(Log2(x) * y) (:472-477), Exp2's clamp (:335-346) and
// the legacy Exp2 / Log2 approximations (:300-314, :388-402).  Products and
// sums stay unfused (-ffp-contract=off); RoundInt rounds to nearest even;
// Min / Max are ordered selects.
#pragma once

#include <cstdint>

namespace swr {

__host__ __device__ inline float srgb_bits(uint32_t u) { return __builtin_bit_cast(float, u); }

__host__ __device__ inline float srgb_log2(float x) {
  const uint32_t ix = __builtin_bit_cast(uint32_t, x);
  float x1 = srgb_bits(((ix & 0x7F800000u) >> 8) | 0x3F800000u);  // exponent, as 1.e / 256
  x1 = (x1 - 1.4960938f) * 256.0f;
  const float x0 = srgb_bits((ix & 0x007FFFFFu) | 0x3F800000u);  // mantissa in [1, 2)
  const float x2 = (9.5428179e-2f * x0 + 4.7779095e-1f) * x0 + 1.9782813e-1f;
  const float x3 = ((1.6618466e-2f * x0 + 2.0350508e-1f) * x0 + 2.7382900e-1f) * x0 + 4.0496687e-2f;
  x1 += (x0 - 1.0f) * (x2 / x3);
  return ix == 0x7F800000u ? x : x1;
}

__host__ __device__ inline float srgb_exp2(float x) {
  float x0 = x < 128.0f ? x : 128.0f;
  const float lo = srgb_bits(0xC2FDFFFFu);  // -126.999992
  x0 = x0 > lo ? x0 : lo;
  const int32_t i = (int32_t)__builtin_rintf(x0 - 0.5f);  // in [-127, 128]
  const float ii = srgb_bits((uint32_t)(i + 127) << 23);
  const float f = x0 - (float)i;
  float ff = srgb_bits(0x3AF61905u);
  ff = ff * f + srgb_bits(0x3C134806u);
  ff = ff * f + srgb_bits(0x3D64AA23u);
  ff = ff * f + srgb_bits(0x3E75EAD4u);
  ff = ff * f + srgb_bits(0x3F31727Bu);
  ff = ff * f + 1.0f;
  return ii * ff;
}

__host__ __device__ inline float srgb_pow(float x, float y) { return srgb_exp2(srgb_log2(x) * y); }

__host__ __device__ inline float linear_to_srgb(float c) {
  const float lc = c * 12.92f;
  const float ec = 1.055f * srgb_pow(c, 1.0f / 2.4f) + -0.055f;
  return c < 0.0031308f ? lc : ec;
}

__host__ __device__ inline float srgb_to_linear(float c) {
  const float lc = c * (1.0f / 12.92f);
  const float ec = srgb_pow(c * (1.0f / 1.055f) + 0.055f / 1.055f, 2.4f);
  return c < 0.04045f ? lc : ec;
}

}  // namespace swr
