// swr_device.hpp -- device-side building blocks of the draw path (gfx950).
//
// Numerics follow what SwiftShader's Subzero backend emits on x86-64
// (src/Reactor/SubzeroReactor.cpp), because the results must match its
// images: round-half-even conversions that return 0x80000000 when out of
// range, truncating conversions with the same rule, compare-and-select
// Min/Max (a NaN picks the second operand), unfused multiply-add, and
// correctly rounded division.  Compile with -ffp-contract=off and
// -fgpu-flush-denormals-to-zero (the reference's worker threads run with
// FTZ|DAZ, src/System/SwiftConfig.cpp:136-139).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swr_params.hpp"

namespace swr {

// ---------------------------------------------------------------------------
// Numeric primitives
// ---------------------------------------------------------------------------
__device__ __forceinline__ float rr_min(float x, float y) { return x < y ? x : y; }
__device__ __forceinline__ float rr_max(float x, float y) { return x > y ? x : y; }

// cvtps2dq (round to nearest even; out of range or NaN -> INT_MIN).
__device__ __forceinline__ int32_t rr_round_int(float x) {
  if (!(x >= -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
  return (int32_t)__builtin_rintf(x);
}
__device__ __forceinline__ int32_t rr_round_int_clamped(float x) { return rr_round_int(rr_min(x, 2147483520.0f)); }
// cvttps2dq (truncate; out of range or NaN -> INT_MIN).
__device__ __forceinline__ int32_t rr_trunc_int(float x) {
  if (!(x >= -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
  return (int32_t)__builtin_truncf(x);
}
__device__ __forceinline__ float rr_frac(float x) {
  return rr_min(x - __builtin_floorf(x), __uint_as_float(0x3F7FFFFFu));
}
__device__ __forceinline__ uint32_t mulhi16(uint32_t a, uint32_t b) { return (a * b) >> 16; }

// log2(sqrt(x)) by exponent-bit arithmetic (SamplerCore.cpp:1333-1341).
__device__ __forceinline__ float log2sqrt(float lod) {
  lod = lod * lod;
  lod = (float)(int32_t)__float_as_uint(lod) - 1065353216.0f;
  return lod * __uint_as_float(0x33000000u);
}

// ---------------------------------------------------------------------------
// Sampler (SamplerCore.cpp) -- implicit-LOD 2D sample of one 2x2 quad.
// ---------------------------------------------------------------------------
__device__ __constant__ const uint16_t kCWeight[17] = {0xFFFF, 0xFFFF, 0x8000, 0x5555, 0x4000, 0x3333, 0x2AAA,
                                                      0x2492, 0x2000, 0x1C71, 0x1999, 0x1745, 0x1555, 0x13B1,
                                                      0x1249, 0x1111, 0x1000};

// Row of the anisotropic weight tables for probe count N = RoundInt(aniso).
// The reference addresses the tables at 16 * N / 8 * N bytes in wrapping
// 32-bit arithmetic (SamplerCore.cpp:396-398, 1097-1099): N is in [0, 16]
// except for an exactly singular UV Jacobian (det = -0, anisotropy = -inf,
// N = INT_MIN), whose offsets wrap to row 0.  Anything else (impossible
// with maxAnisotropy <= 16) is also sent to row 0, so the table read always
// stays in bounds.
__device__ __forceinline__ int aniso_row(int N) {
  const int r = (int)((uint32_t)N << 4) >> 4;
  return (r < 0 || r > 16) ? 0 : r;
}
__device__ __forceinline__ float uv_weight(int n) { return n <= 1 ? 1.0f : 1.0f / (float)n; }
__device__ __forceinline__ float uv_start(int n) { return n <= 1 ? -0.0f : -(float)(n - 1) / (float)(2 * n); }

// 0.16 fixed-point coordinate (address, :2399-2435).
// SamplerCore::address (SamplerCore.cpp:2398-2435) for the address modes
// 0 wrap, 1 clamp to edge, 2 mirrored repeat, 3 mirror clamp to edge.
// The mirrored modes exist only in the extended (XF) kernel variants.
template <bool XF>
__device__ __forceinline__ uint32_t address16(float uw, int mode) {
  if (mode == 1) {
    float cl = rr_min(rr_max(uw, 0.0f), 65535.0f / 65536.0f);
    return (uint32_t)rr_trunc_int(cl * 65536.0f) & 0xFFFFu;
  }
  if (XF && mode == 2) {  // bit 16 of the fixed-point coordinate flips the low 16 bits
    const int32_t conv = rr_trunc_int(uw * 65536.0f);
    const int32_t mirror = (int32_t)((uint32_t)conv << 15) >> 31;
    return (uint32_t)(conv ^ mirror) & 0xFFFFu;
  }
  if (XF && mode == 3) {  // |u|, saturated to [0, 0xFFFF] through a signed 16-bit pack
    const float a = __uint_as_float(__float_as_uint(uw * 65536.0f) & 0x7FFFFFFFu);
    int32_t c = (int32_t)((uint32_t)rr_trunc_int(a) - 0x8000u);
    c = c < -32768 ? -32768 : (c > 32767 ? 32767 : c);
    return (uint32_t)(c + 0x8000) & 0xFFFFu;
  }
  return (uint32_t)rr_trunc_int(uw * 65536.0f) & 0xFFFFu;
}

__device__ __forceinline__ uint32_t offset16(uint32_t uvw, uint32_t off, int wrap, int sign) {
  if (wrap) return (sign < 0 ? uvw - off : uvw + off) & 0xFFFFu;
  if (sign < 0) return uvw > off ? uvw - off : 0u;
  uint32_t s = uvw + off;
  return s > 0xFFFFu ? 0xFFFFu : s;
}

struct Texel16 { uint32_t c[4]; };

// One mip level's parameters, read once per sampling call: the level
// varies per lane, so each read of the parameter arrays is a memory load.
struct TexLevel {
  const void *data;
  int width, height, pitch;
  uint32_t uHalf, vHalf;
  const uint16_t *srgb;  // sRGB texels: the decode table, else null
};

__device__ __forceinline__ TexLevel tex_level(const TexParams &T, int lvl) {
  TexLevel L;
  L.data = T.data[lvl];
  L.width = T.width[lvl];
  L.height = T.height[lvl];
  L.pitch = T.pitchP[lvl];
  L.uHalf = T.uHalf[lvl];
  L.vHalf = T.vHalf[lvl];
  L.srgb = T.srgb ? T.srgbLut : nullptr;
  return L;
}

template <bool XF>
__device__ __forceinline__ void fetch16(const TexLevel &L, uint32_t u, uint32_t v, uint32_t out[4]) {
  uint32_t uu = mulhi16(u, (uint32_t)L.width);
  uint32_t vv = mulhi16(v, (uint32_t)L.height);
  int32_t index = (int32_t)(int16_t)uu + (int32_t)(int16_t)vv * (int32_t)(int16_t)L.pitch;
  uint32_t p = reinterpret_cast<const uint32_t *>(L.data)[(uint32_t)index];
  out[0] = (p & 0xFFu) << 8;
  out[1] = ((p >> 8) & 0xFFu) << 8;
  out[2] = ((p >> 16) & 0xFFu) << 8;
  out[3] = ((p >> 24) & 0xFFu) << 8;
  if (XF && L.srgb) {  // sRGBtoLinearFF00 (SamplerCore.cpp:1966-1978, 2670-2680): R, G, B through the table
    out[0] = L.srgb[out[0] >> 8];
    out[1] = L.srgb[out[1] >> 8];
    out[2] = L.srgb[out[2] >> 8];
  }
}

// Sampler state as the sampling code reads it.  FA < 0: as the parameters
// say.  FA >= 0: the draw is known to use the common state -- linear or
// anisotropic filtering, linear mips, a computed LOD, and address mode FA
// (0 REPEAT, 1 CLAMP_TO_EDGE) on both axes -- and those reads are constants.
template <int FA>
__device__ __forceinline__ int tex_filter(const TexParams &T) {
  if constexpr (FA >= 0) return T.filter == FILT_ANISOTROPIC ? FILT_ANISOTROPIC : FILT_LINEAR;
  return T.filter;
}
template <int FA> __device__ __forceinline__ int tex_mipmap(const TexParams &T) { return FA >= 0 ? MIP_LINEAR : T.mipmap; }
template <int FA> __device__ __forceinline__ int tex_addr_u(const TexParams &T) { return FA >= 0 ? FA : T.addrU; }
template <int FA> __device__ __forceinline__ int tex_addr_v(const TexParams &T) { return FA >= 0 ? FA : T.addrV; }
template <int FA> __device__ __forceinline__ bool tex_skip_lod(const TexParams &T) { return FA < 0 && T.skipLod; }

template <int FA = -1>
__device__ __forceinline__ int select_mip(const TexParams &T, float lod, int second) {
  if (tex_mipmap<FA>(T) == MIP_NONE) return 0;
  int ilod = (tex_mipmap<FA>(T) == MIP_POINT) ? rr_round_int(lod) : rr_trunc_int(lod);
  return ilod + second;
}

// sampleQuad2D + bilinearInterpolate, RGBA8 16-bit path (:668-909, :516-666).
template <bool XF, int FA = -1>
__device__ __forceinline__ void sample_quad16(const TexParams &T, const TexLevel &L, float u, float v, float lod,
                                              uint32_t out[4]) {
  uint32_t uu = address16<XF>(u, tex_addr_u<FA>(T));
  uint32_t vv = address16<XF>(v, tex_addr_v<FA>(T));  // offsets wrap only in wrap mode (:720-723)
  if (tex_filter<FA>(T) == FILT_POINT) {
    fetch16<XF>(L, uu, vv, out);
    return;
  }
  uint32_t uh = L.uHalf, vh = L.vHalf;
  if (tex_filter<FA>(T) == FILT_MIN_LINEAR_MAG_POINT) {
    if (lod <= 0.0f) { uh = 0; vh = 0; }        // offset &= CmpNLE(lod, 0)
  } else if (tex_filter<FA>(T) == FILT_MIN_POINT_MAG_LINEAR) {
    if (!(lod <= 0.0f)) { uh = 0; vh = 0; }     // offset &= CmpLE(lod, 0)
  }
  uint32_t u0 = offset16(uu, uh, tex_addr_u<FA>(T) == 0, -1), v0 = offset16(vv, vh, tex_addr_v<FA>(T) == 0, -1);
  uint32_t u1 = offset16(uu, uh, tex_addr_u<FA>(T) == 0, +1), v1 = offset16(vv, vh, tex_addr_v<FA>(T) == 0, +1);
  uint32_t c00[4], c10[4], c01[4], c11[4];
  fetch16<XF>(L, u0, v0, c00);
  fetch16<XF>(L, u1, v0, c10);
  fetch16<XF>(L, u0, v1, c01);
  fetch16<XF>(L, u1, v1, c11);
  uint32_t f0u = (u0 * (uint32_t)L.width) & 0xFFFFu;
  uint32_t f0v = (v0 * (uint32_t)L.height) & 0xFFFFu;
  uint32_t f1u = f0u ^ 0xFFFFu, f1v = f0v ^ 0xFFFFu;
  uint32_t f0u0v = mulhi16(f0u, f0v), f1u0v = mulhi16(f1u, f0v);
  uint32_t f0u1v = mulhi16(f0u, f1v), f1u1v = mulhi16(f1u, f1v);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint32_t a = mulhi16(c00[k], f1u1v) + mulhi16(c10[k], f0u1v);
    uint32_t b = mulhi16(c01[k], f1u0v) + mulhi16(c11[k], f0u0v);
    out[k] = ((a & 0xFFFFu) + (b & 0xFFFFu)) & 0xFFFFu;
  }
}

template <bool XF, int FA = -1>
__device__ __forceinline__ void sample_aniso16(const TexParams &T, float u, float v, float lod, float aniso,
                                               float uDelta, float vDelta, int second, uint32_t out[4]) {
  const TexLevel L = tex_level(T, select_mip<FA>(T, lod, second));
  if (tex_filter<FA>(T) != FILT_ANISOTROPIC) {
    sample_quad16<XF, FA>(T, L, u, v, lod, out);
    return;
  }
  int N = rr_round_int(aniso);
  const int row = aniso_row(N);
  float A = uv_weight(row), B = uv_start(row);
  uint32_t cw = kCWeight[row];
  uint32_t sum[4] = {0, 0, 0, 0};
  float du = uDelta, dv = vDelta;
  float u0 = u + B * du;
  float v0 = v + B * dv;
  du = du * A;
  dv = dv * A;
  int i = 0;
  do {
    uint32_t c[4];
    sample_quad16<XF, FA>(T, L, u0, v0, lod, c);
    u0 = u0 + du;
    v0 = v0 + dv;
#pragma unroll
    for (int k = 0; k < 4; k++) sum[k] = (sum[k] + mulhi16(c[k], cw)) & 0xFFFFu;
    i++;
  } while (i < N);
#pragma unroll
  for (int k = 0; k < 4; k++) out[k] = sum[k];
}

// Float path: address (:2463-2638), sampleFloat2D (:1155-1248).
// computeIndices (SamplerCore.cpp:2480-2630) for one axis; mode as address16.
template <bool XF, int FA = -1>
__device__ __forceinline__ void address_float(const TexParams &T, float uvw, int dim, int mode, float lod, int &x0,
                                              int &x1, float &f) {
  int filterMask;
  if (tex_filter<FA>(T) == FILT_POINT) filterMask = 0;
  else if (tex_filter<FA>(T) == FILT_MIN_LINEAR_MAG_POINT) filterMask = !(lod <= 0.0f) ? -1 : 0;
  else if (tex_filter<FA>(T) == FILT_MIN_POINT_MAG_LINEAR) filterMask = (lod <= 0.0f) ? -1 : 0;
  else filterMask = -1;
  const float one = __uint_as_float(0x3F7FFFFFu);  // just under 1
  float coord;
  if (mode == 1) {
    coord = rr_min(rr_max(uvw, 0.0f), one);
  } else if (XF && mode == 2) {
    const float h = uvw * 0.5f;
    coord = rr_min(2.0f * __uint_as_float(__float_as_uint(h - __builtin_rintf(h)) & 0x7FFFFFFFu), one);
  } else if (XF && mode == 3) {
    coord = rr_min(__uint_as_float(__float_as_uint(uvw) & 0x7FFFFFFFu), one);
  } else {
    coord = rr_frac(uvw);
  }
  coord = coord * (float)dim;
  if (tex_filter<FA>(T) == FILT_POINT) {
    x0 = rr_trunc_int(coord);
    f = 0.0f;
  } else {
    if (tex_filter<FA>(T) == FILT_MIN_POINT_MAG_LINEAR || tex_filter<FA>(T) == FILT_MIN_LINEAR_MAG_POINT)
      coord = coord - __uint_as_float(0x3F000000u & (uint32_t)filterMask);
    else
      coord = coord - 0.5f;
    float fl = __builtin_floorf(coord);
    x0 = rr_trunc_int(fl);
    f = coord - fl;
  }
  x1 = x0 - filterMask;
  if (tex_filter<FA>(T) != FILT_POINT) {
    if (mode != 0) {
      x0 = x0 > 0 ? x0 : 0;
      x1 = x1 < dim - 1 ? x1 : dim - 1;
    } else {
      if (x0 < 0) x0 = dim - 1;
      if (!(x1 < dim)) x1 = 0;
    }
  }
}

__device__ __forceinline__ float4 fetchf(const TexLevel &L, int x, int yoff) {
  return reinterpret_cast<const float4 *>(L.data)[(uint32_t)(x + yoff)];
}

template <bool XF, int FA = -1>
__device__ __forceinline__ float4 sample_float2d(const TexParams &T, const TexLevel &L, float u, float v, float lod) {
  int x0, x1, y0, y1;
  float fu, fv;
  address_float<XF, FA>(T, u, L.width, tex_addr_u<FA>(T), lod, x0, x1, fu);
  address_float<XF, FA>(T, v, L.height, tex_addr_v<FA>(T), lod, y0, y1, fv);
  int pitch = L.pitch;
  y0 = (int)((uint32_t)y0 * (uint32_t)pitch);
  if (tex_filter<FA>(T) == FILT_POINT) return fetchf(L, x0, y0);
  y1 = (int)((uint32_t)y1 * (uint32_t)pitch);
  float4 c00 = fetchf(L, x0, y0), c10 = fetchf(L, x1, y0);
  float4 c01 = fetchf(L, x0, y1), c11 = fetchf(L, x1, y1);
  float4 r;
  float a, b;
  a = c00.x + fu * (c10.x - c00.x); b = c01.x + fu * (c11.x - c01.x); r.x = a + fv * (b - a);
  a = c00.y + fu * (c10.y - c00.y); b = c01.y + fu * (c11.y - c01.y); r.y = a + fv * (b - a);
  a = c00.z + fu * (c10.z - c00.z); b = c01.z + fu * (c11.z - c01.z); r.z = a + fv * (b - a);
  a = c00.w + fu * (c10.w - c00.w); b = c01.w + fu * (c11.w - c01.w); r.w = a + fv * (b - a);
  return r;
}

template <bool XF, int FA = -1>
__device__ __forceinline__ float4 sample_float_aniso(const TexParams &T, float u, float v, float lod, float aniso,
                                                     float uDelta, float vDelta, int second) {
  const TexLevel L = tex_level(T, select_mip<FA>(T, lod, second));
  if (tex_filter<FA>(T) != FILT_ANISOTROPIC) return sample_float2d<XF, FA>(T, L, u, v, lod);
  int N = rr_round_int(aniso);
  const int row = aniso_row(N);
  float A = uv_weight(row), B = uv_start(row);
  float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float du = uDelta, dv = vDelta;
  float u0 = u + B * du;
  float v0 = v + B * dv;
  du = du * A;
  dv = dv * A;
  int i = 0;
  do {
    float4 c = sample_float2d<XF, FA>(T, L, u0, v0, lod);
    u0 = u0 + du;
    v0 = v0 + dv;
    sum.x = sum.x + c.x * A;
    sum.y = sum.y + c.y * A;
    sum.z = sum.z + c.z * A;
    sum.w = sum.w + c.w * A;
    i++;
  } while (i < N);
  return sum;
}

// LOD of a quad from its lane differences (computeLod2D, :1376-1422),
// followed by bias and clamping (sampleTexture128, :93-157).
struct QuadLod {
  float lod, aniso, uDelta, vDelta;
};

template <int FA = -1>
__device__ __forceinline__ QuadLod quad_lod(const TexParams &T, const float u[4], const float v[4]) {
  QuadLod L;
  L.aniso = 0.0f;
  L.uDelta = 0.0f;
  L.vDelta = 0.0f;
  if (tex_skip_lod<FA>(T)) {
    L.lod = T.minLod;
    return L;
  }
  float dudx = u[1] - u[0], dudy = u[2] - u[0], dvdx = v[1] - v[0], dvdy = v[2] - v[0];
  float dUdx = dudx * T.wwhh[0], dUdy = dudy * T.wwhh[1], dVdx = dvdx * T.wwhh[2], dVdy = dvdy * T.wwhh[3];
  float sx = dUdx * dUdx + dVdx * dVdx;
  float sy = dUdy * dUdy + dVdy * dVdy;
  float lod = rr_max(sx, sy);
  if (tex_filter<FA>(T) == FILT_ANISOTROPIC) {
    float det = dUdx * dVdy - dUdy * dVdx;
    det = det > 0.0f ? det : -det;
    bool major = !(sx < sy);
    L.uDelta = major ? dudx : dudy;
    L.vDelta = major ? dvdx : dvdy;
    float an = lod * (1.0f / det);
    an = rr_min(an, T.maxAnisotropy);
    lod = lod * (1.0f / (an * an));
    L.aniso = an;
  }
  lod = log2sqrt(lod);
  lod = lod + T.mipLodBias;
  lod = rr_max(lod, T.minLod);
  lod = rr_min(lod, T.maxLod);
  L.lod = lod;
  return L;
}

// One lane of sampleTexture128 after the LOD is known (FTEX: RGBA32F
// texture, 32-bit float path; otherwise RGBA8, 16-bit fixed-point path).
template <bool FTEX, bool XF, int FA = -1>
__device__ __forceinline__ float4 sample_lane(const TexParams &T, float u, float v, const QuadLod &L) {
  if (FTEX) {
    float4 c = sample_float_aniso<XF, FA>(T, u, v, L.lod, L.aniso, L.uDelta, L.vDelta, 0);
    if (tex_mipmap<FA>(T) == MIP_LINEAR) {
      float4 cc = sample_float_aniso<XF, FA>(T, u, v, L.lod, L.aniso, L.uDelta, L.vDelta, 1);
      float l4 = rr_frac(L.lod);
      c.x = (cc.x - c.x) * l4 + c.x;
      c.y = (cc.y - c.y) * l4 + c.y;
      c.z = (cc.z - c.z) * l4 + c.z;
      c.w = (cc.w - c.w) * l4 + c.w;
    }
    return c;
  }
  uint32_t c[4];
  sample_aniso16<XF, FA>(T, u, v, L.lod, L.aniso, L.uDelta, L.vDelta, 0, c);
  if (tex_mipmap<FA>(T) == MIP_LINEAR) {
    uint32_t cc[4];
    sample_aniso16<XF, FA>(T, u, v, L.lod, L.aniso, L.uDelta, L.vDelta, 1, cc);
    uint32_t utri = (uint32_t)rr_trunc_int(L.lod * 65536.0f) & 0xFFFFu;
    uint32_t inv = utri ^ 0xFFFFu;
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = (mulhi16(c[k], inv) + mulhi16(cc[k], utri)) & 0xFFFFu;
  }
  const float scale = 1.0f / 65280.0f;
  // B8G8R8A8 texels filter channel by channel like R8G8B8A8; only the
  // order of R and B differs (SamplerCore's per-format swizzle).
  const bool bgra = XF && T.bgra;
  const uint32_t r = bgra ? c[2] : c[0], b = bgra ? c[0] : c[2];
  return make_float4((float)r * scale, (float)c[1] * scale, (float)b * scale, (float)c[3] * scale);
}

// ---------------------------------------------------------------------------
// Output merger pieces (PixelRoutine.cpp)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float blend_factor_rgb(const DrawParams &P, int f, float s, float sa, float d, float da,
                                                  int comp) {
  switch (f) {
    case 0: return 0.0f;
    case 1: return 1.0f;
    case 2: return s;
    case 3: return 1.0f - s;
    case 4: return d;
    case 5: return 1.0f - d;
    case 6: return sa;
    case 7: return 1.0f - sa;
    case 8: return da;
    case 9: return 1.0f - da;
    case 14: return rr_min(1.0f - da, sa);
    case 10: return P.blendConst[comp];
    case 12: return P.blendConst[3];
    case 11: return P.invBlendConst[comp];
    default: return P.invBlendConst[3];
  }
}

__device__ __forceinline__ float blend_factor_alpha(const DrawParams &P, int f, float sa, float da) {
  switch (f) {
    case 0: return 0.0f;
    case 1: return 1.0f;
    case 2:
    case 6: return sa;
    case 3:
    case 7: return 1.0f - sa;
    case 4:
    case 8: return da;
    case 5:
    case 9: return 1.0f - da;
    case 14: return 1.0f;
    case 10:
    case 12: return P.blendConst[3];
    default: return P.invBlendConst[3];
  }
}

__device__ __forceinline__ float blend_apply(int op, float s, float sf, float d, float df) {
  switch (op) {
    case 0: return s * sf + d * df;
    case 1: return s * sf - d * df;
    case 2: return d * df - s * sf;
    case 3: return rr_min(s, d);
    case 4: return rr_max(s, d);
    case BLEND_OP_SRC: return s;
    case BLEND_OP_DST: return d;
    default: return 0.0f;
  }
}

// VK_EXT_blend_operation_advanced modes (computeAdvancedBlendMode,
// PixelRoutine.cpp:1395-1615, alpha :1937-1953): factors are ONE; both
// colours divided by their alpha (0 where alpha is 0); per-channel modes or
// the HSL helpers; then the uncorrelated-overlap weighting.  Abs clears the
// sign bit; Min / Max are ordered selects.
__device__ __forceinline__ bool is_advanced_op(int op) {
  return (op >= 1000148012 && op <= 1000148022) || (op >= 1000148031 && op <= 1000148034);
}
__device__ __forceinline__ float adv_lum(float r, float g, float b) { return r * 0.3f + g * 0.59f + b * 0.11f; }
__device__ __forceinline__ float3 adv_set_lum(float3 cb, float3 cl) {
  const float ldiff = adv_lum(cl.x, cl.y, cl.z) - adv_lum(cb.x, cb.y, cb.z);
  const float3 c = make_float3(cb.x + ldiff, cb.y + ldiff, cb.z + ldiff);
  const float lum = adv_lum(c.x, c.y, c.z);
  const float mn = rr_min(rr_min(c.x, c.y), c.z), mx = rr_max(rr_max(c.x, c.y), c.z);
  if (mn < 0.0f) {
    return make_float3(lum + ((c.x - lum) * lum) / (lum - mn), lum + ((c.y - lum) * lum) / (lum - mn),
                       lum + ((c.z - lum) * lum) / (lum - mn));
  }
  if (mx > 1.0f) {
    return make_float3(lum + ((c.x - lum) * (1.0f - lum)) / (mx - lum), lum + ((c.y - lum) * (1.0f - lum)) / (mx - lum),
                       lum + ((c.z - lum) * (1.0f - lum)) / (mx - lum));
  }
  return c;
}
__device__ __forceinline__ float3 adv_set_lum_sat(float3 cb, float3 cs, float3 cl) {
  const float mnb = rr_min(rr_min(cb.x, cb.y), cb.z);
  const float sb = rr_max(rr_max(cb.x, cb.y), cb.z) - mnb;
  const float ss = rr_max(rr_max(cs.x, cs.y), cs.z) - rr_min(rr_min(cs.x, cs.y), cs.z);
  const bool nz = sb > 0.0f;
  const float3 c = make_float3(nz ? (cb.x - mnb) * ss / sb : 0.0f, nz ? (cb.y - mnb) * ss / sb : 0.0f,
                               nz ? (cb.z - mnb) * ss / sb : 0.0f);
  return adv_set_lum(c, cl);
}
__device__ __forceinline__ float adv_channel(int op, float x, float y) {
  switch (op) {
    case 1000148012: return x * y;                                   // MULTIPLY
    case 1000148013: return x + y - (x * y);                         // SCREEN
    case 1000148014:                                                 // OVERLAY
      return y > 0.5f ? 1.0f - (2.0f * (1.0f - x) * (1.0f - y)) : 2.0f * x * y;
    case 1000148015: return rr_min(x, y);                            // DARKEN
    case 1000148016: return rr_max(x, y);                            // LIGHTEN
    case 1000148017:                                                 // COLORDODGE
      return !(y > 0.0f) ? 0.0f : (x < 1.0f ? rr_min(1.0f, y / (1.0f - x)) : 1.0f);
    case 1000148018:                                                 // COLORBURN
      return !(y < 1.0f) ? 1.0f : (x > 0.0f ? 1.0f - rr_min(1.0f, (1.0f - y) / x) : 0.0f);
    case 1000148019:                                                 // HARDLIGHT
      return x > 0.5f ? 1.0f - (2.0f * (1.0f - x) * (1.0f - y)) : 2.0f * x * y;
    case 1000148020:                                                 // SOFTLIGHT
      if (!(x > 0.5f)) return y - ((1.0f - (2.0f * x)) * y * (1.0f - y));
      if (!(y > 0.25f)) return y + (((2.0f * x) - 1.0f) * y * ((((16.0f * y) - 12.0f) * y) + 3.0f));
      return y + (((2.0f * x) - 1.0f) * (__builtin_sqrtf(y) - y));
    case 1000148021: return __uint_as_float(__float_as_uint(x - y) & 0x7FFFFFFFu);  // DIFFERENCE
    default: return x + y - (x * y * 2.0f);                          // EXCLUSION
  }
}
__device__ __noinline__ float4 advanced_blend(int op, float4 src, float4 dst) {
  float4 s = src, d = dst;
  s.x = s.w != 0.0f ? s.x / s.w : 0.0f;
  s.y = s.w != 0.0f ? s.y / s.w : 0.0f;
  s.z = s.w != 0.0f ? s.z / s.w : 0.0f;
  d.x = d.w != 0.0f ? d.x / d.w : 0.0f;
  d.y = d.w != 0.0f ? d.y / d.w : 0.0f;
  d.z = d.w != 0.0f ? d.z / d.w : 0.0f;
  const float3 sc = make_float3(s.x, s.y, s.z), dc = make_float3(d.x, d.y, d.z);
  float3 b;
  if (op == 1000148031) b = adv_set_lum_sat(sc, dc, dc);       // HSL_HUE
  else if (op == 1000148032) b = adv_set_lum_sat(dc, sc, dc);  // HSL_SATURATION
  else if (op == 1000148033) b = adv_set_lum(sc, dc);          // HSL_COLOR
  else if (op == 1000148034) b = adv_set_lum(dc, sc);          // HSL_LUMINOSITY
  else b = make_float3(adv_channel(op, s.x, d.x), adv_channel(op, s.y, d.y), adv_channel(op, s.z, d.z));
  float p = s.w * d.w;
  b.x = b.x * p; b.y = b.y * p; b.z = b.z * p;
  p = s.w * (1.0f - d.w);
  b.x = b.x + s.x * p; b.y = b.y + s.y * p; b.z = b.z + s.z * p;
  p = d.w * (1.0f - s.w);
  b.x = b.x + d.x * p; b.y = b.y + d.y * p; b.z = b.z + d.z * p;
  return make_float4(b.x, b.y, b.z, src.w + dst.w - (src.w * dst.w));
}

// SRC_ALPHA, ONE_MINUS_SRC_ALPHA, ADD on colour and alpha: the operations
// alpha_blend's factor and operation switches perform for that state.
__device__ __forceinline__ float4 blend_src_alpha(float4 s, float4 d) {
  const float sa = s.w, k = 1.0f - sa;
  return make_float4(s.x * sa + d.x * k, s.y * sa + d.y * k, s.z * sa + d.z * k, s.w * sa + d.w * k);
}

template <bool XF>
__device__ __forceinline__ float4 alpha_blend(const DrawParams &P, float4 s, float4 d) {
  if (XF && is_advanced_op(P.colorOp)) return advanced_blend(P.colorOp, s, d);
  if (P.blendKind == 1) return blend_src_alpha(s, d);
  float sfr = blend_factor_rgb(P, P.srcF, s.x, s.w, d.x, d.w, 0);
  float sfg = blend_factor_rgb(P, P.srcF, s.y, s.w, d.y, d.w, 1);
  float sfb = blend_factor_rgb(P, P.srcF, s.z, s.w, d.z, d.w, 2);
  float dfr = blend_factor_rgb(P, P.dstF, s.x, s.w, d.x, d.w, 0);
  float dfg = blend_factor_rgb(P, P.dstF, s.y, s.w, d.y, d.w, 1);
  float dfb = blend_factor_rgb(P, P.dstF, s.z, s.w, d.z, d.w, 2);
  float sfa = blend_factor_alpha(P, P.srcFA, s.w, d.w);
  float dfa = blend_factor_alpha(P, P.dstFA, s.w, d.w);
  float4 o;
  o.x = blend_apply(P.colorOp, s.x, sfr, d.x, dfr);
  o.y = blend_apply(P.colorOp, s.y, sfg, d.y, dfg);
  o.z = blend_apply(P.colorOp, s.z, sfb, d.z, dfb);
  o.w = blend_apply(P.alphaOp, s.w, sfa, d.w, dfa);
  return o;
}

__device__ __forceinline__ float unorm8_to_float(uint32_t b) { return (float)(b * 257u) * (1.0f / 65535.0f); }

__device__ __forceinline__ uint32_t float_to_unorm8(float c) {
  float v = rr_min(rr_max(c, 0.0f), 1.0f);
  int32_t i = rr_round_int(v * 255.0f);
  return (uint32_t)(i < 0 ? 0 : (i > 255 ? 255 : i));
}

__device__ __forceinline__ uint32_t stencil_apply_op(uint32_t v, int op, uint32_t ref) {
  switch (op) {
    case 0: return v;
    case 1: return 0;
    case 2: return ref & 0xFFu;
    case 3: return v == 255u ? 255u : v + 1u;
    case 4: return v == 0u ? 0u : v - 1u;
    case 5: return (~v) & 0xFFu;
    case 6: return (v + 1u) & 0xFFu;
    default: return (v - 1u) & 0xFFu;
  }
}

// (ref & mask) OP (value & mask) through biased signed-byte compares (:405-449).
__device__ __forceinline__ bool stencil_compare(const StencilParams &S, uint32_t value) {
  uint32_t val = S.useCompareMask ? (value & S.compareMask) : value;
  int32_t vs = (int32_t)(int8_t)(uint8_t)(val + 0x80u);
  int32_t rs = S.refSigned;
  switch (S.compareOp) {
    case 7: return true;
    case 0: return false;
    case 1: return vs > rs;
    case 2: return val == S.refMasked;
    case 5: return val != S.refMasked;
    case 3: return vs > rs || val == S.refMasked;
    case 4: return rs > vs;
    default: return !(vs > rs);
  }
}

}  // namespace swr
