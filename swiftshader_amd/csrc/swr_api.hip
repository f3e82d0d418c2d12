// swr_api.hip -- host side of the C ABI declared in include/swr.h.
//
// Derives the per-draw GPU state from a swr_draw_desc the way the
// reference derives DrawData and its processor states
// (src/Device/Renderer.cpp:183-490, src/Device/PixelProcessor.cpp:74-187,
// src/Device/Context.cpp:1090-1311, src/Pipeline/SpirvShaderSampling.cpp:
// 33-135,275-431, src/Vulkan/VkSampler.hpp:33-36,
// src/Vulkan/VkDescriptorSetLayout.cpp:298-334,468-503), owns the scratch
// memory of the tiled pipeline and enqueues the kernels.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/swr.h"
#include "swr_kernels.hip"
#include "swr_grad.hip"

using namespace swr;

namespace {

thread_local std::string g_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

#define SWR_HIP_CHECK(expr)                                                                 \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(SWR_ERROR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// Device buffer that only grows.
struct Scratch {
  void *ptr = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t need) {
    if (need <= bytes) return hipSuccess;
    size_t nb = std::max(need, bytes + bytes / 2);
    nb = (nb + 255) & ~size_t(255);
    if (ptr) {
      hipError_t e = hipFree(ptr);
      if (e != hipSuccess) return e;
      ptr = nullptr;
      bytes = 0;
    }
    hipError_t e = hipMalloc(&ptr, nb);
    if (e == hipSuccess) bytes = nb;
    return e;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <typename T> T *as() const { return reinterpret_cast<T *>(ptr); }
};

// Events of one profiled draw: setup | bin | raster boundaries.
enum { EV_START, EV_SETUP, EV_BIN, EV_RASTER, EV_COUNT };
struct EventSet {
  hipEvent_t ev[EV_COUNT];
  uint64_t pairs;
};

struct DeviceCtx {
  int device = -1;
  Scratch rec, poly, cnt, off, keysA, keysB, valsA, valsB, scanTmp, sortTmp, tStart, tEnd, stats;
  uint32_t *hostTotal = nullptr;  // pinned
  swr_tuning tuning{};
  bool profiling = false;
  std::vector<EventSet> pending, pool;
  swr_profile acc{};
};

hipError_t take_events(DeviceCtx *c, EventSet &es) {
  if (!c->pool.empty()) {
    es = c->pool.back();
    c->pool.pop_back();
    return hipSuccess;
  }
  for (int i = 0; i < EV_COUNT; i++) {
    hipError_t e = hipEventCreate(&es.ev[i]);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

std::mutex g_mutex;
DeviceCtx *g_ctx[64] = {};

int current_ctx(DeviceCtx **out) {
  int dev = 0;
  SWR_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(SWR_ERROR_DEVICE, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_ctx[dev]) {
    DeviceCtx *c = new DeviceCtx();
    c->device = dev;
    c->tuning.tile_width = 16;
    c->tuning.tile_height = 16;
    hipError_t e = hipHostMalloc((void **)&c->hostTotal, sizeof(uint32_t) * 4, hipHostMallocDefault);
    if (e != hipSuccess) {
      delete c;
      return fail(SWR_ERROR_OUT_OF_MEMORY, "hipHostMalloc: %s", hipGetErrorString(e));
    }
    g_ctx[dev] = c;
  }
  *out = g_ctx[dev];
  return SWR_OK;
}

float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

int blend_operation(int op, int sf, int df, bool isFloat) {
  switch (op) {
    case SWR_BLEND_OP_ADD:
      if (sf == SWR_BLEND_FACTOR_ZERO) {
        if (df == SWR_BLEND_FACTOR_ZERO) return BLEND_OP_ZERO;
        if (df == SWR_BLEND_FACTOR_ONE) return BLEND_OP_DST;
      } else if (sf == SWR_BLEND_FACTOR_ONE && df == SWR_BLEND_FACTOR_ZERO) {
        return BLEND_OP_SRC;
      }
      break;
    case SWR_BLEND_OP_SUBTRACT:
      if (sf == SWR_BLEND_FACTOR_ZERO) {
        if (df == SWR_BLEND_FACTOR_ZERO || !isFloat) return BLEND_OP_ZERO;  // negative, clamped to zero
      } else if (sf == SWR_BLEND_FACTOR_ONE && df == SWR_BLEND_FACTOR_ZERO) {
        return BLEND_OP_SRC;
      }
      break;
    case SWR_BLEND_OP_REVERSE_SUBTRACT:
      if (sf == SWR_BLEND_FACTOR_ZERO) {
        if (df == SWR_BLEND_FACTOR_ZERO) return BLEND_OP_ZERO;
        if (df == SWR_BLEND_FACTOR_ONE) return BLEND_OP_DST;
      } else if (df == SWR_BLEND_FACTOR_ZERO && !isFloat) {
        return BLEND_OP_ZERO;
      }
      break;
    default:
      break;
  }
  return op;
}

int stencil_params(const swr_stencil_op_state &s, StencilParams &o) {
  if (s.compare_op > 7 || s.fail_op > 7 || s.pass_op > 7 || s.depth_fail_op > 7)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "invalid stencil op state");
  o.compareOp = (int)s.compare_op;
  o.failOp = (int)s.fail_op;
  o.passOp = (int)s.pass_op;
  o.depthFailOp = (int)s.depth_fail_op;
  o.useCompareMask = s.compare_mask != 0xFFu;
  o.useWriteMask = (s.write_mask & 0xFFu) != 0xFFu;
  o.compareMask = s.compare_mask & 0xFFu;
  o.writeMask = s.write_mask & 0xFFu;
  o.reference = s.reference & 0xFFu;
  o.refMasked = s.reference & s.compare_mask & 0xFFu;
  o.refSigned = (int32_t)(int8_t)(uint8_t)(((s.reference & s.compare_mask) + 0x80u) & 0xFFu);
  return SWR_OK;
}

bool valid_compare(uint32_t op) { return op <= 7; }
bool valid_factor(uint32_t f) { return f <= 14; }
bool valid_blend_op(uint32_t op) { return op <= 4; }

int build_params(const swr_draw_desc *d, const swr_tuning &tuning, DrawParams &P) {
  memset(&P, 0, sizeof(P));
  if (!d) return fail(SWR_ERROR_INVALID_ARGUMENT, "descriptor is NULL");
  if (!d->positions) return fail(SWR_ERROR_INVALID_ARGUMENT, "positions is NULL");
  if (d->samples != 1 && d->samples != 4) return fail(SWR_ERROR_UNSUPPORTED, "samples must be 1 or 4 (got %u)", d->samples);
  if (d->framebuffer_width == 0 || d->framebuffer_height == 0 || d->framebuffer_width > 8192 ||
      d->framebuffer_height > 8192)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "framebuffer %ux%u outside 1..8192", d->framebuffer_width,
                d->framebuffer_height);
  const swr_rect &ra = d->render_area;
  if (ra.x < 0 || ra.y < 0 || (uint64_t)ra.x + ra.width > d->framebuffer_width ||
      (uint64_t)ra.y + ra.height > d->framebuffer_height)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "render area outside the framebuffer");
  if (d->triangle_count > (1u << 26)) return fail(SWR_ERROR_UNSUPPORTED, "too many triangles (%u)", d->triangle_count);
  if (!d->indices && (uint64_t)d->triangle_count * 3 > d->vertex_count)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "non-indexed draw needs 3 vertices per triangle");
  if (d->shader > SWR_SHADER_MODULATE) return fail(SWR_ERROR_INVALID_ARGUMENT, "unknown shader %u", d->shader);
  if (d->shader != SWR_SHADER_SAMPLE && !d->colors)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "shader reads colours but colors is NULL");
  if (d->shader != SWR_SHADER_INTERPOLATE && !d->texcoords)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "shader samples but texcoords is NULL");

  P.pos = d->positions;
  P.col = d->colors;
  P.uv = d->texcoords;
  P.idx = d->indices;
  P.triCount = (int)d->triangle_count;
  if (d->vertex_count == 0 && d->triangle_count) return fail(SWR_ERROR_INVALID_ARGUMENT, "vertex_count is 0");
  P.vertexCount = (int)std::min<uint32_t>(d->vertex_count, INT32_MAX);

  // Viewport (Renderer.cpp:300-331) with 4 sub-pixel bits.
  const swr_viewport &vp = d->viewport;
  float W = 0.5f * vp.width, H = 0.5f * vp.height;
  float X0 = vp.x + W, Y0 = vp.y + H;
  P.WxF = W * 16.0f;
  P.HxF = H * 16.0f;
  P.X0xF = X0 * 16.0f - 16.0f / 2;
  P.Y0xF = Y0 * 16.0f - 16.0f / 2;
  P.depthRange = vp.max_depth - vp.min_depth;
  P.depthNear = vp.min_depth;
  // Scissor clamped to the render area (Renderer.cpp:333-345).
  int x0 = ra.x, y0 = ra.y, x1 = ra.x + (int)ra.width, y1 = ra.y + (int)ra.height;
  auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
  P.sx0 = clampi(d->scissor.x, x0, x1);
  P.sx1 = clampi((int)std::min<int64_t>((int64_t)d->scissor.x + d->scissor.width, INT32_MAX), x0, x1);
  P.sy0 = clampi(d->scissor.y, y0, y1);
  P.sy1 = clampi((int)std::min<int64_t>((int64_t)d->scissor.y + d->scissor.height, INT32_MAX), y0, y1);
  P.width = (int)d->framebuffer_width;
  P.height = (int)d->framebuffer_height;

  if (d->cull_mode > 3 || d->front_face > 1) return fail(SWR_ERROR_INVALID_ARGUMENT, "invalid cull/front face");
  P.cullMode = (int)d->cull_mode;
  P.frontFace = (int)d->front_face;
  P.ms = (int)d->samples;
  P.msMask = (int)(d->sample_mask & (0xFFFFFFFFu >> (32 - d->samples)));

  // Depth / stencil state (PixelProcessor.cpp:99-137, Context.cpp:930-944).
  if (d->depth.data && d->depth.format != SWR_FORMAT_D32_SFLOAT)
    return fail(SWR_ERROR_UNSUPPORTED, "depth format %u (only D32_SFLOAT)", d->depth.format);
  if (d->stencil.data && d->stencil.format != SWR_FORMAT_S8_UINT)
    return fail(SWR_ERROR_UNSUPPORTED, "stencil format %u (only S8_UINT)", d->stencil.format);
  if (!valid_compare(d->depth_compare_op)) return fail(SWR_ERROR_INVALID_ARGUMENT, "invalid depth compare op");
  P.depthTest = (d->depth.data && d->depth_test_enable) ? 1 : 0;
  P.depthWrite = (P.depthTest && d->depth_write_enable) ? 1 : 0;
  P.depthCompare = (int)d->depth_compare_op;
  P.stencilActive = (d->stencil.data && d->stencil_test_enable) ? 1 : 0;
  int r;
  if ((r = stencil_params(d->front, P.front)) != SWR_OK) return r;
  if ((r = stencil_params(d->back, P.back)) != SWR_OK) return r;
  if (P.stencilActive) {
    bool allKeep = d->front.pass_op == 0 && d->front.depth_fail_op == 0 && d->front.fail_op == 0 &&
                   d->back.pass_op == 0 && d->back.depth_fail_op == 0 && d->back.fail_op == 0;
    bool anyWrite = d->front.write_mask != 0 || d->back.write_mask != 0;
    P.stencilWrite = (!allKeep && anyWrite) ? 1 : 0;
  }

  // Colour output and blend state (Context.cpp:1090-1311).
  bool hasColor = d->color.data && d->color.format != SWR_FORMAT_UNDEFINED;
  if (hasColor && d->color.format != SWR_FORMAT_R8G8B8A8_UNORM && d->color.format != SWR_FORMAT_R32G32B32A32_SFLOAT)
    return fail(SWR_ERROR_UNSUPPORTED, "colour format %u", d->color.format);
  P.colorFloat = d->color.format == SWR_FORMAT_R32G32B32A32_SFLOAT;
  const swr_blend_state &b = d->blend;
  if (!valid_factor(b.src_color_blend_factor) || !valid_factor(b.dst_color_blend_factor) ||
      !valid_factor(b.src_alpha_blend_factor) || !valid_factor(b.dst_alpha_blend_factor) ||
      !valid_blend_op(b.color_blend_op) || !valid_blend_op(b.alpha_blend_op))
    return fail(SWR_ERROR_UNSUPPORTED, "blend state outside the supported VkBlendFactor/VkBlendOp range");
  int cop = blend_operation((int)b.color_blend_op, (int)b.src_color_blend_factor, (int)b.dst_color_blend_factor, P.colorFloat);
  int aop = blend_operation((int)b.alpha_blend_op, (int)b.src_alpha_blend_factor, (int)b.dst_alpha_blend_factor, P.colorFloat);
  if (hasColor) P.colorMask = (cop == BLEND_OP_DST && aop == BLEND_OP_DST) ? 0 : (int)(b.color_write_mask & 0xFu);
  P.blendEnable = (hasColor && b.blend_enable && P.colorMask && (cop != BLEND_OP_SRC || aop != BLEND_OP_SRC)) ? 1 : 0;
  auto bf = [](uint32_t op, uint32_t f) {
    return (op == SWR_BLEND_OP_MIN || op == SWR_BLEND_OP_MAX) ? (int)SWR_BLEND_FACTOR_ONE : (int)f;
  };
  if (P.blendEnable) {
    P.srcF = bf(b.color_blend_op, b.src_color_blend_factor);
    P.dstF = bf(b.color_blend_op, b.dst_color_blend_factor);
    P.colorOp = cop;
    P.srcFA = bf(b.alpha_blend_op, b.src_alpha_blend_factor);
    P.dstFA = bf(b.alpha_blend_op, b.dst_alpha_blend_factor);
    P.alphaOp = aop;
  }
  for (int k = 0; k < 4; k++) {
    float c = d->blend_constants[k];
    // PixelProcessor::setBlendConstant: UNORM targets use the clamped constant.
    float cu = P.colorFloat ? c : clampf(c, 0.0f, 1.0f);
    P.blendConst[k] = cu;
    P.invBlendConst[k] = 1.0f - cu;
  }

  P.shader = (int)d->shader;
  P.useColor = d->shader != SWR_SHADER_SAMPLE;
  P.useUV = d->shader != SWR_SHADER_INTERPOLATE;

  // Texture descriptor + sampler state.
  if (P.useUV) {
    const swr_texture &tx = d->texture;
    const swr_sampler_state &s = d->sampler;
    TexParams &T = P.tex;
    if (tx.level_count < 1 || tx.level_count > SWR_MAX_MIP_LEVELS)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "texture level_count %u outside 1..15", tx.level_count);
    if (tx.format != SWR_FORMAT_R8G8B8A8_UNORM && tx.format != SWR_FORMAT_R32G32B32A32_SFLOAT)
      return fail(SWR_ERROR_UNSUPPORTED, "texture format %u", tx.format);
    T.isFloat = tx.format == SWR_FORMAT_R32G32B32A32_SFLOAT;
    uint32_t texelBytes = T.isFloat ? 16 : 4;
    for (int l = 0; l < SWR_MAX_MIP_LEVELS; l++) {
      const swr_image_level &lv = tx.levels[std::min<uint32_t>(l, tx.level_count - 1)];
      if (!lv.data || lv.width == 0 || lv.height == 0 || lv.width > 16384 || lv.height > 16384)
        return fail(SWR_ERROR_INVALID_ARGUMENT, "texture level %d invalid", l);
      if (lv.row_pitch_bytes % texelBytes || lv.row_pitch_bytes / texelBytes < lv.width ||
          lv.row_pitch_bytes / texelBytes > 32767)
        return fail(SWR_ERROR_INVALID_ARGUMENT, "texture level %d row pitch invalid", l);
      T.data[l] = lv.data;
      T.width[l] = (int)lv.width;
      T.height[l] = (int)lv.height;
      T.pitchP[l] = (int)(lv.row_pitch_bytes / texelBytes);
      T.uHalf[l] = (uint16_t)(0x8000u / lv.width);
      T.vHalf[l] = (uint16_t)(0x8000u / lv.height);
    }
    T.wwhh[0] = T.wwhh[1] = (float)tx.levels[0].width;
    T.wwhh[2] = T.wwhh[3] = (float)tx.levels[0].height;
    if (s.mag_filter > 1 || s.min_filter > 1 || s.mipmap_mode > 1)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "invalid filter state");
    if ((s.address_mode_u != SWR_ADDRESS_REPEAT && s.address_mode_u != SWR_ADDRESS_CLAMP_TO_EDGE) ||
        (s.address_mode_v != SWR_ADDRESS_REPEAT && s.address_mode_v != SWR_ADDRESS_CLAMP_TO_EDGE))
      return fail(SWR_ERROR_UNSUPPORTED, "address mode (REPEAT and CLAMP_TO_EDGE are supported)");
    // convertFilterMode (SpirvShaderSampling.cpp:275-324).
    if (s.anisotropy_enable) T.filter = FILT_ANISOTROPIC;
    else if (s.mag_filter == SWR_FILTER_NEAREST) T.filter = (s.min_filter == SWR_FILTER_LINEAR) ? FILT_MIN_LINEAR_MAG_POINT : FILT_POINT;
    else T.filter = (s.min_filter == SWR_FILTER_NEAREST) ? FILT_MIN_POINT_MAG_LINEAR : FILT_LINEAR;
    if (s.anisotropy_enable && !(s.max_anisotropy >= 1.0f && s.max_anisotropy <= 16.0f))
      return fail(SWR_ERROR_INVALID_ARGUMENT, "max_anisotropy must be in [1, 16]");
    T.mipmap = (s.mipmap_mode == SWR_MIPMAP_MODE_LINEAR) ? MIP_LINEAR : MIP_POINT;
    T.addrU = s.address_mode_u == SWR_ADDRESS_CLAMP_TO_EDGE;
    T.addrV = s.address_mode_v == SWR_ADDRESS_CLAMP_TO_EDGE;
    T.mipLodBias = s.mip_lod_bias;
    T.maxAnisotropy = s.max_anisotropy;
    // vk::SamplerState::ClampLod (VkSampler.hpp:33-36).
    T.minLod = clampf(s.min_lod, 0.0f, 13.0f);
    T.maxLod = clampf(s.max_lod, 0.0f, 13.0f);
    // Single-level views skip the LOD (SpirvShaderSampling.cpp:92-99).
    if (tx.level_count == 1 && T.filter != FILT_MIN_POINT_MAG_LINEAR && T.filter != FILT_MIN_LINEAR_MAG_POINT) {
      T.minLod = 0.0f;
      T.maxLod = 0.0f;
    }
    T.skipLod = (T.minLod == T.maxLod) && T.filter != FILT_ANISOTROPIC;
  }

  // Attachments.
  if (hasColor) {
    uint32_t bpp = P.colorFloat ? 16 : 4;
    if (d->color.row_pitch_bytes < d->framebuffer_width * bpp)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "colour row pitch too small");
    if (d->samples > 1 && d->color.sample_pitch_bytes < (uint64_t)d->color.row_pitch_bytes * d->framebuffer_height)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "colour sample pitch too small");
  }
  if (P.depthTest) {
    if (d->depth.row_pitch_bytes < d->framebuffer_width * 4) return fail(SWR_ERROR_INVALID_ARGUMENT, "depth row pitch too small");
    if (d->samples > 1 && d->depth.sample_pitch_bytes < (uint64_t)d->depth.row_pitch_bytes * d->framebuffer_height)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "depth sample pitch too small");
  }
  if (P.stencilActive) {
    if (d->stencil.row_pitch_bytes < d->framebuffer_width) return fail(SWR_ERROR_INVALID_ARGUMENT, "stencil row pitch too small");
    if (d->samples > 1 && d->stencil.sample_pitch_bytes < (uint64_t)d->stencil.row_pitch_bytes * d->framebuffer_height)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "stencil sample pitch too small");
  }
  P.color = (char *)d->color.data;
  P.colorPitch = (int)d->color.row_pitch_bytes;
  P.colorSlice = (long long)d->color.sample_pitch_bytes;
  P.depth = (char *)d->depth.data;
  P.depthPitch = (int)d->depth.row_pitch_bytes;
  P.depthSlice = (long long)d->depth.sample_pitch_bytes;
  P.stencil = (char *)d->stencil.data;
  P.stencilPitch = (int)d->stencil.row_pitch_bytes;
  P.stencilSlice = (long long)d->stencil.sample_pitch_bytes;
  P.occlusion = reinterpret_cast<unsigned long long *>(d->occlusion_count);
  if (d->primitive_id.data) {
    if (!hasColor || !P.colorFloat || d->primitive_id.format != SWR_FORMAT_R32_UINT)
      return fail(SWR_ERROR_UNSUPPORTED, "primitive_id needs an R32_UINT attachment and an RGBA32F colour attachment");
    if (d->primitive_id.row_pitch_bytes < d->framebuffer_width * 4)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "primitive_id row pitch too small");
    if (d->samples > 1 &&
        d->primitive_id.sample_pitch_bytes < (uint64_t)d->primitive_id.row_pitch_bytes * d->framebuffer_height)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "primitive_id sample pitch too small");
    P.primId = (char *)d->primitive_id.data;
    P.primIdPitch = (int)d->primitive_id.row_pitch_bytes;
    P.primIdSlice = (long long)d->primitive_id.sample_pitch_bytes;
  }

  P.tileW = (int)tuning.tile_width;
  P.tileH = (int)tuning.tile_height;
  P.tilesX = (P.width + P.tileW - 1) / P.tileW;
  P.tilesY = (P.height + P.tileH - 1) / P.tileH;
  return SWR_OK;
}

template <int MS, bool FCOL, bool FTEX>
void launch_raster(const DrawParams &P, dim3 grid, dim3 block, hipStream_t s, const TriRecord *rec, const int *poly,
                   const uint32_t *bin, const uint32_t *ts, const uint32_t *te) {
  hipLaunchKernelGGL((raster_kernel<MS, FCOL, FTEX>), grid, block, 0, s, P, rec, poly, bin, ts, te);
}

int draw_one(DeviceCtx *c, const swr_draw_desc *desc, hipStream_t s) {
  DrawParams P;
  int r = build_params(desc, c->tuning, P);
  if (r != SWR_OK) return r;
  const uint32_t T = desc->triangle_count;
  if (T == 0) return SWR_OK;
  const uint32_t tiles = (uint32_t)(P.tilesX * P.tilesY);

  SWR_HIP_CHECK(c->rec.ensure(sizeof(TriRecord) * (size_t)T));
  SWR_HIP_CHECK(c->poly.ensure(sizeof(int) * 2 * kMaxPolygon * (size_t)T));
  SWR_HIP_CHECK(c->cnt.ensure(sizeof(uint32_t) * ((size_t)T + 1)));
  SWR_HIP_CHECK(c->off.ensure(sizeof(uint32_t) * ((size_t)T + 1)));
  size_t scanBytes = 0;
  SWR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, c->cnt.as<uint32_t>(), c->off.as<uint32_t>(),
                                                 (int)(T + 1), s));
  SWR_HIP_CHECK(c->scanTmp.ensure(scanBytes));

  const int threads = 256;
  EventSet es{};
  const bool prof = c->profiling;
  if (prof) {
    SWR_HIP_CHECK(c->stats.ensure(4 * sizeof(unsigned long long)));
    P.stats = c->stats.as<unsigned long long>();
    SWR_HIP_CHECK(take_events(c, es));
    SWR_HIP_CHECK(hipEventRecord(es.ev[EV_START], s));
  }
  hipLaunchKernelGGL(setup_kernel, dim3((T + threads - 1) / threads), dim3(threads), 0, s, P, c->rec.as<TriRecord>(),
                     c->poly.as<int>(), c->cnt.as<uint32_t>());
  hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(1), 0, s, c->cnt.as<uint32_t>() + T, 0u, 1u);
  SWR_HIP_CHECK(hipGetLastError());
  if (prof) SWR_HIP_CHECK(hipEventRecord(es.ev[EV_SETUP], s));
  SWR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(c->scanTmp.ptr, scanBytes, c->cnt.as<uint32_t>(),
                                                 c->off.as<uint32_t>(), (int)(T + 1), s));
  SWR_HIP_CHECK(hipMemcpyAsync(c->hostTotal, c->off.as<uint32_t>() + T, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  SWR_HIP_CHECK(hipStreamSynchronize(s));
  const uint32_t pairs = *c->hostTotal;
  if (pairs == 0) {
    if (prof) c->pool.push_back(es);
    return SWR_OK;
  }

  SWR_HIP_CHECK(c->keysA.ensure(sizeof(uint32_t) * (size_t)pairs));
  SWR_HIP_CHECK(c->keysB.ensure(sizeof(uint32_t) * (size_t)pairs));
  SWR_HIP_CHECK(c->valsA.ensure(sizeof(uint32_t) * (size_t)pairs));
  SWR_HIP_CHECK(c->valsB.ensure(sizeof(uint32_t) * (size_t)pairs));
  SWR_HIP_CHECK(c->tStart.ensure(sizeof(uint32_t) * (size_t)tiles));
  SWR_HIP_CHECK(c->tEnd.ensure(sizeof(uint32_t) * (size_t)tiles));
  int bits = 1;
  while ((1u << bits) < tiles) bits++;
  size_t sortBytes = 0;
  SWR_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, c->keysA.as<uint32_t>(), c->keysB.as<uint32_t>(),
                                                   c->valsA.as<uint32_t>(), c->valsB.as<uint32_t>(), (int)pairs, 0,
                                                   bits, s));
  SWR_HIP_CHECK(c->sortTmp.ensure(sortBytes));

  hipLaunchKernelGGL(bin_emit_kernel, dim3((pairs + threads - 1) / threads), dim3(threads), 0, s, P,
                     c->rec.as<TriRecord>(), c->off.as<uint32_t>(), pairs, c->keysA.as<uint32_t>(),
                     c->valsA.as<uint32_t>());
  SWR_HIP_CHECK(hipGetLastError());
  SWR_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(c->sortTmp.ptr, sortBytes, c->keysA.as<uint32_t>(),
                                                   c->keysB.as<uint32_t>(), c->valsA.as<uint32_t>(),
                                                   c->valsB.as<uint32_t>(), (int)pairs, 0, bits, s));
  hipLaunchKernelGGL(fill_u32_kernel, dim3((tiles + 255) / 256), dim3(256), 0, s, c->tStart.as<uint32_t>(), 0u, tiles);
  hipLaunchKernelGGL(fill_u32_kernel, dim3((tiles + 255) / 256), dim3(256), 0, s, c->tEnd.as<uint32_t>(), 0u, tiles);
  hipLaunchKernelGGL(tile_range_kernel, dim3((pairs + 255) / 256), dim3(256), 0, s, c->keysB.as<uint32_t>(), pairs,
                     c->tStart.as<uint32_t>(), c->tEnd.as<uint32_t>());
  SWR_HIP_CHECK(hipGetLastError());
  if (prof) SWR_HIP_CHECK(hipEventRecord(es.ev[EV_BIN], s));

  dim3 grid(tiles), block(256);  // one workgroup per bin, one wave per 8x8 block
  const TriRecord *rec = c->rec.as<TriRecord>();
  const int *poly = c->poly.as<int>();
  const uint32_t *bin = c->valsB.as<uint32_t>();
  const uint32_t *ts = c->tStart.as<uint32_t>(), *te = c->tEnd.as<uint32_t>();
  const bool ftex = P.useUV && P.tex.isFloat;
  const bool fcol = P.colorFloat != 0;
  if (P.ms == 1) {
    if (!fcol && !ftex) launch_raster<1, false, false>(P, grid, block, s, rec, poly, bin, ts, te);
    else if (!fcol && ftex) launch_raster<1, false, true>(P, grid, block, s, rec, poly, bin, ts, te);
    else if (fcol && !ftex) launch_raster<1, true, false>(P, grid, block, s, rec, poly, bin, ts, te);
    else launch_raster<1, true, true>(P, grid, block, s, rec, poly, bin, ts, te);
  } else {
    if (!fcol && !ftex) launch_raster<4, false, false>(P, grid, block, s, rec, poly, bin, ts, te);
    else if (!fcol && ftex) launch_raster<4, false, true>(P, grid, block, s, rec, poly, bin, ts, te);
    else if (fcol && !ftex) launch_raster<4, true, false>(P, grid, block, s, rec, poly, bin, ts, te);
    else launch_raster<4, true, true>(P, grid, block, s, rec, poly, bin, ts, te);
  }
  SWR_HIP_CHECK(hipGetLastError());
  if (prof) {
    SWR_HIP_CHECK(hipEventRecord(es.ev[EV_RASTER], s));
    es.pairs = pairs;
    c->pending.push_back(es);
  }
  return SWR_OK;
}

int draw_backward(DeviceCtx *c, const swr_draw_desc *desc, const swr_grad_desc *g, hipStream_t s) {
  if (!g) return fail(SWR_ERROR_INVALID_ARGUMENT, "grad descriptor is NULL");
  DrawParams P;
  int r = build_params(desc, c->tuning, P);
  if (r != SWR_OK) return r;
  if (!P.primId) return fail(SWR_ERROR_INVALID_ARGUMENT, "backward needs the draw's primitive_id attachment");
  if (P.blendEnable) return fail(SWR_ERROR_UNSUPPORTED, "backward of a blended draw");
  if (!g->d_color) return fail(SWR_ERROR_INVALID_ARGUMENT, "d_color is NULL");
  if (g->d_color_row_pitch_bytes < desc->framebuffer_width * 16 ||
      (P.ms > 1 && g->d_color_sample_pitch_bytes < (uint64_t)g->d_color_row_pitch_bytes * desc->framebuffer_height))
    return fail(SWR_ERROR_INVALID_ARGUMENT, "d_color pitches too small");
  if (g->sample_state &&
      (g->sample_state_row_pitch_bytes < desc->framebuffer_width * 16 ||
       (P.ms > 1 &&
        g->sample_state_sample_pitch_bytes < (uint64_t)g->sample_state_row_pitch_bytes * desc->framebuffer_height)))
    return fail(SWR_ERROR_INVALID_ARGUMENT, "sample_state pitches too small");
  GradParams G{};
  G.dColor = (const char *)g->d_color;
  G.dColorPitch = (int)g->d_color_row_pitch_bytes;
  G.dColorSlice = (long long)g->d_color_sample_pitch_bytes;
  G.dPos = g->d_positions;
  G.dCol = P.useColor ? g->d_colors : nullptr;
  G.dUV = P.useUV ? g->d_texcoords : nullptr;
  G.state = (char *)g->sample_state;
  G.statePitch = (int)g->sample_state_row_pitch_bytes;
  G.stateSlice = (long long)g->sample_state_sample_pitch_bytes;
  if (P.useUV) {
    if (!P.tex.isFloat) return fail(SWR_ERROR_UNSUPPORTED, "backward samples RGBA32F textures only");
    if (P.tex.filter == FILT_MIN_POINT_MAG_LINEAR || P.tex.filter == FILT_MIN_LINEAR_MAG_POINT)
      return fail(SWR_ERROR_UNSUPPORTED, "backward of mixed min/mag filters");
    const uint32_t n = desc->texture.level_count;
    for (int l = 0; l < kMipLevels; l++) G.dTex[l] = (float *)g->d_texture[std::min<uint32_t>(l, n - 1)];
  }
  const uint32_t T = desc->triangle_count;
  if (T == 0) return SWR_OK;
  SWR_HIP_CHECK(c->rec.ensure(sizeof(TriRecord) * (size_t)T));
  SWR_HIP_CHECK(c->poly.ensure(sizeof(int) * 2 * kMaxPolygon * (size_t)T));
  SWR_HIP_CHECK(c->cnt.ensure(sizeof(uint32_t) * ((size_t)T + 1)));
  hipLaunchKernelGGL(setup_kernel, dim3((T + 255) / 256), dim3(256), 0, s, P, c->rec.as<TriRecord>(),
                     c->poly.as<int>(), c->cnt.as<uint32_t>());
  dim3 grid((P.width + 15) / 16, (P.height + 15) / 16), block(256);
  if (P.ms == 1)
    hipLaunchKernelGGL(backward_kernel<1>, grid, block, 0, s, P, G, (const TriRecord *)c->rec.as<TriRecord>());
  else
    hipLaunchKernelGGL(backward_kernel<4>, grid, block, 0, s, P, G, (const TriRecord *)c->rec.as<TriRecord>());
  SWR_HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

// Validates a mip chain: level k+1 is max(level k / 2, 1) on each axis.
int check_chain(const swr_texture *t, uint32_t texelBytes) {
  if (t->level_count < 1 || t->level_count > SWR_MAX_MIP_LEVELS)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "level_count %u outside 1..15", t->level_count);
  for (uint32_t l = 0; l < t->level_count; l++) {
    const swr_image_level &v = t->levels[l];
    if (!v.data || v.width == 0 || v.height == 0 || v.row_pitch_bytes % texelBytes ||
        v.row_pitch_bytes / texelBytes < v.width)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "level %u invalid", l);
    if (l > 0) {
      const swr_image_level &u = t->levels[l - 1];
      if (v.width != std::max(u.width / 2, 1u) || v.height != std::max(u.height / 2, 1u))
        return fail(SWR_ERROR_INVALID_ARGUMENT, "level %u is not half of level %u", l, l - 1);
    }
  }
  return SWR_OK;
}

}  // namespace

extern "C" {

int swr_draw_backward(const swr_draw_desc *desc, const swr_grad_desc *grad, void *stream) {
  DeviceCtx *c = nullptr;
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  if (!desc) return fail(SWR_ERROR_INVALID_ARGUMENT, "desc is NULL");
  return draw_backward(c, desc, grad, reinterpret_cast<hipStream_t>(stream));
}

int swr_generate_mips(const swr_texture *tex, void *stream) {
  if (!tex) return fail(SWR_ERROR_INVALID_ARGUMENT, "texture is NULL");
  const bool f = tex->format == SWR_FORMAT_R32G32B32A32_SFLOAT;
  if (!f && tex->format != SWR_FORMAT_R8G8B8A8_UNORM) return fail(SWR_ERROR_UNSUPPORTED, "format %u", tex->format);
  const uint32_t tb = f ? 16 : 4;
  int r = check_chain(tex, tb);
  if (r != SWR_OK) return r;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  for (uint32_t l = 1; l < tex->level_count; l++) {
    const swr_image_level &a = tex->levels[l - 1], &b = tex->levels[l];
    dim3 grid((b.width + 255) / 256, b.height), block(256);
    if (f)
      hipLaunchKernelGGL(mip_down_rgba32f_kernel, grid, block, 0, s, (const float4 *)a.data, (int)a.width,
                         (int)a.height, (int)(a.row_pitch_bytes / tb), (float4 *)b.data, (int)b.width, (int)b.height,
                         (int)(b.row_pitch_bytes / tb));
    else
      hipLaunchKernelGGL(mip_down_rgba8_kernel, grid, block, 0, s, (const uint32_t *)a.data, (int)a.width,
                         (int)a.height, (int)(a.row_pitch_bytes / tb), (uint32_t *)b.data, (int)b.width, (int)b.height,
                         (int)(b.row_pitch_bytes / tb));
  }
  SWR_HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_mip_backward(const swr_texture *grad, void *stream) {
  if (!grad) return fail(SWR_ERROR_INVALID_ARGUMENT, "texture is NULL");
  if (grad->format != SWR_FORMAT_R32G32B32A32_SFLOAT) return fail(SWR_ERROR_UNSUPPORTED, "format %u", grad->format);
  int r = check_chain(grad, 16);
  if (r != SWR_OK) return r;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  for (uint32_t l = grad->level_count - 1; l >= 1; l--) {
    const swr_image_level &fine = grad->levels[l - 1], &coarse = grad->levels[l];
    dim3 grid((fine.width + 255) / 256, fine.height), block(256);
    hipLaunchKernelGGL(mip_up_rgba32f_kernel, grid, block, 0, s, (float4 *)fine.data, (int)fine.width,
                       (int)fine.height, (int)(fine.row_pitch_bytes / 16), (const float4 *)coarse.data,
                       (int)coarse.width, (int)coarse.height, (int)(coarse.row_pitch_bytes / 16));
  }
  SWR_HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_abi_version(void) { return SWR_ABI_VERSION; }

const char *swr_last_error(void) { return g_error.c_str(); }

int swr_init(int device) {
  int count = 0;
  SWR_HIP_CHECK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(SWR_ERROR_INVALID_ARGUMENT, "device %d of %d", device, count);
  SWR_HIP_CHECK(hipSetDevice(device));
  DeviceCtx *c = nullptr;
  return current_ctx(&c);
}

int swr_release(void) {
  DeviceCtx *c = nullptr;
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  SWR_HIP_CHECK(hipDeviceSynchronize());
  for (Scratch *s : {&c->rec, &c->poly, &c->cnt, &c->off, &c->keysA, &c->keysB, &c->valsA, &c->valsB, &c->scanTmp,
                     &c->sortTmp, &c->tStart, &c->tEnd})
    s->release();
  return SWR_OK;
}

int swr_draw(const swr_draw_desc *desc, void *stream) {
  DeviceCtx *c = nullptr;
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  return draw_one(c, desc, reinterpret_cast<hipStream_t>(stream));
}

int swr_draw_batch(const swr_draw_desc *descs, uint32_t count, void *stream) {
  if (count && !descs) return fail(SWR_ERROR_INVALID_ARGUMENT, "descs is NULL");
  DeviceCtx *c = nullptr;
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  for (uint32_t i = 0; i < count; i++) {
    r = draw_one(c, &descs[i], reinterpret_cast<hipStream_t>(stream));
    if (r != SWR_OK) return r;
  }
  return SWR_OK;
}

int swr_resolve(const swr_resolve_desc *d, void *stream) {
  if (!d || !d->src || !d->dst) return fail(SWR_ERROR_INVALID_ARGUMENT, "resolve descriptor incomplete");
  if (d->src_samples != 4) return fail(SWR_ERROR_UNSUPPORTED, "resolve supports 4 samples");
  if (d->format != SWR_FORMAT_R8G8B8A8_UNORM && d->format != SWR_FORMAT_R32G32B32A32_SFLOAT)
    return fail(SWR_ERROR_UNSUPPORTED, "resolve format %u", d->format);
  if (d->width == 0 || d->height == 0) return SWR_OK;
  uint32_t bpp = d->format == SWR_FORMAT_R8G8B8A8_UNORM ? 4 : 16;
  if (d->src_row_pitch_bytes < d->width * bpp || d->dst_row_pitch_bytes < d->width * bpp ||
      d->src_sample_pitch_bytes < (uint64_t)d->src_row_pitch_bytes * d->height)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "resolve pitches too small");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  dim3 block(256), grid((d->width + 255) / 256, d->height);
  if (bpp == 4)
    hipLaunchKernelGGL(resolve_rgba8_kernel, grid, block, 0, s, (const char *)d->src, (int)d->src_row_pitch_bytes,
                       (long long)d->src_sample_pitch_bytes, (char *)d->dst, (int)d->dst_row_pitch_bytes,
                       (int)d->width, (int)d->height);
  else
    hipLaunchKernelGGL(resolve_rgba32f_kernel, grid, block, 0, s, (const char *)d->src, (int)d->src_row_pitch_bytes,
                       (long long)d->src_sample_pitch_bytes, (char *)d->dst, (int)d->dst_row_pitch_bytes,
                       (int)d->width, (int)d->height);
  SWR_HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_profile_enable(int enable) {
  DeviceCtx *c = nullptr;
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  if (enable && !c->profiling) {
    SWR_HIP_CHECK(c->stats.ensure(4 * sizeof(unsigned long long)));
    SWR_HIP_CHECK(hipMemset(c->stats.ptr, 0, 4 * sizeof(unsigned long long)));
  }
  c->profiling = enable != 0;
  return SWR_OK;
}

int swr_profile_read(swr_profile *out) {
  if (!out) return fail(SWR_ERROR_INVALID_ARGUMENT, "profile is NULL");
  DeviceCtx *c = nullptr;
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  for (EventSet &es : c->pending) {
    SWR_HIP_CHECK(hipEventSynchronize(es.ev[EV_RASTER]));
    float a = 0, b = 0, d = 0, t = 0;
    SWR_HIP_CHECK(hipEventElapsedTime(&a, es.ev[EV_START], es.ev[EV_SETUP]));
    SWR_HIP_CHECK(hipEventElapsedTime(&b, es.ev[EV_SETUP], es.ev[EV_BIN]));
    SWR_HIP_CHECK(hipEventElapsedTime(&d, es.ev[EV_BIN], es.ev[EV_RASTER]));
    SWR_HIP_CHECK(hipEventElapsedTime(&t, es.ev[EV_START], es.ev[EV_RASTER]));
    c->acc.setup_ms += a;
    c->acc.bin_ms += b;
    c->acc.raster_ms += d;
    c->acc.total_ms += t;
    c->acc.draws += 1;
    c->acc.pairs += es.pairs;
    c->pool.push_back(es);
  }
  c->pending.clear();
  if (c->stats.ptr) {
    unsigned long long st[4] = {0, 0, 0, 0};
    SWR_HIP_CHECK(hipDeviceSynchronize());
    SWR_HIP_CHECK(hipMemcpy(st, c->stats.ptr, sizeof(st), hipMemcpyDeviceToHost));
    SWR_HIP_CHECK(hipMemset(c->stats.ptr, 0, sizeof(st)));
    c->acc.block_tests = st[0];
    c->acc.block_covers = st[1];
    c->acc.shaded = st[2];
  }
  *out = c->acc;
  c->acc = swr_profile{};
  return SWR_OK;
}

int swr_set_tuning(const swr_tuning *t) {
  if (!t) return fail(SWR_ERROR_INVALID_ARGUMENT, "tuning is NULL");
  if (t->tile_width != 16 || t->tile_height != 16)
    return fail(SWR_ERROR_UNSUPPORTED, "tile %ux%u (16x16 supported)", t->tile_width, t->tile_height);
  DeviceCtx *c = nullptr;
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  c->tuning = *t;
  return SWR_OK;
}

int swr_get_tuning(swr_tuning *t) {
  if (!t) return fail(SWR_ERROR_INVALID_ARGUMENT, "tuning is NULL");
  DeviceCtx *c = nullptr;
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  *t = c->tuning;
  return SWR_OK;
}

}  // extern "C"
