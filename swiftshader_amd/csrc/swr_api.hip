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
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/swr.h"
#include "swr_srgb.hpp"
#include "swr_kernels.hip"
#include "swr_batch.hip"
#include "swr_grad.hip"
#include "swr_decode.hip"

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
  uint64_t draws;
};

// Device scratch of the tiled pipeline for one segment at a time.
struct Pipeline {
  Scratch params, aux;                      // DrawParams[n]; blockBase[n], tileBase[n]
  Scratch rec, poly, cnt, off, sums;        // per primitive slot
  Scratch keysA, keysB, valsA, valsB, hist; // per pair
  Scratch first;                            // per block of emitted pairs
  Scratch tStart, tEnd;                     // per tile
  Scratch counters;  // work counters and flags (kOverflowSlot and below)
  void release() {
    for (Scratch *s : {&params, &aux, &rec, &poly, &cnt, &off, &sums, &keysA, &keysB, &valsA, &valsB, &hist, &first,
                       &tStart, &tEnd, &counters})
      s->release();
  }
};

struct DeviceCtx {
  int device = -1;
  int cus = 256;
  Pipeline pipe;
  Scratch stats;
  uint32_t *hostTotal = nullptr;  // pinned
  std::vector<char> staging;      // host image of the params upload
  void *pinned = nullptr;         // pinned staging buffer for the params upload
  size_t pinnedBytes = 0;
  swr_tuning tuning{};
  bool profiling = false;
  std::vector<EventSet> pending, pool;
  swr_profile acc{};
  // Single draws that do not wait on the host: their pair capacity (0:
  // unknown yet), and the read-back of a draw's pair count into
  // hostTotal[2], complete when pairsEvent is.
  uint32_t pairCap = 0;
  hipEvent_t pairsEvent = nullptr;
  bool pairsPending = false;
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
    c->tuning.raster_workgroups_per_cu = ~0u;  // one workgroup per tile: fastest for a single large draw
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
      c->cus = prop.multiProcessorCount;
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

// The sampler's sRGB decode table, as the reference's Constants builds it
// (Constants.cpp:273-276): sw::sRGBtoLinear (Math.hpp:326-336, the host's
// powf) of i / 255, scaled to 0xFF00 and rounded half up.
const uint16_t *srgb_decode_table() {
  static const std::vector<uint16_t> table = [] {
    std::vector<uint16_t> t(256);
    for (int i = 0; i < 256; i++) {
      volatile float c = (float)i / 0xFF;  // keeps powf a call to the host's libm
      const float lin = c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
      t[i] = (uint16_t)(lin * 0xFF00 + 0.5f);
    }
    return t;
  }();
  return table.data();
}

bool valid_compare(uint32_t op) { return op <= 7; }
bool valid_factor(uint32_t f) { return f <= 14; }
bool is_advanced_blend_op(uint32_t op) {
  return (op >= SWR_BLEND_OP_MULTIPLY_EXT && op <= SWR_BLEND_OP_EXCLUSION_EXT) ||
         (op >= SWR_BLEND_OP_HSL_HUE_EXT && op <= SWR_BLEND_OP_HSL_LUMINOSITY_EXT);
}
bool valid_blend_op(uint32_t op) { return op <= 4 || is_advanced_blend_op(op); }

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
  switch (d->topology) {
    case SWR_TOPOLOGY_POINT_LIST: P.topology = 5; break;
    case SWR_TOPOLOGY_TRIANGLE_LIST: P.topology = 0; break;
    case SWR_TOPOLOGY_TRIANGLE_STRIP: P.topology = 1; break;
    case SWR_TOPOLOGY_TRIANGLE_FAN: P.topology = 2; break;
    case SWR_TOPOLOGY_LINE_LIST: P.topology = 3; break;
    case SWR_TOPOLOGY_LINE_STRIP: P.topology = 4; break;
    default: return fail(SWR_ERROR_UNSUPPORTED, "topology %u (points, line and triangle lists and strips, fans)", d->topology);
  }
  if (d->provoking_vertex > 1) return fail(SWR_ERROR_INVALID_ARGUMENT, "provoking_vertex %u", d->provoking_vertex);
  P.provokeLast = d->provoking_vertex == SWR_PROVOKING_VERTEX_LAST;
  P.line = P.topology == 3 || P.topology == 4;
  P.point = P.topology == 5;
  P.psize = d->point_sizes;
  {
    // Lines: width 0 reads as 1; the rectangle or, for the Bresenham mode
    // without samples, the parallelogram (Renderer.cpp:920-1163).
    P.lineWidth = d->line_width == 0.0f ? 1.0f : d->line_width;
    if (!(P.lineWidth > 0.0f) || P.lineWidth > 64.0f)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "line_width %g outside (0, 64]", (double)d->line_width);
    if (d->line_rasterization_mode > SWR_LINE_RASTERIZATION_RECTANGULAR_SMOOTH)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "line_rasterization_mode %u", d->line_rasterization_mode);
    P.bresenham = P.line && d->line_rasterization_mode == SWR_LINE_RASTERIZATION_BRESENHAM;
    if (P.bresenham && d->samples != 1)
      return fail(SWR_ERROR_UNSUPPORTED, "Bresenham lines with samples (the reference turns multisampling off)");
    if ((P.line || P.point) && d->depth_bias_enable &&
        (d->depth_bias_constant_factor != 0.0f || d->depth_bias_slope_factor != 0.0f))
      return fail(SWR_ERROR_UNSUPPORTED, "depth bias of lines or points (the reference applies it to triangles only)");
  }
  {
    const uint64_t T = d->triangle_count;
    const uint64_t need = P.topology == 0   ? T * 3
                          : P.topology == 3 ? T * 2
                          : P.topology == 4 ? T + 1
                          : P.topology == 5 ? T
                                            : T + 2;
    if (!d->indices && d->triangle_count && need > d->vertex_count)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "non-indexed draw needs %llu vertices", (unsigned long long)need);
    if (d->triangle_count && d->vertex_count == 0) return fail(SWR_ERROR_INVALID_ARGUMENT, "vertex_count is 0");
  }
  if (d->shader > SWR_SHADER_MODULATE) return fail(SWR_ERROR_INVALID_ARGUMENT, "unknown shader %u", d->shader);
  if (d->shader != SWR_SHADER_SAMPLE && !d->colors)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "shader reads colours but colors is NULL");
  if (d->shader != SWR_SHADER_INTERPOLATE && !d->texcoords)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "shader samples but texcoords is NULL");

  P.pos = d->positions;
  P.col = d->colors;
  P.uv = d->texcoords;
  P.idx = d->indices;
  switch (d->index_type) {
    case SWR_INDEX_TYPE_UINT16: P.idxBytes = 2; break;
    case SWR_INDEX_TYPE_UINT32: P.idxBytes = 4; break;
    case SWR_INDEX_TYPE_UINT8: P.idxBytes = 1; break;
    default: return fail(SWR_ERROR_INVALID_ARGUMENT, "index_type %u", d->index_type);
  }
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
  if (d->depth.data && d->depth.format != SWR_FORMAT_D32_SFLOAT && d->depth.format != SWR_FORMAT_D16_UNORM)
    return fail(SWR_ERROR_UNSUPPORTED, "depth format %u (D32_SFLOAT or D16_UNORM)", d->depth.format);
  P.depth16 = d->depth.data && d->depth.format == SWR_FORMAT_D16_UNORM;
  // Depth clip follows depth clamp unless set (Context.cpp:647-648, 686-693);
  // depth clamp takes the viewport's range (PixelProcessor.cpp:120-136).
  if (d->depth_clip_control > 2) return fail(SWR_ERROR_INVALID_ARGUMENT, "depth_clip_control %u", d->depth_clip_control);
  P.depthClip = d->depth_clip_control == 0 ? !d->depth_clamp_enable : d->depth_clip_control == 1;
  if (d->depth_clamp_enable) {
    P.depthClampLo = std::min(d->viewport.min_depth, d->viewport.max_depth);
    P.depthClampHi = std::max(d->viewport.min_depth, d->viewport.max_depth);
  } else {
    P.depthClampLo = 0.0f;
    P.depthClampHi = 1.0f;
  }
  if (d->stencil.data && d->stencil.format != SWR_FORMAT_S8_UINT)
    return fail(SWR_ERROR_UNSUPPORTED, "stencil format %u (only S8_UINT)", d->stencil.format);
  if (!valid_compare(d->depth_compare_op)) return fail(SWR_ERROR_INVALID_ARGUMENT, "invalid depth compare op");
  P.depthTest = (d->depth.data && d->depth_test_enable) ? 1 : 0;
  P.depthWrite = (P.depthTest && d->depth_write_enable) ? 1 : 0;
  // depthBoundsTestActive (Context.cpp:946-949).
  P.depthBounds = (d->depth.data && d->depth_bounds_test_enable) ? 1 : 0;
  // Alpha-to-coverage thresholds per sample (Renderer.cpp:391-411).
  if (d->alpha_to_coverage_enable > 1) return fail(SWR_ERROR_INVALID_ARGUMENT, "alpha_to_coverage_enable %u", d->alpha_to_coverage_enable);
  P.alphaToCoverage = (int)d->alpha_to_coverage_enable;
  if (d->samples == 4) {
    P.a2c[0] = 0.2f; P.a2c[1] = 0.4f; P.a2c[2] = 0.6f; P.a2c[3] = 0.8f;
  } else {
    P.a2c[0] = 0.5f;
  }
  P.depthBoundsMin = d->min_depth_bounds;
  P.depthBoundsMax = d->max_depth_bounds;
  P.depthCompare = (int)d->depth_compare_op;
  {
    // VkCompareOp as (predicate, inversion); NaN depths fail <=, < and ==.
    static const int kPred[8] = {3, 0, 2, 1, 1, 2, 0, 3}, kInv[8] = {1, 1, 0, 1, 0, 1, 0, 0};
    const uint32_t op = d->depth_compare_op < 8 ? d->depth_compare_op : 7u;
    P.zPred = kPred[op];
    P.zInv = kInv[op];
  }
  // Depth bias state (Context.cpp:640-661: a NaN clamp is no clamp;
  // SetupProcessor.cpp:75-77 and PixelProcessor.cpp:119: each term only
  // when its factor is non-zero).
  {
    const bool en = d->depth_bias_enable != 0;
    const float cst = en ? d->depth_bias_constant_factor : 0.0f;
    const float slope = en ? d->depth_bias_slope_factor : 0.0f;
    float clampv = en ? d->depth_bias_clamp : 0.0f;
    if (clampv != clampv) clampv = 0.0f;
    P.biasConstant = cst != 0.0f;
    P.biasSlope = slope != 0.0f;
    P.biasClamp = clampv != 0.0f;
    P.depthBias = P.depthTest && (P.biasConstant || P.biasSlope);
    P.constantDepthBias = cst;
    P.slopeDepthBias = slope;
    P.depthBiasClamp = clampv;
  }
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
  if (hasColor && d->color.format != SWR_FORMAT_R8G8B8A8_UNORM && d->color.format != SWR_FORMAT_B8G8R8A8_UNORM &&
      d->color.format != SWR_FORMAT_R8G8B8A8_SRGB && d->color.format != SWR_FORMAT_B8G8R8A8_SRGB &&
      d->color.format != SWR_FORMAT_R32G32B32A32_SFLOAT)
    return fail(SWR_ERROR_UNSUPPORTED, "colour format %u", d->color.format);
  P.colorFloat = d->color.format == SWR_FORMAT_R32G32B32A32_SFLOAT;
  P.colorBGRA = hasColor && (d->color.format == SWR_FORMAT_B8G8R8A8_UNORM || d->color.format == SWR_FORMAT_B8G8R8A8_SRGB);
  P.colorSRGB = hasColor && (d->color.format == SWR_FORMAT_R8G8B8A8_SRGB || d->color.format == SWR_FORMAT_B8G8R8A8_SRGB);
  const swr_blend_state &b = d->blend;
  if (!valid_factor(b.src_color_blend_factor) || !valid_factor(b.dst_color_blend_factor) ||
      !valid_factor(b.src_alpha_blend_factor) || !valid_factor(b.dst_alpha_blend_factor) ||
      !valid_blend_op(b.color_blend_op) || !valid_blend_op(b.alpha_blend_op))
    return fail(SWR_ERROR_UNSUPPORTED, "blend state outside the supported VkBlendFactor/VkBlendOp range");
  if ((is_advanced_blend_op(b.color_blend_op) || is_advanced_blend_op(b.alpha_blend_op)) &&
      b.color_blend_op != b.alpha_blend_op)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "an advanced blend op must be both the colour and the alpha op");
  int cop = blend_operation((int)b.color_blend_op, (int)b.src_color_blend_factor, (int)b.dst_color_blend_factor, P.colorFloat);
  int aop = blend_operation((int)b.alpha_blend_op, (int)b.src_alpha_blend_factor, (int)b.dst_alpha_blend_factor, P.colorFloat);
  if (hasColor) P.colorMask = (cop == BLEND_OP_DST && aop == BLEND_OP_DST) ? 0 : (int)(b.color_write_mask & 0xFu);
  P.blendEnable = (hasColor && b.blend_enable && P.colorMask && (cop != BLEND_OP_SRC || aop != BLEND_OP_SRC)) ? 1 : 0;
  // Blend factors: ONE for MIN, MAX and the advanced modes (Context.cpp:1149-1179).
  auto bf = [](uint32_t op, uint32_t f) {
    return (op == SWR_BLEND_OP_MIN || op == SWR_BLEND_OP_MAX || is_advanced_blend_op(op)) ? (int)SWR_BLEND_FACTOR_ONE
                                                                                          : (int)f;
  };
  if (P.blendEnable) {
    P.srcF = bf(b.color_blend_op, b.src_color_blend_factor);
    P.dstF = bf(b.color_blend_op, b.dst_color_blend_factor);
    P.colorOp = cop;
    P.advancedBlend = is_advanced_blend_op((uint32_t)cop);
    P.srcFA = bf(b.alpha_blend_op, b.src_alpha_blend_factor);
    P.dstFA = bf(b.alpha_blend_op, b.dst_alpha_blend_factor);
    P.alphaOp = aop;
    P.blendKind = (P.srcF == SWR_BLEND_FACTOR_SRC_ALPHA && P.dstF == SWR_BLEND_FACTOR_ONE_MINUS_SRC_ALPHA &&
                   P.colorOp == 0 && P.srcFA == SWR_BLEND_FACTOR_SRC_ALPHA &&
                   P.dstFA == SWR_BLEND_FACTOR_ONE_MINUS_SRC_ALPHA && P.alphaOp == 0)
                      ? 1
                      : 0;
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
    if (tx.format != SWR_FORMAT_R8G8B8A8_UNORM && tx.format != SWR_FORMAT_B8G8R8A8_UNORM &&
        tx.format != SWR_FORMAT_R8G8B8A8_SRGB && tx.format != SWR_FORMAT_B8G8R8A8_SRGB &&
        tx.format != SWR_FORMAT_R32G32B32A32_SFLOAT)
      return fail(SWR_ERROR_UNSUPPORTED, "texture format %u", tx.format);
    T.isFloat = tx.format == SWR_FORMAT_R32G32B32A32_SFLOAT;
    // B8G8R8A8 (e.g. decoded BC1-BC3) keeps R and B exchanged in memory.
    T.bgra = tx.format == SWR_FORMAT_B8G8R8A8_UNORM || tx.format == SWR_FORMAT_B8G8R8A8_SRGB;
    T.srgb = tx.format == SWR_FORMAT_R8G8B8A8_SRGB || tx.format == SWR_FORMAT_B8G8R8A8_SRGB;
    if (T.srgb) memcpy(T.srgbLut, srgb_decode_table(), sizeof(T.srgbLut));
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
    auto addr_ok = [](uint32_t m) {
      return m == SWR_ADDRESS_REPEAT || m == SWR_ADDRESS_CLAMP_TO_EDGE || m == SWR_ADDRESS_MIRRORED_REPEAT ||
             m == SWR_ADDRESS_MIRROR_CLAMP_TO_EDGE;
    };
    if (!addr_ok(s.address_mode_u) || !addr_ok(s.address_mode_v))
      return fail(SWR_ERROR_UNSUPPORTED, "address mode (CLAMP_TO_BORDER is not supported)");
    // convertFilterMode (SpirvShaderSampling.cpp:275-324).
    if (s.anisotropy_enable) T.filter = FILT_ANISOTROPIC;
    else if (s.mag_filter == SWR_FILTER_NEAREST) T.filter = (s.min_filter == SWR_FILTER_LINEAR) ? FILT_MIN_LINEAR_MAG_POINT : FILT_POINT;
    else T.filter = (s.min_filter == SWR_FILTER_NEAREST) ? FILT_MIN_POINT_MAG_LINEAR : FILT_LINEAR;
    if (s.anisotropy_enable && !(s.max_anisotropy >= 1.0f && s.max_anisotropy <= 16.0f))
      return fail(SWR_ERROR_INVALID_ARGUMENT, "max_anisotropy must be in [1, 16]");
    T.mipmap = (s.mipmap_mode == SWR_MIPMAP_MODE_LINEAR) ? MIP_LINEAR : MIP_POINT;
    // Address modes: 0 wrap, 1 clamp to edge, 2 mirrored repeat, 3 mirror clamp to edge.
    auto addr_code = [](uint32_t m) {
      return m == SWR_ADDRESS_CLAMP_TO_EDGE ? 1 : m == SWR_ADDRESS_MIRRORED_REPEAT ? 2
                                                : m == SWR_ADDRESS_MIRROR_CLAMP_TO_EDGE ? 3 : 0;
    };
    T.addrU = addr_code(s.address_mode_u);
    T.addrV = addr_code(s.address_mode_v);
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
    if (d->depth.row_pitch_bytes < d->framebuffer_width * (P.depth16 ? 2u : 4u))
      return fail(SWR_ERROR_INVALID_ARGUMENT, "depth row pitch too small");
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

// ---------------------------------------------------------------------------
// Batched pipeline
// ---------------------------------------------------------------------------
// Raster variant of a draw: samples, float colour target, float texture,
// and whether it uses a format outside the common RGBA8 / RGBA32F / D32F set
// (D16 depth, BGRA8 or sRGB colour, BGRA8 or sRGB texture), a mirrored
// address mode, the depth bounds test or an advanced blend mode, which the common variants compile
// out.
int variant_of(const DrawParams &P) {
  const bool ftex = P.useUV && P.tex.isFloat;
  const bool xf = P.depth16 || P.colorBGRA || P.colorSRGB || P.advancedBlend || P.depthBounds || P.alphaToCoverage ||
                  (P.useUV && (P.tex.bgra || P.tex.srgb || P.tex.addrU > 1 || P.tex.addrV > 1));
  // The common output state (no stencil or depth bias; depth test and write;
  // all colour channels) has kernels of its own with those checks folded
  // (level 1); so does that state with the common shading state: MODULATE,
  // linear or anisotropic filtering, linear mips, blending off or SRC_ALPHA /
  // ONE_MINUS_SRC_ALPHA / ADD, and REPEAT (level 2) or CLAMP_TO_EDGE (level
  // 3) on both axes.
  const bool simple = !xf && !P.stencilActive && !P.stencilWrite && !P.depthBias && P.depthTest && P.depthWrite &&
                      P.colorMask == 15;
  const bool fast = simple && P.shader == SHADER_MODULATE && P.useColor && P.useUV &&
                    (!P.blendEnable || P.blendKind == 1) && !P.tex.skipLod && P.tex.mipmap == MIP_LINEAR &&
                    P.tex.addrU == P.tex.addrV && P.tex.addrU <= 1 &&
                    (P.tex.filter == FILT_LINEAR || P.tex.filter == FILT_ANISOTROPIC);
  const int level = fast ? 2 + P.tex.addrU : simple ? 1 : 0;
  return (level << 4) | (xf ? 8 : 0) | (P.ms > 1 ? 4 : 0) | (P.colorFloat ? 2 : 0) | (ftex ? 1 : 0);
}

// Host plan of a batch: per-draw state in enqueue order, split into
// segments of draws none of which depends on an earlier one of its segment
// (depends_on: no memory one writes meets memory another reads or writes),
// so that the draws of a segment can be rendered concurrently; each
// segment's draws grouped by raster variant.
struct BatchLayout {
  struct Seg {
    int first = 0, count = 0;  // draws [first, first + count) of P
    uint32_t T = 0, blocks = 0, tiles = 0;
    std::vector<std::pair<int, int>> variants;  // [lo, hi) draw ranges (segment-relative)
    uint32_t pairCap = 0;                       // planned pair capacity (plans only)
  };
  std::vector<DrawParams> P;
  std::vector<int> orig;                      // per draw: index in the caller's array
  std::vector<uint32_t> blockBase, tileBase;  // per draw, segment-relative
  std::vector<Seg> segs;
};

// Device byte ranges [lo, hi) a draw writes (its attachments, read and
// written) and only reads (vertex streams, indices, point sizes, texture
// levels).  Draws of one segment run concurrently, so a later draw joins the
// segment only if none of its ranges meets a range an earlier draw of the
// segment writes, and none of its writes meets an earlier draw's reads --
// the ordering of a batch must equal count swr_draw calls on one stream
// (draw tickets, Renderer.cpp:551-662), also when a texture is an earlier
// draw's attachment or two attachments overlap without sharing a base.
struct Extents {
  std::vector<std::pair<uintptr_t, uintptr_t>> writes, reads;
};

void add_range(std::vector<std::pair<uintptr_t, uintptr_t>> &v, const void *p, uint64_t bytes) {
  if (p && bytes) v.push_back({(uintptr_t)p, (uintptr_t)p + (uintptr_t)bytes});
}

uint64_t image_bytes(uint32_t w, uint32_t h, uint32_t samples, uint64_t rowPitch, uint64_t slice, uint32_t bpp) {
  return (uint64_t)(samples - 1) * slice + (uint64_t)(h - 1) * rowPitch + (uint64_t)w * bpp;
}

// An attachment's sample planes: one range each when they are apart (sample
// planes of several images may interleave), else one range for them all.
void add_image(std::vector<std::pair<uintptr_t, uintptr_t>> &v, const void *p, uint32_t w, uint32_t h,
               uint32_t samples, uint64_t rowPitch, uint64_t slice, uint32_t bpp) {
  const uint64_t plane = image_bytes(w, h, 1, rowPitch, 0, bpp);
  if (samples > 1 && slice > plane) {
    for (uint32_t q = 0; q < samples; q++) add_range(v, (const char *)p + q * slice, plane);
  } else {
    add_range(v, p, image_bytes(w, h, samples, rowPitch, slice, bpp));
  }
}

Extents extents_of(const swr_draw_desc *d, const DrawParams &P) {
  Extents e;
  const uint32_t W = d->framebuffer_width, H = d->framebuffer_height, S = d->samples;
  if (P.color) add_image(e.writes, P.color, W, H, S, P.colorPitch, P.colorSlice, P.colorFloat ? 16 : 4);
  if (P.depth) add_image(e.writes, P.depth, W, H, S, P.depthPitch, P.depthSlice, P.depth16 ? 2 : 4);
  if (P.stencil) add_image(e.writes, P.stencil, W, H, S, P.stencilPitch, P.stencilSlice, 1);
  if (P.primId) add_image(e.writes, P.primId, W, H, S, P.primIdPitch, P.primIdSlice, 4);
  const uint64_t V = (uint64_t)d->vertex_count;
  add_range(e.reads, P.pos, V * 16);
  add_range(e.reads, P.col, V * 16);
  add_range(e.reads, P.uv, V * 8);
  add_range(e.reads, P.psize, V * 4);
  if (P.idx) {
    const uint64_t T = (uint64_t)d->triangle_count;
    const uint64_t n = P.topology == 0 ? T * 3 : P.topology == 3 ? T * 2 : P.topology == 4 ? T + 1
                       : P.topology == 5 ? T : T + 2;
    add_range(e.reads, P.idx, n * (uint64_t)P.idxBytes);
  }
  if (P.useUV) {
    const uint32_t tb = P.tex.isFloat ? 16 : 4;
    for (uint32_t l = 0; l < d->texture.level_count && l < (uint32_t)kMipLevels; l++) {
      const swr_image_level &lv = d->texture.levels[l];
      add_range(e.reads, lv.data, image_bytes(lv.width, lv.height, 1, lv.row_pitch_bytes, 0, tb));
    }
  }
  return e;
}

bool ranges_meet(const std::vector<std::pair<uintptr_t, uintptr_t>> &a,
                 const std::vector<std::pair<uintptr_t, uintptr_t>> &b) {
  for (const auto &x : a)
    for (const auto &y : b)
      if (x.first < y.second && y.first < x.second) return true;
  return false;
}

// Whether draw `later` must wait for draw `earlier` of the same batch.
bool depends_on(const Extents &earlier, const Extents &later) {
  return ranges_meet(earlier.writes, later.writes) || ranges_meet(earlier.writes, later.reads) ||
         ranges_meet(earlier.reads, later.writes);
}

int layout_batch(const swr_draw_desc *descs, uint32_t n, const swr_tuning &tuning, BatchLayout &L) {
  L = BatchLayout{};
  std::vector<DrawParams> all(n);
  std::vector<Extents> ext(n);
  for (uint32_t i = 0; i < n; i++) {
    int r = build_params(&descs[i], tuning, all[i]);
    if (r != SWR_OK) return fail(r, "draw %u: %s", i, g_error.c_str());
    ext[i] = extents_of(&descs[i], all[i]);
  }
  uint32_t i = 0;
  while (i < n) {
    // Extend the segment while the next draw depends on none of its draws.
    uint32_t j = i + 1;
    while (j < n) {
      bool clash = false;
      for (uint32_t k = i; k < j && !clash; k++) clash = depends_on(ext[k], ext[j]);
      if (clash) break;
      j++;
    }
    std::vector<int> order;
    for (uint32_t k = i; k < j; k++)
      if (all[k].triCount > 0) order.push_back((int)k);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return variant_of(all[a]) < variant_of(all[b]); });
    BatchLayout::Seg sg;
    sg.first = (int)L.P.size();
    for (size_t k = 0; k < order.size(); k++) {
      DrawParams P = all[order[k]];
      P.triBase = (int)sg.T;
      L.blockBase.push_back(sg.blocks);
      L.tileBase.push_back(sg.tiles);
      sg.T += (uint32_t)P.triCount;
      sg.blocks += (uint32_t)(P.triCount + 255) / 256;
      sg.tiles += (uint32_t)(P.tilesX * P.tilesY);
      if (k == 0 || variant_of(P) != variant_of(L.P.back()))
        sg.variants.push_back({(int)k, (int)k + 1});
      else
        sg.variants.back().second = (int)k + 1;
      L.P.push_back(P);
      L.orig.push_back(order[k]);
    }
    sg.count = (int)order.size();
    if (sg.count) L.segs.push_back(sg);
    i = j;
  }
  if (L.P.size() > (1u << 20)) return fail(SWR_ERROR_UNSUPPORTED, "too many draws in one batch");
  for (const auto &sg : L.segs)
    if ((uint64_t)sg.T > (1u << 28) || (uint64_t)sg.tiles > (1u << 30))
      return fail(SWR_ERROR_UNSUPPORTED, "batch segment too large");
  return SWR_OK;
}

// Uploads the per-draw params and the block / tile bases of a layout into
// `pipe` (through the context's pinned staging buffer).
int upload_layout(DeviceCtx *c, Pipeline &pipe, const BatchLayout &L, hipStream_t s) {
  const size_t n = L.P.size();
  const size_t pb = sizeof(DrawParams) * n, ab = sizeof(uint32_t) * 2 * n;
  if (!n) return SWR_OK;
  SWR_HIP_CHECK(pipe.params.ensure(pb));
  SWR_HIP_CHECK(pipe.aux.ensure(ab));
  if (c->pinnedBytes < pb + ab) {
    SWR_HIP_CHECK(hipStreamSynchronize(s));
    if (c->pinned) SWR_HIP_CHECK(hipHostFree(c->pinned));
    c->pinned = nullptr;
    c->pinnedBytes = 0;
    SWR_HIP_CHECK(hipHostMalloc(&c->pinned, pb + ab, hipHostMallocDefault));
    c->pinnedBytes = pb + ab;
  }
  char *h = (char *)c->pinned;
  memcpy(h, L.P.data(), pb);
  memcpy(h + pb, L.blockBase.data(), sizeof(uint32_t) * n);
  memcpy(h + pb + sizeof(uint32_t) * n, L.tileBase.data(), sizeof(uint32_t) * n);
  SWR_HIP_CHECK(hipMemcpyAsync(pipe.params.ptr, h, pb, hipMemcpyHostToDevice, s));
  SWR_HIP_CHECK(hipMemcpyAsync(pipe.aux.ptr, h + pb, ab, hipMemcpyHostToDevice, s));
  // The staging buffer is reused by the next call: wait for the copies.
  SWR_HIP_CHECK(hipStreamSynchronize(s));
  return SWR_OK;
}

// Device words of Pipeline::counters: the persistent raster workgroups'
// tile counters in [0, 80) (one per raster variant: at most 40 variants),
// the eager draws' overflow word, the segment's split
// slot, the number of pairs the segment emitted and the backward's two work
// counters.
constexpr int kOverflowSlot = 80;
constexpr int kSplitSlot = 81;
constexpr int kEmittedSlot = 82;
constexpr int kBackwardSlot = 84;
constexpr int kCounterWords = 88;

// 8-bit radix passes that order every tile key below the key 0xFFFFFFFF of
// unused capacity and dropped pairs: 2^(8 passes) > tiles, so the sorted
// digits of a real key never all equal the sentinel's.
uint32_t sort_passes(uint32_t tiles) {
  int bits = 1;
  while (bits < 32 && (1ull << bits) <= (uint64_t)tiles) bits++;
  return (uint32_t)((bits + 7) / 8);
}

// Calls f(buffer, bytes) for every per-segment scratch buffer of `pipe`
// sized for T slots, `cap` pairs and `tiles` tiles.
template <typename F>
void pipe_buffers(Pipeline &pipe, uint32_t T, uint32_t cap, uint32_t tiles, F &&f) {
  const size_t slots = (size_t)T + 1;
  const size_t nb = ((size_t)cap + kSortItems - 1) / kSortItems;
  const size_t scanBlocks = std::max((slots + kScanItems - 1) / kScanItems, (256 * nb + kScanItems - 1) / kScanItems);
  const size_t pc = std::max<size_t>(cap, 1);
  f(pipe.counters, sizeof(uint32_t) * kCounterWords);
  f(pipe.rec, sizeof(TriRecord) * (size_t)T);
  f(pipe.poly, sizeof(int) * 2 * kMaxPolygon * (size_t)T);
  f(pipe.cnt, sizeof(uint32_t) * slots);
  f(pipe.off, sizeof(uint32_t) * slots);
  f(pipe.sums, sizeof(uint32_t) * std::max<size_t>(scanBlocks, 1));
  f(pipe.tStart, sizeof(uint32_t) * std::max<uint32_t>(tiles, 1));
  f(pipe.tEnd, sizeof(uint32_t) * std::max<uint32_t>(tiles, 1));
  f(pipe.keysA, sizeof(uint32_t) * pc);
  f(pipe.keysB, sizeof(uint32_t) * pc);
  f(pipe.valsA, sizeof(uint32_t) * pc);
  f(pipe.valsB, sizeof(uint32_t) * pc);
  f(pipe.hist, sizeof(uint32_t) * 256 * std::max<size_t>(nb, 1));
  f(pipe.first, sizeof(uint32_t) * ((size_t)cap / kEmitBlock + 2));
}

// Sizes the per-segment scratch of `pipe` for T slots, `cap` pairs, `tiles` tiles.
int ensure_pipe(Pipeline &pipe, uint32_t T, uint32_t cap, uint32_t tiles) {
  hipError_t err = hipSuccess;
  pipe_buffers(pipe, T, cap, tiles, [&](Scratch &b, size_t bytes) {
    if (err == hipSuccess) err = b.ensure(bytes);
  });
  SWR_HIP_CHECK(err);
  return SWR_OK;
}

// Exclusive prefix sum of n u32 (in may equal out): two launches up to
// kFusedScanBlocks blocks, three above.  With `first`, the emission index of
// pair capacity `capacity` is built as well (emit_index_kernel's work; in
// must then differ from out and n be the slot count + 1).
void enqueue_scan(const uint32_t *in, uint32_t *out, size_t n, uint32_t *sums, hipStream_t s,
                  uint32_t capacity = 0, uint32_t *first = nullptr, uint32_t *split = nullptr) {
  const uint32_t blocks = (uint32_t)((n + kScanItems - 1) / kScanItems);
  hipLaunchKernelGGL(scan_partial_kernel, dim3(blocks), dim3(256), 0, s, in, out, (uint32_t)n, sums);
  if (blocks <= kFusedScanBlocks) {
    hipLaunchKernelGGL(scan_finish_kernel, dim3(blocks), dim3(256), 0, s, out, (uint32_t)n, (const uint32_t *)sums,
                       in, capacity, first, split);
  } else {
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, s, sums, blocks);
    hipLaunchKernelGGL(scan_add_kernel, dim3(blocks), dim3(256), 0, s, out, (uint32_t)n, (const uint32_t *)sums);
    if (first) {
      const uint32_t T = (uint32_t)n - 1;  // >= kScanItems here
      hipLaunchKernelGGL(emit_index_kernel, dim3((T + 255) / 256), dim3(256), 0, s, (const uint32_t *)out, T,
                         capacity, first, split);
    }
  }
}

// One pass of the stable LSD radix sort of *count (key, value) pairs in
// nb blocks of kSortItems, on bits [shift, shift + 8) of the keys.
void enqueue_radix_pass(const uint32_t *kIn, const uint32_t *vIn, const uint32_t *count, int shift, uint32_t nb,
                        uint32_t *hist, uint32_t *sums, uint32_t *kOut, uint32_t *vOut, hipStream_t s) {
  hipLaunchKernelGGL(radix_hist_kernel, dim3(nb), dim3(256), 0, s, kIn, count, shift, hist, nb);
  const size_t n = (size_t)256 * nb;
  const uint32_t blocks = (uint32_t)((n + kScanItems - 1) / kScanItems);
  if (blocks <= kFusedScanBlocks) {
    hipLaunchKernelGGL(scan_partial_kernel, dim3(blocks), dim3(256), 0, s, (const uint32_t *)hist, hist, (uint32_t)n,
                       sums);
    hipLaunchKernelGGL(radix_scatter_kernel, dim3(nb), dim3(256), 0, s, kIn, vIn, count, shift,
                       (const uint32_t *)hist, nb, (const uint32_t *)sums, blocks, kOut, vOut);
  } else {
    enqueue_scan(hist, hist, n, sums, s);
    hipLaunchKernelGGL(radix_scatter_kernel, dim3(nb), dim3(256), 0, s, kIn, vIn, count, shift,
                       (const uint32_t *)hist, nb, (const uint32_t *)nullptr, 0u, kOut, vOut);
  }
}

// Arguments of one raster launch over tiles [tLo, tHi) of draws [lo, hi).
struct RasterArgs {
  const DrawParams *params;
  const uint32_t *tileBase;
  int lo, hi;
  uint32_t tLo, tHi;
  uint32_t *counter;
  const TriRecord *rec;
  const int *poly;
  const uint32_t *bin;  // NULL: the unbinned pass
  const uint32_t *ts, *te;
  const uint32_t *split;
  uint32_t segT;
  int tail;  // with bin: each tile also walks the slots from *split on
};

template <int MS, bool FCOL, bool FTEX, bool XF, int LEVEL>
void launch_raster_batch(dim3 grid, dim3 block, hipStream_t s, const RasterArgs &a) {
  hipLaunchKernelGGL((raster_batch_kernel<MS, FCOL, FTEX, XF, LEVEL>), grid, block, 0, s, a.params, a.tileBase, a.lo,
                     a.hi, a.tLo, a.tHi, a.counter, a.rec, a.poly, a.bin, a.ts, a.te, a.split, a.segT,
                     a.tail);
}

// The raster kernel of variant v (variant_of).
template <bool XF, int LEVEL>
void launch_raster_variant(int v, dim3 grid, dim3 block, hipStream_t s, const RasterArgs &a) {
  switch (v & 7) {
    case 0: launch_raster_batch<1, false, false, XF, LEVEL>(grid, block, s, a); break;
    case 1: launch_raster_batch<1, false, true, XF, LEVEL>(grid, block, s, a); break;
    case 2: launch_raster_batch<1, true, false, XF, LEVEL>(grid, block, s, a); break;
    case 3: launch_raster_batch<1, true, true, XF, LEVEL>(grid, block, s, a); break;
    case 4: launch_raster_batch<4, false, false, XF, LEVEL>(grid, block, s, a); break;
    case 5: launch_raster_batch<4, false, true, XF, LEVEL>(grid, block, s, a); break;
    case 6: launch_raster_batch<4, true, false, XF, LEVEL>(grid, block, s, a); break;
    default: launch_raster_batch<4, true, true, XF, LEVEL>(grid, block, s, a); break;
  }
}

// Setup and tile-count prefix sum of one segment (off[T] = pair count).
// With a pair capacity `cap` > 0 the scan also builds the emission index for
// it (enqueue_back is then told so); `zeroed`: cnt[T] is already zero.
void enqueue_front(Pipeline &pipe, const BatchLayout &L, const BatchLayout::Seg &sg, hipStream_t s,
                   uint32_t cap = 0, bool zeroed = false) {
  const DrawParams *params = pipe.params.as<DrawParams>() + sg.first;
  const uint32_t *blockBase = pipe.aux.as<uint32_t>() + sg.first;
  hipLaunchKernelGGL(setup_batch_kernel, dim3(sg.blocks), dim3(256), 0, s, params, blockBase, sg.count,
                     pipe.rec.as<TriRecord>(), pipe.poly.as<int>(), pipe.cnt.as<uint32_t>());
  if (!zeroed) hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(1), 0, s, pipe.cnt.as<uint32_t>() + sg.T, 0u, 1u);
  enqueue_scan(pipe.cnt.as<uint32_t>(), pipe.off.as<uint32_t>(), (size_t)sg.T + 1, pipe.sums.as<uint32_t>(), s, cap,
               cap ? pipe.first.as<uint32_t>() : nullptr,
               cap ? pipe.counters.as<uint32_t>() + kSplitSlot : nullptr);
}

// Pair emission, sort, ranges and raster of one segment with pair capacity
// `cap`.  When a run needs more pairs than `cap` (a plan whose geometry grew
// since it was built), the total needed goes to *overflowFlag and, with
// `unbinned`, the primitives whose pairs did not fit are rendered by each
// tile after its binned list (in the same raster launch), so that none is
// dropped.
// `indexed`: enqueue_front built the emission index for `cap`; `zeroed`:
// the tile ranges and the raster's work counters are already zero.
int enqueue_back(DeviceCtx *c, Pipeline &pipe, const BatchLayout &L, const BatchLayout::Seg &sg, uint32_t cap,
                 const swr_tuning &tuning, uint32_t *overflowFlag, bool unbinned, EventSet *es, hipStream_t s,
                 bool indexed = false, bool zeroed = false) {
  const DrawParams *params = pipe.params.as<DrawParams>() + sg.first;
  const uint32_t *tileBase = pipe.aux.as<uint32_t>() + L.P.size() + sg.first;
  uint32_t *split = pipe.counters.as<uint32_t>() + kSplitSlot;
  uint32_t *emitted = pipe.counters.as<uint32_t>() + kEmittedSlot;
  if (cap > 0) {
    if (!indexed) {
      if (sg.T > 0)
        hipLaunchKernelGGL(emit_index_kernel, dim3((sg.T + 255) / 256), dim3(256), 0, s,
                           (const uint32_t *)pipe.off.as<uint32_t>(), sg.T, cap, pipe.first.as<uint32_t>(), split);
      else
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(1), 0, s, split, 0u, 1u);
    }
    hipLaunchKernelGGL(bin_emit_batch_kernel, dim3((cap + kEmitBlock - 1) / kEmitBlock), dim3(kEmitBlock), 0, s, params,
                       tileBase, (const TriRecord *)pipe.rec.as<TriRecord>(), (const uint32_t *)pipe.off.as<uint32_t>(),
                       sg.T, (const uint32_t *)pipe.first.as<uint32_t>(), cap, pipe.keysA.as<uint32_t>(),
                       pipe.valsA.as<uint32_t>(), overflowFlag, (const uint32_t *)split, emitted);
    const uint32_t nb = (cap + kSortItems - 1) / kSortItems;
    const uint32_t passes = sort_passes(sg.tiles);
    uint32_t *kIn = pipe.keysA.as<uint32_t>(), *vIn = pipe.valsA.as<uint32_t>();
    uint32_t *kOut = pipe.keysB.as<uint32_t>(), *vOut = pipe.valsB.as<uint32_t>();
    for (uint32_t p = 0; p < passes; p++) {
      enqueue_radix_pass(kIn, vIn, emitted, (int)(8 * p), nb, pipe.hist.as<uint32_t>(), pipe.sums.as<uint32_t>(), kOut,
                         vOut, s);
      std::swap(kIn, kOut);
      std::swap(vIn, vOut);
    }
    const uint32_t tb = (sg.tiles + 255) / 256;
    if (!zeroed) {
      hipLaunchKernelGGL(fill_u32_kernel, dim3(tb), dim3(256), 0, s, pipe.tStart.as<uint32_t>(), 0u, sg.tiles);
      hipLaunchKernelGGL(fill_u32_kernel, dim3(tb), dim3(256), 0, s, pipe.tEnd.as<uint32_t>(), 0u, sg.tiles);
    }
    hipLaunchKernelGGL(tile_range_batch_kernel, dim3((cap + 255) / 256), dim3(256), 0, s, (const uint32_t *)kIn,
                       (const uint32_t *)emitted, sg.tiles, pipe.tStart.as<uint32_t>(), pipe.tEnd.as<uint32_t>());
    SWR_HIP_CHECK(hipGetLastError());
    if (es) SWR_HIP_CHECK(hipEventRecord(es->ev[EV_BIN], s));

    const int wpt = (int)((tuning.tile_width / 8) * (tuning.tile_height / 8));
    dim3 block(64 * wpt);
    uint32_t *counters = pipe.counters.as<uint32_t>();
    const bool persistent = tuning.raster_workgroups_per_cu != ~0u;
    const uint32_t nv = (uint32_t)sg.variants.size();
    if (persistent && !zeroed) hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, counters, 0u, nv);
    // One launch per variant; with `unbinned`, each tile renders its binned
    // list and then the slots left out of it.
    for (uint32_t v = 0; v < nv; v++) {
      const int lo = sg.variants[v].first, hi = sg.variants[v].second;
      const uint32_t tLo = L.tileBase[sg.first + lo];
      const uint32_t tHi = hi < sg.count ? L.tileBase[sg.first + hi] : sg.tiles;
      if (tHi <= tLo) continue;
      uint32_t wgs = tHi - tLo;
      if (persistent) {
        const uint32_t per = tuning.raster_workgroups_per_cu ? tuning.raster_workgroups_per_cu : 4u;
        wgs = std::min<uint32_t>(wgs, (uint32_t)c->cus * per);
      }
      RasterArgs a;
      a.params = params;
      a.tileBase = tileBase;
      a.lo = lo;
      a.hi = hi;
      a.tLo = tLo;
      a.tHi = tHi;
      a.counter = persistent ? counters + v : nullptr;
      a.rec = pipe.rec.as<TriRecord>();
      a.poly = pipe.poly.as<int>();
      a.bin = vIn;
      a.tail = unbinned ? 1 : 0;
      a.ts = pipe.tStart.as<uint32_t>();
      a.te = pipe.tEnd.as<uint32_t>();
      a.split = split;
      a.segT = sg.T;
      const int var = variant_of(L.P[sg.first + lo]);
      switch (var & 8 ? -1 : var >> 4) {
        case -1: launch_raster_variant<true, 0>(var, dim3(wgs), block, s, a); break;
        case 3: launch_raster_variant<false, 3>(var, dim3(wgs), block, s, a); break;
        case 2: launch_raster_variant<false, 2>(var, dim3(wgs), block, s, a); break;
        case 1: launch_raster_variant<false, 1>(var, dim3(wgs), block, s, a); break;
        default: launch_raster_variant<false, 0>(var, dim3(wgs), block, s, a); break;
      }
    }
  } else if (es) {
    SWR_HIP_CHECK(hipEventRecord(es->ev[EV_BIN], s));
  }
  SWR_HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int check_tuning(const swr_tuning &t) {
  if ((t.tile_width != 8 && t.tile_width != 16) || (t.tile_height != 8 && t.tile_height != 16))
    return fail(SWR_ERROR_UNSUPPORTED, "tile %ux%u (8 or 16 pixels per side)", t.tile_width, t.tile_height);
  if (t.raster_workgroups_per_cu != ~0u && t.raster_workgroups_per_cu > 64)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "raster_workgroups_per_cu %u", t.raster_workgroups_per_cu);
  return SWR_OK;
}

__global__ void store_params_kernel(DrawParams P, DrawParams *dst, uint32_t *aux, uint32_t *cntEnd, uint32_t *counters,
                                    uint32_t nCounters, uint32_t *ts, uint32_t *te, uint32_t tiles);

// store_params_kernel's launch for segment sg of pipe, which also clears
// what the front and back ends would otherwise clear with fills: cnt[T],
// the tile ranges and the raster's work counters.
void enqueue_store_params(const Pipeline &pipe, const DrawParams &P, const BatchLayout::Seg &sg, uint32_t nCounters,
                          hipStream_t s) {
  const uint32_t blocks = std::max<uint32_t>((sg.tiles + 255) / 256, 1u);
  hipLaunchKernelGGL(store_params_kernel, dim3(blocks), dim3(256), 0, s, P, pipe.params.as<DrawParams>(),
                     pipe.aux.as<uint32_t>(), pipe.cnt.as<uint32_t>() + sg.T, pipe.counters.as<uint32_t>(), nCounters,
                     pipe.tStart.as<uint32_t>(), pipe.tEnd.as<uint32_t>(), sg.tiles);
}

// A single draw with a known pair capacity: parameters written by a kernel
// (no staging buffer to wait for), pairs beyond the capacity rendered by
// the unbinned raster pass (no count to wait for); the draw's count is read
// back without a wait and grows the capacity for the draws after it.
int draw_nowait(DeviceCtx *c, const BatchLayout &L, const swr_tuning &tuning, hipStream_t s) {
  const BatchLayout::Seg &sg = L.segs[0];
  Pipeline &pipe = c->pipe;
  SWR_HIP_CHECK(pipe.params.ensure(sizeof(DrawParams)));
  SWR_HIP_CHECK(pipe.aux.ensure(sizeof(uint32_t) * 2));
  int r = ensure_pipe(pipe, sg.T, c->pairCap, sg.tiles);
  if (r != SWR_OK) return r;
  enqueue_store_params(pipe, L.P[0], sg, 1, s);
  enqueue_front(pipe, L, sg, s, c->pairCap, true);
  SWR_HIP_CHECK(hipGetLastError());
  r = enqueue_back(c, pipe, L, sg, c->pairCap, tuning, pipe.counters.as<uint32_t>() + kOverflowSlot, true, nullptr, s,
                   true, true);
  if (r != SWR_OK) return r;
  if (!c->pairsPending) {
    if (!c->pairsEvent) SWR_HIP_CHECK(hipEventCreateWithFlags(&c->pairsEvent, hipEventDisableTiming));
    SWR_HIP_CHECK(hipMemcpyAsync(c->hostTotal + 2, pipe.off.as<uint32_t>() + sg.T, sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, s));
    SWR_HIP_CHECK(hipEventRecord(c->pairsEvent, s));
    c->pairsPending = true;
  }
  return SWR_OK;
}

// The pair capacity for no-wait single draws after a draw needing `pairs`.
void grow_pair_cap(DeviceCtx *c, uint64_t pairs) {
  c->pairCap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(c->pairCap, pairs + pairs / 8 + 1024), 0xFFFFFF00ull);
}

// Eager draws: each segment reads its pair count back before sizing the
// pair buffers (one host wait per segment) -- except a single draw once a
// pair capacity is known (draw_nowait).
int draw_eager(DeviceCtx *c, const swr_draw_desc *descs, uint32_t n, hipStream_t s) {
  BatchLayout L;
  swr_tuning tuning = c->tuning;
  int r = layout_batch(descs, n, tuning, L);
  if (r != SWR_OK) return r;
  if (L.segs.empty()) return SWR_OK;
  const bool prof = c->profiling;
  if (c->pairsPending) {
    const hipError_t q = hipEventQuery(c->pairsEvent);
    if (q == hipSuccess) {
      grow_pair_cap(c, c->hostTotal[2]);
      c->pairsPending = false;
    } else {
      (void)hipGetLastError();  // not ready yet: no error to keep
    }
  }
  if (n == 1 && !prof && L.segs.size() == 1 && c->pairCap) return draw_nowait(c, L, tuning, s);
  if (prof) {
    SWR_HIP_CHECK(c->stats.ensure(4 * sizeof(unsigned long long)));
    for (DrawParams &P : L.P) P.stats = c->stats.as<unsigned long long>();
  }
  Pipeline &pipe = c->pipe;
  r = upload_layout(c, pipe, L, s);
  if (r != SWR_OK) return r;
  for (const auto &sg : L.segs) {
    EventSet es{};
    if (prof) {
      SWR_HIP_CHECK(take_events(c, es));
      SWR_HIP_CHECK(hipEventRecord(es.ev[EV_START], s));
    }
    r = ensure_pipe(pipe, sg.T, 0, sg.tiles);
    if (r != SWR_OK) return r;
    enqueue_front(pipe, L, sg, s);
    SWR_HIP_CHECK(hipGetLastError());
    if (prof) SWR_HIP_CHECK(hipEventRecord(es.ev[EV_SETUP], s));
    SWR_HIP_CHECK(hipMemcpyAsync(c->hostTotal, pipe.off.as<uint32_t>() + sg.T, sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, s));
    SWR_HIP_CHECK(hipStreamSynchronize(s));
    const uint32_t pairs = *c->hostTotal;
    if (n == 1) grow_pair_cap(c, pairs);
    r = ensure_pipe(pipe, sg.T, pairs, sg.tiles);
    if (r != SWR_OK) return r;
    r = enqueue_back(c, pipe, L, sg, pairs, tuning, pipe.counters.as<uint32_t>() + kOverflowSlot, false,
                     prof ? &es : nullptr, s);
    if (r != SWR_OK) return r;
    if (prof) {
      SWR_HIP_CHECK(hipEventRecord(es.ev[EV_RASTER], s));
      es.pairs = pairs;
      es.draws = (uint64_t)sg.count;
      c->pending.push_back(es);
    }
  }
  return SWR_OK;
}

// ---- swr_draw_noalloc: one draw whose scratch is a fixed device array ----
// Callers that cannot allocate device memory render through `scratch`, a
// file-scope device array, instead of the grow-on-demand pipeline buffers.
// Host code never uses the array's name as an address (on the host it is a
// stand-in, not device memory): each device's copy is located once with
// hipGetSymbolAddress, with that device current, and its address kept in
// g_static.  The pair capacity is whatever the array leaves after the
// per-primitive buffers; pairs beyond it go through the unbinned raster
// pass, so no host wait is needed and nothing is allocated per call.
constexpr size_t kStaticScratchBytes = SWR_NOALLOC_SCRATCH_BYTES;
constexpr uint32_t kStaticMinPairs = 65536;  // smallest pair capacity a draw may be left with

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

struct StaticDev {
  char *base = nullptr;  // this device's copy of `scratch`
  int cus = 0;
};
StaticDev g_static[64];  // guarded by g_mutex

// Locates device `dev`'s copy of the array (once); `dev` must be current.
int static_setup(int dev) {
  if (dev < 0 || dev >= 64) return fail(SWR_ERROR_DEVICE, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_mutex);
  StaticDev &sd = g_static[dev];
  if (sd.base) return SWR_OK;
  void *p = nullptr;
  SWR_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(SWR_ERROR_DEVICE, "no device address for the static scratch array");
  int cus = 0;
  SWR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  sd.cus = cus > 0 ? cus : 256;
  sd.base = (char *)p;
  return SWR_OK;
}

// Places the buffers of one segment (n draws, T slots, `cap` pairs, `tiles`
// tiles) at `base`, 256-byte aligned; returns the bytes they span (`base`
// NULL: sizes only).
size_t place_pipe(char *base, Pipeline &pipe, uint32_t n, uint32_t T, uint32_t cap, uint32_t tiles) {
  size_t at = 0;
  auto put = [&](Scratch &b, size_t bytes) {
    at = (at + 255) & ~size_t(255);
    b.ptr = base ? base + at : nullptr;
    b.bytes = bytes;
    at += bytes;
  };
  put(pipe.params, sizeof(DrawParams) * n);
  put(pipe.aux, sizeof(uint32_t) * 2 * n);
  pipe_buffers(pipe, T, cap, tiles, put);
  return at;
}

// Writes a one-draw segment's parameters and bases into the scratch (a
// kernel, so that nothing goes through host memory).
static_assert(sizeof(DrawParams) % 4 == 0 && sizeof(DrawParams) + 32 <= 4096, "DrawParams as a kernel argument");
__global__ void store_params_kernel(DrawParams P, DrawParams *dst, uint32_t *aux, uint32_t *cntEnd, uint32_t *counters,
                                    uint32_t nCounters, uint32_t *ts, uint32_t *te, uint32_t tiles) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < tiles) {
    ts[t] = 0u;
    te[t] = 0u;
  }
  if (blockIdx.x != 0) return;
  const uint32_t *src = reinterpret_cast<const uint32_t *>(&P);
  uint32_t *d = reinterpret_cast<uint32_t *>(dst);
  for (uint32_t i = threadIdx.x; i < sizeof(DrawParams) / 4; i += blockDim.x) d[i] = src[i];
  if (threadIdx.x < nCounters) counters[threadIdx.x] = 0u;
  if (threadIdx.x == 0) {
    aux[0] = 0u;  // blockBase
    aux[1] = 0u;  // tileBase
    *cntEnd = 0u;
  }
}

int draw_noalloc(const swr_draw_desc *desc, hipStream_t s) {
  int dev = 0;
  SWR_HIP_CHECK(hipGetDevice(&dev));
  int r = static_setup(dev);
  if (r != SWR_OK) return r;
  swr_tuning tuning{};
  tuning.tile_width = 16;
  tuning.tile_height = 16;
  tuning.raster_workgroups_per_cu = ~0u;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_ctx[dev]) tuning = g_ctx[dev]->tuning;  // as swr_set_tuning left it for swr_draw
  }
  BatchLayout L;
  r = layout_batch(desc, 1, tuning, L);
  if (r != SWR_OK) return r;
  if (L.segs.empty()) return SWR_OK;
  const BatchLayout::Seg &sg = L.segs[0];
  Pipeline pipe;
  const size_t fixed = place_pipe(nullptr, pipe, 1, sg.T, 0, sg.tiles);
  uint32_t cap = 0;
  if (fixed < kStaticScratchBytes) {
    uint64_t lo = 0, hi = (kStaticScratchBytes - fixed) / 16 + 1;  // place_pipe(cap = lo) fits, hi does not
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) / 2;
      if (place_pipe(nullptr, pipe, 1, sg.T, (uint32_t)mid, sg.tiles) <= kStaticScratchBytes) lo = mid;
      else hi = mid;
    }
    cap = (uint32_t)std::min<uint64_t>(lo, 0xFFFFFF00ull);
  }
  if (cap < kStaticMinPairs)
    return fail(SWR_ERROR_UNSUPPORTED,
                "swr_draw_noalloc: %u primitives and %u tiles leave room for %u (tile, primitive) pairs in the "
                "%zu-byte scratch (at least %u needed)",
                sg.T, sg.tiles, cap, kStaticScratchBytes, kStaticMinPairs);
  const StaticDev &sd = g_static[dev];
  place_pipe(sd.base, pipe, 1, sg.T, cap, sg.tiles);
  enqueue_store_params(pipe, L.P[0], sg, 1, s);
  SWR_HIP_CHECK(hipGetLastError());
  enqueue_front(pipe, L, sg, s, cap, true);
  SWR_HIP_CHECK(hipGetLastError());
  DeviceCtx dc;  // enqueue_back reads only the CU count from it
  dc.device = dev;
  dc.cus = sd.cus;
  return enqueue_back(&dc, pipe, L, sg, cap, tuning, pipe.counters.as<uint32_t>() + kOverflowSlot, true, nullptr, s,
                      true, true);
}

// Validates a gradient descriptor against its draw and derives GradParams.
int grad_params(const swr_draw_desc *desc, const DrawParams &P, const swr_grad_desc *g, GradParams &G) {
  G = GradParams{};
  if (!g) return fail(SWR_ERROR_INVALID_ARGUMENT, "grad descriptor is NULL");
  if (!P.primId) return fail(SWR_ERROR_INVALID_ARGUMENT, "backward needs the draw's primitive_id attachment");
  if (P.line || P.point) return fail(SWR_ERROR_UNSUPPORTED, "backward of lines and points");
  if (P.useUV && (P.tex.addrU > 1 || P.tex.addrV > 1))
    return fail(SWR_ERROR_UNSUPPORTED, "backward of mirrored address modes");
  if (P.depthBounds) return fail(SWR_ERROR_UNSUPPORTED, "backward of a draw with the depth bounds test");
  if (P.alphaToCoverage) return fail(SWR_ERROR_UNSUPPORTED, "backward of a draw with alpha-to-coverage");
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
  return SWR_OK;
}

int draw_backward(DeviceCtx *c, const swr_draw_desc *desc, const swr_grad_desc *g, hipStream_t s) {
  DrawParams P;
  int r = build_params(desc, c->tuning, P);
  if (r != SWR_OK) return r;
  GradParams G;
  r = grad_params(desc, P, g, G);
  if (r != SWR_OK) return r;
  const uint32_t T = desc->triangle_count;
  if (T == 0) return SWR_OK;
  Pipeline &pipe = c->pipe;
  SWR_HIP_CHECK(pipe.rec.ensure(sizeof(TriRecord) * (size_t)T));
  SWR_HIP_CHECK(pipe.poly.ensure(sizeof(int) * 2 * kMaxPolygon * (size_t)T));
  SWR_HIP_CHECK(pipe.cnt.ensure(sizeof(uint32_t) * ((size_t)T + 1)));
  hipLaunchKernelGGL(setup_kernel, dim3((T + 255) / 256), dim3(256), 0, s, P, pipe.rec.as<TriRecord>(),
                     pipe.poly.as<int>(), pipe.cnt.as<uint32_t>());
  dim3 grid((P.width + 15) / 16, (P.height + 15) / 16), block(256);
  if (P.ms == 1)
    hipLaunchKernelGGL(backward_kernel<1>, grid, block, 0, s, P, G, (const TriRecord *)pipe.rec.as<TriRecord>());
  else
    hipLaunchKernelGGL(backward_kernel<4>, grid, block, 0, s, P, G, (const TriRecord *)pipe.rec.as<TriRecord>());
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

struct swr_batch_plan {
  int device = -1;
  std::vector<swr_draw_desc> descs;
  uint32_t headroom = 10;
  swr_tuning tuning{};
  BatchLayout L;
  Pipeline pipe;
  Scratch flags;  // per segment: pairs needed when a run overflowed
  // Backward: per draw (layout order) gradient targets and 16x16-block bases.
  std::vector<swr_grad_desc> gradDescs;  // caller order
  Scratch gparams, gbase;
  std::vector<uint32_t> bwBase;
  bool hasGrads = false;
  // Binned texel scatter, per segment (see BinArgs): whether it is used,
  // its bucket count, record capacity and offsets into the plan's tables.
  struct SegBins {
    bool on = false;
    uint32_t nBuckets = 0, cap = 0, groupOff = 0, bucketOff = 0, drawOff = 0;
  };
  std::vector<SegBins> bins;
  Scratch binGroups, binBuckets, binDrawGroup;                  // tables of every segment
  Scratch binRec, binKeyA, binKeyB, binValA, binValB, binHist;  // record slots and their sort (largest segment)
  Scratch binStart, binEnd, binChunks, binChunkBase, binSums;   // per bucket (largest segment)
  Scratch binCounters;                                          // slots reserved, scatter work counter
  void releaseBins() {
    for (Scratch *b : {&binGroups, &binBuckets, &binDrawGroup, &binRec, &binKeyA, &binKeyB, &binValA, &binValB,
                       &binHist, &binStart, &binEnd, &binChunks, &binChunkBase, &binSums, &binCounters})
      b->release();
  }
};

namespace {

// (Re)builds a plan for its current tuning: layout, upload, one setup pass
// per segment to size the pair buffers.
int plan_build(DeviceCtx *c, swr_batch_plan *pl, hipStream_t s) {
  int r = layout_batch(pl->descs.data(), (uint32_t)pl->descs.size(), pl->tuning, pl->L);
  if (r != SWR_OK) return r;
  r = upload_layout(c, pl->pipe, pl->L, s);
  if (r != SWR_OK) return r;
  uint32_t maxT = 0, maxCap = 0, maxTiles = 0;
  for (auto &sg : pl->L.segs) {
    r = ensure_pipe(pl->pipe, sg.T, 0, sg.tiles);
    if (r != SWR_OK) return r;
    enqueue_front(pl->pipe, pl->L, sg, s);
    SWR_HIP_CHECK(hipGetLastError());
    SWR_HIP_CHECK(hipMemcpyAsync(c->hostTotal, pl->pipe.off.as<uint32_t>() + sg.T, sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, s));
    SWR_HIP_CHECK(hipStreamSynchronize(s));
    const uint64_t need = *c->hostTotal;
    const uint64_t cap = need + need * pl->headroom / 100 + 256;
    if (cap > 0xFFFFFF00ull) return fail(SWR_ERROR_UNSUPPORTED, "batch segment needs %llu pairs", (unsigned long long)cap);
    sg.pairCap = (uint32_t)cap;
    maxT = std::max(maxT, sg.T);
    maxCap = std::max(maxCap, sg.pairCap);
    maxTiles = std::max(maxTiles, sg.tiles);
  }
  r = ensure_pipe(pl->pipe, maxT, maxCap, maxTiles);
  if (r != SWR_OK) return r;
  SWR_HIP_CHECK(pl->flags.ensure(sizeof(uint32_t) * std::max<size_t>(pl->L.segs.size(), 1)));
  SWR_HIP_CHECK(hipMemsetAsync(pl->flags.ptr, 0, sizeof(uint32_t) * std::max<size_t>(pl->L.segs.size(), 1), s));
  SWR_HIP_CHECK(hipStreamSynchronize(s));
  return SWR_OK;
}

int plan_enqueue(DeviceCtx *c, swr_batch_plan *pl, hipStream_t s) {
  for (size_t i = 0; i < pl->L.segs.size(); i++) {
    const auto &sg = pl->L.segs[i];
    enqueue_front(pl->pipe, pl->L, sg, s, sg.pairCap);
    int r = enqueue_back(c, pl->pipe, pl->L, sg, sg.pairCap, pl->tuning, pl->flags.as<uint32_t>() + i, true, nullptr, s,
                         sg.pairCap > 0);
    if (r != SWR_OK) return r;
  }
  SWR_HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int plan_upload_grads(DeviceCtx *c, swr_batch_plan *pl, hipStream_t s);

int plan_ctx(swr_batch_plan *pl, DeviceCtx **c) {
  if (!pl) return fail(SWR_ERROR_INVALID_ARGUMENT, "plan is NULL");
  int dev = 0;
  SWR_HIP_CHECK(hipGetDevice(&dev));
  if (dev != pl->device) return fail(SWR_ERROR_INVALID_ARGUMENT, "plan belongs to device %d (current %d)", pl->device, dev);
  return current_ctx(c);
}

}  // namespace

extern "C" {

int swr_batch_plan_create(const swr_draw_desc *descs, uint32_t count, uint32_t pair_headroom_percent,
                          swr_batch_plan **plan, void *stream) {
  if (!plan) return fail(SWR_ERROR_INVALID_ARGUMENT, "plan is NULL");
  *plan = nullptr;
  if (count && !descs) return fail(SWR_ERROR_INVALID_ARGUMENT, "descs is NULL");
  if (pair_headroom_percent > 1000) return fail(SWR_ERROR_INVALID_ARGUMENT, "headroom above 1000%%");
  DeviceCtx *c = nullptr;
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  swr_batch_plan *pl = new swr_batch_plan();
  pl->device = c->device;
  pl->descs.assign(descs, descs + count);
  pl->headroom = pair_headroom_percent;
  pl->tuning = c->tuning;
  r = plan_build(c, pl, reinterpret_cast<hipStream_t>(stream));
  if (r != SWR_OK) {
    pl->pipe.release();
    pl->flags.release();
    delete pl;
    return r;
  }
  *plan = pl;
  return SWR_OK;
}

int swr_batch_plan_run(swr_batch_plan *plan, void *stream) {
  DeviceCtx *c = nullptr;
  int r = plan_ctx(plan, &c);
  if (r != SWR_OK) return r;
  return plan_enqueue(c, plan, reinterpret_cast<hipStream_t>(stream));
}

int swr_batch_plan_check(swr_batch_plan *plan, uint32_t *pairs_needed) {
  DeviceCtx *c = nullptr;
  int r = plan_ctx(plan, &c);
  if (r != SWR_OK) return r;
  SWR_HIP_CHECK(hipDeviceSynchronize());
  std::vector<uint32_t> f(std::max<size_t>(plan->L.segs.size(), 1), 0);
  SWR_HIP_CHECK(hipMemcpy(f.data(), plan->flags.ptr, sizeof(uint32_t) * f.size(), hipMemcpyDeviceToHost));
  uint32_t m = 0;
  for (uint32_t v : f) m = std::max(m, v);
  if (pairs_needed) *pairs_needed = m;
  return SWR_OK;
}

int swr_batch_plan_autotune(swr_batch_plan *plan, swr_tuning *best, void *stream) {
  DeviceCtx *c = nullptr;
  int r = plan_ctx(plan, &c);
  if (r != SWR_OK) return r;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const uint32_t kPerTile = ~0u;
  struct Cand { uint32_t w, h, wg; };
  const Cand cands[] = {{16, 16, kPerTile}, {16, 16, 2}, {16, 16, 3}, {16, 16, 4}, {16, 8, kPerTile}, {16, 8, 4},
                        {16, 8, 6},         {8, 8, kPerTile}, {8, 8, 8},  {8, 8, 12}};
  hipEvent_t e0, e1;
  SWR_HIP_CHECK(hipEventCreate(&e0));
  SWR_HIP_CHECK(hipEventCreate(&e1));
  float bestMs = 1e30f;
  swr_tuning bestT = plan->tuning;
  for (const Cand &cd : cands) {
    swr_tuning t{};
    t.tile_width = cd.w;
    t.tile_height = cd.h;
    t.raster_workgroups_per_cu = cd.wg;
    plan->tuning = t;
    r = plan_build(c, plan, s);
    if (r == SWR_OK) r = plan_enqueue(c, plan, s);  // warm-up
    float ms = 0.0f;
    if (r == SWR_OK) {
      SWR_HIP_CHECK(hipEventRecord(e0, s));
      for (int k = 0; k < 3 && r == SWR_OK; k++) r = plan_enqueue(c, plan, s);
      SWR_HIP_CHECK(hipEventRecord(e1, s));
      SWR_HIP_CHECK(hipEventSynchronize(e1));
      SWR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    }
    if (r != SWR_OK) break;
    if (ms < bestMs) {
      bestMs = ms;
      bestT = t;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (r != SWR_OK) return r;
  plan->tuning = bestT;
  r = plan_build(c, plan, s);
  if (r != SWR_OK) return r;
  if (plan->hasGrads) {
    r = plan_upload_grads(c, plan, s);
    if (r != SWR_OK) return r;
  }
  if (best) *best = bestT;
  return SWR_OK;
}

}  // extern "C"

namespace {

// Uploads the gradient targets of a plan in its layout order.
// Draws whose texel gradients scatter alike: the same gradient buffers and
// the texture parameters the taps depend on (level sizes and pitches,
// filter, mip mode, address modes).
bool same_tap_target(const DrawParams &A, const GradParams &Ga, const DrawParams &B, const GradParams &Gb) {
  const TexParams &a = A.tex, &b = B.tex;
  if (a.filter != b.filter || a.mipmap != b.mipmap || a.addrU != b.addrU || a.addrV != b.addrV) return false;
  for (int l = 0; l < kMipLevels; l++)
    if (Ga.dTex[l] != Gb.dTex[l] || a.width[l] != b.width[l] || a.height[l] != b.height[l] ||
        a.pitchP[l] != b.pitchP[l])
      return false;
  return true;
}

// Per segment: the texture groups of its draws, their bucket tables and the
// record slots: one per sample of every draw with texture gradients (a
// bound on its fragments) plus the last reservation of every wave the
// gather can launch (a segment with more groups than kMaxBinGroups, or more
// slots than 32-bit indices hold, keeps the texel tables).
int plan_build_bins(swr_batch_plan *pl, const std::vector<GradParams> &G, int cus, hipStream_t s) {
  pl->bins.assign(pl->L.segs.size(), swr_batch_plan::SegBins{});
  std::vector<BinGroup> groups;
  std::vector<BinBucket> buckets;
  std::vector<int8_t> drawGroup;
  uint64_t maxCap = 0, maxBuckets = 0;
  for (size_t si = 0; si < pl->L.segs.size(); si++) {
    const auto &sg = pl->L.segs[si];
    auto &sb = pl->bins[si];
    std::vector<int> rep;  // a draw (segment-relative) per group
    std::vector<int8_t> dg(sg.count, -1);
    uint64_t cap = 0;
    bool ok = true;
    for (int k = 0; k < sg.count && ok; k++) {
      const int i = sg.first + k;
      const DrawParams &P = pl->L.P[i];
      if (!P.useUV || !G[i].dTex[0]) continue;
      int gi = -1;
      for (size_t j = 0; j < rep.size() && gi < 0; j++)
        if (same_tap_target(pl->L.P[sg.first + rep[j]], G[sg.first + rep[j]], P, G[i])) gi = (int)j;
      if (gi < 0) {
        if ((int)rep.size() >= kMaxBinGroups) ok = false;
        gi = (int)rep.size();
        rep.push_back(k);
      }
      dg[k] = (int8_t)gi;
      cap += (uint64_t)P.width * (uint64_t)P.height * (uint64_t)P.ms;
    }
    // Up to two gather launches (one per sample count) of cus * 2 workgroups of 4 waves.
    cap += 2ull * (uint64_t)cus * 2u * 4u * kBinReserve;
    if (!ok || rep.empty() || cap >= (1ull << 31)) continue;
    sb.on = true;
    sb.cap = (uint32_t)cap;
    sb.groupOff = (uint32_t)groups.size();
    sb.bucketOff = (uint32_t)buckets.size();
    sb.drawOff = (uint32_t)drawGroup.size();
    uint32_t nb = 0;
    for (size_t j = 0; j < rep.size(); j++) {
      const TexParams &T = pl->L.P[sg.first + rep[j]].tex;
      BinGroup bg{};
      bg.draw = rep[j];
      for (int l = 0; l < kMipLevels; l++) {
        const uint32_t tx = ((uint32_t)T.width[l] + kBinTile - 1) / kBinTile;
        const uint32_t ty = ((uint32_t)T.height[l] + kBinTile - 1) / kBinTile;
        bg.levelBase[l] = nb;
        bg.tilesX[l] = tx;
        for (uint32_t y = 0; y < ty; y++)
          for (uint32_t x = 0; x < tx; x++)
            buckets.push_back(BinBucket{(uint8_t)j, (uint8_t)l, (uint16_t)x, (uint16_t)y});
        nb += tx * ty;
      }
      groups.push_back(bg);
    }
    sb.nBuckets = nb;
    drawGroup.insert(drawGroup.end(), dg.begin(), dg.end());
    maxCap = std::max<uint64_t>(maxCap, cap);
    maxBuckets = std::max<uint64_t>(maxBuckets, nb);
  }
  if (groups.empty()) return SWR_OK;
  SWR_HIP_CHECK(pl->binGroups.ensure(sizeof(BinGroup) * groups.size()));
  SWR_HIP_CHECK(pl->binBuckets.ensure(sizeof(BinBucket) * buckets.size()));
  SWR_HIP_CHECK(pl->binDrawGroup.ensure(drawGroup.size()));
  SWR_HIP_CHECK(pl->binRec.ensure(sizeof(TapRecord) * maxCap));
  for (Scratch *b : {&pl->binKeyA, &pl->binKeyB, &pl->binValA, &pl->binValB})
    SWR_HIP_CHECK(b->ensure(sizeof(uint32_t) * maxCap));
  const size_t nbSort = (maxCap + kSortItems - 1) / kSortItems;
  SWR_HIP_CHECK(pl->binHist.ensure(sizeof(uint32_t) * 256 * nbSort));
  const size_t nb1 = maxBuckets + 1;
  SWR_HIP_CHECK(pl->binStart.ensure(sizeof(uint32_t) * nb1));
  SWR_HIP_CHECK(pl->binEnd.ensure(sizeof(uint32_t) * nb1));
  SWR_HIP_CHECK(pl->binChunks.ensure(sizeof(uint32_t) * nb1));
  SWR_HIP_CHECK(pl->binChunkBase.ensure(sizeof(uint32_t) * nb1));
  const size_t scanItems = std::max(nb1, 256 * nbSort);
  SWR_HIP_CHECK(pl->binSums.ensure(sizeof(uint32_t) * ((scanItems + kScanItems - 1) / kScanItems + 1)));
  SWR_HIP_CHECK(pl->binCounters.ensure(sizeof(uint32_t) * 2));
  SWR_HIP_CHECK(hipMemcpyAsync(pl->binGroups.ptr, groups.data(), sizeof(BinGroup) * groups.size(),
                               hipMemcpyHostToDevice, s));
  SWR_HIP_CHECK(hipMemcpyAsync(pl->binBuckets.ptr, buckets.data(), sizeof(BinBucket) * buckets.size(),
                               hipMemcpyHostToDevice, s));
  SWR_HIP_CHECK(hipMemcpyAsync(pl->binDrawGroup.ptr, drawGroup.data(), drawGroup.size(), hipMemcpyHostToDevice, s));
  SWR_HIP_CHECK(hipStreamSynchronize(s));
  return SWR_OK;
}

int plan_upload_grads(DeviceCtx *c, swr_batch_plan *pl, hipStream_t s) {
  const size_t n = pl->L.P.size();
  std::vector<GradParams> G(n);
  pl->bwBase.assign(n, 0);
  for (const auto &sg : pl->L.segs) {
    uint32_t base = 0;
    for (int k = 0; k < sg.count; k++) {
      const int i = sg.first + k;
      const DrawParams &P = pl->L.P[i];
      const int o = pl->L.orig[i];
      int r = grad_params(&pl->descs[o], P, &pl->gradDescs[o], G[i]);
      if (r != SWR_OK) return fail(r, "draw %d: %s", o, g_error.c_str());
      pl->bwBase[i] = base;
      base += (uint32_t)(((P.width + 15) / 16) * ((P.height + 15) / 16));
    }
  }
  if (!n) return SWR_OK;
  SWR_HIP_CHECK(pl->gparams.ensure(sizeof(GradParams) * n));
  SWR_HIP_CHECK(pl->gbase.ensure(sizeof(uint32_t) * n));
  SWR_HIP_CHECK(hipMemcpyAsync(pl->gparams.ptr, G.data(), sizeof(GradParams) * n, hipMemcpyHostToDevice, s));
  SWR_HIP_CHECK(hipMemcpyAsync(pl->gbase.ptr, pl->bwBase.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
  SWR_HIP_CHECK(hipStreamSynchronize(s));
  return plan_build_bins(pl, G, c->cus, s);
}

int plan_enqueue_backward(swr_batch_plan *pl, int cus, hipStream_t s) {
  Pipeline &pipe = pl->pipe;
  for (size_t si = 0; si < pl->L.segs.size(); si++) {
    const auto &sg = pl->L.segs[si];
    const auto &sb = pl->bins.size() > si ? pl->bins[si] : swr_batch_plan::SegBins{};
    const DrawParams *params = pipe.params.as<DrawParams>() + sg.first;
    const uint32_t *blockBase = pipe.aux.as<uint32_t>() + sg.first;
    hipLaunchKernelGGL(setup_batch_kernel, dim3(sg.blocks), dim3(256), 0, s, params, blockBase, sg.count,
                       pipe.rec.as<TriRecord>(), pipe.poly.as<int>(), pipe.cnt.as<uint32_t>());
    const GradParams *grads = pl->gparams.as<GradParams>() + sg.first;
    const uint32_t *bb = pl->gbase.as<uint32_t>() + sg.first;
    BinArgs bin{};
    uint32_t *binCtr = pl->binCounters.as<uint32_t>();
    if (sb.on) {
      bin.rec = pl->binRec.as<TapRecord>();
      bin.key = pl->binKeyA.as<uint32_t>();
      bin.val = pl->binValA.as<uint32_t>();
      bin.reserved = binCtr;
      bin.cap = sb.cap;
      bin.drawGroup = pl->binDrawGroup.as<int8_t>() + sb.drawOff;
      bin.groups = pl->binGroups.as<BinGroup>() + sb.groupOff;
      hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, binCtr, 0u, 2u);
    }
    // Draws of one sample count are contiguous (the layout groups variants);
    // each group gets its own work counter (counters[kBackwardSlot + group]).
    uint32_t *ctr = pipe.counters.as<uint32_t>() + kBackwardSlot;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, ctr, 0u, 2u);
    int group = 0;
    int lo = 0;
    while (lo < sg.count) {
      const int ms = pl->L.P[sg.first + lo].ms;
      int hi = lo + 1;
      while (hi < sg.count && pl->L.P[sg.first + hi].ms == ms) hi++;
      const uint32_t bLo = pl->bwBase[sg.first + lo];
      const DrawParams &Ph = pl->L.P[sg.first + hi - 1];
      const uint32_t bHi = pl->bwBase[sg.first + hi - 1] + (uint32_t)(((Ph.width + 15) / 16) * ((Ph.height + 15) / 16));
      if (bHi > bLo) {
        const uint32_t wgs = std::min<uint32_t>(bHi - bLo, (uint32_t)cus * 2u);
        const TriRecord *rec = pipe.rec.as<TriRecord>();
        if (ms == 1 && sb.on)
          hipLaunchKernelGGL((backward_batch_kernel<1, true>), dim3(wgs), dim3(256), 0, s, params, grads, bb, lo, hi,
                             bLo, bHi, ctr + group, rec, bin);
        else if (ms == 1)
          hipLaunchKernelGGL((backward_batch_kernel<1, false>), dim3(wgs), dim3(256), 0, s, params, grads, bb, lo, hi,
                             bLo, bHi, ctr + group, rec, bin);
        else if (sb.on)
          hipLaunchKernelGGL((backward_batch_kernel<4, true>), dim3(wgs), dim3(256), 0, s, params, grads, bb, lo, hi,
                             bLo, bHi, ctr + group, rec, bin);
        else
          hipLaunchKernelGGL((backward_batch_kernel<4, false>), dim3(wgs), dim3(256), 0, s, params, grads, bb, lo, hi,
                             bLo, bHi, ctr + group, rec, bin);
      }
      group++;
      lo = hi;
    }
    if (sb.on) {
      // The reserved slots sorted by bucket (stable radix sort of (key,
      // slot); unused slots carry a key above every bucket and sort last),
      // each bucket's range of the sorted slots, the work items, the scatter.
      const uint32_t nbSort = (sb.cap + kSortItems - 1) / kSortItems;
      uint32_t *kIn = pl->binKeyA.as<uint32_t>(), *vIn = pl->binValA.as<uint32_t>();
      uint32_t *kOut = pl->binKeyB.as<uint32_t>(), *vOut = pl->binValB.as<uint32_t>();
      uint32_t *hist = pl->binHist.as<uint32_t>(), *sums = pl->binSums.as<uint32_t>();
      const uint32_t *count = binCtr;
      for (uint32_t p = 0; p < sort_passes(sb.nBuckets); p++) {
        enqueue_radix_pass(kIn, vIn, count, (int)(8 * p), nbSort, hist, sums, kOut, vOut, s);
        std::swap(kIn, kOut);
        std::swap(vIn, vOut);
      }
      const uint32_t nb1 = sb.nBuckets + 1;
      uint32_t *start = pl->binStart.as<uint32_t>(), *end = pl->binEnd.as<uint32_t>();
      uint32_t *chunks = pl->binChunks.as<uint32_t>(), *chunkBase = pl->binChunkBase.as<uint32_t>();
      hipLaunchKernelGGL(fill_u32_kernel, dim3((nb1 + 255) / 256), dim3(256), 0, s, start, 0u, nb1);
      hipLaunchKernelGGL(fill_u32_kernel, dim3((nb1 + 255) / 256), dim3(256), 0, s, end, 0u, nb1);
      hipLaunchKernelGGL(tile_range_batch_kernel, dim3((sb.cap + 255) / 256), dim3(256), 0, s, (const uint32_t *)kIn,
                         count, sb.nBuckets, start, end);
      hipLaunchKernelGGL(tap_chunks_kernel, dim3((nb1 + 255) / 256), dim3(256), 0, s, (const uint32_t *)start,
                         (const uint32_t *)end, sb.nBuckets, chunks);
      enqueue_scan(chunks, chunkBase, nb1, sums, s);
      hipLaunchKernelGGL(tap_scatter_kernel, dim3((uint32_t)cus * 2u), dim3(kScatterThreads), 0, s,
                         (const TapRecord *)bin.rec, (const uint32_t *)vIn, (const uint32_t *)start,
                         (const uint32_t *)end, (const uint32_t *)chunkBase, sb.nBuckets,
                         (const BinBucket *)pl->binBuckets.as<BinBucket>() + sb.bucketOff, bin.groups, params, grads,
                         binCtr + 1);
    }
  }
  SWR_HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

}  // namespace

extern "C" {

int swr_batch_plan_set_grads(swr_batch_plan *plan, const swr_grad_desc *grads, uint32_t count, void *stream) {
  DeviceCtx *c = nullptr;
  int r = plan_ctx(plan, &c);
  if (r != SWR_OK) return r;
  if (!grads || count != plan->descs.size())
    return fail(SWR_ERROR_INVALID_ARGUMENT, "need one grad descriptor per planned draw (%zu)", plan->descs.size());
  plan->gradDescs.assign(grads, grads + count);
  plan->hasGrads = false;
  r = plan_upload_grads(c, plan, reinterpret_cast<hipStream_t>(stream));
  if (r != SWR_OK) return r;
  plan->hasGrads = true;
  return SWR_OK;
}

int swr_batch_plan_backward(swr_batch_plan *plan, void *stream) {
  DeviceCtx *c = nullptr;
  int r = plan_ctx(plan, &c);
  if (r != SWR_OK) return r;
  if (!plan->hasGrads) return fail(SWR_ERROR_INVALID_ARGUMENT, "swr_batch_plan_set_grads has not been called");
  return plan_enqueue_backward(plan, c->cus, reinterpret_cast<hipStream_t>(stream));
}

int swr_batch_plan_destroy(swr_batch_plan *plan) {
  if (!plan) return SWR_OK;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && dev == plan->device) (void)hipDeviceSynchronize();
  plan->pipe.release();
  plan->flags.release();
  plan->gparams.release();
  plan->gbase.release();
  plan->releaseBins();
  delete plan;
  return SWR_OK;
}

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
  if (!f && tex->format != SWR_FORMAT_R8G8B8A8_UNORM && tex->format != SWR_FORMAT_B8G8R8A8_UNORM)
    return fail(SWR_ERROR_UNSUPPORTED, "format %u", tex->format);  // the box filter is byte-wise
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
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  return static_setup(device);
}

int swr_release(void) {
  DeviceCtx *c = nullptr;
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  SWR_HIP_CHECK(hipDeviceSynchronize());
  c->pipe.release();
  c->pairCap = 0;
  c->pairsPending = false;
  return SWR_OK;
}

int swr_draw(const swr_draw_desc *desc, void *stream) {
  DeviceCtx *c = nullptr;
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  if (!desc) return fail(SWR_ERROR_INVALID_ARGUMENT, "descriptor is NULL");
  return draw_eager(c, desc, 1, reinterpret_cast<hipStream_t>(stream));
}

int swr_draw_noalloc(const swr_draw_desc *desc, void *stream) {
  if (!desc) return fail(SWR_ERROR_INVALID_ARGUMENT, "descriptor is NULL");
  return draw_noalloc(desc, reinterpret_cast<hipStream_t>(stream));
}

int swr_draw_batch(const swr_draw_desc *descs, uint32_t count, void *stream) {
  if (count && !descs) return fail(SWR_ERROR_INVALID_ARGUMENT, "descs is NULL");
  DeviceCtx *c = nullptr;
  int r = current_ctx(&c);
  if (r != SWR_OK) return r;
  return draw_eager(c, descs, count, reinterpret_cast<hipStream_t>(stream));
}

int swr_resolve(const swr_resolve_desc *d, void *stream) {
  if (!d || !d->src || !d->dst) return fail(SWR_ERROR_INVALID_ARGUMENT, "resolve descriptor incomplete");
  if (d->src_samples != 4) return fail(SWR_ERROR_UNSUPPORTED, "resolve supports 4 samples");
  // B8G8R8A8 takes the same byte-wise path as R8G8B8A8 (Blitter.cpp:2142).
  // sRGB attachments resolve in linear space, as the reference's blit routine does.
  const bool srgb = d->format == SWR_FORMAT_R8G8B8A8_SRGB || d->format == SWR_FORMAT_B8G8R8A8_SRGB;
  if (d->format != SWR_FORMAT_R8G8B8A8_UNORM && d->format != SWR_FORMAT_B8G8R8A8_UNORM && !srgb &&
      d->format != SWR_FORMAT_R32G32B32A32_SFLOAT)
    return fail(SWR_ERROR_UNSUPPORTED, "resolve format %u", d->format);
  if (d->width == 0 || d->height == 0) return SWR_OK;
  uint32_t bpp = d->format == SWR_FORMAT_R32G32B32A32_SFLOAT ? 16 : 4;
  if (d->src_row_pitch_bytes < d->width * bpp || d->dst_row_pitch_bytes < d->width * bpp ||
      d->src_sample_pitch_bytes < (uint64_t)d->src_row_pitch_bytes * d->height)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "resolve pitches too small");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  dim3 block(256), grid((d->width + 255) / 256, d->height);
  if (srgb)
    hipLaunchKernelGGL(resolve_srgb8_kernel, grid, block, 0, s, (const char *)d->src, (int)d->src_row_pitch_bytes,
                       (long long)d->src_sample_pitch_bytes, (char *)d->dst, (int)d->dst_row_pitch_bytes,
                       (int)d->width, (int)d->height);
  else if (bpp == 4)
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

namespace {
// The clear value's packed form for `format` (fastClear, Blitter.cpp:
// 170-230, after the clamp of Blitter::clear, :64-86; sw::clamp lets a NaN
// through, and x86-64 converts a NaN to 0 here).
uint32_t unorm8_clear(float c) {
  if (c < 0.0f) c = 0.0f;
  if (c > 1.0f) c = 1.0f;
  volatile float t = 255.0f * c;  // two roundings, as the reference's unfused expression
  const float u = t + 0.5f;
  return u == u ? (uint32_t)u : 0u;
}

// A D16 clear runs the blit routine from a D32F pixel (Blitter.cpp:64-86,
// 1427-1484, 1085-1087): the clamp above, the 0xFFFF scale, Min(., 0xFFFF)
// and Max(., 0) as ordered selects (a NaN becomes 0xFFFF), then RoundInt,
// to nearest even.
uint32_t d16_clear(float d) {
  if (d < 0.0f) d = 0.0f;
  if (d > 1.0f) d = 1.0f;
  float v = d * 65535.0f;
  v = v < 65535.0f ? v : 65535.0f;
  v = v > 0.0f ? v : 0.0f;
  return (uint32_t)std::nearbyint(v) & 0xFFFFu;
}

// An sRGB clear runs the blit routine from an RGBA32F pixel with sRGB
// conversion (Blitter.cpp:64-86, 1427-1484, 647-661, 676-690): the clamp
// above, linearToSRGB of R, G, B, the 255 scale, Min(., 255) and Max(., 0)
// as ordered selects, then RoundInt to nearest even.
uint32_t srgb8_clear(float c, int k) {
  if (c < 0.0f) c = 0.0f;
  if (c > 1.0f) c = 1.0f;
  float v = k < 3 ? swr::linear_to_srgb(c) : c;
  v = v * 255.0f;
  v = v < 255.0f ? v : 255.0f;
  v = v > 0.0f ? v : 0.0f;
  return (uint32_t)std::nearbyint(v);
}

// A validated clear: the packed value, element size and the area (w <= 0:
// nothing to clear).
int clear_prepare(const swr_clear_desc *d, ClearRect &r) {
  r = ClearRect{};
  if (!d || !d->target.data) return fail(SWR_ERROR_INVALID_ARGUMENT, "clear descriptor incomplete");
  if (d->samples != 1 && d->samples != 4) return fail(SWR_ERROR_UNSUPPORTED, "clear supports 1 or 4 samples");
  uint32_t bpp = 0;
  uint4 v = make_uint4(0, 0, 0, 0);
  switch (d->target.format) {
    case SWR_FORMAT_R8G8B8A8_UNORM:
      bpp = 4;
      v.x = unorm8_clear(d->value.color[0]) | (unorm8_clear(d->value.color[1]) << 8) |
            (unorm8_clear(d->value.color[2]) << 16) | (unorm8_clear(d->value.color[3]) << 24);
      break;
    case SWR_FORMAT_B8G8R8A8_UNORM:
      bpp = 4;
      v.x = unorm8_clear(d->value.color[2]) | (unorm8_clear(d->value.color[1]) << 8) |
            (unorm8_clear(d->value.color[0]) << 16) | (unorm8_clear(d->value.color[3]) << 24);
      break;
    case SWR_FORMAT_R8G8B8A8_SRGB:
    case SWR_FORMAT_B8G8R8A8_SRGB: {
      bpp = 4;
      const int bgra = d->target.format == SWR_FORMAT_B8G8R8A8_SRGB;
      for (int k = 0; k < 4; k++) v.x |= srgb8_clear(d->value.color[k], k) << (8 * (bgra && (k & 1) == 0 ? 2 - k : k));
      break;
    }
    case SWR_FORMAT_R32G32B32A32_SFLOAT:
      bpp = 16;
      memcpy(&v, d->value.color, 16);
      break;
    case SWR_FORMAT_D16_UNORM:
      bpp = 2;
      v.x = d16_clear(d->value.depth_stencil.depth);
      break;
    case SWR_FORMAT_D32_SFLOAT:
      bpp = 4;
      memcpy(&v.x, &d->value.depth_stencil.depth, 4);
      break;
    case SWR_FORMAT_S8_UINT:
      bpp = 1;
      v.x = d->value.depth_stencil.stencil & 0xFFu;
      break;
    default:
      return fail(SWR_ERROR_UNSUPPORTED, "clear format %u", d->target.format);
  }
  if (d->target.row_pitch_bytes < d->width * bpp ||
      (d->samples > 1 && d->target.sample_pitch_bytes < (uint64_t)d->target.row_pitch_bytes * d->height))
    return fail(SWR_ERROR_INVALID_ARGUMENT, "clear pitches too small");
  int64_t x0 = 0, y0 = 0, x1 = d->width, y1 = d->height;
  if (d->use_area) {
    x0 = std::max<int64_t>(x0, d->area.x);
    y0 = std::max<int64_t>(y0, d->area.y);
    x1 = std::min<int64_t>(x1, (int64_t)d->area.x + d->area.width);
    y1 = std::min<int64_t>(y1, (int64_t)d->area.y + d->area.height);
  }
  if (x1 <= x0 || y1 <= y0) return SWR_OK;
  r.base = (char *)d->target.data;
  r.pitch = (int)d->target.row_pitch_bytes;
  r.slice = (long long)d->target.sample_pitch_bytes;
  r.x0 = (int)x0;
  r.y0 = (int)y0;
  r.w = (int)(x1 - x0);
  r.h = (int)(y1 - y0);
  r.samples = (int)d->samples;
  r.bytes = (int)bpp;
  r.value = v;
  return SWR_OK;
}
}  // namespace

int swr_clear(const swr_clear_desc *d, void *stream) {
  ClearRect r;
  const int rc = clear_prepare(d, r);
  if (rc != SWR_OK) return rc;
  if (r.w <= 0) return SWR_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  dim3 block(256), grid((r.w + 255) / 256, r.h, r.samples);
  if (r.bytes == 16)
    hipLaunchKernelGGL(clear_rect_kernel<16>, grid, block, 0, s, r.base, r.pitch, r.slice, r.x0, r.y0, r.w, r.h, r.value);
  else if (r.bytes == 4)
    hipLaunchKernelGGL(clear_rect_kernel<4>, grid, block, 0, s, r.base, r.pitch, r.slice, r.x0, r.y0, r.w, r.h, r.value);
  else if (r.bytes == 2)
    hipLaunchKernelGGL(clear_rect_kernel<2>, grid, block, 0, s, r.base, r.pitch, r.slice, r.x0, r.y0, r.w, r.h, r.value);
  else
    hipLaunchKernelGGL(clear_rect_kernel<1>, grid, block, 0, s, r.base, r.pitch, r.slice, r.x0, r.y0, r.w, r.h, r.value);
  SWR_HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_clear_batch(const swr_clear_desc *descs, uint32_t count, void *stream) {
  if (count && !descs) return fail(SWR_ERROR_INVALID_ARGUMENT, "clear batch without descriptors");
  std::vector<ClearRect> rects;
  rects.reserve(count);
  std::vector<std::pair<uintptr_t, uintptr_t>> ranges;
  for (uint32_t i = 0; i < count; i++) {
    ClearRect r;
    const int rc = clear_prepare(&descs[i], r);
    if (rc != SWR_OK) return fail(rc, "clear %u: %s", i, g_error.c_str());
    if (r.w <= 0) continue;
    // The span of bytes the clear writes, plane by plane (from the area's
    // first byte to its last: side-by-side areas of the same rows count as
    // overlapping).
    for (int q = 0; q < r.samples; q++) {
      const char *lo = r.base + (long long)q * r.slice + (long long)r.y0 * r.pitch + (long long)r.x0 * r.bytes;
      ranges.push_back({(uintptr_t)lo, (uintptr_t)(lo + (long long)(r.h - 1) * r.pitch + (long long)r.w * r.bytes)});
    }
    rects.push_back(r);
  }
  // The clears run concurrently: their byte ranges must be disjoint.
  std::sort(ranges.begin(), ranges.end());
  for (size_t k = 1; k < ranges.size(); k++)
    if (ranges[k].first < ranges[k - 1].second)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "clear batch targets overlap");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  for (int bytes : {16, 4, 2, 1}) {
    ClearBatch cb{};
    auto launch = [&]() {
      if (!cb.n) return;
      dim3 block(256), grid(cb.start[cb.n]);
      if (bytes == 16) hipLaunchKernelGGL(clear_batch_kernel<16>, grid, block, 0, s, cb);
      else if (bytes == 4) hipLaunchKernelGGL(clear_batch_kernel<4>, grid, block, 0, s, cb);
      else if (bytes == 2) hipLaunchKernelGGL(clear_batch_kernel<2>, grid, block, 0, s, cb);
      else hipLaunchKernelGGL(clear_batch_kernel<1>, grid, block, 0, s, cb);
      cb = ClearBatch{};
    };
    for (const ClearRect &r : rects) {
      if (r.bytes != bytes) continue;
      const uint64_t blocks = ((uint64_t)r.w * (uint64_t)r.h * (uint64_t)r.samples + kClearPixels - 1) / kClearPixels;
      if (cb.n == kClearBatch || (uint64_t)cb.start[cb.n] + blocks > 0x7FFFFFFFull) launch();
      cb.r[cb.n] = r;
      cb.start[cb.n + 1] = cb.start[cb.n] + (uint32_t)blocks;
      cb.n++;
    }
    launch();
  }
  SWR_HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_decode_bc(const swr_bc_desc *d, void *stream) {
  if (!d || !d->src || !d->dst) return fail(SWR_ERROR_INVALID_ARGUMENT, "BC descriptor incomplete");
  if (d->n < 1 || d->n > 7) return fail(SWR_ERROR_INVALID_ARGUMENT, "BC%u is not a BC format", d->n);
  if (d->width == 0 || d->height == 0) return SWR_OK;
  if (d->width > 16384 || d->height > 16384) return fail(SWR_ERROR_INVALID_ARGUMENT, "image larger than 16384");
  if (reinterpret_cast<uintptr_t>(d->src) & 7u) return fail(SWR_ERROR_INVALID_ARGUMENT, "src is not 8-byte aligned");
  const uint32_t bpp = d->dst_bytes_per_pixel;
  if (d->n <= 3 || d->n == 7) {
    if (bpp != 4) return fail(SWR_ERROR_INVALID_ARGUMENT, "BC%u writes 4-byte pixels", d->n);
    if ((reinterpret_cast<uintptr_t>(d->dst) | d->dst_row_pitch_bytes) & 3u)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "BC%u needs a 4-byte aligned destination and pitch", d->n);
  } else if (d->n == 6) {
    if (bpp != 8) return fail(SWR_ERROR_INVALID_ARGUMENT, "BC6H writes 8-byte pixels (four half floats)");
    if ((reinterpret_cast<uintptr_t>(d->dst) | d->dst_row_pitch_bytes) & 1u)
      return fail(SWR_ERROR_INVALID_ARGUMENT, "BC6H needs a 2-byte aligned destination and pitch");
  } else if (bpp < d->n - 3u) {
    return fail(SWR_ERROR_INVALID_ARGUMENT, "BC%u needs %u bytes per pixel", d->n, d->n - 3u);
  }
  if ((uint64_t)d->dst_row_pitch_bytes < (uint64_t)d->width * bpp)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "destination pitch too small");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  dim3 block(256), grid((d->width + 255) / 256, d->height);
  if (d->n >= 6)
    hipLaunchKernelGGL(decode_bptc_kernel, grid, block, 0, s, (const uint8_t *)d->src, (uint8_t *)d->dst,
                       (int)d->width, (int)d->height, (int)d->dst_row_pitch_bytes, (int)bpp, (int)d->n,
                       (int)(d->n == 6 && !d->no_alpha_or_unsigned));
  else
    hipLaunchKernelGGL(decode_bc_kernel, grid, block, 0, s, (const uint8_t *)d->src, (uint8_t *)d->dst, (int)d->width,
                       (int)d->height, (int)d->dst_row_pitch_bytes, (int)bpp, (int)d->n, (int)d->no_alpha_or_unsigned);
  SWR_HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_decode_etc(const swr_etc_desc *d, void *stream) {
  if (!d || !d->src || !d->dst) return fail(SWR_ERROR_INVALID_ARGUMENT, "ETC descriptor incomplete");
  if (d->input_type > SWR_ETC_RGBA) return fail(SWR_ERROR_INVALID_ARGUMENT, "ETC input type %u", d->input_type);
  if (d->width == 0 || d->height == 0) return SWR_OK;
  if (d->width > 16384 || d->height > 16384) return fail(SWR_ERROR_INVALID_ARGUMENT, "image larger than 16384");
  if (reinterpret_cast<uintptr_t>(d->src) & 7u) return fail(SWR_ERROR_INVALID_ARGUMENT, "src is not 8-byte aligned");
  const uint32_t t = d->input_type;
  const bool eac = t <= SWR_ETC_RG_UNSIGNED;
  const uint32_t bpp = eac ? (t <= SWR_ETC_R_UNSIGNED ? 2u : 4u) : 4u;
  if (d->dst_bytes_per_pixel != bpp)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "ETC input type %u writes %u-byte pixels", t, bpp);
  const uintptr_t align = eac ? 1u : 3u;
  if ((reinterpret_cast<uintptr_t>(d->dst) | d->dst_row_pitch_bytes) & align)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "destination or pitch not %u-byte aligned", (unsigned)align + 1);
  if ((uint64_t)d->dst_row_pitch_bytes < (uint64_t)d->width * bpp)
    return fail(SWR_ERROR_INVALID_ARGUMENT, "destination pitch too small");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  dim3 block(256), grid((d->width + 255) / 256, d->height);
  hipLaunchKernelGGL(decode_etc_kernel, grid, block, 0, s, (const uint8_t *)d->src, (uint8_t *)d->dst, (int)d->width,
                     (int)d->height, (int)d->dst_row_pitch_bytes, (int)bpp, (int)t);
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
    c->acc.draws += es.draws;
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
  int r0 = check_tuning(*t);
  if (r0 != SWR_OK) return r0;
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
