// swr_params.hpp -- POD state shared by the host launcher and the kernels.
//
// DrawParams is the GPU form of what the reference keeps in DrawData and
// the SetupProcessor / PixelProcessor states (src/Device/Renderer.hpp:58-113,
// src/Device/SetupProcessor.hpp:38-63, src/Device/PixelProcessor.hpp:36-108),
// derived once per draw on the host.  TriRecord is the per-primitive setup
// output the tile rasteriser consumes (the reference's sw::Primitive,
// src/Device/Primitive.hpp:30-70, without its O(height) span outline: spans
// are re-derived per pixel from the edges).
#pragma once

#include <stdint.h>

namespace swr {

enum : int { FILT_POINT = 0, FILT_MIN_POINT_MAG_LINEAR, FILT_MIN_LINEAR_MAG_POINT, FILT_LINEAR, FILT_ANISOTROPIC };
enum : int { MIP_NONE = 0, MIP_POINT, MIP_LINEAR };
// Canonical blend operations beyond VkBlendOp (Context.cpp:1181-1271).
enum : int { BLEND_OP_SRC = 1000148000, BLEND_OP_DST = 1000148001, BLEND_OP_ZERO = 1000148002 };
enum : int { SHADER_INTERPOLATE = 0, SHADER_SAMPLE = 1, SHADER_MODULATE = 2 };

constexpr int kMipLevels = 15;
constexpr int kMaxPolygon = 16;
// A binned (tile, primitive) pair's value: the primitive's record slot in
// bits 0-27, the 8x8 blocks of the tile it may cover in bits 28-31.
constexpr uint32_t kSlotMask = (1u << 28) - 1u;

struct TexParams {
  const void *data[kMipLevels];
  int width[kMipLevels];
  int height[kMipLevels];
  int pitchP[kMipLevels];
  uint32_t uHalf[kMipLevels];
  uint32_t vHalf[kMipLevels];
  float wwhh[4];
  int isFloat;
  int bgra;  // B8G8R8A8 texels: R and B exchanged in memory
  int srgb;  // sRGB texels, decoded through srgbLut
  uint16_t srgbLut[256];  // sRGBtoLinearFF_FF00 (Constants.cpp:273-276)
  int filter, mipmap, addrU, addrV;  // addr: 0 wrap, 1 clamp
  int skipLod;
  float mipLodBias, maxAnisotropy, minLod, maxLod;
};

struct StencilParams {
  int compareOp, failOp, passOp, depthFailOp;
  int useCompareMask, useWriteMask;
  uint32_t compareMask, writeMask, reference, refMasked;
  int32_t refSigned;
};

struct DrawParams {
  // Inputs.
  const float *pos;
  const float *col;
  const float *uv;
  const void *idx;
  int idxBytes;  // 1, 2 or 4 (VkIndexType UINT8_EXT, UINT16, UINT32)
  int triCount;
  int vertexCount;
  int topology;     // 0 triangle list, 1 triangle strip, 2 fan, 3 line list, 4 line strip, 5 points
  int line, bresenham;  // lines; set up as Bresenham parallelograms
  float lineWidth;
  int point;
  const float *psize;  // PointSize per vertex, or null (1)
  int provokeLast;  // VK_PROVOKING_VERTEX_MODE_LAST_VERTEX
  // Viewport / scissor (Renderer.cpp:300-345).
  float WxF, HxF, X0xF, Y0xF, depthRange, depthNear;
  int sx0, sx1, sy0, sy1;
  int width, height;
  // Rasterisation.
  int cullMode, frontFace;
  int ms, msMask;
  // Depth / stencil.
  int depthTest, depthWrite, depthCompare;
  // The depth compare as a predicate of the stored (zb) and fragment (z)
  // depths, inverted when zInv: 0 zb <= z, 1 zb < z, 2 zb == z, 3 true.
  int zPred, zInv;
  int depth16;  // D16_UNORM depth attachment (else D32_SFLOAT)
  int depthClip;                    // clip against the near and far planes
  int depthBounds;                  // depth bounds test on the stored depth
  int alphaToCoverage;              // sample q kept where the shaded alpha is not below a2c[q]
  float a2c[4];
  float depthBoundsMin, depthBoundsMax;
  float depthClampLo, depthClampHi;  // every sample's depth range
  // Depth bias: which terms the setup applies, and whether samples get it.
  int biasConstant, biasSlope, biasClamp, depthBias;
  float constantDepthBias, slopeDepthBias, depthBiasClamp;
  int stencilActive, stencilWrite;
  StencilParams front, back;
  // Colour.
  int colorMask;   // canonical 4-bit write mask; 0 = no colour output
  int colorFloat;  // RGBA32F target
  int colorBGRA;   // B8G8R8A8 target (R and B exchanged in memory)
  int colorSRGB;   // sRGB-encoded 8-bit target
  int blendEnable;
  int advancedBlend;  // colorOp (= alphaOp) is a VK_EXT_blend_operation_advanced mode
  int srcF, dstF, colorOp, srcFA, dstFA, alphaOp;
  int blendKind;  // 1: SRC_ALPHA, ONE_MINUS_SRC_ALPHA, ADD for colour and alpha
  float blendConst[4], invBlendConst[4];
  int shader, useColor, useUV;
  TexParams tex;
  // Attachments.
  char *color;
  int colorPitch;
  long long colorSlice;
  char *depth;
  int depthPitch;
  long long depthSlice;
  char *stencil;
  int stencilPitch;
  long long stencilSlice;
  char *primId;  // R32_UINT visibility record (RGBA32F targets only) or NULL
  int primIdPitch;
  long long primIdSlice;
  unsigned long long *occlusion;
  unsigned long long *stats;  // [block tests, block covers, shaded] or NULL
  // Tiling.
  int tileW, tileH, tilesX, tilesY;
  int triBase;  // first record slot of this draw's primitives in its batch
};

// Per-primitive setup output.  n == 0 marks a culled primitive.  The first
// 80 bytes are what the coverage and depth passes read per lane; the rest
// (w plane and attribute planes) is read when fragments are shaded.
struct TriRecord {
  int nflags;       // bits 0-7: polygon vertex count n (3 unless clipped; 0: culled);
                    // bit 8: d (winding); bit 9: front facing
  int yMin, yMax;   // rows with possible coverage (scissor-clamped)
  int xMin, xMax;   // pixel columns with possible coverage
  int X[3], Y[3];   // 28.4 polygon vertices when n == 3
  float zlo;        // lower bound of the evaluated depth (bias included) over the bounding box
  float x0, y0;
  float zA, zB, zC;
  float zBias;      // depth bias added to every sample's depth (0 when off)
  float wA, wB, wC;
  float V[6][3];    // colour rgba, uv
  int draw;         // index of the primitive's draw in its batch
};

static_assert(sizeof(TriRecord) == 160, "TriRecord layout");

}  // namespace swr
