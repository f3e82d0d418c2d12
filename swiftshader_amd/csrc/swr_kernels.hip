// swr_kernels.hip -- HIP kernels of the fixed-function draw path (gfx950).
//
// Kernels of a draw (the front end between setup and raster, binning and
// the sort, is in swr_batch.hip):
//   setup_kernel     one thread per triangle: vertex tail (clip flags,
//                    28.4 projection), trivial reject, clipping, culling,
//                    plane equations, pixel/tile bounding box (setup_one,
//                    shared with the batch front end).
//   raster_batch_kernel  one workgroup per screen tile, one wave per 8x8
//                    block, one pixel per lane (a 2x2 quad in lanes
//                    4k..4k+3); the block's depth / stencil / colour live in
//                    registers while the wave walks the tile's primitive
//                    list in order: coverage, stencil, depth, shading, blend.
//   clear_rect_kernel, resolve_*_kernel  clears and the 4x -> 1x resolve
//                    (compressed texture decoding is in swr_decode.hip).
//
// Reference semantics: src/Pipeline/VertexRoutine.cpp:116-154,601-611,
// src/Device/Clipper.cpp, src/Pipeline/SetupRoutine.cpp,
// src/Device/QuadRasterizer.cpp, src/Pipeline/PixelRoutine.cpp,
// src/Pipeline/PixelProgram.cpp, src/Pipeline/SamplerCore.cpp,
// src/Device/Blitter.cpp:1524-1545,2076-2200.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swr_device.hpp"
#include "swr_params.hpp"
#include "swr_srgb.hpp"

namespace swr {

enum : int { CLIP_RIGHT = 1, CLIP_TOP = 2, CLIP_FAR = 4, CLIP_LEFT = 8, CLIP_BOTTOM = 16, CLIP_NEAR = 32,
             CLIP_FRUSTUM = 63, CLIP_SIDES = 27, CLIP_FINITE = 128 };

// Standard 4x sample locations relative to the pixel centre
// (Constants.hpp:26-47) and their 28.4 offsets Xf/Yf (Constants.cpp:291-297).
__device__ __constant__ const float kSampleX[4] = {-0.125f, 0.375f, -0.375f, 0.125f};
__device__ __constant__ const float kSampleY[4] = {-0.375f, -0.125f, 0.125f, 0.375f};
__device__ __constant__ const int kXf[4] = {-2, 6, -6, 2};
__device__ __constant__ const int kYf[4] = {-6, -2, 2, 6};

// ---------------------------------------------------------------------------
// Setup
// ---------------------------------------------------------------------------
struct ClipVtx { float v[4]; };

__device__ __forceinline__ float clip_dist(const ClipVtx &p, int plane) {
  switch (plane) {
    case 0: return p.v[2];
    case 1: return p.v[3] - p.v[2];
    case 2: return p.v[3] + p.v[0];
    case 3: return p.v[3] - p.v[0];
    case 4: return p.v[3] - p.v[1];
    default: return p.v[3] + p.v[1];
  }
}

__device__ __forceinline__ ClipVtx clip_edge(const ClipVtx &Vi, const ClipVtx &Vj, float di, float dj) {
  float D = 1.0f / (dj - di);
  ClipVtx o;
#pragma unroll
  for (int k = 0; k < 4; k++) o.v[k] = (dj * Vi.v[k] - di * Vj.v[k]) * D;
  return o;
}

// Sutherland-Hodgman against near, far, left, right, top, bottom in that
// order, stopping below 3 vertices (Clipper.cpp:271-299).  Returns the
// number of clip stages applied (the reference's Polygon::i).  tmp: the
// caller's second vertex buffer.
__device__ int clip_polygon(ClipVtx (&poly)[kMaxPolygon], int &n, int flagsOr, ClipVtx (&tmp)[kMaxPolygon]) {
  const int bits[6] = {CLIP_NEAR, CLIP_FAR, CLIP_LEFT, CLIP_RIGHT, CLIP_TOP, CLIP_BOTTOM};
  int stages = 0;
  for (int k = 0; k < 6; k++) {
    if (k > 0 && n < 3) break;
    if (!(flagsOr & bits[k])) continue;
    int t = 0;
    for (int i = 0; i < n; i++) {
      int j = (i == n - 1) ? 0 : i + 1;
      float di = clip_dist(poly[i], k), dj = clip_dist(poly[j], k);
      if (di >= 0) {
        if (t < kMaxPolygon) tmp[t] = poly[i];
        t++;
        if (dj < 0) {
          if (t < kMaxPolygon) tmp[t] = clip_edge(poly[i], poly[j], di, dj);
          t++;
        }
      } else if (dj > 0) {
        if (t < kMaxPolygon) tmp[t] = clip_edge(poly[j], poly[i], dj, di);
        t++;
      }
    }
    n = t < kMaxPolygon ? t : kMaxPolygon;
    for (int i = 0; i < n; i++) poly[i] = tmp[i];
    stages++;
  }
  return stages;
}

__device__ __forceinline__ int wadd(int a, int b) { return (int)((unsigned)a + (unsigned)b); }
__device__ __forceinline__ int wsub(int a, int b) { return (int)((unsigned)a - (unsigned)b); }

// D16 compares and stores Min(Max(Round(z * 0xFFFF), 0), 0xFFFF) of the
// clamped depth (PixelRoutine.cpp:506-510, 686-711; Round is to nearest even).
__device__ __forceinline__ float d16_value(float z) {
  return rr_min(rr_max(__builtin_rintf(z * 65535.0f), 0.0f), 65535.0f);
}

// B8G8R8A8 keeps R and B exchanged in memory (PixelRoutine.cpp:1091-1110,
// 2130-2143): the tile holds R8G8B8A8 order.
__device__ __forceinline__ uint32_t bgra_swap(int bgra, uint32_t v) {
  return bgra ? (v & 0xFF00FF00u) | ((v >> 16) & 0xFFu) | ((v & 0xFFu) << 16) : v;
}

// Index p of the draw's index buffer, of its VkIndexType.
__device__ __forceinline__ uint32_t load_index(const DrawParams &P, uint32_t p) {
  if (P.idxBytes == 2) return reinterpret_cast<const uint16_t *>(P.idx)[p];
  if (P.idxBytes == 1) return reinterpret_cast<const uint8_t *>(P.idx)[p];
  return reinterpret_cast<const uint32_t *>(P.idx)[p];
}

// Vertex indices of primitive t in the reference's batch order
// (setBatchIndices, Renderer.cpp:50-145); out-of-range indices read the
// last vertex instead of faulting.
__device__ __forceinline__ void tri_vertices(const DrawParams &P, uint32_t t, uint32_t vi[3]) {
  const uint32_t f = P.provokeLast ? 0u : 1u;  // 1: first-vertex mode
  uint32_t p0, p1, p2;
  if (P.topology == 5) {
    p0 = p1 = p2 = t;  // a point's every setup vertex is its one
  } else if (P.topology >= 3) {
    // A line's vertices, exchanged in last-vertex mode; the setup's third
    // vertex is the second (SetupRoutine.cpp:52-54).
    const uint32_t b = P.topology == 3 ? 2u * t : t;
    p0 = b + (f ? 0u : 1u);
    p1 = b + f;
    p2 = p1;
  } else if (P.topology == 1) {
    p0 = t + (f ? 0u : 2u);
    p1 = t + (t & 1u) + f;
    p2 = t + (~t & 1u) + f;
  } else if (P.topology == 2) {
    const uint32_t a = t + 1u, b = t + 2u;
    p0 = f ? a : b;
    p1 = f ? b : 0u;
    p2 = f ? 0u : a;
  } else {
    p0 = 3u * t + (f ? 0u : 2u);
    p1 = 3u * t + (f ? 1u : 0u);
    p2 = 3u * t + (f ? 2u : 1u);
  }
  vi[0] = P.idx ? load_index(P, p0) : p0;
  vi[1] = P.idx ? load_index(P, p1) : p1;
  vi[2] = P.idx ? load_index(P, p2) : p2;
#pragma unroll
  for (int k = 0; k < 3; k++) vi[k] = vi[k] < (uint32_t)P.vertexCount ? vi[k] : (uint32_t)P.vertexCount - 1u;
}

// DrawCall::setupLine (Renderer.cpp:920-1163): the rectangle of width
// P.lineWidth centred on the segment P0-P1, or, in the Bresenham mode, the
// parallelogram joining P0's and P1's diamonds of half-width lineWidth / 2
// along the major axis, in clip space.  Returns false for a rejected line.
__device__ __forceinline__ bool line_polygon(const DrawParams &P, const float (&P0)[4], const float (&P1)[4],
                                             ClipVtx (&cp)[kMaxPolygon]) {
  if (P0[3] <= 0.0f && P1[3] <= 0.0f) return false;
  const float W = P.WxF * (1.0f / 16.0f);
  const float H = P.HxF * (1.0f / 16.0f);
  float dx = W * (P1[0] / P1[3] - P0[0] / P0[3]);
  float dy = H * (P1[1] / P1[3] - P0[1] / P0[3]);
  if (dx == 0.0f && dy == 0.0f) return false;
  if (P.bresenham) {
    const float dx0 = P.lineWidth * 0.5f * P0[3] / W, dy0 = P.lineWidth * 0.5f * P0[3] / H;
    const float dx1 = P.lineWidth * 0.5f * P1[3] / W, dy1 = P.lineWidth * 0.5f * P1[3] / H;
    // Corners kept per major direction: right and down run P0, P1, P1, P0
    // along one diamond axis; up and left run P0, P0, P1, P1.
    const bool alongFirst = dx > -dy;
    const bool xMajor = dx > dy;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const bool first = alongFirst ? (k == 0 || k == 3) : (k < 2);
      for (int j = 0; j < 4; j++) cp[k].v[j] = first ? P0[j] : P1[j];
    }
    if (alongFirst && xMajor) {  // right: up and down corners
      cp[0].v[1] += +dy0; cp[1].v[1] += +dy1; cp[2].v[1] += -dy1; cp[3].v[1] += -dy0;
    } else if (alongFirst) {     // down: left and right corners
      cp[0].v[0] += -dx0; cp[1].v[0] += -dx1; cp[2].v[0] += +dx1; cp[3].v[0] += +dx0;
    } else if (xMajor) {         // up: left and right corners
      cp[0].v[0] += -dx0; cp[1].v[0] += +dx0; cp[2].v[0] += +dx1; cp[3].v[0] += -dx1;
    } else {                     // left: up and down corners
      cp[0].v[1] += +dy0; cp[1].v[1] += -dy0; cp[2].v[1] += -dy1; cp[3].v[1] += +dy1;
    }
  } else {
    const float scale = P.lineWidth * 0.5f / sqrtf(dx * dx + dy * dy);
    dx = dx * scale;
    dy = dy * scale;
    const float dx0h = dx * P0[3] / H, dy0w = dy * P0[3] / W;
    const float dx1h = dx * P1[3] / H, dy1w = dy * P1[3] / W;
#pragma unroll
    for (int k = 0; k < 4; k++)
      for (int j = 0; j < 4; j++) cp[k].v[j] = (k == 0 || k == 3) ? P0[j] : P1[j];
    cp[0].v[0] += -dy0w; cp[0].v[1] += +dx0h;
    cp[1].v[0] += -dy1w; cp[1].v[1] += +dx1h;
    cp[2].v[0] += +dy1w; cp[2].v[1] += -dx1h;
    cp[3].v[0] += +dy0w; cp[3].v[1] += -dx0h;
  }
  return true;
}

// DrawCall::setupPoint (Renderer.cpp:1166-1210): the square of side
// PointSize (clamped to [1, 1023] by two compares, so a NaN passes) centred
// on the vertex, offsets pSize * w * halfPixel, halfPixel = 0.5 / (viewport
// extent / 2).
__device__ __forceinline__ void point_polygon(const DrawParams &P, const float (&V)[4], float size,
                                              ClipVtx (&cp)[kMaxPolygon]) {
  float pSize = size;
  if (pSize < 1.0f) pSize = 1.0f;
  if (pSize > 1023.0f) pSize = 1023.0f;
  const float halfPixelX = 0.5f / (P.WxF * (1.0f / 16.0f));
  const float halfPixelY = 0.5f / (P.HxF * (1.0f / 16.0f));
  const float X = pSize * V[3] * halfPixelX;
  const float Y = pSize * V[3] * halfPixelY;
#pragma unroll
  for (int k = 0; k < 4; k++)
    for (int j = 0; j < 4; j++) cp[k].v[j] = V[j];
  cp[0].v[0] -= X; cp[0].v[1] += Y;
  cp[1].v[0] += X; cp[1].v[1] += Y;
  cp[2].v[0] += X; cp[2].v[1] -= Y;
  cp[3].v[0] -= X; cp[3].v[1] -= Y;
}

template <typename T>
__device__ __forceinline__ T pick3(T a, T b, T c, int i) { return i == 0 ? a : (i == 1 ? b : c); }

// Vertex arrays of the polygon path (clipped triangles, lines, points),
// indexed at run time.
struct PolyWork {
  ClipVtx cp[kMaxPolygon], tmp[kMaxPolygon];
  int X[kMaxPolygon], Y[kMaxPolygon];
};

// Setup of primitive t of the draw P, written to record slot g (the
// primitive's index in the batch); `draw` is the draw's index in the batch;
// w holds the polygon path's arrays.  Without POLY, the primitives that need
// the polygon path are left alone and false is returned (w is not touched),
// so that a kernel can give that path workspace in LDS for a few lanes at a
// time instead of scratch memory for every thread.
template <bool POLY = true>
__device__ __forceinline__ bool setup_one(const DrawParams &P, int t, size_t g, int draw, TriRecord *__restrict__ rec,
                                          int *__restrict__ poly, uint32_t *__restrict__ tileCount, PolyWork &w) {
  if (!POLY && (P.point || P.line)) return false;

  uint32_t vi[3];
  tri_vertices(P, (uint32_t)t, vi);

  float pos[3][4], attr[3][6];
  int flags[3], PX[3], PY[3];
  float pz[3], pw[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float4 p = reinterpret_cast<const float4 *>(P.pos)[vi[k]];
    pos[k][0] = p.x; pos[k][1] = p.y; pos[k][2] = p.z; pos[k][3] = p.w;
    if (P.col) {
      float4 c = reinterpret_cast<const float4 *>(P.col)[vi[k]];
      attr[k][0] = c.x; attr[k][1] = c.y; attr[k][2] = c.z; attr[k][3] = c.w;
    } else {
      attr[k][0] = attr[k][1] = attr[k][2] = attr[k][3] = 0.0f;
    }
    if (P.uv) {
      float2 u = reinterpret_cast<const float2 *>(P.uv)[vi[k]];
      attr[k][4] = u.x; attr[k][5] = u.y;
    } else {
      attr[k][4] = attr[k][5] = 0.0f;
    }
    // Clip flags (VertexRoutine.cpp:116-154).
    float x = p.x, y = p.y, z = p.z, w = p.w;
    int f = 0;
    if (w < x) f |= CLIP_RIGHT;
    if (w < y) f |= CLIP_TOP;
    if (!(-w <= x)) f |= CLIP_LEFT;
    if (!(-w <= y)) f |= CLIP_BOTTOM;
    if (P.depthClip) {  // near and far only with depth clipping (VertexRoutine.cpp:137-143)
      if (w < z) f |= CLIP_FAR;
      if (!(0.0f <= z)) f |= CLIP_NEAR;
    }
    const float maxPos = __uint_as_float(0x7F7FFFFFu);
    if (__builtin_fabsf(x) <= maxPos && __builtin_fabsf(y) <= maxPos && __builtin_fabsf(z) <= maxPos) f |= CLIP_FINITE;
    flags[k] = f;
    // Projection (VertexRoutine.cpp:601-611).
    float ww = __uint_as_float(__float_as_uint(w) | ((w == 0.0f) ? 0x3F800000u : 0u));
    float rhw = 1.0f / ww;
    PX[k] = rr_round_int_clamped(P.X0xF + (x * rhw) * P.WxF);
    PY[k] = rr_round_int_clamped(P.Y0xF + (y * rhw) * P.HxF);
    pz[k] = z * rhw;
    pw[k] = rhw;
  }

  TriRecord &R = rec[g];
  tileCount[g] = 0;
  R.nflags = 0;
  R.draw = draw;
  ClipVtx (&cp)[kMaxPolygon] = w.cp;
  int n = 3;
  int stages = 0;
  int d = 1;
  int frontFacing = 1;
  if (P.point) {
    // setupPoint: the point's square, clipped against every frustum plane;
    // no culling, front-facing, d = 1 (SetupRoutine.cpp:115-119).
    point_polygon(P, pos[0], P.psize ? P.psize[vi[0]] : 1.0f, cp);
    n = 4;
    stages = clip_polygon(cp, n, P.depthClip ? CLIP_FRUSTUM : CLIP_SIDES, w.tmp);
    if (n < 3) return true;
  } else if (P.line) {
    // setupLine: the line's polygon, clipped against every frustum plane;
    // no culling, front-facing, d = 1 (SetupRoutine.cpp:115-119).
    if (!line_polygon(P, pos[0], pos[1], cp)) return true;
    n = 4;
    stages = clip_polygon(cp, n, P.depthClip ? CLIP_FRUSTUM : CLIP_SIDES, w.tmp);
    if (n < 3) return true;
  } else {
    // setupSolidTriangles (Renderer.cpp:733-776).
    if ((flags[0] & flags[1] & flags[2]) != CLIP_FINITE) return true;
    int flagsOr = flags[0] | flags[1] | flags[2];
    if (flagsOr != CLIP_FINITE) {
      if (!POLY) return false;
#pragma unroll
      for (int k = 0; k < 3; k++)
        for (int c = 0; c < 4; c++) cp[k].v[c] = pos[k][c];
      stages = clip_polygon(cp, n, flagsOr, w.tmp);
      if (n < 3) return true;
    }
    // Culling and facing (SetupRoutine.cpp:71-120).
    float x0 = (float)PX[0], x1 = (float)PX[1], x2 = (float)PX[2];
    float y0 = (float)PY[0], y1 = (float)PY[1], y2 = (float)PY[2];
    float A = (y0 - y2) * x1 + (y2 - y1) * x0 + (y1 - y0) * x2;
    int w012 = (int)(__float_as_uint(pos[0][3]) ^ __float_as_uint(pos[1][3]) ^ __float_as_uint(pos[2][3]));
    if (w012 < 0) A = -A;
    frontFacing = (P.frontFace == 0) ? (A >= 0.0f) : (A <= 0.0f);
    if ((P.cullMode & 1) && frontFacing) return true;
    if ((P.cullMode & 2) && !frontFacing) return true;
    d = (A > 0.0f) ? 1 : 0;
  }

  // Polygon in 28.4 (reprojected when clipped or a line,
  // SetupRoutine.cpp:122-143).
  // An unclipped triangle keeps its three vertices in registers; only a
  // clipped polygon, a line or a point goes through the n-vertex arrays.
  const bool isPolygon = POLY && (stages != 0 || P.line || P.point);
  int (&X)[kMaxPolygon] = w.X;
  int (&Y)[kMaxPolygon] = w.Y;
  int yMin, yMax, xLo, xHi;
  if (isPolygon) {
    for (int i = 0; i < n; i++) {
      float w = cp[i].v[3];
      float rhw = (w != 0.0f) ? 1.0f / w : 1.0f;
      X[i] = rr_round_int(P.X0xF + (cp[i].v[0] * rhw) * P.WxF);
      Y[i] = rr_round_int(P.Y0xF + (cp[i].v[1] * rhw) * P.HxF);
    }
    yMin = Y[0]; yMax = Y[0]; xLo = X[0]; xHi = X[0];
    for (int i = 1; i < n; i++) {
      yMin = min(yMin, Y[i]); yMax = max(yMax, Y[i]);
      xLo = min(xLo, X[i]); xHi = max(xHi, X[i]);
    }
  } else {
    yMin = min(min(PY[0], PY[1]), PY[2]);
    yMax = max(max(PY[0], PY[1]), PY[2]);
    xLo = min(min(PX[0], PX[1]), PX[2]);
    xHi = max(max(PX[0], PX[1]), PX[2]);
  }
  if (P.ms > 1) {
    yMin = wadd(yMin, 9) >> 4;   // yMinMultiSampleOffset (Constants.hpp:51)
    yMax = wadd(yMax, 21) >> 4;  // yMaxMultiSampleOffset (Constants.hpp:52)
  } else {
    yMin = wadd(yMin, 15) >> 4;
    yMax = wadd(yMax, 15) >> 4;
  }
  yMin = max(yMin, P.sy0);
  yMax = min(yMax, P.sy1);
  if (yMin >= yMax) return true;
  // Columns that can hold coverage: left >= ceil(min X / 16), right <= ceil(max X / 16).
  const int sxo = (P.ms > 1) ? 6 : 0;
  int xMin = max(wadd(wsub(xLo, sxo), 15) >> 4, P.sx0);
  int xMax = min(wadd(wadd(xHi, sxo), 15) >> 4, P.sx1);
  if (xMin >= xMax) return true;

  // Sort by minimum y, then by maximum w (SetupRoutine.cpp:270-294),
  // triangles only.
  int o0 = 0, o1 = 1, o2 = 2;
  if (!P.line && !P.point) {
    float y0 = pos[0][1], y1 = pos[1][1], y2 = pos[2][1];
    float ym = rr_min(rr_min(y0, y1), y2);
    if (ym == y1) { int tt = o0; o0 = o1; o1 = o2; o2 = tt; }
    if (ym == y2) { int tt = o2; o2 = o1; o1 = o0; o0 = tt; }
  }
  if (!P.line && !P.point) {
    float w0 = pick3(pos[0][3], pos[1][3], pos[2][3], o0), w1 = pick3(pos[0][3], pos[1][3], pos[2][3], o1),
          w2 = pick3(pos[0][3], pos[1][3], pos[2][3], o2);
    float wm = rr_max(rr_max(w0, w1), w2);
    if (wm == w1) { int tt = o0; o0 = o1; o1 = o2; o2 = tt; }
    if (wm == w2) { int tt = o2; o2 = o1; o1 = o0; o0 = tt; }
  }
  // Plane equations (SetupRoutine.cpp:296-548).
  // The rotated vertices are picked by selects, not by indexing the
  // per-vertex arrays with a run-time index (which would place them in
  // scratch memory).
  float w1 = pick3(pos[0][3], pos[1][3], pos[2][3], o1), w2 = pick3(pos[0][3], pos[1][3], pos[2][3], o2);
  float rhw0 = pick3(pw[0], pw[1], pw[2], o0);
  int X0 = pick3(PX[0], PX[1], PX[2], o0), X1 = pick3(PX[0], PX[1], PX[2], o1), X2 = pick3(PX[0], PX[1], PX[2], o2);
  int Y0 = pick3(PY[0], PY[1], PY[2], o0), Y1 = pick3(PY[0], PY[1], PY[2], o1), Y2 = pick3(PY[0], PY[1], PY[2], o2);
  if (P.line) {  // a third point across the line (SetupRoutine.cpp:317-321)
    X2 = wsub(wadd(X1, Y1), Y0);
    Y2 = wsub(wadd(Y1, X0), X1);
  }
  R.x0 = (float)X0 * (1.0f / 16.0f);
  R.y0 = (float)Y0 * (1.0f / 16.0f);
  X1 = wsub(X1, X0); Y1 = wsub(Y1, Y0);
  X2 = wsub(X2, X0); Y2 = wsub(Y2, Y0);
  float x1 = w1 * (1.0f / 16.0f) * (float)X1;
  float y1 = w1 * (1.0f / 16.0f) * (float)Y1;
  float x2 = w2 * (1.0f / 16.0f) * (float)X2;
  float y2 = w2 * (1.0f / 16.0f) * (float)Y2;
  float a = x1 * y2 - x2 * y1;
  float M00 = 0.0f, M01 = 0.0f, M02 = rhw0, M10 = 0.0f, M11 = 0.0f, M20 = 0.0f, M21 = 0.0f;
  if (a != 0.0f) {
    float A = 1.0f / a;
    float D = A * rhw0;
    M00 = (y1 * w2 - y2 * w1) * D;
    M01 = (x2 * w1 - x1 * w2) * D;
    M10 = y2 * A;
    M11 = -x2 * A;
    M20 = -y1 * A;
    M21 = x1 * A;
  }
  R.wA = (M00 + M10) + M20;
  R.wB = (M01 + M11) + M21;
  R.wC = (M02 + 0.0f) + 0.0f;
  if (P.depthTest) {
    float z0 = pick3(pz[0], pz[1], pz[2], o0), z1 = pick3(pz[0], pz[1], pz[2], o1) - z0,
          z2 = pick3(pz[0], pz[1], pz[2], o2) - z0;
    float fx1 = (float)X1 * (1.0f / 16.0f), fy1 = (float)Y1 * (1.0f / 16.0f);
    float fx2 = (float)X2 * (1.0f / 16.0f), fy2 = (float)Y2 * (1.0f / 16.0f);
    if (!P.point) {
      float D = P.depthRange / (fx1 * fy2 - fx2 * fy1);
      R.zA = (fy2 * z1 - fy1 * z2) * D;
      R.zB = (fx1 * z2 - fx2 * z1) * D;
    } else {  // a point's depth is constant (SetupRoutine.cpp:393-409)
      R.zA = 0.0f;
      R.zB = 0.0f;
    }
    R.zC = z0 * P.depthRange + P.depthNear;
    // Depth bias (SetupRoutine.cpp:417-475), floating-point depth buffer:
    // r = 2^(e - 23) for the largest exponent e among the plane constant and
    // z1, z2 mapped through the depth range -- z1, z2 being the differences
    // from z0 formed above, as in the reference.
    float bias = 0.0f;
    if (P.biasConstant) {
      const float Z1 = z1 * P.depthRange + P.depthNear, Z2 = z2 * P.depthRange + P.depthNear;
      const int e0 = __float_as_int(R.zC) & 0x7F800000, e1 = __float_as_int(Z1) & 0x7F800000;
      const int e2 = __float_as_int(Z2) & 0x7F800000;
      const int e = max(max(e0, e1), e2);
      // D16: minimumResolvableDepthDifference (Renderer.cpp:427-430).
      const float r = P.depth16 ? 1.01f / 65535.0f : __int_as_float(e) * (1.0f / (float)(1 << 23));
      bias = r * P.constantDepthBias;
    }
    if (P.biasSlope) {
      const float aA = R.zA > 0.0f ? R.zA : -R.zA, aB = R.zB > 0.0f ? R.zB : -R.zB;  // Reactor's Abs
      const float m = rr_max(aA, aB);
      bias = bias + m * P.slopeDepthBias;
    }
    if ((P.biasConstant || P.biasSlope) && P.biasClamp)
      bias = P.depthBiasClamp > 0.0f ? rr_min(bias, P.depthBiasClamp) : rr_max(bias, P.depthBiasClamp);
    R.zBias = bias;
  } else {
    R.zA = R.zB = R.zC = 0.0f;
    R.zBias = 0.0f;
  }
#pragma unroll
  for (int k = 0; k < 6; k++) {
    float i0 = pick3(attr[0][k], attr[1][k], attr[2][k], o0), i1 = pick3(attr[0][k], attr[1][k], attr[2][k], o1),
          i2 = pick3(attr[0][k], attr[1][k], attr[2][k], o2);
    R.V[k][0] = (i0 * M00 + i1 * M10) + i2 * M20;
    R.V[k][1] = (i0 * M01 + i1 * M11) + i2 * M21;
    R.V[k][2] = (i0 * M02 + i1 * 0.0f) + i2 * 0.0f;
  }

  // Lower bound of z as the rasteriser evaluates it, x*A + (C + y*B) in
  // float, over the bounding box widened by half a pixel (sample offsets):
  // the affine minimum at a corner, less a margin of 8 ulp of the largest
  // term.  Lets a tile skip primitives that cannot pass its depth test.
  float zlo = -__builtin_inff();
  if (P.depthTest && __builtin_isfinite(R.zA) && __builtin_isfinite(R.zB) && __builtin_isfinite(R.zC) &&
      __builtin_isfinite(R.zBias)) {
    const double xa = (double)xMin - 0.5 - (double)R.x0, xb = (double)(xMax - 1) + 0.5 - (double)R.x0;
    const double ya = (double)yMin - 0.5 - (double)R.y0, yb = (double)(yMax - 1) + 0.5 - (double)R.y0;
    const double A = R.zA, B = R.zB, C = R.zC;
    const double Zb = P.depthBias ? (double)R.zBias : 0.0;
    const double m =
        fmin(fmin(A * xa + B * ya, A * xb + B * ya), fmin(A * xa + B * yb, A * xb + B * yb)) + C + Zb;
    const double mag =
        fabs(A) * fmax(fabs(xa), fabs(xb)) + fabs(C) + 2.0 * fabs(B) * fmax(fabs(ya), fabs(yb)) + 2.0 * fabs(Zb);
    const double lo = m - mag * 0x1p-21 - 1e-30;
    if (__builtin_isfinite(lo)) {
      float f = (float)lo;
      if ((double)f > lo) {  // step one float towards -inf
        const uint32_t u = __float_as_uint(f);
        f = (f > 0.0f) ? __uint_as_float(u - 1) : (f < 0.0f ? __uint_as_float(u + 1) : -1.17549435e-38f);
      }
      zlo = f;
    }
  }
  R.zlo = zlo;
  R.yMin = yMin;
  R.yMax = yMax;
  R.xMin = xMin;
  R.xMax = xMax;
  if (!isPolygon) {
#pragma unroll
    for (int i = 0; i < 3; i++) { R.X[i] = PX[i]; R.Y[i] = PY[i]; }
  } else if (n == 3) {
    for (int i = 0; i < 3; i++) { R.X[i] = X[i]; R.Y[i] = Y[i]; }
  } else {
    int *pp = poly + g * (2 * kMaxPolygon);
    for (int i = 0; i < n; i++) { pp[i] = X[i]; pp[kMaxPolygon + i] = Y[i]; }
  }
  R.nflags = n | (d << 8) | (frontFacing ? 1 << 9 : 0);
  // Tiles are 8 or 16 pixels a side and xMin, yMin >= 0 (scissor-clamped to
  // a render area at x, y >= 0): shifts divide exactly.
  const int sw = __builtin_ctz((uint32_t)P.tileW), sh = __builtin_ctz((uint32_t)P.tileH);
  int tx0 = xMin >> sw, tx1 = (xMax - 1) >> sw;
  int ty0 = yMin >> sh, ty1 = (yMax - 1) >> sh;
  tileCount[g] = (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1));
  return true;
}

// Polygon-path workspaces per wave of the setup kernels (640 B each).
constexpr int kPolySlots = 16;

// Setup of primitive t (slot g, draw `draw`) by the calling lane, t < 0 for
// none: every lane first runs the setup without the polygon path; the lanes
// whose primitive needs it (clipped triangles, lines, points) then run it
// kPolySlots at a time per wave, each with a workspace in `work` (the
// wave's kPolySlots entries in LDS), so that no thread uses scratch memory.
// Called by every lane of the wave.
__device__ __forceinline__ void setup_lanes(const DrawParams &P, int t, size_t g, int draw, TriRecord *__restrict__ rec,
                                            int *__restrict__ poly, uint32_t *__restrict__ tileCount,
                                            PolyWork *work) {
  const int lane = threadIdx.x & 63;
  bool later = t >= 0 && !setup_one<false>(P, t, g, draw, rec, poly, tileCount, work[0]);
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (uint64_t m = __ballot(later); m; m = __ballot(later)) {  // wave-uniform
    const uint32_t rank = (uint32_t)__builtin_popcountll(m & below);
    const bool mine = later && rank < (uint32_t)kPolySlots;
    if (mine) setup_one<true>(P, t, g, draw, rec, poly, tileCount, work[rank]);
    later = later && !mine;
    // The next round's lanes reuse these workspaces.
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  }
}

// One draw: one thread per primitive.
__global__ __launch_bounds__(256) void setup_kernel(DrawParams P, TriRecord *__restrict__ rec,
                                                    int *__restrict__ poly, uint32_t *__restrict__ tileCount) {
  __shared__ PolyWork work[4][kPolySlots];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  setup_lanes(P, t < P.triCount ? t : -1, (size_t)t, 0, rec, poly, tileCount, work[threadIdx.x >> 6]);
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Tile rasteriser
// ---------------------------------------------------------------------------
// Work decomposition: one workgroup per 16x16 bin, one wavefront per 8x8
// block of the bin, one pixel per lane (lanes 4k..4k+3 form a 2x2 quad).
// Each wave walks the bin's triangle list in API order, 64 entries at a
// time: every lane computes the exact 8x8 coverage mask of one triangle
// (triangle-parallel), a ballot compacts the triangles that cover the
// block, and the wave then applies them one by one (pixel-parallel) to the
// block's depth / stencil / colour, which stay in registers throughout.
//
// Coverage is the reference's span rule evaluated per sample point.  The
// reference walks each edge with an exact integer DDA that stores
// x = ceil(Xe(y) / 16) for rows [ceil(Y1/16), ceil(Y2/16)) into the left or
// right span, later edges overwriting earlier ones (SetupRoutine.cpp:
// 550-621), and a sample is covered iff left <= x < right
// (QuadRasterizer.cpp:181-206).  With
//   E(x, y) = DY * (16 x - X1) - DX * (16 y - Y1),   DY = Y2 - Y1 > 0,
// "ceil(Xe/16) <= x" is exactly E >= 0 and "x < ceil(Xe/16)" is exactly
// E < 0, so the span test becomes two integer sign tests per sample.  E at
// the block origin is formed in 64 bits and clamped to +-2^30: beyond that
// its sign cannot change inside the block, and the 32-bit steps are exact.

struct LaneEdge {
  int e;      // E at the block origin (clamped)
  int ex;     // dE / dx (16 DY > 0)
  int ey;     // dE / dy (-16 DX)
  int y1, y2; // rows the edge spans
  int right;  // writes the right span
};

__device__ __forceinline__ LaneEdge lane_edge(int Xa, int Ya, int Xb, int Yb, int ox, int oy) {
  LaneEdge E;
  const bool swap = Yb < Ya;
  const int X1 = swap ? Xb : Xa, X2 = swap ? Xa : Xb;
  const int Y1 = swap ? Yb : Ya, Y2 = swap ? Ya : Yb;
  const int DX = X2 - X1, DY = Y2 - Y1;
  E.y1 = (Y1 + 15) >> 4;  // a horizontal edge gets y1 == y2 and spans no row
  E.y2 = (Y2 + 15) >> 4;
  long long v = (long long)DY * (long long)(16 * ox - X1) - (long long)DX * (long long)(16 * oy - Y1);
  const long long lim = 1ll << 30;
  E.e = (int)(v > lim ? lim : (v < -lim ? -lim : v));
  E.ex = 16 * DY;
  E.ey = -16 * DX;
  E.right = swap ? 1 : 0;
  return E;
}

// Coverage mask (bit r*8+c = pixel (ox+c, oy+r)) of polygon (X, Y, n) for
// the sample at 28.4 offset (xo, yo), before the scissor mask is applied.
template <typename VFetch>
__device__ __forceinline__ uint64_t block_coverage(const VFetch &vtx, int n, int d, int xo, int yo, int ox, int oy) {
  int Lv[8], Lx[8], Rv[8], Rx[8];
#pragma unroll
  for (int r = 0; r < 8; r++) { Lv[r] = -1; Lx[r] = 0; Rv[r] = 0; Rx[r] = 0; }
  for (int i = 0; i < n; i++) {
    int ia = i + 1 - d, ib = i + d;
    ia = ia >= n ? ia - n : ia;
    ib = ib >= n ? ib - n : ib;
    int Xa, Ya, Xb, Yb;
    vtx(ia, Xa, Ya);
    vtx(ib, Xb, Yb);
    const LaneEdge E = lane_edge(Xa - xo, Ya - yo, Xb - xo, Yb - yo, ox, oy);
    // Rows of the block the edge spans, as a bit mask, routed to the left or
    // right span; each row then takes (E, dE/dx) under a select mask
    // (bitfield insert), so the update has no branches.
    const int r1 = min(max(E.y1 - oy, 0), 8), r2 = min(max(E.y2 - oy, 0), 8);
    const uint32_t rows = (0xFFu << r1) & ~(0xFFu << r2) & 0xFFu;
    const uint32_t rowsR = E.right ? rows : 0u, rowsL = rows ^ rowsR;
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const int v = E.e + E.ey * r;
      const uint32_t ml = 0u - ((rowsL >> r) & 1u), mr = 0u - ((rowsR >> r) & 1u);
      Lv[r] = (int)(((uint32_t)v & ml) | ((uint32_t)Lv[r] & ~ml));
      Lx[r] = (int)(((uint32_t)E.ex & ml) | ((uint32_t)Lx[r] & ~ml));
      Rv[r] = (int)(((uint32_t)v & mr) | ((uint32_t)Rv[r] & ~mr));
      Rx[r] = (int)(((uint32_t)E.ex & mr) | ((uint32_t)Rx[r] & ~mr));
    }
  }
  uint64_t mask = 0;
#pragma unroll
  for (int r = 0; r < 8; r++) {
    // Covered iff lv >= 0 && rv < 0, i.e. the sign bit of (rv & ~lv).
    int lv = Lv[r] + 7 * Lx[r], rv = Rv[r] + 7 * Rx[r];
    uint32_t bits = 0;
#pragma unroll
    for (int c = 7; c >= 0; c--) {
      bits = (bits << 1) | ((uint32_t)(rv & ~lv) >> 31);
      lv -= Lx[r];
      rv -= Rx[r];
    }
    mask |= (uint64_t)bits << (8 * r);
  }
  return mask;
}

template <int MS, bool FCOL>
struct PixelState {
  float zb[MS];
  uint32_t sb[MS];
  uint32_t cb[FCOL ? 1 : MS];
  float4 cf[FCOL ? MS : 1];
  uint32_t id[FCOL ? MS : 1];  // primitive_id record (float targets only)
};

// Pixel of lane `l` inside its 8x8 block: lanes 4k..4k+3 form the 2x2 quad
// k, quads in row-major order over the block.
__device__ __forceinline__ int lane_cx(int l) { return 2 * ((l >> 2) & 3) + (l & 1); }
__device__ __forceinline__ int lane_cy(int l) { return 2 * (l >> 4) + ((l >> 1) & 1); }

__device__ __forceinline__ float rdl(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

// Interpolated attribute k at (xf, yf): (x*A + (C + y*B)) * rhw, the
// reference's operation order (QuadRasterizer.cpp:235-250).
__device__ __forceinline__ float interp(const TriRecord &R, int k, float xf, float yf, float rhw) {
  return (xf * R.V[k][0] + (R.V[k][2] + yf * R.V[k][1])) * rhw;
}

__device__ __forceinline__ float quad_rhw(const TriRecord &R, float xf, float yf) {
  const float Dw = R.wC + yf * R.wB;
  return 1.0f / (xf * R.wA + Dw);
}

// Texture LOD state of the 2x2 quad holding pixel (px, py): the lane
// differences of u, v over the quad's pixels (0,0), (1,0), (0,1), helper
// lanes included, evaluated from the plane equations (computeLod2D,
// SamplerCore.cpp:1376-1422).
template <int FA = -1>
__device__ __forceinline__ QuadLod quad_state(const DrawParams &P, const TriRecord &R, int px, int py) {
  const int qx = px & ~1, qy = py & ~1;
  float uq[4], vq[4];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const float xi = (float)(qx + (i & 1)) - R.x0, yi = (float)(qy + (i >> 1)) - R.y0;
    const float ri = quad_rhw(R, xi, yi);
    uq[i] = interp(R, 4, xi, yi, ri);
    vq[i] = interp(R, 5, xi, yi, ri);
  }
  uq[3] = 0.0f;
  vq[3] = 0.0f;
  return quad_lod<FA>(P.tex, uq, vq);
}

// Shades one queued fragment: the fixed shader family at pixel (px, py)
// of primitive R, with the texture LOD of the 2x2 quad the pixel belongs
// to (helper lanes evaluated from the plane equations).  Output clamped
// for UNORM targets (PixelProgram::clampColor).
// FA >= 0: the draw's shader is MODULATE and its sampler state the common
// one with address mode FA (tex_filter and the other readers of the sampler
// state).
template <bool FCOL, bool FTEX, bool XF, int FA = -1>
__device__ __forceinline__ float4 shade(const DrawParams &P, const TriRecord &R, int px, int py,
                                        bool clampUnorm = true) {
  const int shader = FA >= 0 ? (int)SHADER_MODULATE : P.shader;
  const float xf = (float)px - R.x0, yf = (float)py - R.y0;
  const float rhw = quad_rhw(R, xf, yf);
  float4 col = make_float4(0, 0, 0, 0);
  if (shader != SHADER_SAMPLE) {
    col.x = interp(R, 0, xf, yf, rhw);
    col.y = interp(R, 1, xf, yf, rhw);
    col.z = interp(R, 2, xf, yf, rhw);
    col.w = interp(R, 3, xf, yf, rhw);
  }
  if (shader != SHADER_INTERPOLATE) {
    const float u = interp(R, 4, xf, yf, rhw), v = interp(R, 5, xf, yf, rhw);
    const QuadLod Lq = quad_state<FA>(P, R, px, py);
    float4 c = sample_lane<FTEX, XF, FA>(P.tex, u, v, Lq);
    if (shader == SHADER_MODULATE) {
      c.x = c.x * col.x;
      c.y = c.y * col.y;
      c.z = c.z * col.z;
      c.w = c.w * col.w;
    }
    col = c;
  }
  if (!FCOL && clampUnorm) {
    col.x = rr_min(rr_max(col.x, 0.0f), 1.0f);
    col.y = rr_min(rr_max(col.y, 0.0f), 1.0f);
    col.z = rr_min(rr_max(col.z, 0.0f), 1.0f);
    col.w = rr_min(rr_max(col.w, 0.0f), 1.0f);
  }
  return col;
}

// Blend a shaded colour into the samples of this lane's pixel selected by
// xMask, then write (PixelRoutine.cpp:1617-2936).
// FAST: blending, when on, is SRC_ALPHA, ONE_MINUS_SRC_ALPHA, ADD.
template <int MS, bool FCOL, bool XF, bool FAST = false>
__device__ __forceinline__ void blend_write(const DrawParams &P, PixelState<MS, FCOL> &S, float4 s, int xMask,
                                            uint32_t tri) {
#pragma unroll
  for (int q = 0; q < MS; q++) {
    if (!((xMask >> q) & 1)) continue;
    float4 dst;
    if (FCOL) {
      dst = S.cf[FCOL ? q : 0];
    } else {
      const uint32_t p = S.cb[FCOL ? 0 : q];
      dst = make_float4(unorm8_to_float(p & 0xFFu), unorm8_to_float((p >> 8) & 0xFFu),
                        unorm8_to_float((p >> 16) & 0xFFu), unorm8_to_float(p >> 24));
      if ((XF && P.colorSRGB) && P.blendEnable) {  // blended in linear space (PixelRoutine.cpp:1811-1827)
        dst.x = srgb_to_linear(dst.x);
        dst.y = srgb_to_linear(dst.y);
        dst.z = srgb_to_linear(dst.z);
      }
    }
    float4 o = P.blendEnable ? (FAST ? blend_src_alpha(s, dst) : alpha_blend<XF>(P, s, dst)) : s;
    if (!FCOL && (XF && P.colorSRGB)) {  // writeColor encodes R, G, B first (PixelRoutine.cpp:1963-1969)
      o.x = linear_to_srgb(o.x);
      o.y = linear_to_srgb(o.y);
      o.z = linear_to_srgb(o.z);
    }
    // Masked channels keep their value (selects, not branches).
    if (FCOL) {
      float4 &c = S.cf[FCOL ? q : 0];
      c.x = (P.colorMask & 1) ? o.x : c.x;
      c.y = (P.colorMask & 2) ? o.y : c.y;
      c.z = (P.colorMask & 4) ? o.z : c.z;
      c.w = (P.colorMask & 8) ? o.w : c.w;
      S.id[FCOL ? q : 0] = tri - (uint32_t)P.triBase + 1u;
    } else {
      const uint32_t wm = ((P.colorMask & 1) ? 0x000000FFu : 0u) | ((P.colorMask & 2) ? 0x0000FF00u : 0u) |
                          ((P.colorMask & 4) ? 0x00FF0000u : 0u) | ((P.colorMask & 8) ? 0xFF000000u : 0u);
      const uint32_t o8 = float_to_unorm8(o.x) | (float_to_unorm8(o.y) << 8) | (float_to_unorm8(o.z) << 16) |
                          (float_to_unorm8(o.w) << 24);
      S.cb[FCOL ? 0 : q] = (S.cb[FCOL ? 0 : q] & ~wm) | (o8 & wm);
    }
  }
}

constexpr int kQueue = 64;  // deferred-shading queue entries per wave

// One screen tile of draw P: `bin` is the tile's index in the draw's tile
// grid, [start, end) its range of the sorted (tile, primitive) list.  The
// workgroup has one wave per 8x8 block of the tile.
// LEVEL 1: the draw has no stencil and no depth bias, tests and writes
// depth and writes all four colour channels (the checks below fold away).
// LEVEL 2 (3): also the common shading state with REPEAT (CLAMP_TO_EDGE)
// addressing (shade's FA, blend_write's FAST).
template <int MS, bool FCOL, bool FTEX, bool XF, int LEVEL>
__device__ __forceinline__ void raster_tile(const DrawParams &P, int bin, uint32_t start, uint32_t end0,
                                            const TriRecord *__restrict__ rec, const int *__restrict__ poly,
                                            const uint32_t *__restrict__ binTri) {
  constexpr bool SIMPLE = LEVEL >= 1, FAST = LEVEL >= 2;
  constexpr int FA = LEVEL >= 2 ? LEVEL - 2 : -1;
  const bool stencilActive = !SIMPLE && P.stencilActive, stencilWrite = !SIMPLE && P.stencilWrite;
  const bool depthBias = !SIMPLE && P.depthBias, depthTest = SIMPLE || P.depthTest;
  const bool depthWrite = SIMPLE || P.depthWrite;
  const int colorMask = SIMPLE ? 15 : P.colorMask;
  // Per-wave queue of fragments that passed the tests and await shading,
  // in API order: primitive, owner lane | sample mask << 8, shaded colour.
  __shared__ uint32_t qTri[4][kQueue];
  __shared__ uint32_t qPix[4][kQueue];
  __shared__ float4 qCol[4][kQueue];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int bx = bin % P.tilesX, by = bin / P.tilesX;
  const int bw = P.tileW >> 3;  // 8x8 blocks per tile row
  const int ox = bx * P.tileW + 8 * (wave % bw), oy = by * P.tileH + 8 * (wave / bw);
  const int cx = lane_cx(lane), cy = lane_cy(lane);
  const int px = ox + cx, py = oy + cy;
  const int myBit = cy * 8 + cx;

  // A block outside the scissor walks no list (no early return: the
  // workgroup reduces its counters together at the end).
  const bool waveActive = !(ox >= P.sx1 || oy >= P.sy1 || ox + 8 <= P.sx0 || oy + 8 <= P.sy0);
  const uint32_t end = waveActive ? end0 : start;

  uint64_t scissorMask = 0;
  {
    uint32_t cols = 0;
#pragma unroll
    for (int c = 0; c < 8; c++)
      if (ox + c >= P.sx0 && ox + c < P.sx1) cols |= 1u << c;
#pragma unroll
    for (int r = 0; r < 8; r++)
      if (oy + r >= P.sy0 && oy + r < P.sy1) scissorMask |= (uint64_t)cols << (8 * r);
  }
  const bool inside = (scissorMask >> myBit) & 1;

  PixelState<MS, FCOL> S;
#pragma unroll
  for (int q = 0; q < MS; q++) {
    S.zb[q] = 0.0f;
    S.sb[q] = 0;
    if (!FCOL) S.cb[FCOL ? 0 : q] = 0;
    if (FCOL) S.cf[FCOL ? q : 0] = make_float4(0, 0, 0, 0);
    if (FCOL) S.id[FCOL ? q : 0] = 0;
    if (!inside) continue;
    if (FCOL && P.primId)
      S.id[FCOL ? q : 0] =
          *reinterpret_cast<const uint32_t *>(P.primId + q * P.primIdSlice + (long long)py * P.primIdPitch + 4ll * px);
    if (depthTest || (XF && P.depthBounds))  // D16 holds the float of its integer (readDepth16, PixelRoutine.cpp:466-481)
      S.zb[q] = (XF && P.depth16) ? (float)*reinterpret_cast<const uint16_t *>(P.depth + q * P.depthSlice +
                                                                        (long long)py * P.depthPitch + 2ll * px)
                          : *reinterpret_cast<const float *>(P.depth + q * P.depthSlice +
                                                             (long long)py * P.depthPitch + 4ll * px);
    if (stencilActive) S.sb[q] = (uint8_t)P.stencil[q * P.stencilSlice + (long long)py * P.stencilPitch + px];
    if (colorMask) {
      if (FCOL)
        S.cf[FCOL ? q : 0] =
            *reinterpret_cast<const float4 *>(P.color + q * P.colorSlice + (long long)py * P.colorPitch + 16ll * px);
      else
        S.cb[FCOL ? 0 : q] = bgra_swap(
            (XF && P.colorBGRA),
            *reinterpret_cast<const uint32_t *>(P.color + q * P.colorSlice + (long long)py * P.colorPitch + 4ll * px));
    }
  }

  // Hierarchical depth: a primitive whose evaluated depth is bounded below
  // by zlo cannot pass LESS (LESS_OR_EQUAL) anywhere in a block whose
  // largest stored depth is <= zlo (< zlo).  Only when a failing fragment
  // has no side effect (no stencil writes).
  const bool hiz = depthTest && !stencilWrite && (P.depthCompare == 1 || P.depthCompare == 3);

  unsigned long long occl = 0;
  uint32_t nTests = 0, nCovers = 0, nShaded = 0;  // statistics (profiling only)
  int qCount = 0;           // wave-uniform queue length
  uint64_t myEntries = 0;   // queue slots holding fragments of this lane's pixel

  auto flush = [&]() __attribute__((always_inline)) {
    // Shade the queued fragments, one per lane...
    if (lane < qCount) {
      const uint32_t t = qTri[wave][lane];
      const uint32_t pix = qPix[wave][lane];
      const int l = (int)(pix & 63u);
      qCol[wave][lane] = shade<FCOL, FTEX, XF, FA>(P, rec[t], ox + lane_cx(l), oy + lane_cy(l));
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    // ...then blend them into their pixels in queue (API) order.
    uint64_t m = myEntries;
    while (m) {
      const int k = __builtin_ctzll(m);
      m &= m - 1;
      blend_write<MS, FCOL, XF, FAST>(P, S, qCol[wave][k], (int)(qPix[wave][k] >> 8), qTri[wave][k]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    myEntries = 0;
    qCount = 0;
  };

  // Largest stored depth of the block (+inf when any is NaN): a primitive
  // whose depth is bounded below by zl cannot pass where zl >= it (LESS) or
  // zl > it (LESS_OR_EQUAL).
  auto block_zmax = [&]() __attribute__((always_inline)) {
    float v = -__builtin_inff();
    bool nan = false;
    if (inside) {
#pragma unroll
      for (int q = 0; q < MS; q++) {
        v = fmaxf(v, S.zb[q]);
        nan |= S.zb[q] != S.zb[q];
      }
    }
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return __ballot(nan) ? __builtin_inff() : v;
  };

  // One batch of up to 64 candidates in list order: each lane computes the
  // exact coverage of its primitive; then the covering primitives are
  // applied in list order.  A candidate waited in the ring while earlier
  // ones wrote depth, so its depth bound zl is tested again first.
  auto process_chunk = [&](bool act, uint32_t tid, float zl) __attribute__((always_inline)) {
    if (hiz) {
      const float zmax = block_zmax();
      act = act && !(P.depthCompare == 1 ? zl >= zmax : zl > zmax);
    }
    // ---- triangle-parallel coverage of this block ----
    uint32_t mlo[MS], mhi[MS];
    int flags = 0;
    float x0 = 0, y0 = 0, zA = 0, zB = 0, zC = 0, zBias = 0;
#pragma unroll
    for (int q = 0; q < MS; q++) { mlo[q] = 0; mhi[q] = 0; }
    if (act) {
      const int4 *rp = reinterpret_cast<const int4 *>(rec + tid);
      const int4 h0 = rp[0];  // nflags, yMin, yMax, xMin
      const int4 h1 = rp[1];  // xMax, X0, X1, X2
      const int4 h2 = rp[2];  // Y0, Y1, Y2, zlo
      const float4 h3 = reinterpret_cast<const float4 *>(rp)[3];  // x0, y0, zA, zB
      const float4 h4 = reinterpret_cast<const float4 *>(rp)[4];  // zC, zBias, wA, wB
      const int n = h0.x & 0xFF, d = (h0.x >> 8) & 1;
      flags = h0.x >> 8;
      x0 = h3.x; y0 = h3.y; zA = h3.z; zB = h3.w; zC = h4.x; zBias = h4.y;
      const int *pp = poly + (size_t)tid * (2 * kMaxPolygon);
      auto vtx = [&](int i, int &Xo, int &Yo) {
        if (n == 3) {
          // Mask selects: a conditional pick among the components of one
          // vector would become an indexed access through scratch memory.
          const int m0 = -(int)(i == 0), m1 = -(int)(i == 1), m2 = -(int)(i == 2);
          Xo = (h1.y & m0) | (h1.z & m1) | (h1.w & m2);
          Yo = (h2.x & m0) | (h2.y & m1) | (h2.z & m2);
        } else {
          Xo = pp[i];
          Yo = pp[kMaxPolygon + i];
        }
      };
#pragma unroll
      for (int q = 0; q < MS; q++) {
        if (!((P.msMask >> q) & 1)) continue;
        const int xo = (MS > 1) ? kXf[q] : 0, yo = (MS > 1) ? kYf[q] : 0;
        const uint64_t m = block_coverage(vtx, n, d, xo, yo, ox, oy) & scissorMask;
        mlo[q] = (uint32_t)m;
        mhi[q] = (uint32_t)(m >> 32);
      }
    }
    uint32_t anyBits = 0;
#pragma unroll
    for (int q = 0; q < MS; q++) anyBits |= mlo[q] | mhi[q];
    uint64_t todo = __ballot(anyBits != 0);
    nCovers += anyBits != 0 ? 1u : 0u;

    // ---- pixel-parallel tests, in list order ----
    while (todo) {
      const int j = __builtin_ctzll(todo);
      todo &= todo - 1;
      int cMask = 0;  // bit q: sample q of this lane's pixel is covered
#pragma unroll
      for (int q = 0; q < MS; q++) {
        const uint32_t lo = __builtin_amdgcn_readlane(mlo[q], j), hi = __builtin_amdgcn_readlane(mhi[q], j);
        const uint32_t word = myBit < 32 ? lo : hi;
        cMask |= (int)((word >> (myBit & 31)) & 1u) << q;
      }
      const int fl = __builtin_amdgcn_readlane(flags, j);
      const StencilParams &SP = (fl & 2) ? P.front : P.back;

      int sMask = cMask, zMask = cMask;
      if (stencilActive && cMask) {
#pragma unroll
        for (int q = 0; q < MS; q++)
          if (((cMask >> q) & 1) && !stencil_compare(SP, S.sb[q])) sMask &= ~(1 << q);
      }
      if (XF && P.alphaToCoverage) {
        // The reference shades before the depth test and keeps sample q
        // only where !(alpha < a2c[q]) (alphaTest / alphaToCoverage,
        // PixelProgram.cpp:243-259, PixelRoutine.cpp:317-337, 643-657); the
        // alpha is clamped when the colour attachment is UNORM.  The
        // fragment is shaded again, identically, if it reaches the queue.
        if (cMask) {
          const uint32_t tj = __builtin_amdgcn_readlane(tid, j);
          const float a = shade<FCOL, FTEX, XF, FA>(P, rec[tj], px, py, P.color != nullptr).w;
          int aMask = 0;
#pragma unroll
          for (int q = 0; q < MS; q++)
            if (!(a < P.a2c[q])) aMask |= 1 << q;
          cMask &= aMask;
        }
        sMask &= cMask;
        zMask &= cMask;
      }
      float zq[MS];
      if (depthTest) {
        const float xf = (float)px - rdl(x0, j);  // xFragment (PixelRoutine.cpp:125)
        const float yf = (float)py - rdl(y0, j);  // yFragment (QuadRasterizer.cpp:126)
        const float a = rdl(zA, j), b = rdl(zB, j), c = rdl(zC, j);
        const float zbias = depthBias ? rdl(zBias, j) : 0.0f;
        zMask = 0;
#pragma unroll
        for (int q = 0; q < MS; q++) {
          const float yy = (MS > 1) ? yf + kSampleY[q] : yf;
          const float xx = (MS > 1) ? xf - kSampleX[q] : xf;
          const float Dz = c + yy * b;
          float z = xx * a + Dz;
          if (depthBias) z = z + zbias;  // PixelRoutine.cpp:141-144, before the clamp
          z = rr_min(rr_max(z, P.depthClampLo), P.depthClampHi);  // clampDepth
          if ((XF && P.depth16)) z = d16_value(z);
          zq[q] = z;
          const float zb = S.zb[q];
          // The compare op without branches: one of zb <= z, zb < z, zb == z
          // or true, inverted for LESS, LESS_OR_EQUAL, NOT_EQUAL and NEVER.
          const bool le = zb <= z, lt = zb < z, eq = zb == z;
          const bool pr = P.zPred == 0 ? le : (P.zPred == 1 ? lt : (P.zPred == 2 ? eq : true));
          if (pr != (P.zInv != 0)) zMask |= 1 << q;
        }
        zMask &= cMask;
        if (stencilActive) zMask &= sMask;
      }
      if (XF && P.depthBounds && cMask) {
        // depthBoundsTest (PixelRoutine.cpp:576-641) on the stored depth, D16
        // as its code / 0xFFFF. With the depth test off the coverage loses
        // the sample while zMask keeps it, so the occlusion count does too.
        int bMask = 0;
#pragma unroll
        for (int q = 0; q < MS; q++) {
          const float zs = P.depth16 ? S.zb[q] * (1.0f / 65535.0f) : S.zb[q];
          if (P.depthBoundsMin <= zs && zs <= P.depthBoundsMax) bMask |= 1 << q;
        }
        if (depthTest) zMask &= bMask;
        else cMask &= bMask;
      }
      if (stencilWrite && cMask) {
#pragma unroll
        for (int q = 0; q < MS; q++) {
          if (!((cMask >> q) & 1)) continue;
          int op;
          if (!((sMask >> q) & 1)) op = SP.failOp;
          else if (depthTest && !((zMask >> q) & 1)) op = SP.depthFailOp;
          else op = SP.passOp;
          const uint32_t bv = S.sb[q];
          uint32_t nv = stencil_apply_op(bv, op, SP.reference);
          if (SP.useWriteMask) nv = (nv & SP.writeMask) | (bv & ~SP.writeMask);
          S.sb[q] = nv & 0xFFu;
        }
      }
      if (depthWrite) {
#pragma unroll
        for (int q = 0; q < MS; q++)
          if ((zMask >> q) & 1) S.zb[q] = zq[q];
      }
      occl += __builtin_popcount((unsigned)(zMask & sMask));
      if (!colorMask) continue;
      int xMask = depthTest ? zMask : cMask;
      if (stencilActive) xMask &= sMask;
      const uint64_t need = __ballot(xMask != 0);
      if (!need) continue;
      // Queue the fragments that write colour; shading is deferred so that
      // it runs with every lane busy.
      const int cnt = __builtin_popcountll(need);
      if (qCount + cnt > kQueue) flush();
      if (xMask) {
        const int pos = qCount + __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0));
        qTri[wave][pos] = __builtin_amdgcn_readlane(tid, j);
        qPix[wave][pos] = (uint32_t)lane | ((uint32_t)xMask << 8);
        myEntries |= 1ull << pos;
        nShaded++;
      }
      qCount += cnt;
    }
  };

  // Candidates (box meets the block, hierarchical depth test passed) are
  // compacted, in list order, into a per-wave ring, and the exact coverage
  // runs on full batches of 64: a 64-entry chunk of the list yields about a
  // third of a wave of candidates.
  __shared__ uint32_t ring[4][128];
  __shared__ float ringZ[4][128];
  uint32_t ringHead = 0, ringCount = 0;  // wave-uniform

  for (uint32_t base = start; base < end; base += 64) {
    // ---- block depth bound for the hierarchical test ----
    const float zmax = hiz ? block_zmax() : __builtin_inff();
    // ---- candidate test: box overlap, then the hierarchical depth bound ----
    const uint32_t e = base + lane;
    bool test = false;
    uint32_t tid = 0;
    float zlo = -__builtin_inff();
    // A binned pair carries the blocks of its tile that the primitive's
    // edges may reach (bits 28-31); the box test completes the overlap test.
    bool reach = true;
    if (e < end) {
      if (binTri) {
        const uint32_t v = binTri[e];
        tid = v & kSlotMask;
        reach = (v >> (28 + wave)) & 1u;
      } else {
        tid = e;
      }
    }
    if (e < end && reach) {
      const int4 *rp = reinterpret_cast<const int4 *>(rec + tid);
      const int4 h0 = rp[0];  // nflags, yMin, yMax, xMin
      const int4 h1 = rp[1];  // xMax, ...
      test = h0.y < oy + 8 && h0.z > oy && h0.w < ox + 8 && h1.x > ox;
      if (hiz && test) {
        const float zlo0 = __int_as_float(rp[2].w);                  // zlo
        const float4 h3 = reinterpret_cast<const float4 *>(rp)[3];  // x0, y0, zA, zB
        const float4 h4 = reinterpret_cast<const float4 *>(rp)[4];  // zC, zBias, ...
        const float x0 = h3.x, y0 = h3.y, zA = h3.z, zB = h3.w, zC = h4.x;
        const float zBias = depthBias ? h4.y : 0.0f;
        // Lower bound of the evaluated plane over this block (sample offsets
        // included), tightened with the primitive's bound over its box.
        const float xa = (float)ox - 0.5f - x0, xb = (float)(ox + 7) + 0.5f - x0;
        const float ya = (float)oy - 0.5f - y0, yb = (float)(oy + 7) + 0.5f - y0;
        const float m =
            fminf(fminf(zA * xa + zB * ya, zA * xb + zB * ya), fminf(zA * xa + zB * yb, zA * xb + zB * yb)) + zC +
            zBias;
        const float mag = fabsf(zA) * fmaxf(fabsf(xa), fabsf(xb)) + fabsf(zC) +
                          2.0f * fabsf(zB) * fmaxf(fabsf(ya), fabsf(yb)) + 2.0f * fabsf(zBias);
        float zl = fmaxf(zlo0, m - mag * 0x1p-19f);
        if (!(zl == zl)) zl = -__builtin_inff();
        zl = fminf(zl, P.depthClampHi);
        if ((XF && P.depth16)) zl = d16_value(fmaxf(zl, 0.0f));  // the bound in stored units (monotone map)
        test = !(P.depthCompare == 1 ? zl >= zmax : zl > zmax);
        zlo = zl;
      }
    }
    nTests += test ? 1u : 0u;
    const uint64_t tb = __ballot(test);
    if (test) {
      const uint32_t pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(tb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)tb, 0));
      ring[wave][(ringHead + ringCount + pos) & 127] = tid;
      ringZ[wave][(ringHead + ringCount + pos) & 127] = zlo;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    ringCount += (uint32_t)__builtin_popcountll(tb);
    if (ringCount >= 64) {
      const uint32_t t2 = ring[wave][(ringHead + lane) & 127];
      const float z2 = ringZ[wave][(ringHead + lane) & 127];
      ringHead = (ringHead + 64) & 127;
      ringCount -= 64;
      process_chunk(true, t2, z2);
    }
  }
  while (ringCount > 0) {
    const uint32_t nb = ringCount < 64 ? ringCount : 64;
    const bool act = (uint32_t)lane < nb;
    const uint32_t t2 = act ? ring[wave][(ringHead + lane) & 127] : 0u;
    const float z2 = act ? ringZ[wave][(ringHead + lane) & 127] : -__builtin_inff();
    ringHead = (ringHead + nb) & 127;
    ringCount -= nb;
    process_chunk(act, t2, z2);
  }
  if (qCount) flush();

  // Write the block back.
  if (inside) {
#pragma unroll
    for (int q = 0; q < MS; q++) {
      if (P.depthWrite) {
        if ((XF && P.depth16))
          *reinterpret_cast<uint16_t *>(P.depth + q * P.depthSlice + (long long)py * P.depthPitch + 2ll * px) =
              (uint16_t)S.zb[q];
        else
          *reinterpret_cast<float *>(P.depth + q * P.depthSlice + (long long)py * P.depthPitch + 4ll * px) = S.zb[q];
      }
      if (P.stencilWrite) P.stencil[q * P.stencilSlice + (long long)py * P.stencilPitch + px] = (char)S.sb[q];
      if (FCOL && P.primId)
        *reinterpret_cast<uint32_t *>(P.primId + q * P.primIdSlice + (long long)py * P.primIdPitch + 4ll * px) =
            S.id[FCOL ? q : 0];
      if (P.colorMask) {
        if (FCOL)
          *reinterpret_cast<float4 *>(P.color + q * P.colorSlice + (long long)py * P.colorPitch + 16ll * px) =
              S.cf[FCOL ? q : 0];
        else
          *reinterpret_cast<uint32_t *>(P.color + q * P.colorSlice + (long long)py * P.colorPitch + 4ll * px) =
              bgra_swap((XF && P.colorBGRA), S.cb[FCOL ? 0 : q]);
      }
    }
  }

  // Counters: reduce over the wave, then over the workgroup, so that one
  // atomic per counter and workgroup reaches the (single) global address.
  __shared__ unsigned long long red[4][4];
  if (P.occlusion || P.stats) {
    unsigned long long v[4] = {occl, nTests, nCovers, nShaded};
#pragma unroll
    for (int k = 0; k < 4; k++)
      for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 4; k++) red[wave][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      const int k = threadIdx.x;
      const int nw = (int)(blockDim.x >> 6);
      unsigned long long t = red[0][k];
      for (int w = 1; w < nw; w++) t += red[w][k];
      if (k == 0) {
        if (P.occlusion && t) atomicAdd(P.occlusion, t);
      } else if (P.stats) {
        atomicAdd(P.stats + (k - 1), t);
      }
    }
  }
}

// A batch of draws: persistent workgroups take global tiles [tileLo, tileHi)
// from an atomic counter (or, with counter == NULL, workgroup b renders tile
// tileLo + b once); tile t belongs to the draw d in [drawLo, drawHi) with
// tileBase[d] <= t < tileBase[d + 1].
//
// binTri == NULL selects the unbinned pass: every tile walks the record
// slots of its draw from *split on (the primitives the binned pass left out
// because their pairs exceeded the planned capacity) in slot order, testing
// each against its blocks as it would a binned list.  It does nothing when
// *split >= segT (everything was binned).  With binTri and `tail`, each tile
// walks its binned list and then those slots, so that one launch does both
// passes (the tiles of one launch never share pixels, so only the order
// within a tile matters).
// Occupancy: 4 waves per SIMD (128 VGPRs) for the single-sample kernels
// (the float-colour ones spill 12 bytes there and still run faster than at
// 3 waves), 3 for the multisample ones.
template <int MS, bool FCOL, bool FTEX, bool XF, int LEVEL>
__global__ __launch_bounds__(256, MS == 1 ? 4 : 3) void raster_batch_kernel(
                                                           const DrawParams *__restrict__ params,
                                                           const uint32_t *__restrict__ tileBase, int drawLo,
                                                           int drawHi, uint32_t tileLo, uint32_t tileHi,
                                                           uint32_t *__restrict__ counter,
                                                           const TriRecord *__restrict__ rec,
                                                           const int *__restrict__ poly,
                                                           const uint32_t *__restrict__ binTri,
                                                           const uint32_t *__restrict__ tileStart,
                                                           const uint32_t *__restrict__ tileEnd,
                                                           const uint32_t *__restrict__ split, uint32_t segT,
                                                           int tail) {
  __shared__ uint32_t sTile;
  const uint32_t from = (binTri && !tail) ? segT : *split;
  if (!binTri && from >= segT) return;
  // Passes per tile: 0 = the binned list, 1 = the slots from *split on.
  const int pass0 = binTri ? 0 : 1, pass1 = from < segT ? 2 : 1;
  for (uint32_t iter = 0;; iter++) {
    uint32_t t;
    if (counter) {
      if (threadIdx.x == 0) sTile = tileLo + atomicAdd(counter, 1u);
      __syncthreads();
      t = __builtin_amdgcn_readfirstlane(sTile);
      __syncthreads();
    } else {
      if (iter) break;
      t = tileLo + blockIdx.x;
    }
    if (t >= tileHi) break;
    int lo = drawLo, hi = drawHi;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (tileBase[mid] <= t) lo = mid;
      else hi = mid;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
#pragma nounroll
    for (int pass = pass0; pass < pass1; pass++) {
      uint32_t s, e;
      if (pass == 0) {
        s = tileStart[t];
        e = tileEnd[t];
      } else {
        const uint32_t b = (uint32_t)params[lo].triBase;
        s = from > b ? from : b;
        e = b + (uint32_t)params[lo].triCount;
      }
      if (s < e)
        raster_tile<MS, FCOL, FTEX, XF, LEVEL>(params[lo], (int)(t - tileBase[lo]), s, e, rec, poly,
                                               pass == 0 ? binTri : nullptr);
    }
  }
}

// ---------------------------------------------------------------------------
// MSAA resolve (Blitter.cpp:2076-2200 for RGBA8, :1524-1545 for float)
// ---------------------------------------------------------------------------
// Fills a rectangle of every sample plane with one BYTES-wide value (the
// packed clear value).  Grid: x over the rectangle's width, y over its
// rows, z over the sample planes.
template <int BYTES>
__global__ void clear_rect_kernel(char *base, int rowPitch, long long slice, int x0, int y0, int w, int h,
                                  uint4 value) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w || y >= h) return;
  char *p = base + (long long)blockIdx.z * slice + (long long)(y0 + y) * rowPitch + (long long)BYTES * (x0 + x);
  if (BYTES == 16) *reinterpret_cast<uint4 *>(p) = value;
  else if (BYTES == 4) *reinterpret_cast<uint32_t *>(p) = value.x;
  else if (BYTES == 2) *reinterpret_cast<uint16_t *>(p) = (uint16_t)value.x;
  else *reinterpret_cast<uint8_t *>(p) = (uint8_t)value.x;
}

// Up to kClearBatch clears of one element size in one launch (a kernel
// argument, so nothing is uploaded): rect k owns blocks [start[k],
// start[k + 1]), each block kClearPixels consecutive pixels of the rect's
// pixels in (sample, row, column) order, kClearPixels / 256 per thread.
constexpr int kClearBatch = 24;
constexpr int kClearPixels = 1024;
struct ClearRect {
  char *base;
  long long slice;
  int pitch, x0, y0, w, h, samples, bytes;
  uint4 value;
};
struct ClearBatch {
  ClearRect r[kClearBatch];
  uint32_t start[kClearBatch + 1];
  int n;
};

template <int BYTES>
__global__ void clear_batch_kernel(ClearBatch cb) {
  const uint32_t b = blockIdx.x;
  int k = 0;
  while (k + 1 < cb.n && cb.start[k + 1] <= b) k++;
  const ClearRect &r = cb.r[k];
  const uint64_t plane = (uint64_t)r.w * (uint64_t)r.h, total = plane * (uint64_t)r.samples;
  const uint64_t first = (uint64_t)(b - cb.start[k]) * kClearPixels;
#pragma unroll
  for (int j = 0; j < kClearPixels / 256; j++) {
    const uint64_t i = first + (uint64_t)threadIdx.x + 256u * (uint64_t)j;
    if (i >= total) break;
    const uint64_t q = i / plane, rem = i - q * plane;
    const uint32_t y = (uint32_t)(rem / (uint64_t)r.w), x = (uint32_t)(rem - (uint64_t)y * (uint64_t)r.w);
    char *p = r.base + (long long)q * r.slice + (long long)(r.y0 + (int)y) * r.pitch + (long long)BYTES * (r.x0 + (int)x);
    if (BYTES == 16) *reinterpret_cast<uint4 *>(p) = r.value;
    else if (BYTES == 4) *reinterpret_cast<uint32_t *>(p) = r.value.x;
    else if (BYTES == 2) *reinterpret_cast<uint16_t *>(p) = (uint16_t)r.value.x;
    else *reinterpret_cast<uint8_t *>(p) = (uint8_t)r.value.x;
  }
}

// 4x resolve of an sRGB attachment: the reference's blit routine (Blitter.cpp
// 1505-1545, 1427-1484, 676-690) decodes each sample to linear, averages in
// linear space and encodes the average: per R, G, B channel
// byte * (1 / 255) -> sRGBtoLinear -> * 255, summed in sample order, * 0.25,
// * (1 / 255) -> linearToSRGB -> * 255, rounded to nearest even and
// saturated; alpha takes the same path without the conversions.
__global__ void resolve_srgb8_kernel(const char *src, int srcPitch, long long slice, char *dst, int dstPitch, int w,
                                     int h) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w || y >= h) return;
  uint32_t s[4];
#pragma unroll
  for (int q = 0; q < 4; q++) s[q] = *reinterpret_cast<const uint32_t *>(src + q * slice + (long long)y * srcPitch + 4ll * x);
  const float un = 1.0f / 255.0f;
  uint32_t o = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    float acc = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      float v = (float)((s[q] >> (8 * k)) & 0xFFu) * un;
      if (k < 3) v = srgb_to_linear(v);
      v = v * 255.0f;
      acc = q == 0 ? v : acc + v;
    }
    float c = acc * 0.25f;
    c = c * un;
    if (k < 3) c = linear_to_srgb(c);
    c = c * 255.0f;
    const int32_t i = rr_round_int(c);
    o |= (uint32_t)(i < 0 ? 0 : (i > 255 ? 255 : i)) << (8 * k);
  }
  *reinterpret_cast<uint32_t *>(dst + (long long)y * dstPitch + 4ll * x) = o;
}

__global__ void resolve_rgba8_kernel(const char *src, int srcPitch, long long slice, char *dst, int dstPitch, int w,
                                     int h) {
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w || y >= h) return;
  uint32_t s[4];
#pragma unroll
  for (int q = 0; q < 4; q++) s[q] = *reinterpret_cast<const uint32_t *>(src + q * slice + (long long)y * srcPitch + 4ll * x);
  uint32_t o = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint32_t a = (s[0] >> (8 * k)) & 0xFF, b = (s[1] >> (8 * k)) & 0xFF, c = (s[2] >> (8 * k)) & 0xFF,
             dd = (s[3] >> (8 * k)) & 0xFF;
    uint32_t ab = (a + b + 1) >> 1, cd = (c + dd + 1) >> 1;
    o |= ((ab + cd + 1) >> 1) << (8 * k);
  }
  *reinterpret_cast<uint32_t *>(dst + (long long)y * dstPitch + 4ll * x) = o;
}

__global__ void resolve_rgba32f_kernel(const char *src, int srcPitch, long long slice, char *dst, int dstPitch, int w,
                                       int h) {
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w || y >= h) return;
  float4 s[4];
#pragma unroll
  for (int q = 0; q < 4; q++) s[q] = *reinterpret_cast<const float4 *>(src + q * slice + (long long)y * srcPitch + 16ll * x);
  float4 o;
  o.x = (((s[0].x + s[1].x) + s[2].x) + s[3].x) * 0.25f;
  o.y = (((s[0].y + s[1].y) + s[2].y) + s[3].y) * 0.25f;
  o.z = (((s[0].z + s[1].z) + s[2].z) + s[3].z) * 0.25f;
  o.w = (((s[0].w + s[1].w) + s[2].w) + s[3].w) * 0.25f;
  *reinterpret_cast<float4 *>(dst + (long long)y * dstPitch + 16ll * x) = o;
}


#define SWR_RASTER_INSTANCE(MS, FCOL, FTEX, XF, LEVEL)                                                           \
  template __global__ void raster_batch_kernel<MS, FCOL, FTEX, XF, LEVEL>(                                      \
      const DrawParams *, const uint32_t *, int, int, uint32_t, uint32_t, uint32_t *, const TriRecord *, const int *, \
      const uint32_t *, const uint32_t *, const uint32_t *, const uint32_t *, uint32_t, int);
SWR_RASTER_INSTANCE(1, false, false, false, 0)
SWR_RASTER_INSTANCE(1, false, true, false, 0)
SWR_RASTER_INSTANCE(1, true, false, false, 0)
SWR_RASTER_INSTANCE(1, true, true, false, 0)
SWR_RASTER_INSTANCE(4, false, false, false, 0)
SWR_RASTER_INSTANCE(4, false, true, false, 0)
SWR_RASTER_INSTANCE(4, true, false, false, 0)
SWR_RASTER_INSTANCE(4, true, true, false, 0)
SWR_RASTER_INSTANCE(1, false, false, true, 0)
SWR_RASTER_INSTANCE(1, false, true, true, 0)
SWR_RASTER_INSTANCE(1, true, false, true, 0)
SWR_RASTER_INSTANCE(1, true, true, true, 0)
SWR_RASTER_INSTANCE(4, false, false, true, 0)
SWR_RASTER_INSTANCE(4, false, true, true, 0)
SWR_RASTER_INSTANCE(4, true, false, true, 0)
SWR_RASTER_INSTANCE(4, true, true, true, 0)
SWR_RASTER_INSTANCE(1, false, false, false, 1)
SWR_RASTER_INSTANCE(1, false, true, false, 1)
SWR_RASTER_INSTANCE(1, true, false, false, 1)
SWR_RASTER_INSTANCE(1, true, true, false, 1)
SWR_RASTER_INSTANCE(4, false, false, false, 1)
SWR_RASTER_INSTANCE(4, false, true, false, 1)
SWR_RASTER_INSTANCE(4, true, false, false, 1)
SWR_RASTER_INSTANCE(4, true, true, false, 1)
SWR_RASTER_INSTANCE(1, false, false, false, 2)
SWR_RASTER_INSTANCE(1, false, true, false, 2)
SWR_RASTER_INSTANCE(1, true, false, false, 2)
SWR_RASTER_INSTANCE(1, true, true, false, 2)
SWR_RASTER_INSTANCE(4, false, false, false, 2)
SWR_RASTER_INSTANCE(4, false, true, false, 2)
SWR_RASTER_INSTANCE(4, true, false, false, 2)
SWR_RASTER_INSTANCE(4, true, true, false, 2)
SWR_RASTER_INSTANCE(1, false, false, false, 3)
SWR_RASTER_INSTANCE(1, false, true, false, 3)
SWR_RASTER_INSTANCE(1, true, false, false, 3)
SWR_RASTER_INSTANCE(1, true, true, false, 3)
SWR_RASTER_INSTANCE(4, false, false, false, 3)
SWR_RASTER_INSTANCE(4, false, true, false, 3)
SWR_RASTER_INSTANCE(4, true, false, false, 3)
SWR_RASTER_INSTANCE(4, true, true, false, 3)
#undef SWR_RASTER_INSTANCE

}  // namespace swr
