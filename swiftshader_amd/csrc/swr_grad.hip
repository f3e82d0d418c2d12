// swr_grad.hip -- gradients of a draw, and the mip chain with its adjoint.
//
// Differentiable rendering has no counterpart in the reference; the forward
// function being differentiated is the reference's fixed-function draw
// (PixelProgram.cpp:138-364 shading, SamplerCore.cpp float path) with:
//   * coverage and visibility held fixed: each covered sample's colour comes
//     from the primitive recorded in the primitive_id attachment;
//   * the texture's sampling state held fixed at the forward pass's values:
//     LOD (hence mip levels and trilinear weight), anisotropic probe count
//     and axis -- recomputed here bit-for-bit from the same plane equations;
//   * attributes interpolated at the pixel centre with perspective-correct
//     barycentrics of the UNSNAPPED projected vertices, so that positions
//     have a gradient (the forward's 28.4 snapping has zero derivative).
// The barycentric chain and texel coordinates run in FP64 (screen positions
// near 2^10 pixels and texel coordinates near 2^12 lose too much in FP32 for
// 1e-4 gradient agreement); per-tap products run in FP32, and the results
// are summed per workgroup in LDS, then added to FP32 buffers with float
// atomics.
//
//   backward_kernel    one thread per pixel; per distinct owning primitive
//                      of the pixel's samples: re-derive the fragment,
//                      scatter dL/dtexel over the filter footprint, and
//                      dL/d(colour, uv, x, y, w) of the three vertices.
//   mip_down_kernel    2x2 box filter, level k -> k+1.
//   mip_up_kernel      its adjoint, level k+1 -> k (accumulating).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swr_device.hpp"
#include "swr_params.hpp"

namespace swr {

struct GradParams {
  const char *dColor;
  int dColorPitch;
  long long dColorSlice;
  float *dPos;
  float *dCol;
  float *dUV;
  float *dTex[kMipLevels];  // per level (levels past the view repeat the last)
  char *state;
  int statePitch;
  long long stateSlice;
};

// ---- binned texel scatter (batch plans) -------------------------------------
// A batch plan's backward splits the texel scatter from the rest.  The
// per-fragment pass gathers (the filter value and its u, v derivatives) and
// writes one footprint record per fragment, keyed by the kBinTile-texel tile
// of its finer mip level within its texture group (draws that share
// texture-gradient buffers and sampler state), into slots its wave reserves
// kBinReserve at a time.  The (key, slot) pairs of every draw of the segment
// are then sorted by key with the stable radix sort of the binning front
// end, and each tile's records are scattered, one group at a time, into LDS
// copies of the tile (with a halo for the probes that cross its edges) at
// both mip levels, which reach global memory once per group: all views that
// hit one tile share one set of atomics.
constexpr int kBinTile = 32;                          // tile side at the finer level (texels)
constexpr int kBinHalo0 = 6;                          // halo at the finer level (an 8-probe footprint)
constexpr int kBinHalo1 = 4;                          // halo at the coarser level
constexpr int kBinR0 = kBinTile + 2 * kBinHalo0;      // LDS region side, finer level
constexpr int kBinR1 = kBinTile / 2 + 2 * kBinHalo1;  // LDS region side, coarser level
constexpr int kBinCells = kBinR0 * kBinR0 + kBinR1 * kBinR1;  // LDS cells of both regions
constexpr uint32_t kBinChunk = 8192;                  // records per scatter work item
constexpr int kScatterThreads = 512;                   // tap_scatter_kernel workgroup (two per CU fit in LDS)
constexpr int kMaxBinGroups = 8;
constexpr uint32_t kBinReserve = 512;                 // record slots a wave reserves at a time
constexpr uint32_t kNoBucket = 0xFFFFFFFFu;           // key of a reserved slot left unused

// Footprint of one fragment: texture coordinate, frozen sampling state and
// the upstream gradient of the filtered value.
struct TapRecord {
  double u, v;
  QuadLod L;
  float4 g;
};
static_assert(sizeof(TapRecord) == 48, "TapRecord layout");

// Bucket layout of a texture group: tile (tx, ty) of level l is bucket
// levelBase[l] + ty * tilesX[l] + tx; `draw` is a draw of the group (its
// texture parameters and gradient buffers describe the whole group).
struct BinGroup {
  uint32_t levelBase[kMipLevels];
  uint32_t tilesX[kMipLevels];
  int draw;
};

// Bucket b's tile: group, level and tile coordinates.
struct BinBucket {
  uint8_t group, level;
  uint16_t tx, ty;
};

struct BinArgs {
  TapRecord *rec;           // records, by slot
  uint32_t *key;            // per slot: bucket (kNoBucket: unused)
  uint32_t *val;            // per slot: the slot's index (the sort's payload)
  uint32_t *reserved;       // slots reserved so far
  uint32_t cap;             // slots allocated: the segment's fragment bound plus every wave's last reservation
  const int8_t *drawGroup;  // per draw of the segment: its texture group, -1 for none
  const BinGroup *groups;
};

// A wave's current reservation (wave-uniform): next slot and slots left.
struct SlotRange {
  uint32_t base, left;
};

// One axis of the float-path address computation (SamplerCore.cpp:2463-2638)
// in FP64: tap indices x0/x1, the filter fraction f and df/d(uvw).
__device__ __forceinline__ void axis64(int filter, double uvw, int dim, int clampMode, int &x0, int &x1, double &f,
                                       double &dfdu) {
  double c, dc;
  if (clampMode) {
    const double hi = 1.0 - 0x1p-24;  // 0x3F7FFFFF
    if (!(uvw > 0.0)) { c = 0.0; dc = 0.0; }
    else if (uvw > hi) { c = hi; dc = 0.0; }
    else { c = uvw; dc = 1.0; }
  } else {
    c = uvw - floor(uvw);
    dc = 1.0;
    if (!(c >= 0.0 && c <= 1.0)) {  // NaN / infinite coordinate: texel 0, no gradient
      c = 0.0;
      dc = 0.0;
    }
  }
  double coord = c * (double)dim;
  if (filter == FILT_POINT) {
    x0 = (int)coord;
    x0 = x0 < dim - 1 ? x0 : dim - 1;
    x1 = x0;
    f = 0.0;
    dfdu = 0.0;
    return;
  }
  coord -= 0.5;
  const double fl = floor(coord);
  x0 = (int)fl;
  f = coord - fl;
  dfdu = dc * (double)dim;
  x1 = x0 + 1;
  if (clampMode) {
    x0 = x0 > 0 ? x0 : 0;
    x1 = x1 < dim - 1 ? x1 : dim - 1;
  } else {
    if (x0 < 0) x0 = dim - 1;
    if (x1 >= dim) x1 = 0;
  }
}

struct Acc4 { double x, y, z, w; };

// Float add to global memory through a global (not flat) atomic: a flat
// atomic also counts against the LDS wait counter, so every later LDS
// access would wait for it.
typedef __attribute__((address_space(1))) float GlobalFloat;
__device__ __forceinline__ void gadd(float *p, float v) {
  (void)__builtin_amdgcn_global_atomic_fadd_f32((GlobalFloat *)p, v);
}

// ---- aggregation of the gradient scatter --------------------------------------
// Float atomics to global memory cost one request per (instruction, cache
// line): lanes that hit the same line in one instruction share a request,
// lanes that hit different lines do not.  Texel gradients are therefore
// first summed in LDS, in one hash table per wave, and reach global memory
// in flushes where 4 adjacent lanes carry the 4 channels of a texel and 16
// adjacent lanes the 4 texels of an aligned group (x = 4k..4k+3 of a row):
// the hash keeps the low two bits of the texel index as the low two bits of
// the slot and probes in steps of 4, so a group's texels sit in 4 adjacent
// slots.  A wave flushes its own table (no workgroup barrier) whenever it
// passes kFlushAt entries before a bilinear footprint, so a table never
// fills; a key that still finds no slot within kProbes goes straight to
// global memory.  Primitive gradients (27 values per primitive) are summed
// in one table per workgroup.
constexpr int kWaves = 4;                // waves per workgroup (256 threads)
constexpr int kWaveSlots = 512;          // texel entries per wave (key: level << 27 | texel index)
constexpr uint32_t kFlushAt = 256;       // at most 256 new keys per footprint step
constexpr int kTriSlots = 64;            // primitive entries per workgroup
constexpr int kTriVals = 27;             // per primitive: 3 x (dx, dy, dw), 3 x rgba, 3 x uv
constexpr int kProbes = 8;
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;

// Channel c of slot s sits at val[c * kValStride + s]: one array per
// channel (a lane's four adds hit four banks that neighbouring slots do not
// share), padded so that the four channels of one slot fall in different
// banks when a flush reads them from adjacent lanes.
constexpr int kValStride = kWaveSlots + 8;

struct WaveTexTable {
  uint32_t key[kWaveSlots];
  float val[4 * kValStride];
  uint32_t count;  // occupied slots
};

// LDS of a backward workgroup: the texture-gradient buffers of the draw being
// processed and the primitive table; one texel table per wave unless the
// texel scatter is binned (BIN: the records go to tap_scatter_kernel).
struct GradLdsBase {
  float *dTex[kMipLevels];  // texture-gradient buffers of the draw being processed, per level
  uint32_t triKey[kTriSlots];
  float triVal[kTriSlots][kTriVals];
  uint32_t triCount;  // occupied slots
};
template <bool BIN> struct GradLds : GradLdsBase {
  WaveTexTable tex[kWaves];
};
template <> struct GradLds<true> : GradLdsBase {};

__device__ __forceinline__ WaveTexTable &wave_table(GradLds<false> &S) { return S.tex[threadIdx.x >> 6]; }

// Clears the wave's texel table (the whole wave).
__device__ __forceinline__ void wave_tex_init(WaveTexTable &W) {
  const int lane = threadIdx.x & 63;
  for (int i = lane; i < kWaveSlots; i += 64) {
    W.key[i] = kEmptyKey;
#pragma unroll
    for (int c = 0; c < 4; c++) W.val[c * kValStride + i] = 0.0f;
  }
  if (lane == 0) W.count = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Adds the wave's texel table to the gradient buffers dTex (per level, in
// LDS) and clears it.
// Runs on the wave's active lanes, whichever they are: lane j of n active
// lanes takes the (slot, channel) pairs j, j + n, ... so that 4 adjacent
// lanes write the channels of one texel and 16 the texels of one group.
// LDS operations of a wave complete in order, so the key reads precede the
// clearing writes.
__device__ __forceinline__ void wave_tex_flush(WaveTexTable &W, float *const *dTex) {
  const uint64_t act = __ballot(1);
  const uint32_t n = (uint32_t)__builtin_popcountll(act);
  const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0));
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // Each value is read and zeroed in one exchange (empty slots hold zeros).
#pragma unroll 4
  for (uint32_t j = rank; j < (uint32_t)kWaveSlots * 4; j += n) {
    const uint32_t key = W.key[j >> 2];
    const float v = atomicExch(&W.val[(j & 3) * kValStride + (j >> 2)], 0.0f);
    if (key != kEmptyKey) {
      float *g = dTex[key >> 27];
      gadd(g + 4 * (size_t)(key & ((1u << 27) - 1)) + (j & 3), v);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (uint32_t j = rank; j < (uint32_t)kWaveSlots; j += n) W.key[j] = kEmptyKey;
  if (rank == 0) W.count = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void lds_init_tri(GradLdsBase &S) {
  for (int i = threadIdx.x; i < kTriSlots; i += blockDim.x) S.triKey[i] = kEmptyKey;
  for (int i = threadIdx.x; i < kTriSlots * kTriVals; i += blockDim.x) (&S.triVal[0][0])[i] = 0.0f;
  if (threadIdx.x == 0) S.triCount = 0;
}

template <bool BIN>
__device__ __forceinline__ void lds_init(GradLds<BIN> &S, const GradParams &G) {
  if constexpr (!BIN) wave_tex_init(wave_table(S));
  lds_init_tri(S);
  if (threadIdx.x < kMipLevels) S.dTex[threadIdx.x] = G.dTex[threadIdx.x];
  __syncthreads();
}

__device__ __forceinline__ uint32_t texel_slot(uint32_t key) {
  return ((((key >> 2) * 2654435761u) >> 25) << 2) | (key & 3u);
}

// Adds (a, b, c, d) to texel `key` of the table, probing from the second
// slot of its sequence (the first was taken); global memory when no slot
// is found.
__device__ __forceinline__ void texel_add_slow(WaveTexTable &W, float *g, uint32_t key, float a, float b, float c,
                                            float d) {
  const uint32_t h = texel_slot(key);
  for (int p = 1; p < kProbes; p++) {
    const uint32_t sl = (h + 4u * (uint32_t)p) & (kWaveSlots - 1);
    const uint32_t old = atomicCAS(&W.key[sl], kEmptyKey, key);
    if (old == kEmptyKey) atomicAdd(&W.count, 1u);
    if (old == kEmptyKey || old == key) {
      atomicAdd(&W.val[sl], a);
      atomicAdd(&W.val[kValStride + sl], b);
      atomicAdd(&W.val[2 * kValStride + sl], c);
      atomicAdd(&W.val[3 * kValStride + sl], d);
      return;
    }
  }
  float *o = g + 4 * (size_t)(key & ((1u << 27) - 1));
  gadd(o + 0, a);
  gadd(o + 1, b);
  gadd(o + 2, c);
  gadd(o + 3, d);
}

// Adds gTex * w[k] to texel idx[k] of level lvl for the 4 taps of a
// bilinear footprint.  The first probe of all four keys is issued at once
// (one LDS round trip rather than four); a tap whose key is absent (w = 0,
// or an index beyond the key range) probes with the empty key, which
// changes nothing, and adds zeros.  Taps whose first slot holds another
// key continue in texel_add_slow.
__device__ __forceinline__ void texels_add4(WaveTexTable &W, float *g, int lvl, const uint32_t idx[4],
                                            const float w[4], const float4 &gTex) {
  uint32_t key[4], sl[4], old[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const bool want = w[k] != 0.0f && idx[k] < (1u << 27);
    key[k] = want ? (((uint32_t)lvl << 27) | idx[k]) : kEmptyKey;
    sl[k] = texel_slot(key[k]) & (kWaveSlots - 1);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) old[k] = atomicCAS(&W.key[sl[k]], kEmptyKey, key[k]);
  uint32_t fresh = 0;
  bool ok[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    fresh += (key[k] != kEmptyKey && old[k] == kEmptyKey) ? 1u : 0u;
    ok[k] = key[k] != kEmptyKey && (old[k] == kEmptyKey || old[k] == key[k]);
  }
  if (fresh) atomicAdd(&W.count, fresh);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const float m = ok[k] ? w[k] : 0.0f;
    atomicAdd(&W.val[sl[k]], gTex.x * m);
    atomicAdd(&W.val[kValStride + sl[k]], gTex.y * m);
    atomicAdd(&W.val[2 * kValStride + sl[k]], gTex.z * m);
    atomicAdd(&W.val[3 * kValStride + sl[k]], gTex.w * m);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (w[k] != 0.0f && !ok[k]) {
      if (idx[k] < (1u << 27)) {
        texel_add_slow(W, g, key[k], gTex.x * w[k], gTex.y * w[k], gTex.z * w[k], gTex.w * w[k]);
      } else {
        float *o = g + 4 * (size_t)idx[k];
        gadd(o + 0, gTex.x * w[k]);
        gadd(o + 1, gTex.y * w[k]);
        gadd(o + 2, gTex.z * w[k]);
        gadd(o + 3, gTex.w * w[k]);
      }
    }
  }
}

// Slot of primitive t in the workgroup table, or NULL when the table is full.
__device__ __forceinline__ float *tri_slot(GradLdsBase &S, uint32_t t) {
  const uint32_t h = (t * 2654435761u) >> 26;
  for (int p = 0; p < kProbes; p++) {
    const uint32_t sl = (h + (uint32_t)p) & (kTriSlots - 1);
    const uint32_t old = atomicCAS(&S.triKey[sl], kEmptyKey, t);
    if (old == kEmptyKey) atomicAdd(&S.triCount, 1u);
    if (old == kEmptyKey || old == t) return S.triVal[sl];
  }
  return nullptr;
}

// Adds the per-primitive values v[27] of primitive t to the gradient buffers.
__device__ __forceinline__ void tri_scatter_global(const DrawParams &P, const GradParams &G, uint32_t t,
                                                   const float *v) {
  uint32_t vi[3];
  tri_vertices(P, t, vi);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (G.dPos) {
      float *o = G.dPos + 4 * (size_t)vi[k];
      gadd(o + 0, v[3 * k + 0]);
      gadd(o + 1, v[3 * k + 1]);
      gadd(o + 3, v[3 * k + 2]);
    }
    if (G.dCol) {
      float *o = G.dCol + 4 * (size_t)vi[k];
      gadd(o + 0, v[9 + 4 * k + 0]);
      gadd(o + 1, v[9 + 4 * k + 1]);
      gadd(o + 2, v[9 + 4 * k + 2]);
      gadd(o + 3, v[9 + 4 * k + 3]);
    }
    if (G.dUV) {
      float *o = G.dUV + 2 * (size_t)vi[k];
      gadd(o + 0, v[21 + 2 * k + 0]);
      gadd(o + 1, v[21 + 2 * k + 1]);
    }
  }
}

__device__ __forceinline__ void lds_flush_tri(GradLdsBase &S, const DrawParams &P, const GradParams &G) {
  for (int i = threadIdx.x; i < kTriSlots; i += blockDim.x) {
    const uint32_t t = S.triKey[i];
    if (t != kEmptyKey) tri_scatter_global(P, G, t, S.triVal[i]);
  }
}

// Flushes both tables (all threads).
template <bool BIN>
__device__ __forceinline__ void lds_flush(GradLds<BIN> &S, const DrawParams &P, const GradParams &G) {
  if constexpr (!BIN) wave_tex_flush(wave_table(S), S.dTex);
  __syncthreads();
  lds_flush_tri(S, P, G);
}

// One mip level of the texture and of its gradient, read once per fragment
// (the level varies per lane: indexing the parameter arrays costs a load).
struct LevelRef {
  const float4 *tex;
  float *grad;
  int width, height, pitch, lvl;
};

__device__ __forceinline__ LevelRef level_ref(const TexParams &T, const GradParams &G, int lvl) {
  LevelRef r;
  r.tex = reinterpret_cast<const float4 *>(T.data[lvl]);
  r.grad = G.dTex[lvl];
  r.width = T.width[lvl];
  r.height = T.height[lvl];
  r.pitch = T.pitchP[lvl];
  r.lvl = lvl;
  return r;
}

// One bilinear (or point) footprint at level R with weight `wgt`: adds
// wgt * filter(u, v) to `val`, wgt * d(filter . gTex)/du, /dv to dU, dV, and
// scatters gTex * wgt * tap weight into the level's texel gradient.  The
// texel coordinate and its fraction are formed in FP64; the per-tap
// products in FP32.
template <bool SCATTER>
__device__ __forceinline__ void tap_backward(WaveTexTable *W, const TexParams &T, const LevelRef &R, double u,
                                             double v, double wgt, const float4 &gTex, float4 &val, float &dU,
                                             float &dV) {
  int x0, x1, y0, y1;
  double fu, fv, dfu, dfv;
  axis64(T.filter, u, R.width, T.addrU, x0, x1, fu, dfu);
  axis64(T.filter, v, R.height, T.addrV, y0, y1, fv, dfv);
  const int i00 = x0 + y0 * R.pitch, i10 = x1 + y0 * R.pitch, i01 = x0 + y1 * R.pitch, i11 = x1 + y1 * R.pitch;
  const float4 c00 = R.tex[i00], c10 = R.tex[i10], c01 = R.tex[i01], c11 = R.tex[i11];
  const float w00 = (float)(wgt * ((1.0 - fu) * (1.0 - fv))), w10 = (float)(wgt * (fu * (1.0 - fv)));
  const float w01 = (float)(wgt * ((1.0 - fu) * fv)), w11 = (float)(wgt * (fu * fv));
  val.x += w00 * c00.x + w10 * c10.x + w01 * c01.x + w11 * c11.x;
  val.y += w00 * c00.y + w10 * c10.y + w01 * c01.y + w11 * c11.y;
  val.z += w00 * c00.z + w10 * c10.z + w01 * c01.z + w11 * c11.z;
  val.w += w00 * c00.w + w10 * c10.w + w01 * c01.w + w11 * c11.w;
  const float d00 = gTex.x * c00.x + gTex.y * c00.y + gTex.z * c00.z + gTex.w * c00.w;
  const float d10 = gTex.x * c10.x + gTex.y * c10.y + gTex.z * c10.z + gTex.w * c10.w;
  const float d01 = gTex.x * c01.x + gTex.y * c01.y + gTex.z * c01.z + gTex.w * c01.w;
  const float d11 = gTex.x * c11.x + gTex.y * c11.y + gTex.z * c11.z + gTex.w * c11.w;
  const float ffu = (float)fu, ffv = (float)fv;
  dU += (float)(wgt * dfu) * ((1.0f - ffv) * (d10 - d00) + ffv * (d11 - d01));
  dV += (float)(wgt * dfv) * ((1.0f - ffu) * (d01 - d00) + ffu * (d11 - d10));
  if constexpr (SCATTER) {
    float *g = R.grad;
    if (g) {
      const uint32_t idx[4] = {(uint32_t)i00, (uint32_t)i10, (uint32_t)i01, (uint32_t)i11};
      const float w[4] = {w00, w10, w01, w11};
      texels_add4(*W, g, R.lvl, idx, w, gTex);
    }
  }
}

// The forward filter at (u, v) with the frozen state L: anisotropic probes
// along (uDelta, vDelta) (sampleFloatAniso, SamplerCore.cpp:1085-1141),
// each trilinear between two levels (sampleTexture128, :200-253).  With BIN
// the taps are only gathered (the value and its u, v derivatives): the
// texel scatter is left to tap_scatter_kernel.
template <bool BIN>
__device__ __forceinline__ void texture_backward(GradLds<BIN> &S, const TexParams &T, const GradParams &G,
                                                 const QuadLod &L, double u, double v, const float4 &gTex,
                                                 float4 &val, float &dU, float &dV) {
  WaveTexTable *W = nullptr;
  if constexpr (!BIN) W = &wave_table(S);
  int N = 1;
  double A = 1.0, B = 0.0;
  if (T.filter == FILT_ANISOTROPIC) {
    N = rr_round_int(L.aniso);
    A = (double)uv_weight(aniso_row(N));
    B = (double)uv_start(aniso_row(N));
    N = N > 1 ? N : 1;  // the forward's probe loop runs at least once (a singular Jacobian gives N = INT_MIN)
  }
  const int lvl0 = select_mip(T, L.lod, 0);
  double wl0 = 1.0, wl1 = 0.0;
  if (T.mipmap == MIP_LINEAR) {
    wl1 = (double)rr_frac(L.lod);
    wl0 = 1.0 - wl1;
  }
  // Level-major: consecutive probes on one level overlap, so the wave's
  // table sees repeated keys close together.  The loops run in step over
  // the wave's active lanes (a lane with fewer probes idles), so that a
  // table flush, checked once per step, always has the whole set of active
  // lanes to spread its slots over.
  const bool anyGrad = G.dTex[0] != nullptr;
  for (int lv = 0; lv < 2; lv++) {
    const bool levelOn = lv == 0 || wl1 != 0.0;
    if (!__ballot(levelOn)) break;
    const LevelRef R = level_ref(T, G, levelOn ? lvl0 + lv : lvl0);
    const double wl = lv ? wl1 : wl0;
    for (int i = 0;; i++) {
      const bool on = levelOn && i < N;
      if (!__ballot(on)) break;
      // The active lanes read the same count: the branch is uniform over them.
      if constexpr (!BIN)
        if (anyGrad && __builtin_amdgcn_readfirstlane(W->count) > kFlushAt) wave_tex_flush(*W, S.dTex);
      if (on) {
        const double t = B + (double)i * A;
        const double ui = u + t * (double)L.uDelta, vi = v + t * (double)L.vDelta;
        tap_backward<!BIN>(W, T, R, ui, vi, A * wl, gTex, val, dU, dV);
      }
    }
  }
}

// Gradient of one fragment: primitive t at pixel (px, py), upstream g.
// With BIN the texel gradient is not scattered: the fragment's footprint
// record goes to `rec` and the function returns true.  The primitive's 27
// attribute gradients go to `val` (all zero when none are wanted or the
// fragment has none); the caller adds them to the primitive table.
template <bool BIN>
__device__ bool fragment_backward(GradLds<BIN> &S, const DrawParams &P, const GradParams &G, uint32_t t, int px,
                                  int py, const Acc4 &g, const QuadLod &L, TapRecord &rec, float (&val)[kTriVals]) {
  bool has = false;
  uint32_t vi[3];
  tri_vertices(P, t, vi);
  double X[3], Y[3], rhw[3], x[3], y[3], w[3];
  const double X0 = (double)P.X0xF / 16.0, Y0 = (double)P.Y0xF / 16.0;
  const double Wh = (double)P.WxF / 16.0, Hh = (double)P.HxF / 16.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float4 p = reinterpret_cast<const float4 *>(P.pos)[vi[k]];
    x[k] = p.x;
    y[k] = p.y;
    w[k] = p.w;
    rhw[k] = p.w != 0.0f ? 1.0 / (double)p.w : 1.0;
    X[k] = X0 + x[k] * rhw[k] * Wh;
    Y[k] = Y0 + y[k] * rhw[k] * Hh;
  }
  const double D = (Y[1] - Y[2]) * (X[0] - X[2]) + (X[2] - X[1]) * (Y[0] - Y[2]);
  if (!(D != 0.0)) return false;
  const double iD = 1.0 / D;
  const double gx[3] = {(Y[1] - Y[2]) * iD, (Y[2] - Y[0]) * iD, (Y[0] - Y[1]) * iD};
  const double gy[3] = {(X[2] - X[1]) * iD, (X[0] - X[2]) * iD, (X[1] - X[0]) * iD};
  const double ex = (double)px - X[2], ey = (double)py - Y[2];
  double lam[3];
  lam[0] = gx[0] * ex + gy[0] * ey;
  lam[1] = gx[1] * ex + gy[1] * ey;
  lam[2] = 1.0 - lam[0] - lam[1];
  double r[3], Ssum = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    r[k] = lam[k] * rhw[k];
    Ssum += r[k];
  }
  if (!(Ssum != 0.0)) return false;
  const double iS = 1.0 / Ssum;
  double b[3];
#pragma unroll
  for (int k = 0; k < 3; k++) b[k] = r[k] * iS;

  // Attributes at the fragment.
  float4 vc[3];
  float2 vt[3];
  float4 col = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
  double u = 0.0, v = 0.0;
  if (P.useColor) {
    double cx = 0.0, cy = 0.0, cz = 0.0, cw = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      vc[k] = reinterpret_cast<const float4 *>(P.col)[vi[k]];
      cx += b[k] * vc[k].x;
      cy += b[k] * vc[k].y;
      cz += b[k] * vc[k].z;
      cw += b[k] * vc[k].w;
    }
    col = make_float4((float)cx, (float)cy, (float)cz, (float)cw);
  }
  if (P.useUV) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      vt[k] = reinterpret_cast<const float2 *>(P.uv)[vi[k]];
      u += b[k] * vt[k].x;
      v += b[k] * vt[k].y;
    }
  }

  // Shader: out = colour | texture | texture * colour.
  const float4 gf = make_float4((float)g.x, (float)g.y, (float)g.z, (float)g.w);
  float4 dcol = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float dU = 0.0f, dV = 0.0f;
  const bool attrGrads = G.dPos || G.dCol || G.dUV;
  if (P.useUV) {
    const float4 gTex = P.useColor ? make_float4(gf.x * col.x, gf.y * col.y, gf.z * col.z, gf.w * col.w) : gf;
    float4 tex = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // With the texel scatter binned, the sampled value and its u, v
    // derivatives serve only the vertex-attribute gradients.
    if (!BIN || attrGrads) texture_backward<BIN>(S, P.tex, G, L, u, v, gTex, tex, dU, dV);
    if constexpr (BIN) {
      if (G.dTex[0]) {
        rec.u = u;
        rec.v = v;
        rec.L = L;
        rec.g = gTex;
        has = true;
      }
    }
    if (P.useColor) dcol = make_float4(gf.x * tex.x, gf.y * tex.y, gf.z * tex.z, gf.w * tex.w);
  } else {
    dcol = gf;
  }
  if (!attrGrads) return has;

  // Attribute gradients and dL/db.
  double Gk[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    Gk[k] = 0.0;
    if (P.useColor) {
      const float bk = (float)b[k];
      val[9 + 4 * k + 0] = bk * dcol.x;
      val[9 + 4 * k + 1] = bk * dcol.y;
      val[9 + 4 * k + 2] = bk * dcol.z;
      val[9 + 4 * k + 3] = bk * dcol.w;
      Gk[k] += (double)(dcol.x * vc[k].x + dcol.y * vc[k].y + dcol.z * vc[k].z + dcol.w * vc[k].w);
    }
    if (P.useUV) {
      val[21 + 2 * k + 0] = (float)(b[k] * (double)dU);
      val[21 + 2 * k + 1] = (float)(b[k] * (double)dV);
      Gk[k] += (double)dU * vt[k].x + (double)dV * vt[k].y;
    }
  }
  if (G.dPos) {
    // b = r / sum(r), r_k = lambda_k * rhw_k.
    const double Abar = b[0] * Gk[0] + b[1] * Gk[1] + b[2] * Gk[2];
    double dlam[3], drhw[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double dr = (Gk[k] - Abar) * iS;
      dlam[k] = dr * rhw[k];
      drhw[k] = dr * lam[k];
    }
    // d lambda_i / d X_j = -lambda_j * d lambda_i / d px (likewise for Y).
    const double Gx = dlam[0] * gx[0] + dlam[1] * gx[1] + dlam[2] * gx[2];
    const double Gy = dlam[0] * gy[0] + dlam[1] * gy[1] + dlam[2] * gy[2];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double dX = -lam[k] * Gx, dY = -lam[k] * Gy;
      const double dr = drhw[k] + dX * x[k] * Wh + dY * y[k] * Hh;
      val[3 * k + 0] = (float)(dX * rhw[k] * Wh);
      val[3 * k + 1] = (float)(dY * rhw[k] * Hh);
      val[3 * k + 2] = (float)(w[k] != 0.0 ? -dr * rhw[k] * rhw[k] : 0.0);
    }
  }
  return has;
}

// Adds each lane's primitive gradients val (primitive t; t = kEmptyKey: none)
// to the primitive table.  Lanes of one primitive usually sit next to each
// other in the batch: runs of equal t are summed first (a segmented scan
// across the wave), and the last lane of each run adds the run's sums, so a
// run of L lanes costs one set of float LDS atomics instead of L.
template <bool BIN>
__device__ __forceinline__ void tri_add_wave(GradLds<BIN> &S, const DrawParams &P, const GradParams &G, uint32_t t,
                                             float (&val)[kTriVals]) {
  const int lane = threadIdx.x & 63;
  const uint64_t any = __ballot(t != kEmptyKey);
  if (!any) return;
  const uint32_t prev = (uint32_t)__shfl_up((int)t, 1, 64);
  const uint64_t starts = __ballot(lane == 0 || prev != t);
  bool add = t != kEmptyKey;
  if (starts != ~0ull) {
    const uint64_t upTo = lane == 63 ? ~0ull : (2ull << lane) - 1ull;
    const int runStart = 63 - __builtin_clzll(starts & upTo);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const bool take = lane - d >= runStart;
#pragma unroll
      for (int k = 0; k < kTriVals; k++) {
        const float up = __shfl_up(val[k], d, 64);
        if (take) val[k] += up;
      }
    }
    add = add && (lane == 63 || ((starts >> (lane + 1)) & 1ull));
  }
  if (!add) return;
  float *slot = tri_slot(S, t);
  if (!slot) {
    tri_scatter_global(P, G, t, val);
    return;
  }
#pragma unroll
  for (int k = 0; k < kTriVals; k++)
    if (val[k] != 0.0f) atomicAdd(slot + k, val[k]);
}

// Per-wave ring of fragments: one entry per (pixel, distinct owning
// primitive), at most MS per pixel.  A block adds at most 64 * MS entries to
// fewer than 64 left over from the blocks before.
constexpr uint32_t kRing = 512;

struct FragQueue {
  uint32_t xy[kWaves][kRing];    // x | y << 16
  uint32_t tri[kWaves][kRing];   // primitive index
  uint32_t mask[kWaves][kRing];  // samples the fragment owns
};

// Wave-uniform position of a wave's ring.
struct RingPos {
  uint32_t head, count;
};

// Appends the fragments of the wave's pixels (px, py) of draw P: the
// samples of a pixel owned by one primitive share one fragment, shaded once
// at the pixel centre, whose upstream gradient is the sum of theirs.
template <int MS>
__device__ __forceinline__ void enqueue_fragments(FragQueue &Q, RingPos &rp, const DrawParams &P, int px, int py) {
  const int wave = threadIdx.x >> 6;
  uint32_t id[MS];
  const bool inside = px < P.width && py < P.height;
#pragma unroll
  for (int q = 0; q < MS; q++)
    id[q] = inside ? *reinterpret_cast<const uint32_t *>(P.primId + q * P.primIdSlice + (long long)py * P.primIdPitch +
                                                         4ll * px)
                   : 0u;
#pragma unroll
  for (int q = 0; q < MS; q++) {
    const uint32_t t1 = id[q];
    bool first = t1 != 0u && t1 <= (uint32_t)P.triCount;
#pragma unroll
    for (int q2 = 0; q2 < q; q2++) first &= id[q2] != t1;
    uint32_t mask = 0;
#pragma unroll
    for (int q2 = q; q2 < MS; q2++) mask |= id[q2] == t1 ? 1u << q2 : 0u;
    const uint64_t bal = __ballot(first);
    if (first) {
      const uint32_t pos = (rp.head + rp.count + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0))) &
                           (kRing - 1);
      Q.xy[wave][pos] = (uint32_t)px | ((uint32_t)py << 16);
      Q.tri[wave][pos] = t1 - 1u;
      Q.mask[wave][pos] = mask;
    }
    rp.count += (uint32_t)__builtin_popcountll(bal);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bucket of a footprint record: the kBinTile-texel tile, at the finer of its
// mip levels, of the texel its coordinate falls in (wrapped or clamped as
// the address mode does).
__device__ __forceinline__ uint32_t tap_bucket(const TexParams &T, const BinGroup &Bg, double u, double v, float lod) {
  int l0 = select_mip(T, lod, 0);
  l0 = l0 < 0 ? 0 : (l0 > kMipLevels - 1 ? kMipLevels - 1 : l0);
  double cu, cv;
  if (T.addrU) cu = u > 0.0 ? (u < 1.0 ? u : 1.0 - 0x1p-24) : 0.0;
  else cu = u - floor(u);
  if (T.addrV) cv = v > 0.0 ? (v < 1.0 ? v : 1.0 - 0x1p-24) : 0.0;
  else cv = v - floor(v);
  if (!(cu >= 0.0 && cu < 1.0)) cu = 0.0;  // NaN / infinite coordinates
  if (!(cv >= 0.0 && cv < 1.0)) cv = 0.0;
  const int W = T.width[l0], H = T.height[l0];
  int x = (int)(cu * (double)W), y = (int)(cv * (double)H);
  x = x < W - 1 ? x : W - 1;
  y = y < H - 1 ? y : H - 1;
  return Bg.levelBase[l0] + (uint32_t)(y / kBinTile) * Bg.tilesX[l0] + (uint32_t)(x / kBinTile);
}

// Marks the slots left in the wave's reservation as unused (all lanes).
__device__ __forceinline__ void release_slots(const BinArgs &bin, SlotRange &sr) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t i = lane; i < sr.left; i += 64) {
    const uint32_t q = sr.base + i;
    if (q < bin.cap) {
      bin.key[q] = kNoBucket;
      bin.val[q] = q;
    }
  }
  sr.base += sr.left;
  sr.left = 0;
}

// Writes the records of the wave's lanes that have one (all lanes call
// this together) into the wave's reserved slots; when they do not all fit,
// the first ones fill the reservation and the rest start a new one of
// kBinReserve slots (one atomic per reservation).  No slot is skipped, so
// the slots reserved never exceed the records plus one reservation per
// wave.
__device__ __forceinline__ void emit_records(const BinArgs &bin, SlotRange &sr, bool has, const TapRecord &r,
                                             uint32_t key) {
  const int lane = threadIdx.x & 63;
  const uint64_t want = __ballot(has);
  if (!want) return;
  const uint32_t cnt = (uint32_t)__builtin_popcountll(want);
  uint32_t fresh = 0;
  if (cnt > sr.left) {
    if (lane == 0) fresh = atomicAdd(bin.reserved, kBinReserve);
    fresh = __builtin_amdgcn_readfirstlane(fresh);
  }
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint32_t rank = (uint32_t)__builtin_popcountll(want & lt);
  const uint32_t p = rank < sr.left ? sr.base + rank : fresh + (rank - sr.left);
  if (has && p < bin.cap) {
    bin.rec[p] = r;
    bin.key[p] = key;
    bin.val[p] = p;
  }
  if (cnt > sr.left) {
    sr.base = fresh + (cnt - sr.left);
    sr.left = kBinReserve - (cnt - sr.left);
  } else {
    sr.base += cnt;
    sr.left -= cnt;
  }
}

// Runs the wave's queued fragments of draw P (rec holds the draw's records:
// slot t = primitive t) 64 at a time, and the rest as well when `all`, so
// that edge pixels with several primitives cost extra lanes rather than
// extra passes, and partial batches occur only at draw changes.  With BIN
// the texel gradients leave as footprint records of draw group `grp`.
template <int MS, bool BIN>
__device__ __forceinline__ void drain_fragments(GradLds<BIN> &S, FragQueue &Q, RingPos &rp, const DrawParams &P,
                                                const GradParams &G, const TriRecord *__restrict__ rec, bool all,
                                                const BinArgs &bin, SlotRange &sr, int grp) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  while (rp.count >= 64 || (all && rp.count > 0)) {
    const uint32_t n = rp.count < 64 ? rp.count : 64;
    const uint32_t e = (rp.head + (uint32_t)lane) & (kRing - 1);
    bool has = false;
    TapRecord r;
    uint32_t key = 0;
    float val[kTriVals];
#pragma unroll
    for (int k = 0; k < kTriVals; k++) val[k] = 0.0f;
    uint32_t tv = kEmptyKey;  // primitive whose gradients val holds
    if ((uint32_t)lane < n) {
      const uint32_t xy = Q.xy[wave][e], t = Q.tri[wave][e], mask = Q.mask[wave][e];
      const int x = (int)(xy & 0xFFFFu), y = (int)(xy >> 16);
      Acc4 g = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < MS; q++) {
        if (!(mask >> q & 1u)) continue;
        const float4 d = *reinterpret_cast<const float4 *>(G.dColor + q * G.dColorSlice +
                                                           (long long)y * G.dColorPitch + 16ll * x);
        g.x += d.x;
        g.y += d.y;
        g.z += d.z;
        g.w += d.w;
      }
      QuadLod L = {0.0f, 0.0f, 0.0f, 0.0f};
      if (P.useUV) L = quad_state(P, rec[t], x, y);
      if (G.state) {
#pragma unroll
        for (int q = 0; q < MS; q++)
          if (mask >> q & 1u)
            *reinterpret_cast<float4 *>(G.state + q * G.stateSlice + (long long)y * G.statePitch + 16ll * x) =
                make_float4(L.lod, L.aniso, L.uDelta, L.vDelta);
      }
      has = fragment_backward<BIN>(S, P, G, t, x, y, g, L, r, val);
      if (BIN && has) key = tap_bucket(P.tex, bin.groups[grp], r.u, r.v, r.L.lod);
      tv = t;
    }
    if (G.dPos || G.dCol || G.dUV) tri_add_wave(S, P, G, tv, val);
    if constexpr (BIN) emit_records(bin, sr, has, r, key);
    rp.head = (rp.head + n) & (kRing - 1);
    rp.count -= n;
  }
}

// One draw: 16x16 pixels per workgroup.
template <int MS>
__global__ __launch_bounds__(256) void backward_kernel(DrawParams P, GradParams G, const TriRecord *__restrict__ rec) {
  __shared__ GradLds<false> S;
  __shared__ FragQueue Q;
  lds_init(S, G);
  RingPos rp = {0u, 0u};
  const BinArgs none = {};
  SlotRange sr = {0u, 0u};
  enqueue_fragments<MS>(Q, rp, P, blockIdx.x * 16 + (threadIdx.x & 15), blockIdx.y * 16 + (threadIdx.x >> 4));
  drain_fragments<MS, false>(S, Q, rp, P, G, rec, true, none, sr, 0);
  lds_flush(S, P, G);
}

// A batch segment, persistent: workgroups take 16x16-pixel blocks
// [blockLo, blockHi) from an atomic counter (block b belongs to the draw d in
// [drawLo, drawHi) with blockBase[d] <= b < blockBase[d + 1], in row-major
// order over its framebuffer) and keep their LDS tables across blocks: each
// wave's texel table flushes itself when it fills and is flushed when the
// next block writes other texture-gradient buffers; the primitive table is
// flushed when it passes half full or the next block belongs to another
// draw; both at the end.  Views that share a texture thus reach its
// gradient once per table rather than once per block.  Each wave carries
// the fragments that do not fill a batch of 64 over to its next block of
// the same draw.  With BIN there are no texel tables: the draws' texel
// gradients leave as footprint records (bin), which tap_scatter_kernel
// scatters.
template <int MS, bool BIN>
__global__ __launch_bounds__(256) void backward_batch_kernel(const DrawParams *__restrict__ params,
                                                             const GradParams *__restrict__ grads,
                                                             const uint32_t *__restrict__ blockBase, int drawLo,
                                                             int drawHi, uint32_t blockLo, uint32_t blockHi,
                                                             uint32_t *__restrict__ counter,
                                                             const TriRecord *__restrict__ rec, BinArgs bin) {
  __shared__ GradLds<BIN> S;
  __shared__ FragQueue Q;
  __shared__ uint32_t sBlock;
  lds_init(S, grads[drawLo]);
  int cur = -1;  // draw whose gradients the tables hold (wave-uniform)
  RingPos rp = {0u, 0u};
  SlotRange sr = {0u, 0u};
  for (;;) {
    if (threadIdx.x == 0) sBlock = blockLo + atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t b = __builtin_amdgcn_readfirstlane(sBlock);
    const uint32_t triCount = __builtin_amdgcn_readfirstlane(S.triCount);
    __syncthreads();
    int lo = -1;
    if (b < blockHi) {
      int l = drawLo, hi = drawHi;
      while (hi - l > 1) {
        const int mid = (l + hi) >> 1;
        if (blockBase[mid] <= b) l = mid;
        else hi = mid;
      }
      lo = __builtin_amdgcn_readfirstlane(l);
    }
    if (cur >= 0 && lo != cur) {
      // The rest of draw cur's fragments, before its tables are flushed (the
      // primitive table is shared: every wave finishes adding first).
      drain_fragments<MS, BIN>(S, Q, rp, params[cur], grads[cur], rec + params[cur].triBase, true, bin, sr,
                               BIN ? bin.drawGroup[cur] : 0);
      __syncthreads();
    }
    if (cur >= 0) {
      if constexpr (!BIN) {
        bool texSwitch = lo < 0;
        for (int k = 0; k < kMipLevels && !texSwitch; k++) texSwitch = grads[lo].dTex[k] != grads[cur].dTex[k];
        if (texSwitch) wave_tex_flush(wave_table(S), S.dTex);
      }
      const bool fr = lo != cur || triCount > kTriSlots / 2;
      if (fr) {
        lds_flush_tri(S, params[cur], grads[cur]);
        __syncthreads();
        lds_init_tri(S);
        __syncthreads();
      }
    }
    if (lo < 0) break;
    if (lo != cur) {
      if (threadIdx.x < kMipLevels) S.dTex[threadIdx.x] = grads[lo].dTex[threadIdx.x];
      __syncthreads();
    }
    cur = lo;
    const DrawParams &P = params[lo];
    const uint32_t local = b - blockBase[lo];
    const uint32_t bx = (uint32_t)(P.width + 15) >> 4;
    const int px = (int)(local % bx) * 16 + (threadIdx.x & 15), py = (int)(local / bx) * 16 + (threadIdx.x >> 4);
    enqueue_fragments<MS>(Q, rp, P, px, py);
    drain_fragments<MS, BIN>(S, Q, rp, P, grads[lo], rec + P.triBase, false, bin, sr, BIN ? bin.drawGroup[lo] : 0);
    __syncthreads();
  }
  if constexpr (BIN) release_slots(bin, sr);
}

// A wave-uniform value pinned in a scalar register (readfirstlane cannot be
// moved into a loop, so neither can the load that feeds it).
__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float *uniform_ptr(float *p) {
  const uint64_t a = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a), hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  return (float *)(((uint64_t)hi << 32) | lo);
}

// Work items of the scatter per bucket: ceil(records / kBinChunk), the
// bucket's records being [start[b], end[b]) of the sorted slots.
__global__ void tap_chunks_kernel(const uint32_t *__restrict__ start, const uint32_t *__restrict__ end,
                                  uint32_t nBuckets, uint32_t *__restrict__ chunks) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nBuckets) chunks[b] = (end[b] - start[b] + kBinChunk - 1) / kBinChunk;
  else if (b == nBuckets) chunks[b] = 0u;
}

// The LDS regions hold 64-bit fixed point: float adds in LDS retire about
// one lane every three clocks, 64-bit integer adds about six lanes a clock.
// A work item scales its values by S = 2^s, chosen from the largest record
// gradient m of the item so that m * S < 2^46; every product g * w (w <= 1)
// then rounds to an integer below 2^46 in magnitude, a cell's sum over the
// item (at most kBinChunk records, total weight <= 1 each per level) stays
// below 2^62, and the rounding adds at most 2^-47 * m per product -- far
// below the float rounding of the products themselves.  Integer sums do not
// depend on the order of the adds.
//
// The regions are channel-planar: channel c of cell i is word c * kBinCells
// + i, so that lanes adding to different texels of one channel hit
// different banks.
//
// A region is side x side cells from (ox, oy) of a w x h level.  On a
// wrapping axis (REPEAT) texel x stands for every x + k * w, so it takes the
// cell of its first copy at or after the origin: the taps that wrap past a
// border of the level, and the probes of a footprint that leave [0, 1),
// stay in the region of the tile they came from.
__device__ __forceinline__ int bin_cell_axis(int x, int o, int dim, bool wrap) {
  int r = x - o;
  if (wrap) {
    if (r < 0) r += dim;
    else if (r >= dim) r -= dim;
  }
  return r;
}

// x (|x| < 2^51, an already scaled product) rounded to the nearest integer:
// adding 1.5 * 2^52 leaves the integer in the low mantissa bits.
__device__ __forceinline__ long long fixed_round(float x) {
  const double t = (double)x + 0x1.8p52;
  return __double_as_longlong(t) - 0x4338000000000000ll;
}

__device__ __forceinline__ void lds_add_fixed(long long *c, float x) {
  atomicAdd((unsigned long long *)c, (unsigned long long)fixed_round(x));
}

// Adds the scaled products v (g * w * S) to texel (x, y) of a level: into
// region cell `cell` when the texel has one (cell >= 0), else (unscaled,
// v * invS) straight to the gradient buffer.
__device__ __forceinline__ void bin_add(long long *regions, int cell, float *g, int pitch, float invS, int x, int y,
                                        const float4 &v) {
  if (v.x == 0.0f && v.y == 0.0f && v.z == 0.0f && v.w == 0.0f) return;
  if (cell >= 0) {
    long long *c = regions + cell;
    lds_add_fixed(c, v.x);
    lds_add_fixed(c + kBinCells, v.y);
    lds_add_fixed(c + 2 * kBinCells, v.z);
    lds_add_fixed(c + 3 * kBinCells, v.w);
  } else if (g) {
    float *o = g + 4 * ((size_t)x + (size_t)y * (size_t)pitch);
    if (v.x != 0.0f) gadd(o + 0, v.x * invS);
    if (v.y != 0.0f) gadd(o + 1, v.y * invS);
    if (v.z != 0.0f) gadd(o + 2, v.z * invS);
    if (v.w != 0.0f) gadd(o + 3, v.w * invS);
  }
}

// Texel of cell coordinate r of an axis (the inverse of bin_cell_axis; -1:
// no texel, for a cell outside a clamped level).
__device__ __forceinline__ int bin_texel_axis(int r, int o, int dim, bool wrap) {
  int x = o + r;
  if (wrap) {
    x %= dim;
    return x < 0 ? x + dim : x;
  }
  return x >= 0 && x < dim ? x : -1;
}

// Adds the nonzero cells of a region (divided by S: times invS) to the
// gradient buffer and clears them.
__device__ __forceinline__ void bin_flush(long long *regions, int base, int side, int ox, int oy, int w, int h,
                                          bool wrapU, bool wrapV, float *g, int pitch, double invS) {
  for (int i = threadIdx.x; i < side * side; i += blockDim.x) {
    long long *c = regions + base + i;
    const long long q0 = c[0], q1 = c[kBinCells], q2 = c[2 * kBinCells], q3 = c[3 * kBinCells];
    if ((q0 | q1 | q2 | q3) == 0) continue;
    c[0] = 0;
    c[kBinCells] = 0;
    c[2 * kBinCells] = 0;
    c[3 * kBinCells] = 0;
    const int x = bin_texel_axis(i % side, ox, w, wrapU), y = bin_texel_axis(i / side, oy, h, wrapV);
    if (!g || x < 0 || y < 0) continue;
    float *o = g + 4 * ((size_t)x + (size_t)y * (size_t)pitch);
    if (q0) gadd(o + 0, (float)((double)q0 * invS));
    if (q1) gadd(o + 1, (float)((double)q1 * invS));
    if (q2) gadd(o + 2, (float)((double)q2 * invS));
    if (q3) gadd(o + 3, (float)((double)q3 * invS));
  }
}

// Largest |channel| of the record gradients over the workgroup (blockDim.x
// = kScatterThreads).
__device__ __forceinline__ float block_max(float v, float *red) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  float m = red[0];
  for (int w = 1; w < kScatterThreads / 64; w++) m = fmaxf(m, red[w]);
  __syncthreads();
  return m;
}

// The texel scatter of the sorted records, persistent: a workgroup takes
// work items (up to kBinChunk records of one bucket) from an atomic counter;
// each thread walks its records' taps exactly as texture_backward places and
// weights them (FP64 coordinates, FP32 products) and adds them to the
// bucket's tile at both levels in LDS; the tiles then reach the gradient
// buffers with one atomic per touched texel and channel.  Bucket b's records
// are rec[slot[k]] for k in [start[b], end[b]).  An item whose gradients are
// not all finite adds in float, straight to the gradient buffers.
__global__ __launch_bounds__(kScatterThreads) void tap_scatter_kernel(const TapRecord *__restrict__ rec,
                                                                      const uint32_t *__restrict__ slot,
                                                                      const uint32_t *__restrict__ start,
                                                                      const uint32_t *__restrict__ end,
                                                                      const uint32_t *__restrict__ chunkBase,
                                                                      uint32_t nBuckets,
                                                                      const BinBucket *__restrict__ buckets,
                                                                      const BinGroup *__restrict__ groups,
                                                                      const DrawParams *__restrict__ params,
                                                                      const GradParams *__restrict__ grads,
                                                                      uint32_t *__restrict__ counter) {
  constexpr int kR1At = kBinR0 * kBinR0;  // the coarser level's region follows the finer one's
  __shared__ long long Rq[4 * kBinCells];
  __shared__ float red[kScatterThreads / 64];
  __shared__ uint32_t sItem;
  for (int i = threadIdx.x; i < 4 * kBinCells; i += blockDim.x) Rq[i] = 0;
  const uint32_t items = chunkBase[nBuckets];
  for (;;) {
    if (threadIdx.x == 0) sItem = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t c = __builtin_amdgcn_readfirstlane(sItem);
    __syncthreads();
    if (c >= items) break;
    // Bucket of item c: the largest b with chunkBase[b] <= c.
    uint32_t lo = 0, hi = nBuckets;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (chunkBase[mid] <= c) lo = mid;
      else hi = mid;
    }
    const uint32_t bk = __builtin_amdgcn_readfirstlane(lo);
    const BinBucket B = buckets[bk];
    const BinGroup &Bg = groups[B.group];
    const TexParams &T = params[Bg.draw].tex;
    const GradParams &G = grads[Bg.draw];
    const int l0 = B.level, l1 = l0 + 1 < kMipLevels ? l0 + 1 : l0;
    // The item's sampler and level parameters live in scalar registers for
    // the whole record loop.
    const int filter = uniform_int(T.filter), addrU = uniform_int(T.addrU), addrV = uniform_int(T.addrV);
    const bool mipLinear = uniform_int(T.mipmap) == MIP_LINEAR;
    const int W0 = uniform_int(T.width[l0]), H0 = uniform_int(T.height[l0]), P0 = uniform_int(T.pitchP[l0]);
    const int W1 = uniform_int(T.width[l1]), H1 = uniform_int(T.height[l1]), P1 = uniform_int(T.pitchP[l1]);
    float *g0 = uniform_ptr(G.dTex[l0]), *g1 = uniform_ptr(G.dTex[l1]);
    const int ox0 = (int)B.tx * kBinTile - kBinHalo0, oy0 = (int)B.ty * kBinTile - kBinHalo0;
    const int ox1 = (int)B.tx * (kBinTile / 2) - kBinHalo1, oy1 = (int)B.ty * (kBinTile / 2) - kBinHalo1;
    const uint32_t first = start[bk] + (c - chunkBase[bk]) * kBinChunk;
    const uint32_t last = min(end[bk], first + kBinChunk);
    // Neighbouring records come from neighbouring pixels and hit the same
    // texels: thread t walks records first + t * steps + j, so that the lanes
    // of a wave add to texels of records far apart in the chunk.
    const uint32_t steps = (last - first + blockDim.x - 1) / blockDim.x;
    float m = 0.0f;
    for (uint32_t j = 0; j < steps; j++) {
      const uint32_t k = first + threadIdx.x * steps + j;
      if (k >= last) break;
      const float4 gr = rec[slot[k]].g;
      m = fmaxf(m, fmaxf(fmaxf(fabsf(gr.x), fabsf(gr.y)), fmaxf(fabsf(gr.z), fabsf(gr.w))));
      if (!(gr.x - gr.x == 0.0f && gr.y - gr.y == 0.0f && gr.z - gr.z == 0.0f && gr.w - gr.w == 0.0f))
        m = __int_as_float(0x7f800000);  // not finite
    }
    m = block_max(m, red);
    const bool fixed = m < __int_as_float(0x7f800000);
    // S = 2^(46 - e) with m < 2^e; at most 2^126 so that w * S stays finite.
    int e = 0;
    if (fixed && m > 0.0f) (void)frexpf(m, &e);
    const int sExp = fixed ? min(46 - e, 126) : 0;
    const float S = ldexpf(1.0f, sExp), invS = ldexpf(1.0f, -sExp);
    long long *const R = fixed ? Rq : nullptr;
    for (uint32_t j = 0; j < steps; j++) {
      const uint32_t k = first + threadIdx.x * steps + j;
      if (k >= last) break;
      const TapRecord r = rec[slot[k]];
      int N = 1;
      double A = 1.0, Bs = 0.0;
      if (filter == FILT_ANISOTROPIC) {
        N = rr_round_int(r.L.aniso);
        A = (double)uv_weight(aniso_row(N));
        Bs = (double)uv_start(aniso_row(N));
        N = N > 1 ? N : 1;  // at least one probe, as in the forward
      }
      double wl0 = 1.0, wl1 = 0.0;
      if (mipLinear) {
        wl1 = (double)rr_frac(r.L.lod);
        wl0 = 1.0 - wl1;
      }
      for (int lv = 0; lv < 2; lv++) {
        if (lv == 1 && wl1 == 0.0) break;
        const int W = lv ? W1 : W0, H = lv ? H1 : H0, pitch = lv ? P1 : P0;
        float *g = lv ? g1 : g0;
        const int base = lv ? kR1At : 0;
        const int side = fixed ? (lv ? kBinR1 : kBinR0) : 0, ox = lv ? ox1 : ox0, oy = lv ? oy1 : oy0;
        const double wgt = A * (lv ? wl1 : wl0);
        for (int i = 0; i < N; i++) {
          const double t = Bs + (double)i * A;
          const double ui = r.u + t * (double)r.L.uDelta, vi = r.v + t * (double)r.L.vDelta;
          int x0, x1, y0, y1;
          double fu, fv, dfu, dfv;
          axis64(filter, ui, W, addrU, x0, x1, fu, dfu);
          axis64(filter, vi, H, addrV, y0, y1, fv, dfv);
          // The weights as texture_backward rounds them, then scaled by S
          // (exact: a power of two).
          const float w00 = (float)(wgt * ((1.0 - fu) * (1.0 - fv))) * S, w10 = (float)(wgt * (fu * (1.0 - fv))) * S;
          const float w01 = (float)(wgt * ((1.0 - fu) * fv)) * S, w11 = (float)(wgt * (fu * fv)) * S;
          // The footprint's cells: rows and columns placed once for its four taps.
          const int rx0 = bin_cell_axis(x0, ox, W, addrU == 0), rx1 = bin_cell_axis(x1, ox, W, addrU == 0);
          const int ry0 = bin_cell_axis(y0, oy, H, addrV == 0), ry1 = bin_cell_axis(y1, oy, H, addrV == 0);
          const bool cx0 = (unsigned)rx0 < (unsigned)side, cx1 = (unsigned)rx1 < (unsigned)side;
          const bool cy0 = (unsigned)ry0 < (unsigned)side, cy1 = (unsigned)ry1 < (unsigned)side;
          const int row0 = base + ry0 * side, row1 = base + ry1 * side;
          bin_add(R, cx0 && cy0 ? row0 + rx0 : -1, g, pitch, invS, x0, y0,
                  make_float4(r.g.x * w00, r.g.y * w00, r.g.z * w00, r.g.w * w00));
          bin_add(R, cx1 && cy0 ? row0 + rx1 : -1, g, pitch, invS, x1, y0,
                  make_float4(r.g.x * w10, r.g.y * w10, r.g.z * w10, r.g.w * w10));
          bin_add(R, cx0 && cy1 ? row1 + rx0 : -1, g, pitch, invS, x0, y1,
                  make_float4(r.g.x * w01, r.g.y * w01, r.g.z * w01, r.g.w * w01));
          bin_add(R, cx1 && cy1 ? row1 + rx1 : -1, g, pitch, invS, x1, y1,
                  make_float4(r.g.x * w11, r.g.y * w11, r.g.z * w11, r.g.w * w11));
        }
      }
    }
    __syncthreads();
    if (fixed) {
      bin_flush(Rq, 0, kBinR0, ox0, oy0, W0, H0, addrU == 0, addrV == 0, g0, P0, (double)invS);
      bin_flush(Rq, kR1At, kBinR1, ox1, oy1, W1, H1, addrU == 0, addrV == 0, g1, P1, (double)invS);
    }
    __syncthreads();
  }
}

// 2x2 box filter (odd sizes: floor, the second texel clamped).
__global__ void mip_down_rgba32f_kernel(const float4 *src, int sw, int sh, int sp, float4 *dst, int dw, int dh, int dp) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dw || y >= dh) return;
  const int x0 = 2 * x < sw - 1 ? 2 * x : sw - 1, x1 = 2 * x + 1 < sw - 1 ? 2 * x + 1 : sw - 1;
  const int y0 = 2 * y < sh - 1 ? 2 * y : sh - 1, y1 = 2 * y + 1 < sh - 1 ? 2 * y + 1 : sh - 1;
  const float4 a = src[x0 + (size_t)y0 * sp], b = src[x1 + (size_t)y0 * sp];
  const float4 c = src[x0 + (size_t)y1 * sp], d = src[x1 + (size_t)y1 * sp];
  dst[x + (size_t)y * dp] = make_float4(((a.x + b.x) + (c.x + d.x)) * 0.25f, ((a.y + b.y) + (c.y + d.y)) * 0.25f,
                                        ((a.z + b.z) + (c.z + d.z)) * 0.25f, ((a.w + b.w) + (c.w + d.w)) * 0.25f);
}

__global__ void mip_down_rgba8_kernel(const uint32_t *src, int sw, int sh, int sp, uint32_t *dst, int dw, int dh,
                                      int dp) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dw || y >= dh) return;
  const int x0 = 2 * x < sw - 1 ? 2 * x : sw - 1, x1 = 2 * x + 1 < sw - 1 ? 2 * x + 1 : sw - 1;
  const int y0 = 2 * y < sh - 1 ? 2 * y : sh - 1, y1 = 2 * y + 1 < sh - 1 ? 2 * y + 1 : sh - 1;
  const uint32_t a = src[x0 + (size_t)y0 * sp], b = src[x1 + (size_t)y0 * sp];
  const uint32_t c = src[x0 + (size_t)y1 * sp], d = src[x1 + (size_t)y1 * sp];
  uint32_t o = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t s = ((a >> (8 * k)) & 0xFFu) + ((b >> (8 * k)) & 0xFFu) + ((c >> (8 * k)) & 0xFFu) +
                       ((d >> (8 * k)) & 0xFFu) + 2u;
    o |= (s >> 2) << (8 * k);
  }
  dst[x + (size_t)y * dp] = o;
}

// Adjoint of mip_down_rgba32f_kernel: texel (x, y) of the finer level
// receives 0.25 * (its multiplicity in the box) of its parent's gradient.
__global__ void mip_up_rgba32f_kernel(float4 *fine, int fw, int fh, int fp, const float4 *coarse, int cw, int ch,
                                      int cp) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= fw || y >= fh) return;
  const int mx = fw == 1 ? 2 : (x < 2 * cw ? 1 : 0);
  const int my = fh == 1 ? 2 : (y < 2 * ch ? 1 : 0);
  if (!mx || !my) return;
  const float4 g = coarse[(x >> 1) + (size_t)(y >> 1) * cp];
  const float s = 0.25f * (float)(mx * my);
  float4 &f = fine[x + (size_t)y * fp];
  f.x += g.x * s;
  f.y += g.y * s;
  f.z += g.z * s;
  f.w += g.w * s;
}

template __global__ void backward_kernel<1>(DrawParams, GradParams, const TriRecord *);
template __global__ void backward_kernel<4>(DrawParams, GradParams, const TriRecord *);
template __global__ void backward_batch_kernel<1, false>(const DrawParams *, const GradParams *, const uint32_t *, int,
                                                         int, uint32_t, uint32_t, uint32_t *, const TriRecord *,
                                                         BinArgs);
template __global__ void backward_batch_kernel<4, false>(const DrawParams *, const GradParams *, const uint32_t *, int,
                                                         int, uint32_t, uint32_t, uint32_t *, const TriRecord *,
                                                         BinArgs);
template __global__ void backward_batch_kernel<1, true>(const DrawParams *, const GradParams *, const uint32_t *, int,
                                                        int, uint32_t, uint32_t, uint32_t *, const TriRecord *,
                                                        BinArgs);
template __global__ void backward_batch_kernel<4, true>(const DrawParams *, const GradParams *, const uint32_t *, int,
                                                        int, uint32_t, uint32_t, uint32_t *, const TriRecord *,
                                                        BinArgs);

}  // namespace swr
