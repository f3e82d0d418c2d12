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

struct GradLds {
  WaveTexTable tex[kWaves];
  float *dTex[kMipLevels];  // texture-gradient buffers of the draw being processed, per level
  uint32_t triKey[kTriSlots];
  float triVal[kTriSlots][kTriVals];
  uint32_t triCount;  // occupied slots
};

__device__ __forceinline__ WaveTexTable &wave_table(GradLds &S) { return S.tex[threadIdx.x >> 6]; }

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
      unsafeAtomicAdd(g + 4 * (size_t)(key & ((1u << 27) - 1)) + (j & 3), v);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (uint32_t j = rank; j < (uint32_t)kWaveSlots; j += n) W.key[j] = kEmptyKey;
  if (rank == 0) W.count = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void lds_init_tri(GradLds &S) {
  for (int i = threadIdx.x; i < kTriSlots; i += blockDim.x) S.triKey[i] = kEmptyKey;
  for (int i = threadIdx.x; i < kTriSlots * kTriVals; i += blockDim.x) (&S.triVal[0][0])[i] = 0.0f;
  if (threadIdx.x == 0) S.triCount = 0;
}

__device__ __forceinline__ void lds_init(GradLds &S, const GradParams &G) {
  wave_tex_init(wave_table(S));
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
  unsafeAtomicAdd(o + 0, a);
  unsafeAtomicAdd(o + 1, b);
  unsafeAtomicAdd(o + 2, c);
  unsafeAtomicAdd(o + 3, d);
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
        unsafeAtomicAdd(o + 0, gTex.x * w[k]);
        unsafeAtomicAdd(o + 1, gTex.y * w[k]);
        unsafeAtomicAdd(o + 2, gTex.z * w[k]);
        unsafeAtomicAdd(o + 3, gTex.w * w[k]);
      }
    }
  }
}

// Slot of primitive t in the workgroup table, or NULL when the table is full.
__device__ __forceinline__ float *tri_slot(GradLds &S, uint32_t t) {
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
      unsafeAtomicAdd(o + 0, v[3 * k + 0]);
      unsafeAtomicAdd(o + 1, v[3 * k + 1]);
      unsafeAtomicAdd(o + 3, v[3 * k + 2]);
    }
    if (G.dCol) {
      float *o = G.dCol + 4 * (size_t)vi[k];
      unsafeAtomicAdd(o + 0, v[9 + 4 * k + 0]);
      unsafeAtomicAdd(o + 1, v[9 + 4 * k + 1]);
      unsafeAtomicAdd(o + 2, v[9 + 4 * k + 2]);
      unsafeAtomicAdd(o + 3, v[9 + 4 * k + 3]);
    }
    if (G.dUV) {
      float *o = G.dUV + 2 * (size_t)vi[k];
      unsafeAtomicAdd(o + 0, v[21 + 2 * k + 0]);
      unsafeAtomicAdd(o + 1, v[21 + 2 * k + 1]);
    }
  }
}

__device__ __forceinline__ void lds_flush_tri(GradLds &S, const DrawParams &P, const GradParams &G) {
  for (int i = threadIdx.x; i < kTriSlots; i += blockDim.x) {
    const uint32_t t = S.triKey[i];
    if (t != kEmptyKey) tri_scatter_global(P, G, t, S.triVal[i]);
  }
}

// Flushes both tables (all threads).
__device__ __forceinline__ void lds_flush(GradLds &S, const DrawParams &P, const GradParams &G) {
  wave_tex_flush(wave_table(S), S.dTex);
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
__device__ __forceinline__ void tap_backward(WaveTexTable &W, const TexParams &T, const LevelRef &R, double u,
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
  float *g = R.grad;
  if (g) {
    const uint32_t idx[4] = {(uint32_t)i00, (uint32_t)i10, (uint32_t)i01, (uint32_t)i11};
    const float w[4] = {w00, w10, w01, w11};
    texels_add4(W, g, R.lvl, idx, w, gTex);
  }
}

// The forward filter at (u, v) with the frozen state L: anisotropic probes
// along (uDelta, vDelta) (sampleFloatAniso, SamplerCore.cpp:1085-1141),
// each trilinear between two levels (sampleTexture128, :200-253).
__device__ __forceinline__ void texture_backward(GradLds &S, const TexParams &T, const GradParams &G,
                                                 const QuadLod &L, double u, double v, const float4 &gTex,
                                                 float4 &val, float &dU, float &dV) {
  WaveTexTable &W = wave_table(S);
  int N = 1;
  double A = 1.0, B = 0.0;
  if (T.filter == FILT_ANISOTROPIC) {
    N = rr_round_int(L.aniso);
    A = (double)uv_weight(aniso_row(N));
    B = (double)uv_start(aniso_row(N));
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
      if (anyGrad && __builtin_amdgcn_readfirstlane(W.count) > kFlushAt) wave_tex_flush(W, S.dTex);
      if (on) {
        const double t = B + (double)i * A;
        const double ui = u + t * (double)L.uDelta, vi = v + t * (double)L.vDelta;
        tap_backward(W, T, R, ui, vi, A * wl, gTex, val, dU, dV);
      }
    }
  }
}

// Gradient of one fragment: primitive t at pixel (px, py), upstream g.
__device__ void fragment_backward(GradLds &S, const DrawParams &P, const GradParams &G, uint32_t t, int px, int py,
                                  const Acc4 &g, const QuadLod &L) {
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
  if (!(D != 0.0)) return;
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
  if (!(Ssum != 0.0)) return;
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
  if (P.useUV) {
    const float4 gTex = P.useColor ? make_float4(gf.x * col.x, gf.y * col.y, gf.z * col.z, gf.w * col.w) : gf;
    float4 tex = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    texture_backward(S, P.tex, G, L, u, v, gTex, tex, dU, dV);
    if (P.useColor) dcol = make_float4(gf.x * tex.x, gf.y * tex.y, gf.z * tex.z, gf.w * tex.w);
  } else {
    dcol = gf;
  }

  // Attribute gradients and dL/db.
  float val[kTriVals];
#pragma unroll
  for (int k = 0; k < kTriVals; k++) val[k] = 0.0f;
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
  if (!G.dPos && !G.dCol && !G.dUV) return;
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

// Runs the wave's queued fragments of draw P (rec holds the draw's records:
// slot t = primitive t) 64 at a time, and the rest as well when `all`, so
// that edge pixels with several primitives cost extra lanes rather than
// extra passes, and partial batches occur only at draw changes.
template <int MS>
__device__ __forceinline__ void drain_fragments(GradLds &S, FragQueue &Q, RingPos &rp, const DrawParams &P,
                                                const GradParams &G, const TriRecord *__restrict__ rec, bool all) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  while (rp.count >= 64 || (all && rp.count > 0)) {
    const uint32_t n = rp.count < 64 ? rp.count : 64;
    const uint32_t e = (rp.head + (uint32_t)lane) & (kRing - 1);
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
      fragment_backward(S, P, G, t, x, y, g, L);
    }
    rp.head = (rp.head + n) & (kRing - 1);
    rp.count -= n;
  }
}

// One draw: 16x16 pixels per workgroup.
template <int MS>
__global__ __launch_bounds__(256) void backward_kernel(DrawParams P, GradParams G, const TriRecord *__restrict__ rec) {
  __shared__ GradLds S;
  __shared__ FragQueue Q;
  lds_init(S, G);
  RingPos rp = {0u, 0u};
  enqueue_fragments<MS>(Q, rp, P, blockIdx.x * 16 + (threadIdx.x & 15), blockIdx.y * 16 + (threadIdx.x >> 4));
  drain_fragments<MS>(S, Q, rp, P, G, rec, true);
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
// the same draw.
template <int MS>
__global__ __launch_bounds__(256) void backward_batch_kernel(const DrawParams *__restrict__ params,
                                                             const GradParams *__restrict__ grads,
                                                             const uint32_t *__restrict__ blockBase, int drawLo,
                                                             int drawHi, uint32_t blockLo, uint32_t blockHi,
                                                             uint32_t *__restrict__ counter,
                                                             const TriRecord *__restrict__ rec) {
  __shared__ GradLds S;
  __shared__ FragQueue Q;
  __shared__ uint32_t sBlock;
  lds_init(S, grads[drawLo]);
  int cur = -1;  // draw whose gradients the tables hold (wave-uniform)
  RingPos rp = {0u, 0u};
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
      drain_fragments<MS>(S, Q, rp, params[cur], grads[cur], rec + params[cur].triBase, true);
      __syncthreads();
    }
    if (cur >= 0) {
      bool texSwitch = lo < 0;
      for (int k = 0; k < kMipLevels && !texSwitch; k++) texSwitch = grads[lo].dTex[k] != grads[cur].dTex[k];
      const bool fr = lo != cur || triCount > kTriSlots / 2;
      if (texSwitch) wave_tex_flush(wave_table(S), S.dTex);
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
    drain_fragments<MS>(S, Q, rp, P, grads[lo], rec + P.triBase, false);
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
template __global__ void backward_batch_kernel<1>(const DrawParams *, const GradParams *, const uint32_t *, int, int,
                                                  uint32_t, uint32_t, uint32_t *, const TriRecord *);
template __global__ void backward_batch_kernel<4>(const DrawParams *, const GradParams *, const uint32_t *, int, int,
                                                  uint32_t, uint32_t, uint32_t *, const TriRecord *);

}  // namespace swr
