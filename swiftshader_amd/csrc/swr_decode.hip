// swr_decode.hip -- compressed texture decoding: BC1-BC7 (BC_Decoder::Decode,
// src/Device/BC_Decoder.cpp:1746-1771) and ETC2 / EAC (ETC_Decoder::Decode,
// src/Device/ETC_Decoder.cpp:670-734).
//
// One thread per pixel.  A pixel's thread reads its 4x4 block (8 or 16
// bytes; the threads of one block read the same words, which the cache
// serves once) and decodes only its own texel: bit positions inside a block
// are closed forms of the texel index (index-stream offsets count the
// shorter anchor indices before the texel), so no thread walks the other
// 15 texels and nothing is indexed through scratch memory.  A row of pixels
// is written by consecutive threads at consecutive addresses.
//
// Outputs follow the reference's layouts: BC1-BC3 and BC7 write B8G8R8A8
// pixels, BC4 / BC5 one / two bytes, BC6H four half floats (alpha 1.0);
// ETC2 RGB, RGB with punch-through alpha and RGBA write B8G8R8A8, EAC R11
// one and RG11 two 16-bit values per pixel.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swr {

// ---------------------------------------------------------------------------
// BC1-BC5
// ---------------------------------------------------------------------------
// BC_Decoder.cpp:31-217, 1662-1777.  Colour endpoints widen 565 by bit
// replication, blue in byte 0; 4-colour mode (c0 > c1, or a separate alpha
// block) takes thirds, 3-colour mode halves and black, transparent when the
// format has alpha.
// Channel blocks take sevenths (c0 > c1) or fifths plus the format's
// extremes; divisions truncate toward zero, as C's do.
__device__ __forceinline__ int bc_channel_value(uint64_t blk, int t, bool isSigned) {
  const int c0 = isSigned ? (int)(int8_t)(blk & 0xFF) : (int)(blk & 0xFF);
  const int c1 = isSigned ? (int)(int8_t)((blk >> 8) & 0xFF) : (int)((blk >> 8) & 0xFF);
  const int i = (int)((blk >> (16 + 3 * t)) & 7);
  if (i == 0) return c0;
  if (i == 1) return c1;
  if (c0 > c1) return ((8 - i) * c0 + (i - 1) * c1) / 7;
  if (i < 6) return ((6 - i) * c0 + (i - 1) * c1) / 5;
  return i == 6 ? (isSigned ? -128 : 0) : (isSigned ? 127 : 255);
}

__device__ __forceinline__ uint32_t bc_color_value(uint64_t blk, int t, bool hasAlpha, bool separateAlpha) {
  const uint32_t c0 = (uint32_t)(blk & 0xFFFF), c1 = (uint32_t)((blk >> 16) & 0xFFFF);
  const int i = (int)((blk >> (32 + 2 * t)) & 3);
  const bool four = separateAlpha || c0 > c1;
  uint32_t out = 0;
#pragma unroll
  for (int ch = 0; ch < 3; ch++) {
    int e0, e1;
    if (ch == 0) {
      e0 = (int)(((c0 & 0x1Fu) << 3) | ((c0 & 0x1Cu) >> 2));
      e1 = (int)(((c1 & 0x1Fu) << 3) | ((c1 & 0x1Cu) >> 2));
    } else if (ch == 1) {
      e0 = (int)(((c0 & 0x7E0u) >> 3) | ((c0 & 0x600u) >> 9));
      e1 = (int)(((c1 & 0x7E0u) >> 3) | ((c1 & 0x600u) >> 9));
    } else {
      e0 = (int)(((c0 & 0xF800u) >> 8) | ((c0 & 0xE000u) >> 13));
      e1 = (int)(((c1 & 0xF800u) >> 8) | ((c1 & 0xE000u) >> 13));
    }
    int v;
    if (i == 0) v = e0;
    else if (i == 1) v = e1;
    else if (i == 2) v = four ? (2 * e0 + e1) / 3 : (e0 + e1) >> 1;
    else v = four ? (2 * e1 + e0) / 3 : 0;
    out |= (uint32_t)v << (8 * ch);
  }
  const bool transparent = i == 3 && !four && hasAlpha;
  return out | (transparent ? 0u : 0xFF000000u);
}

__global__ void decode_bc_kernel(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, int w, int h, int pitch,
                                 int bpp, int n, int noAlphaU) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w || y >= h) return;
  const int bw = (w + 3) >> 2;
  const int t = (y & 3) * 4 + (x & 3);
  const long long b = (long long)(y >> 2) * bw + (x >> 2);
  uint8_t *px = dst + (long long)y * pitch + (long long)x * bpp;
  if (n == 1 || n == 4) {
    const uint64_t blk = reinterpret_cast<const uint64_t *>(src)[b];
    if (n == 1) *reinterpret_cast<uint32_t *>(px) = bc_color_value(blk, t, !noAlphaU, false);
    else px[0] = (uint8_t)bc_channel_value(blk, t, !noAlphaU);
    return;
  }
  const uint64_t lo = reinterpret_cast<const uint64_t *>(src)[2 * b], hi = reinterpret_cast<const uint64_t *>(src)[2 * b + 1];
  if (n == 5) {
    px[0] = (uint8_t)bc_channel_value(lo, t, !noAlphaU);
    px[1] = (uint8_t)bc_channel_value(hi, t, !noAlphaU);
    return;
  }
  uint32_t v = bc_color_value(hi, t, false, true);
  const uint32_t a = n == 2 ? (uint32_t)((lo >> (4 * t)) & 0xF) * 0x11u : (uint32_t)(uint8_t)bc_channel_value(lo, t, false);
  *reinterpret_cast<uint32_t *>(px) = (v & 0x00FFFFFFu) | (a << 24);
}

// ---------------------------------------------------------------------------
// BC6H and BC7 (BPTC): a 128-bit little-endian bit stream read LSB first.
// ---------------------------------------------------------------------------
// Partition tables of the BPTC formats.  Two subsets: bit t = subset of
// texel t (t = x + 4 y).  Three subsets: bits 2t+1..2t.  Anchor texels of
// subset 1 of 2 (bits 3..0), subset 1 of 3 (7..4) and subset 2 of 3 (11..8);
// texel 0 anchors subset 0.
__constant__ const uint16_t kBptcPart2[64] = {
    0xcccc, 0x8888, 0xeeee, 0xecc8, 0xc880, 0xfeec, 0xfec8, 0xec80, 0xc800, 0xffec, 0xfe80, 0xe800, 0xffe8,
    0xff00, 0xfff0, 0xf000, 0xf710, 0x008e, 0x7100, 0x08ce, 0x008c, 0x7310, 0x3100, 0x8cce, 0x088c, 0x3110,
    0x6666, 0x366c, 0x17e8, 0x0ff0, 0x718e, 0x399c, 0xaaaa, 0xf0f0, 0x5a5a, 0x33cc, 0x3c3c, 0x55aa, 0x9696,
    0xa55a, 0x73ce, 0x13c8, 0x324c, 0x3bdc, 0x6996, 0xc33c, 0x9966, 0x0660, 0x0272, 0x04e4, 0x4e40, 0x2720,
    0xc936, 0x936c, 0x39c6, 0x639c, 0x9336, 0x9cc6, 0x817e, 0xe718, 0xccf0, 0x0fcc, 0x7744, 0xee22,
};
__constant__ const uint32_t kBptcPart3[64] = {
    0xaa685050, 0x6a5a5040, 0x5a5a4200, 0x5450a0a8, 0xa5a50000, 0xa0a05050, 0x5555a0a0, 0x5a5a5050,
    0xaa550000, 0xaa555500, 0xaaaa5500, 0x90909090, 0x94949494, 0xa4a4a4a4, 0xa9a59450, 0x2a0a4250,
    0xa5945040, 0x0a425054, 0xa5a5a500, 0x55a0a0a0, 0xa8a85454, 0x6a6a4040, 0xa4a45000, 0x1a1a0500,
    0x0050a4a4, 0xaaa59090, 0x14696914, 0x69691400, 0xa08585a0, 0xaa821414, 0x50a4a450, 0x6a5a0200,
    0xa9a58000, 0x5090a0a8, 0xa8a09050, 0x24242424, 0x00aa5500, 0x24924924, 0x24499224, 0x50a50a50,
    0x500aa550, 0xaaaa4444, 0x66660000, 0xa5a0a5a0, 0x50a050a0, 0x69286928, 0x44aaaa44, 0x66666600,
    0xaa444444, 0x54a854a8, 0x95809580, 0x96969600, 0xa85454a8, 0x80959580, 0xaa141414, 0x96960000,
    0xaaaa1414, 0xa05050a0, 0xa0a5a5a0, 0x96000000, 0x40804080, 0xa9a8a9a8, 0xaaaaaa44, 0x2a4a5254,
};
__constant__ const uint16_t kBptcAnchor[64] = {
    0xf3f, 0x83f, 0x8ff, 0x3ff, 0xf8f, 0xf3f, 0x3ff, 0x8ff, 0xf8f, 0xf8f, 0xf6f, 0xf6f, 0xf6f, 0xf5f, 0xf3f, 0x83f,
    0xf3f, 0x832, 0xf88, 0x3f2, 0xf32, 0x838, 0xf68, 0x8af, 0x352, 0xf88, 0x682, 0xa62, 0xf88, 0xf58, 0xaf2, 0x8f2,
    0xf8f, 0x3ff, 0xf36, 0xa58, 0xa62, 0x8a8, 0x98f, 0xaff, 0x6f2, 0xf38, 0x8f2, 0xf52, 0x3f2, 0x6ff, 0x6ff, 0x8f6,
    0xf36, 0x3f2, 0xf56, 0xf58, 0xf5f, 0xf8f, 0xf52, 0xfa2, 0xf5f, 0xfaf, 0xf8f, 0xfdf, 0x3ff, 0xfc2, 0xf32, 0x83f,
};

// Interpolation weights (out of 64) for 2-, 3- and 4-bit indices, packed
// 4 (2-bit), 8 (3-bit) or 16 (4-bit) bytes: weight i of b-bit indices is
// byte i of kBptcWeights[b - 2].
__constant__ const uint8_t kBptcWeights[3][16] = {
    {0, 21, 43, 64},
    {0, 9, 18, 27, 37, 46, 55, 64},
    {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64},
};

// n bits (0..32) of the stream from bit `off` (0..127).
__device__ __forceinline__ uint32_t bptc_bits(uint64_t lo, uint64_t hi, int off, int n) {
  if (n == 0) return 0u;
  uint64_t v;
  if (off >= 64) v = hi >> (off - 64);
  else if (off == 0) v = lo;
  else v = (lo >> off) | (hi << (64 - off));
  return (uint32_t)(v & ((1ull << n) - 1ull));
}

// BC7 modes: subsets, partition bits, rotation bits, index-selection bit,
// colour bits, alpha bits, per-endpoint and shared p-bits, primary and
// secondary index bits.
struct Bc7Mode {
  uint8_t ns, pb, rb, isb, cb, ab, epb, spb, ib, ib2;
};
__constant__ const Bc7Mode kBc7Modes[8] = {
    {3, 4, 0, 0, 4, 0, 1, 0, 3, 0}, {2, 6, 0, 0, 6, 0, 0, 1, 3, 0}, {3, 6, 0, 0, 5, 0, 0, 0, 2, 0},
    {2, 6, 0, 0, 7, 0, 1, 0, 2, 0}, {1, 0, 2, 1, 5, 6, 0, 0, 2, 3}, {1, 0, 2, 0, 7, 8, 0, 0, 2, 2},
    {1, 0, 0, 0, 7, 7, 1, 0, 4, 0}, {2, 6, 0, 0, 5, 5, 1, 0, 2, 0},
};

// A b-bit endpoint widened to 8 bits by replicating its high bits.
__device__ __forceinline__ int bc7_widen(int v, int b) {
  v = (v << (8 - b)) & 0xFF;
  return v | (v >> b);
}

__device__ __forceinline__ int bptc_lerp(int e0, int e1, int idx, int bits) {
  const int w = kBptcWeights[bits - 2][idx];
  return ((64 - w) * e0 + w * e1 + 32) >> 6;
}

// Texel t of a BC7 block, packed B8G8R8A8.  Mode m is the position of the
// lowest set bit of the first byte; no set bit: (0, 0, 0, 0).
__device__ uint32_t bc7_texel(uint64_t lo, uint64_t hi, int t) {
  const uint32_t first = (uint32_t)(lo & 0xFF);
  if (!first) return 0u;
  const int m = __builtin_ctz(first);
  const Bc7Mode M = kBc7Modes[m];
  const int ns = M.ns, ne = 2 * ns;
  int pos = m + 1;
  const int part = (int)bptc_bits(lo, hi, pos, M.pb);
  pos += M.pb;
  const int rot = (int)bptc_bits(lo, hi, pos, M.rb);
  pos += M.rb;
  const int sel = (int)bptc_bits(lo, hi, pos, M.isb);
  pos += M.isb;
  // Subset of the texel and the anchors below it.
  int s = 0, a1 = 16, a2 = 16;
  if (ns == 2) {
    s = (kBptcPart2[part] >> t) & 1;
    a1 = kBptcAnchor[part] & 15;
  } else if (ns == 3) {
    s = (int)((kBptcPart3[part] >> (2 * t)) & 3u);
    a1 = (kBptcAnchor[part] >> 4) & 15;
    a2 = (kBptcAnchor[part] >> 8) & 15;
  }
  // Stream layout: R of every endpoint, then G, B, then A, then p-bits,
  // then the primary and the secondary index streams.
  const int colorAt = pos, alphaAt = colorAt + 3 * ne * M.cb, pbitAt = alphaAt + ne * M.ab;
  const int primAt = pbitAt + (M.epb ? ne : (M.spb ? 2 : 0));
  const int secAt = primAt + 16 * M.ib - ns;
  const int e0 = 2 * s, e1 = 2 * s + 1;
  const int pb0 = M.epb ? (int)bptc_bits(lo, hi, pbitAt + e0, 1) : (M.spb ? (int)bptc_bits(lo, hi, pbitAt + s, 1) : 0);
  const int pb1 = M.epb ? (int)bptc_bits(lo, hi, pbitAt + e1, 1) : pb0;
  const bool hasP = M.epb || M.spb;
  const int cbits = M.cb + (hasP ? 1 : 0), abits = M.ab + M.epb;
  int c0[4], c1[4];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    int v0 = (int)bptc_bits(lo, hi, colorAt + (c * ne + e0) * M.cb, M.cb);
    int v1 = (int)bptc_bits(lo, hi, colorAt + (c * ne + e1) * M.cb, M.cb);
    if (hasP) {
      v0 = v0 << 1 | pb0;
      v1 = v1 << 1 | pb1;
    }
    c0[c] = bc7_widen(v0, cbits);
    c1[c] = bc7_widen(v1, cbits);
  }
  if (M.ab) {
    int v0 = (int)bptc_bits(lo, hi, alphaAt + e0 * M.ab, M.ab), v1 = (int)bptc_bits(lo, hi, alphaAt + e1 * M.ab, M.ab);
    if (M.epb) {
      v0 = v0 << 1 | pb0;
      v1 = v1 << 1 | pb1;
    }
    c0[3] = bc7_widen(v0, abits);
    c1[3] = bc7_widen(v1, abits);
  } else {
    c0[3] = c1[3] = 255;
  }
  // Index offsets: every texel before t takes ib bits, the anchors below it
  // one bit fewer; an anchor texel's own index is one bit shorter.
  const int below = (t > 0 ? 1 : 0) + (a1 < t ? 1 : 0) + (a2 < t ? 1 : 0);
  const int anchor = (t == 0 || t == a1 || t == a2) ? 1 : 0;
  const int pIdx = (int)bptc_bits(lo, hi, primAt + M.ib * t - below, M.ib - anchor);
  const int sIdx = M.ib2 ? (int)bptc_bits(lo, hi, secAt + M.ib2 * t - (t > 0 ? 1 : 0), M.ib2 - (t == 0 ? 1 : 0)) : 0;
  // Colour takes the secondary index when the selection bit is set, alpha
  // when there is a secondary index and the bit is clear.
  const bool cSec = M.ib2 && sel, aSec = M.ib2 && !sel;
  const int cIdx = cSec ? sIdx : pIdx, cBits = cSec ? M.ib2 : M.ib;
  const int aIdx = aSec ? sIdx : pIdx, aBits = aSec ? M.ib2 : M.ib;
  int r = bptc_lerp(c0[0], c1[0], cIdx, cBits), g = bptc_lerp(c0[1], c1[1], cIdx, cBits);
  int b = bptc_lerp(c0[2], c1[2], cIdx, cBits), a = bptc_lerp(c0[3], c1[3], aIdx, aBits);
  // Rotation exchanges alpha with R (1), G (2) or B (3).
  if (rot == 1) { const int x = a; a = r; r = x; }
  else if (rot == 2) { const int x = a; a = g; g = x; }
  else if (rot == 3) { const int x = a; a = b; b = x; }
  return (uint32_t)b | (uint32_t)g << 8 | (uint32_t)r << 16 | (uint32_t)a << 24;
}

// BC6H modes: the mode value (2 or 5 bits), its bit count, regions,
// endpoint bits, whether endpoints 1-3 are deltas from endpoint 0 and the
// delta bits per channel, then the header's fields in stream order after
// the mode bits.  A field code is
//   dest << 9 | first << 5 | (count - 1) << 1 | descending
// where dest is endpoint * 3 + channel (R, G, B) or 12 for the partition;
// an ascending field puts its count stream bits at value bits first,
// first + 1, ...; a descending one at first, first - 1, ... (the format
// tables' reversed ranges, e.g. bits 10..15 of endpoint 0 stored bit 15
// first).
struct Bc6Mode {
  uint8_t mode, modeBits, regions, epBits, delta, dBits[3], nFields;
  uint16_t field[24];
};
__constant__ const Bc6Mode kBc6Modes[14] = {
    {0, 2, 2, 10, 1, {5, 5, 5}, 20,
     {0x0e80, 0x1080, 0x1680, 0x0012, 0x0212, 0x0412, 0x0608, 0x1480, 0x0e06, 0x0808, 0x1600, 0x1406, 0x0a08, 0x1620,
      0x1006, 0x0c08, 0x1640, 0x1208, 0x1660, 0x1808}},
    {1, 2, 2, 7, 1, {6, 6, 6}, 22,
     {0x0ea0, 0x1482, 0x000c, 0x1602, 0x1080, 0x020c, 0x10a0, 0x1640, 0x0e80, 0x040c, 0x1660, 0x16a0, 0x1680, 0x060a,
      0x0e06, 0x080a, 0x1406, 0x0a0a, 0x1006, 0x0c0a, 0x120a, 0x1808}},
    {2, 5, 2, 11, 1, {5, 4, 4}, 19,
     {0x0012, 0x0212, 0x0412, 0x0608, 0x0140, 0x0e06, 0x0806, 0x0340, 0x1600, 0x1406, 0x0a06, 0x0540, 0x1620, 0x1006,
      0x0c08, 0x1640, 0x1208, 0x1660, 0x1808}},
    {6, 5, 2, 11, 1, {4, 5, 4}, 21,
     {0x0012, 0x0212, 0x0412, 0x0606, 0x0140, 0x1480, 0x0e06, 0x0808, 0x0340, 0x1406, 0x0a06, 0x0540, 0x1620, 0x1006,
      0x0c06, 0x1600, 0x1640, 0x1206, 0x0e80, 0x1660, 0x1808}},
    {10, 5, 2, 11, 1, {4, 4, 5}, 21,
     {0x0012, 0x0212, 0x0412, 0x0606, 0x0140, 0x1080, 0x0e06, 0x0806, 0x0340, 0x1600, 0x1406, 0x0a08, 0x0540, 0x1006,
      0x0c06, 0x1620, 0x1640, 0x1206, 0x1680, 0x1660, 0x1808}},
    {14, 5, 2, 9, 1, {5, 5, 5}, 20,
     {0x0010, 0x1080, 0x0210, 0x0e80, 0x0410, 0x1680, 0x0608, 0x1480, 0x0e06, 0x0808, 0x1600, 0x1406, 0x0a08, 0x1620,
      0x1006, 0x0c08, 0x1640, 0x1208, 0x1660, 0x1808}},
    {18, 5, 2, 8, 1, {6, 5, 5}, 20,
     {0x000e, 0x1480, 0x1080, 0x020e, 0x1640, 0x0e80, 0x040e, 0x1660, 0x1680, 0x060a, 0x0e06, 0x0808, 0x1600, 0x1406,
      0x0a08, 0x1620, 0x1006, 0x0c0a, 0x120a, 0x1808}},
    {22, 5, 2, 8, 1, {5, 6, 5}, 22,
     {0x000e, 0x1600, 0x1080, 0x020e, 0x0ea0, 0x0e80, 0x040e, 0x14a0, 0x1680, 0x0608, 0x1480, 0x0e06, 0x080a, 0x1406,
      0x0a08, 0x1620, 0x1006, 0x0c08, 0x1640, 0x1208, 0x1660, 0x1808}},
    {26, 5, 2, 8, 1, {5, 5, 6}, 22,
     {0x000e, 0x1620, 0x1080, 0x020e, 0x10a0, 0x0e80, 0x040e, 0x16a0, 0x1680, 0x0608, 0x1480, 0x0e06, 0x0808, 0x1600,
      0x1406, 0x0a0a, 0x1006, 0x0c08, 0x1640, 0x1208, 0x1660, 0x1808}},
    {30, 5, 2, 6, 0, {6, 6, 6}, 24,
     {0x000a, 0x1480, 0x1600, 0x1620, 0x1080, 0x020a, 0x0ea0, 0x10a0, 0x1640, 0x0e80, 0x040a, 0x14a0, 0x1660, 0x16a0,
      0x1680, 0x060a, 0x0e06, 0x080a, 0x1406, 0x0a0a, 0x1006, 0x0c0a, 0x120a, 0x1808}},
    {3, 5, 1, 10, 0, {10, 10, 10}, 6, {0x0012, 0x0212, 0x0412, 0x0612, 0x0812, 0x0a12}},
    {7, 5, 1, 11, 1, {9, 9, 9}, 9, {0x0012, 0x0212, 0x0412, 0x0610, 0x0140, 0x0810, 0x0340, 0x0a10, 0x0540}},
    {11, 5, 1, 12, 1, {8, 8, 8}, 9, {0x0012, 0x0212, 0x0412, 0x060e, 0x0163, 0x080e, 0x0363, 0x0a0e, 0x0563}},
    {15, 5, 1, 16, 1, {4, 4, 4}, 9, {0x0012, 0x0212, 0x0412, 0x0606, 0x01eb, 0x0806, 0x03eb, 0x0a06, 0x05eb}},
};

// The low `bits` bits of v sign-extended, as a 16-bit pattern.
__device__ __forceinline__ int bc6_sext16(int v, int bits) {
  const int m = 1 << (bits - 1);
  return (int)(uint16_t)(int16_t)((v ^ m) - m);
}

// Endpoint value (a 16-bit pattern) to the 16-bit interpolation domain.
__device__ __forceinline__ int bc6_unquantize(int v, int bits, bool isSigned) {
  if (!isSigned) {
    if (bits >= 15 || v == 0) return v;
    if (v == (1 << bits) - 1) return 0xFFFF;
    return (int)((((uint32_t)v << 16) + 0x8000u) >> bits);
  }
  if (bits >= 16 || v == 0) return v;
  const int sv = (int)(int16_t)(uint16_t)v;
  const bool neg = sv < 0;
  const int mag = neg ? -sv : sv;
  const int r = mag >= (1 << (bits - 1)) - 1 ? 0x7FFF : ((mag << 15) + 0x4000) >> (bits - 1);
  return (int)(uint16_t)(neg ? -r : r);
}

// R, G, B half floats of texel t of a BC6H block (UF16, or SF16 when
// isSigned).  An invalid or reserved mode decodes to 0.
__device__ void bc6_texel(uint64_t lo, uint64_t hi, int t, bool isSigned, uint16_t out[3]) {
  out[0] = out[1] = out[2] = 0;
  const int modeVal = (lo & 2) == 0 ? (int)(lo & 3) : (int)(lo & 31);
  int mi = -1;
#pragma unroll
  for (int i = 0; i < 14; i++)
    if (kBc6Modes[i].mode == modeVal) mi = i;
  if (mi < 0) return;
  const Bc6Mode &M = kBc6Modes[mi];
  // Header: 12 endpoint values and the partition, accumulated by selects
  // so that every destination stays in a register.
  uint32_t v[13];
#pragma unroll
  for (int d = 0; d < 13; d++) v[d] = 0u;
  int pos = M.modeBits;
  for (int f = 0; f < M.nFields; f++) {
    const uint32_t code = M.field[f];
    const int dest = (int)(code >> 9), firstBit = (int)((code >> 5) & 15), n = (int)((code >> 1) & 15) + 1;
    const uint32_t chunk = bptc_bits(lo, hi, pos, n);
    pos += n;
    const uint32_t val = (code & 1u) ? (__builtin_bitreverse32(chunk) >> (32 - n)) << (firstBit - n + 1)
                                     : chunk << firstBit;
#pragma unroll
    for (int d = 0; d < 13; d++) v[d] |= dest == d ? val : 0u;
  }
  const int part = (int)v[12];
  const int ne = 2 * M.regions;
  int ep[4][3], eb[4][3];
#pragma unroll
  for (int e = 0; e < 4; e++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      ep[e][c] = (int)v[3 * e + c];
      eb[e][c] = (e == 0 || !M.delta) ? M.epBits : M.dBits[c];
    }
  // Sign extension (every endpoint of SF16, the deltas of UF16), deltas
  // applied modulo the endpoint width, then the unquantization.
#pragma unroll
  for (int e = 0; e < 4; e++)
#pragma unroll
    for (int c = 0; c < 3; c++)
      if (e < ne && (isSigned || (M.delta && e > 0))) ep[e][c] = bc6_sext16(ep[e][c], eb[e][c]);
  if (M.delta) {
#pragma unroll
    for (int e = 1; e < 4; e++)
#pragma unroll
      for (int c = 0; c < 3; c++)
        if (e < ne) {
          ep[e][c] = (ep[0][c] + ep[e][c]) & ((1 << M.epBits) - 1);
          eb[e][c] = M.epBits;
          if (isSigned) ep[e][c] = bc6_sext16(ep[e][c], M.epBits);
        }
  }
  const int s = M.regions == 2 ? (kBptcPart2[part] >> t) & 1 : 0;
  const int a1 = M.regions == 2 ? (kBptcAnchor[part] & 15) : 16;
  const int ib = M.regions == 1 ? 4 : 3;
  const int idxAt = 128 - (16 * ib - M.regions);
  const int below = (t > 0 ? 1 : 0) + (a1 < t ? 1 : 0);
  const int anchor = (t == 0 || t == a1) ? 1 : 0;
  const int idx = (int)bptc_bits(lo, hi, idxAt + ib * t - below, ib - anchor);
  const int w = kBptcWeights[ib - 2][idx];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const int sa = s ? ep[2][c] : ep[0][c], sz = s ? ep[3][c] : ep[1][c];
    const int ba = s ? eb[2][c] : eb[0][c], bz = s ? eb[3][c] : eb[1][c];
    int a = bc6_unquantize(sa, ba, isSigned), z = bc6_unquantize(sz, bz, isSigned);
    if (isSigned) {
      a = (int)(int16_t)(uint16_t)a;
      z = (int)(int16_t)(uint16_t)z;
    }
    uint32_t r = (uint32_t)((a * (64 - w) + z * w + 32) >> 6);
    if (isSigned) {
      // Magnitude scaled by 31/32 with the sign in bit 15; -0 becomes +0.
      r = (r & 0x80000000u) ? ((((~r + 1u) * 31u) >> 5) | 0x8000u) : ((r * 31u) >> 5);
      if (r == 0x8000u) r = 0u;
    } else {
      r = (r * 31u) >> 6;
    }
    out[c] = (uint16_t)r;
  }
}

// n = 6: four halfs per pixel (bpp >= 8); n = 7: B8G8R8A8 (bpp >= 4).
__global__ void decode_bptc_kernel(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, int w, int h,
                                   int pitch, int bpp, int n, int isSigned) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w || y >= h) return;
  const int bw = (w + 3) >> 2;
  const long long b = (long long)(y >> 2) * bw + (x >> 2);
  const uint64_t lo = reinterpret_cast<const uint64_t *>(src)[2 * b], hi = reinterpret_cast<const uint64_t *>(src)[2 * b + 1];
  const int t = (x & 3) + 4 * (y & 3);
  uint8_t *px = dst + (long long)y * pitch + (long long)x * bpp;
  if (n == 7) {
    *reinterpret_cast<uint32_t *>(px) = bc7_texel(lo, hi, t);
  } else {
    uint16_t rgb[3];
    bc6_texel(lo, hi, t, isSigned != 0, rgb);
    uint16_t *o = reinterpret_cast<uint16_t *>(px);
    o[0] = rgb[0];
    o[1] = rgb[1];
    o[2] = rgb[2];
    o[3] = 0x3C00;  // 1.0
  }
}

// ---------------------------------------------------------------------------
// ETC2 / EAC: 64-bit big-endian block words (byte 0 holds bits 63..56).
// ---------------------------------------------------------------------------
enum : int { ETC_R_SIGNED = 0, ETC_R_UNSIGNED, ETC_RG_SIGNED, ETC_RG_UNSIGNED, ETC_RGB, ETC_RGB_A1, ETC_RGBA };

__device__ __forceinline__ uint64_t etc_word(const uint8_t *p) {
  const uint2 v = *reinterpret_cast<const uint2 *>(p);
  return (uint64_t)__builtin_bswap32(v.x) << 32 | (uint64_t)__builtin_bswap32(v.y);
}
__device__ __forceinline__ int etc_bits(uint64_t w, int hi, int lo) {
  return (int)((w >> lo) & ((1ull << (hi - lo + 1)) - 1ull));
}
__device__ __forceinline__ int etc_sext3(int v) { return (v ^ 4) - 4; }
__device__ __forceinline__ int etc_clamp(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }
__device__ __forceinline__ int etc_ext4(int x) { return (x << 4) | x; }
__device__ __forceinline__ int etc_ext5(int x) { return (x << 3) | (x >> 2); }
__device__ __forceinline__ int etc_ext6(int x) { return (x << 2) | (x >> 4); }
__device__ __forceinline__ int etc_ext7(int x) { return (x << 1) | (x >> 6); }

// Intensity modifier tables (by codeword: the small and the large step),
// T / H mode distances and the EAC modifier tables.
__constant__ const uint8_t kEtcModifier[8][2] = {{2, 8}, {5, 17}, {9, 29}, {13, 42}, {18, 60}, {24, 80}, {33, 106}, {47, 183}};
__constant__ const uint8_t kEtcDistance[8] = {3, 6, 11, 16, 23, 32, 41, 64};
__constant__ const int8_t kEacModifier[16][8] = {
    {-3, -6, -9, -15, 2, 5, 8, 14}, {-3, -7, -10, -13, 2, 6, 9, 12}, {-2, -5, -8, -13, 1, 4, 7, 12},
    {-2, -4, -6, -13, 1, 3, 5, 12}, {-3, -6, -8, -12, 2, 5, 7, 11}, {-3, -7, -9, -11, 2, 6, 8, 10},
    {-4, -7, -8, -11, 3, 6, 7, 10}, {-3, -5, -8, -11, 2, 4, 7, 10}, {-2, -6, -8, -10, 1, 5, 7, 9},
    {-2, -5, -8, -10, 1, 4, 7, 9},  {-2, -4, -8, -10, 1, 3, 7, 9},  {-2, -5, -7, -10, 1, 4, 6, 9},
    {-3, -4, -7, -10, 2, 3, 6, 9},  {-1, -2, -3, -10, 0, 1, 2, 9},  {-4, -6, -8, -9, 3, 5, 7, 8},
    {-3, -5, -7, -9, 2, 4, 6, 8}};

// Pixel index (0..3) of pixel (x, y): MSB at bit 16 + 4x + y, LSB at 4x + y.
__device__ __forceinline__ int etc_pixel_index(uint64_t w, int x, int y) {
  const int k = 4 * x + y;
  return (int)(((w >> (16 + k)) & 1u) << 1 | ((w >> k) & 1u));
}

// Pixel (x, y) of an ETC2 RGB block: (r, g, b) and whether a
// punch-through block makes it transparent.
__device__ void etc2_rgb_texel(uint64_t w, bool punch, int x, int y, int &r, int &g, int &b, bool &transparent) {
  const int diff = (int)((w >> 33) & 1), flip = (int)((w >> 32) & 1);
  const bool nonOpaque = punch && !diff;
  const int idx = etc_pixel_index(w, x, y);
  transparent = false;
  const int half = flip ? (y >= 2) : (x >= 2);
  const int cw = half ? etc_bits(w, 36, 34) : etc_bits(w, 39, 37);
  // Individual and differential modes: base colour + modifier by index
  // (0: +small, 1: +large, 2: -small, 3: -large; small is 0 in a
  // non-opaque block).
  const int small = nonOpaque ? 0 : kEtcModifier[cw][0], large = kEtcModifier[cw][1];
  const int mod = idx == 0 ? small : idx == 1 ? large : idx == 2 ? -small : -large;
  if (!diff && !punch) {
    const int sh = half ? 0 : 4;
    r = etc_clamp(etc_ext4(etc_bits(w, 59 + sh, 56 + sh)) + mod);
    g = etc_clamp(etc_ext4(etc_bits(w, 51 + sh, 48 + sh)) + mod);
    b = etc_clamp(etc_ext4(etc_bits(w, 43 + sh, 40 + sh)) + mod);
    return;
  }
  const int R = etc_bits(w, 63, 59), G = etc_bits(w, 55, 51), B = etc_bits(w, 47, 43);
  const int r2 = R + etc_sext3(etc_bits(w, 58, 56)), g2 = G + etc_sext3(etc_bits(w, 50, 48));
  const int b2 = B + etc_sext3(etc_bits(w, 42, 40));
  int p0, p1, p2;  // the paint colour of this pixel's index
  if (r2 < 0 || r2 > 31) {
    // T mode: colour 1, colour 2 +d, colour 2, colour 2 -d.
    const int d = kEtcDistance[etc_bits(w, 35, 34) << 1 | etc_bits(w, 32, 32)];
    const int c1[3] = {etc_ext4(etc_bits(w, 60, 59) << 2 | etc_bits(w, 57, 56)), etc_ext4(etc_bits(w, 55, 52)),
                       etc_ext4(etc_bits(w, 51, 48))};
    const int c2[3] = {etc_ext4(etc_bits(w, 47, 44)), etc_ext4(etc_bits(w, 43, 40)), etc_ext4(etc_bits(w, 39, 36))};
    const int o = idx == 1 ? d : idx == 3 ? -d : 0;
    p0 = idx == 0 ? c1[0] : etc_clamp(c2[0] + o);
    p1 = idx == 0 ? c1[1] : etc_clamp(c2[1] + o);
    p2 = idx == 0 ? c1[2] : etc_clamp(c2[2] + o);
  } else if (g2 < 0 || g2 > 31) {
    // H mode: colour 1 +-d, colour 2 +-d; d's low bit orders the colours.
    const int c1[3] = {etc_ext4(etc_bits(w, 62, 59)), etc_ext4(etc_bits(w, 58, 56) << 1 | etc_bits(w, 52, 52)),
                       etc_ext4(etc_bits(w, 51, 51) << 3 | etc_bits(w, 49, 48) << 1 | etc_bits(w, 47, 47))};
    const int c2[3] = {etc_ext4(etc_bits(w, 46, 43)), etc_ext4(etc_bits(w, 42, 40) << 1 | etc_bits(w, 39, 39)),
                       etc_ext4(etc_bits(w, 38, 35))};
    const int v1 = c1[0] << 16 | c1[1] << 8 | c1[2], v2 = c2[0] << 16 | c2[1] << 8 | c2[2];
    const int d = kEtcDistance[etc_bits(w, 34, 34) << 2 | etc_bits(w, 32, 32) << 1 | (v1 >= v2 ? 1 : 0)];
    const int o = (idx & 1) ? -d : d;
    p0 = etc_clamp((idx < 2 ? c1[0] : c2[0]) + o);
    p1 = etc_clamp((idx < 2 ? c1[1] : c2[1]) + o);
    p2 = etc_clamp((idx < 2 ? c1[2] : c2[2]) + o);
  } else if (b2 < 0 || b2 > 31) {
    // Planar mode: origin, horizontal and vertical colours.
    const int O[3] = {etc_ext6(etc_bits(w, 62, 57)), etc_ext7(etc_bits(w, 56, 56) << 6 | etc_bits(w, 54, 49)),
                      etc_ext6(etc_bits(w, 48, 48) << 5 | etc_bits(w, 44, 43) << 3 | etc_bits(w, 41, 40) << 1 |
                               etc_bits(w, 39, 39))};
    const int Hc[3] = {etc_ext6(etc_bits(w, 38, 34) << 1 | etc_bits(w, 32, 32)), etc_ext7(etc_bits(w, 31, 25)),
                       etc_ext6(etc_bits(w, 24, 19))};
    const int V[3] = {etc_ext6(etc_bits(w, 18, 13)), etc_ext7(etc_bits(w, 12, 6)), etc_ext6(etc_bits(w, 5, 0))};
    r = etc_clamp(((x * (Hc[0] - O[0]) + y * (V[0] - O[0]) + 2) >> 2) + O[0]);
    g = etc_clamp(((x * (Hc[1] - O[1]) + y * (V[1] - O[1]) + 2) >> 2) + O[1]);
    b = etc_clamp(((x * (Hc[2] - O[2]) + y * (V[2] - O[2]) + 2) >> 2) + O[2]);
    return;
  } else {
    // Differential mode: a 5-bit base colour and a 3-bit signed delta.
    r = etc_clamp(etc_ext5(half ? r2 : R) + mod);
    g = etc_clamp(etc_ext5(half ? g2 : G) + mod);
    b = etc_clamp(etc_ext5(half ? b2 : B) + mod);
    transparent = nonOpaque && idx == 2;
    return;
  }
  r = p0;
  g = p1;
  b = p2;
  transparent = nonOpaque && idx == 2;
}

// One channel of pixel (x, y) of an ETC2 alpha or EAC block: base +
// modifier * multiplier (8-bit alpha), or base * 8 + 4 + modifier *
// multiplier * 8 (EAC, 11-bit; multiplier 0 adds the modifier alone).
__device__ __forceinline__ int etc_channel(uint64_t w, int x, int y, bool isSigned, bool eac) {
  const int base = isSigned ? (int)(int8_t)etc_bits(w, 63, 56) : etc_bits(w, 63, 56);
  const int mult = etc_bits(w, 55, 52), table = etc_bits(w, 51, 48);
  const int k = 4 * x + y;
  const int m = kEacModifier[table][etc_bits(w, 47 - 3 * k, 45 - 3 * k)];
  if (!eac) return base + m * mult;
  return mult == 0 ? base * 8 + 4 + m : base * 8 + 4 + m * mult * 8;
}

// EAC value clamped to [0, 2047] (signed: [-1023, 1023]), shifted into 16 bits.
__device__ __forceinline__ uint16_t eac16(int v, bool isSigned) {
  const int lo = isSigned ? -1023 : 0, hi = isSigned ? 1023 : 2047;
  v = v < lo ? lo : (v > hi ? hi : v);
  return (uint16_t)(int16_t)(v << 5);
}

// The reference's output layout: block (bx, by) starts at
// by * 4 * pitch + bx * 4 * bpp; inside it, pixel (x, y) of a colour type
// sits at y * pitch + 4 x, channel c of an EAC type at y * pitch +
// (x * channels + c) * 2.
__global__ void decode_etc_kernel(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, int w, int h, int pitch,
                                  int bpp, int type) {
  const int px = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y;
  if (px >= w || py >= h) return;
  const int x = px & 3, y = py & 3;
  const int bw = (w + 3) >> 2;
  const long long b = (long long)(py >> 2) * bw + (px >> 2);
  const bool wide = type == ETC_RG_SIGNED || type == ETC_RG_UNSIGNED || type == ETC_RGBA;
  const uint8_t *blk = src + b * (wide ? 16 : 8);
  uint8_t *out = dst + (long long)(py >> 2) * 4 * pitch + (long long)(px >> 2) * 4 * bpp + (long long)y * pitch;
  if (type <= ETC_RG_UNSIGNED) {
    const bool isSigned = type == ETC_R_SIGNED || type == ETC_RG_SIGNED;
    const int nch = type <= ETC_R_UNSIGNED ? 1 : 2;
    uint16_t *o = reinterpret_cast<uint16_t *>(out + (long long)x * nch * 2);
    o[0] = eac16(etc_channel(etc_word(blk), x, y, isSigned, true), isSigned);
    if (nch == 2) o[1] = eac16(etc_channel(etc_word(blk + 8), x, y, isSigned, true), isSigned);
    return;
  }
  int a = 255;
  if (type == ETC_RGBA) {
    a = etc_clamp(etc_channel(etc_word(blk), x, y, false, false));
    blk += 8;
  }
  int r, g, bl;
  bool transparent;
  etc2_rgb_texel(etc_word(blk), type == ETC_RGB_A1, x, y, r, g, bl, transparent);
  *reinterpret_cast<uint32_t *>(out + 4ll * x) =
      transparent ? 0u : ((uint32_t)bl | (uint32_t)g << 8 | (uint32_t)r << 16 | (uint32_t)a << 24);
}

}  // namespace swr
