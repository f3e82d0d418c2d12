/*
 * swr_oracle_etc.c -- CPU oracle for ETC2 / EAC block decoding.
 *
 * TEST INFRASTRUCTURE ONLY, like swr_oracle.c (linked into the same
 * libswr_oracle.so): the checker the GPU decoder is compared against.
 *
 * Restates ETC_Decoder::Decode (src/Device/ETC_Decoder.cpp:670-734), which
 * vk::Image::decodeETC2 calls per mip level (src/Vulkan/VkImage.cpp), from
 * the ETC2 / EAC block formats (OpenGL ES 3.0, Khronos Data Format
 * Specification "ETC2 compressed texture image formats").  A block is read
 * as one 64-bit big-endian word w (byte 0 holds bits 63..56), the
 * specification's own bit numbering; the reference reads the same bits
 * through little-endian bitfield structs.  Where the reference departs from
 * the specification this file follows the reference:
 *   - EAC signed values add the +4 rounding term like unsigned ones, and a
 *     base codeword of -128 is used as is (ETC_Decoder.cpp:617-622);
 *   - EAC results are clamped to [0, 2047] (signed: [-1023, 1023]) and
 *     shifted left by 5 into 16 bits (clampEAC, :29-34).
 * Output layouts (ETC_Decoder.cpp:87-140, 146-176): RGB, RGB with
 * punch-through alpha and RGBA write B8G8R8A8 pixels; R11 writes one 16-bit
 * value, RG11 two (R then G) per pixel.  Pixels outside the image are not
 * written.
 */
#include <stdint.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

enum { ETC_R_SIGNED, ETC_R_UNSIGNED, ETC_RG_SIGNED, ETC_RG_UNSIGNED, ETC_RGB, ETC_RGB_A1, ETC_RGBA };

static uint64_t be64(const uint8_t *p) {
  uint64_t w = 0;
  for (int i = 0; i < 8; i++) w = (w << 8) | p[i];
  return w;
}
static int bits(uint64_t w, int hi, int lo) { return (int)((w >> lo) & ((1ull << (hi - lo + 1)) - 1)); }
static int sext(int v, int n) { return (v ^ (1 << (n - 1))) - (1 << (n - 1)); }
static int clamp255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }
static int ext4(int x) { return (x << 4) | x; }
static int ext5(int x) { return (x << 3) | (x >> 2); }
static int ext6(int x) { return (x << 2) | (x >> 4); }
static int ext7(int x) { return (x << 1) | (x >> 6); }

/* 2-bit pixel index of pixel (x, y): MSB at bit 16 + (4x + y), LSB at 4x + y. */
static int pixel_index(uint64_t w, int x, int y) {
  const int k = 4 * x + y;
  return (int)(((w >> (16 + k)) & 1) << 1 | ((w >> k) & 1));
}

/* Intensity modifiers per table codeword, ordered by pixel index
 * (ETC2 "intensity modifier" table with the index mapping applied). */
static const int kModifier[8][2] = {{2, 8}, {5, 17}, {9, 29}, {13, 42}, {18, 60}, {24, 80}, {33, 106}, {47, 183}};
static int modifier(int cw, int idx, int nonOpaque) {
  const int small = nonOpaque ? 0 : kModifier[cw][0], large = kModifier[cw][1];
  switch (idx) {
    case 0: return small;
    case 1: return large;
    case 2: return -small;
    default: return -large;
  }
}
static const int kDistance[8] = {3, 6, 11, 16, 23, 32, 41, 64};

/* RGB(A1) block -> rgb[16][3], transparent[16] (pixel i = x + 4 y). */
static void etc2_rgb(uint64_t w, int punch, int rgb[16][3], int transparent[16]) {
  const int diff = (int)((w >> 33) & 1), flip = (int)((w >> 32) & 1);
  const int nonOpaque = punch && !diff;
  memset(transparent, 0, 16 * sizeof(int));
  if (!diff && !punch) {
    /* individual mode: two 4-bit base colours */
    const int c[2][3] = {{ext4(bits(w, 63, 60)), ext4(bits(w, 55, 52)), ext4(bits(w, 47, 44))},
                         {ext4(bits(w, 59, 56)), ext4(bits(w, 51, 48)), ext4(bits(w, 43, 40))}};
    const int cw[2] = {bits(w, 39, 37), bits(w, 36, 34)};
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        const int s = flip ? (y >= 2) : (x >= 2);
        const int m = modifier(cw[s], pixel_index(w, x, y), 0);
        for (int k = 0; k < 3; k++) rgb[x + 4 * y][k] = clamp255(c[s][k] + m);
      }
    return;
  }
  const int R = bits(w, 63, 59), G = bits(w, 55, 51), B = bits(w, 47, 43);
  const int r2 = R + sext(bits(w, 58, 56), 3), g2 = G + sext(bits(w, 50, 48), 3), b2 = B + sext(bits(w, 42, 40), 3);
  int paint[4][3];
  int paintMode = 1;
  if (r2 < 0 || r2 > 31) {
    /* T mode */
    const int c1[3] = {ext4(bits(w, 60, 59) << 2 | bits(w, 57, 56)), ext4(bits(w, 55, 52)), ext4(bits(w, 51, 48))};
    const int c2[3] = {ext4(bits(w, 47, 44)), ext4(bits(w, 43, 40)), ext4(bits(w, 39, 36))};
    const int d = kDistance[bits(w, 35, 34) << 1 | bits(w, 32, 32)];
    for (int k = 0; k < 3; k++) {
      paint[0][k] = c1[k];
      paint[1][k] = clamp255(c2[k] + d);
      paint[2][k] = c2[k];
      paint[3][k] = clamp255(c2[k] - d);
    }
  } else if (g2 < 0 || g2 > 31) {
    /* H mode */
    const int c1[3] = {ext4(bits(w, 62, 59)), ext4(bits(w, 58, 56) << 1 | bits(w, 52, 52)),
                       ext4(bits(w, 51, 51) << 3 | bits(w, 49, 48) << 1 | bits(w, 47, 47))};
    const int c2[3] = {ext4(bits(w, 46, 43)), ext4(bits(w, 42, 40) << 1 | bits(w, 39, 39)), ext4(bits(w, 38, 35))};
    const int v1 = c1[0] << 16 | c1[1] << 8 | c1[2], v2 = c2[0] << 16 | c2[1] << 8 | c2[2];
    const int d = kDistance[bits(w, 34, 34) << 2 | bits(w, 32, 32) << 1 | (v1 >= v2 ? 1 : 0)];
    for (int k = 0; k < 3; k++) {
      paint[0][k] = clamp255(c1[k] + d);
      paint[1][k] = clamp255(c1[k] - d);
      paint[2][k] = clamp255(c2[k] + d);
      paint[3][k] = clamp255(c2[k] - d);
    }
  } else if (b2 < 0 || b2 > 31) {
    /* planar mode: origin, horizontal and vertical colours */
    const int O[3] = {ext6(bits(w, 62, 57)), ext7(bits(w, 56, 56) << 6 | bits(w, 54, 49)),
                      ext6(bits(w, 48, 48) << 5 | bits(w, 44, 43) << 3 | bits(w, 41, 40) << 1 | bits(w, 39, 39))};
    const int Hc[3] = {ext6(bits(w, 38, 34) << 1 | bits(w, 32, 32)), ext7(bits(w, 31, 25)), ext6(bits(w, 24, 19))};
    const int V[3] = {ext6(bits(w, 18, 13)), ext7(bits(w, 12, 6)), ext6(bits(w, 5, 0))};
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++)
        for (int k = 0; k < 3; k++)
          rgb[x + 4 * y][k] = clamp255(((x * (Hc[k] - O[k]) + y * (V[k] - O[k]) + 2) >> 2) + O[k]);
    return;
  } else {
    /* differential mode: a 5-bit base colour and a 3-bit signed delta */
    paintMode = 0;
    const int c[2][3] = {{ext5(R), ext5(G), ext5(B)}, {ext5(r2), ext5(g2), ext5(b2)}};
    const int cw[2] = {bits(w, 39, 37), bits(w, 36, 34)};
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        const int s = flip ? (y >= 2) : (x >= 2);
        const int idx = pixel_index(w, x, y);
        const int m = modifier(cw[s], idx, nonOpaque);
        for (int k = 0; k < 3; k++) rgb[x + 4 * y][k] = clamp255(c[s][k] + m);
        transparent[x + 4 * y] = nonOpaque && idx == 2;
      }
  }
  if (paintMode)
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        const int idx = pixel_index(w, x, y);
        for (int k = 0; k < 3; k++) rgb[x + 4 * y][k] = paint[idx][k];
        transparent[x + 4 * y] = nonOpaque && idx == 2;
      }
}

/* EAC / ETC2-alpha modifier table (16 tables x 8 indices). */
static const int kEacTable[16][8] = {
    {-3, -6, -9, -15, 2, 5, 8, 14}, {-3, -7, -10, -13, 2, 6, 9, 12}, {-2, -5, -8, -13, 1, 4, 7, 12},
    {-2, -4, -6, -13, 1, 3, 5, 12}, {-3, -6, -8, -12, 2, 5, 7, 11}, {-3, -7, -9, -11, 2, 6, 8, 10},
    {-4, -7, -8, -11, 3, 6, 7, 10}, {-3, -5, -8, -11, 2, 4, 7, 10}, {-2, -6, -8, -10, 1, 5, 7, 9},
    {-2, -5, -8, -10, 1, 4, 7, 9},  {-2, -4, -8, -10, 1, 3, 7, 9},  {-2, -5, -7, -10, 1, 4, 6, 9},
    {-3, -4, -7, -10, 2, 3, 6, 9},  {-1, -2, -3, -10, 0, 1, 2, 9},  {-4, -6, -8, -9, 3, 5, 7, 8},
    {-3, -5, -7, -9, 2, 4, 6, 8}};

/* One channel of pixel (x, y): base + modifier * multiplier (ETC2 alpha,
 * 8-bit), or the EAC 11-bit form base * 8 + 4 + modifier * multiplier * 8
 * (multiplier 0: the modifier alone). */
static int single_channel(uint64_t w, int x, int y, int isSigned, int eac) {
  const int base = isSigned ? (int)(int8_t)bits(w, 63, 56) : bits(w, 63, 56);
  const int mult = bits(w, 55, 52), table = bits(w, 51, 48);
  const int k = 4 * x + y;
  const int m = kEacTable[table][bits(w, 47 - 3 * k, 45 - 3 * k)];
  if (!eac) return base + m * mult;
  return mult == 0 ? base * 8 + 4 + m : base * 8 + 4 + m * mult * 8;
}

static int16_t eac16(int v, int isSigned) {
  const int lo = isSigned ? -1023 : 0, hi = isSigned ? 1023 : 2047;
  v = v < lo ? lo : (v > hi ? hi : v);
  return (int16_t)(v << 5);
}

EXPORT int swo_decode_etc(const uint8_t *src, uint8_t *dst, int w, int h, int dstPitch, int dstBpp, int type) {
  if (type < ETC_R_SIGNED || type > ETC_RGBA) return 0;
  const int blockBytes = (type == ETC_RG_SIGNED || type == ETC_RG_UNSIGNED || type == ETC_RGBA) ? 16 : 8;
  const int bw = (w + 3) / 4;
  for (int by = 0; by < (h + 3) / 4; by++)
    for (int bx = 0; bx < bw; bx++) {
      const uint8_t *blk = src + (size_t)(by * bw + bx) * blockBytes;
      uint8_t *out = dst + (size_t)by * 4 * dstPitch + (size_t)bx * 4 * dstBpp;
      if (type <= ETC_RG_UNSIGNED) {
        const int nch = type <= ETC_R_UNSIGNED ? 1 : 2;
        const int isSigned = type == ETC_R_SIGNED || type == ETC_RG_SIGNED;
        for (int y = 0; y < 4 && 4 * by + y < h; y++)
          for (int x = 0; x < 4 && 4 * bx + x < w; x++)
            for (int c = 0; c < nch; c++) {
              const int16_t v = eac16(single_channel(be64(blk + 8 * c), x, y, isSigned, 1), isSigned);
              memcpy(out + (size_t)y * dstPitch + (size_t)(x * nch + c) * 2, &v, 2);
            }
        continue;
      }
      int alpha[16];
      for (int i = 0; i < 16; i++) alpha[i] = 255;
      const uint8_t *color = blk;
      if (type == ETC_RGBA) {
        const uint64_t a = be64(blk);
        for (int y = 0; y < 4; y++)
          for (int x = 0; x < 4; x++) alpha[x + 4 * y] = clamp255(single_channel(a, x, y, 0, 0));
        color = blk + 8;
      }
      int rgb[16][3], transparent[16];
      etc2_rgb(be64(color), type == ETC_RGB_A1, rgb, transparent);
      for (int y = 0; y < 4 && 4 * by + y < h; y++)
        for (int x = 0; x < 4 && 4 * bx + x < w; x++) {
          uint8_t *p = out + (size_t)y * dstPitch + (size_t)x * 4;
          const int i = x + 4 * y;
          if (transparent[i]) {
            p[0] = p[1] = p[2] = p[3] = 0;
          } else {
            p[0] = (uint8_t)rgb[i][2];
            p[1] = (uint8_t)rgb[i][1];
            p[2] = (uint8_t)rgb[i][0];
            p[3] = (uint8_t)alpha[i];
          }
        }
    }
  return 1;
}
