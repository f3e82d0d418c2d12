/*
 * swr_oracle_bptc.c -- CPU oracle for BC6H and BC7 block decoding.
 *
 * TEST INFRASTRUCTURE ONLY, like swr_oracle.c (linked into the same
 * libswr_oracle.so): the checker the GPU decoder is compared against.
 *
 * Restates BC_Decoder::Decode for n = 6 and 7 (src/Device/BC_Decoder.cpp:
 * 224-1090 BC6H, 1092-1658 BC7, dispatch 1746-1771) from the BPTC block
 * formats (ARB_texture_compression_bptc; the BC6H and BC7 format pages of
 * the Direct3D 11 documentation).  A block is a 128-bit little-endian bit
 * stream read LSB first.  Where the reference departs from the format
 * definition this file follows the reference:
 *   - an invalid BC7 mode (first byte 0) decodes to (0, 0, 0, 0), an
 *     invalid or reserved BC6H mode to (0, 0, 0, 1.0);
 *   - BC6H endpoints keep their width when already 15 (unsigned) or 16
 *     (signed) bits or zero, before interpolation (:486-540);
 *   - the BC6H output's alpha is 1.0 (0x3C00) and a negative zero is
 *     written as +0 (:632-646).
 * Output: BC7 writes B8G8R8A8 pixels; BC6H writes four 16-bit half floats
 * (R, G, B, A = 1.0) per pixel.  Pixels outside the image are not written.
 *
 * The partition and anchor tables are the formats' own constants (the
 * same values as the reference's PartitionTable2/3 and AnchorTable2/3a/3b),
 * held here packed: one bit (2 subsets) or two bits (3 subsets) per texel,
 * and the three anchor texels of a partition in one 12-bit word.
 */
#include <stdint.h>
#include <string.h>

/* 2-subset partitions: bit t is the subset of texel t (t = x + 4 y). */
static const uint16_t kPart2[64] = {
    0xcccc, 0x8888, 0xeeee, 0xecc8, 0xc880, 0xfeec, 0xfec8, 0xec80,
    0xc800, 0xffec, 0xfe80, 0xe800, 0xffe8, 0xff00, 0xfff0, 0xf000,
    0xf710, 0x008e, 0x7100, 0x08ce, 0x008c, 0x7310, 0x3100, 0x8cce,
    0x088c, 0x3110, 0x6666, 0x366c, 0x17e8, 0x0ff0, 0x718e, 0x399c,
    0xaaaa, 0xf0f0, 0x5a5a, 0x33cc, 0x3c3c, 0x55aa, 0x9696, 0xa55a,
    0x73ce, 0x13c8, 0x324c, 0x3bdc, 0x6996, 0xc33c, 0x9966, 0x0660,
    0x0272, 0x04e4, 0x4e40, 0x2720, 0xc936, 0x936c, 0x39c6, 0x639c,
    0x9336, 0x9cc6, 0x817e, 0xe718, 0xccf0, 0x0fcc, 0x7744, 0xee22,
};
/* 3-subset partitions: bits 2t+1..2t are the subset of texel t. */
static const uint32_t kPart3[64] = {
    0xaa685050, 0x6a5a5040, 0x5a5a4200, 0x5450a0a8, 0xa5a50000, 0xa0a05050, 0x5555a0a0, 0x5a5a5050,
    0xaa550000, 0xaa555500, 0xaaaa5500, 0x90909090, 0x94949494, 0xa4a4a4a4, 0xa9a59450, 0x2a0a4250,
    0xa5945040, 0x0a425054, 0xa5a5a500, 0x55a0a0a0, 0xa8a85454, 0x6a6a4040, 0xa4a45000, 0x1a1a0500,
    0x0050a4a4, 0xaaa59090, 0x14696914, 0x69691400, 0xa08585a0, 0xaa821414, 0x50a4a450, 0x6a5a0200,
    0xa9a58000, 0x5090a0a8, 0xa8a09050, 0x24242424, 0x00aa5500, 0x24924924, 0x24499224, 0x50a50a50,
    0x500aa550, 0xaaaa4444, 0x66660000, 0xa5a0a5a0, 0x50a050a0, 0x69286928, 0x44aaaa44, 0x66666600,
    0xaa444444, 0x54a854a8, 0x95809580, 0x96969600, 0xa85454a8, 0x80959580, 0xaa141414, 0x96960000,
    0xaaaa1414, 0xa05050a0, 0xa0a5a5a0, 0x96000000, 0x40804080, 0xa9a8a9a8, 0xaaaaaa44, 0x2a4a5254,
};
/* Anchor texels: bits 3..0 subset 1 of 2, 7..4 subset 1 of 3, 11..8 subset 2 of 3. */
static const uint16_t kAnchor[64] = {
    0xf3f, 0x83f, 0x8ff, 0x3ff, 0xf8f, 0xf3f, 0x3ff, 0x8ff,
    0xf8f, 0xf8f, 0xf6f, 0xf6f, 0xf6f, 0xf5f, 0xf3f, 0x83f,
    0xf3f, 0x832, 0xf88, 0x3f2, 0xf32, 0x838, 0xf68, 0x8af,
    0x352, 0xf88, 0x682, 0xa62, 0xf88, 0xf58, 0xaf2, 0x8f2,
    0xf8f, 0x3ff, 0xf36, 0xa58, 0xa62, 0x8a8, 0x98f, 0xaff,
    0x6f2, 0xf38, 0x8f2, 0xf52, 0x3f2, 0x6ff, 0x6ff, 0x8f6,
    0xf36, 0x3f2, 0xf56, 0xf58, 0xf5f, 0xf8f, 0xf52, 0xfa2,
    0xf5f, 0xfaf, 0xf8f, 0xfdf, 0x3ff, 0xfc2, 0xf32, 0x83f,
};

static const int kW2[4] = {0, 21, 43, 64};
static const int kW3[8] = {0, 9, 18, 27, 37, 46, 55, 64};
static const int kW4[16] = {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64};

static int subset_of(int ns, int part, int t) {
  if (ns == 2) return (kPart2[part] >> t) & 1;
  if (ns == 3) return (int)((kPart3[part] >> (2 * t)) & 3);
  return 0;
}
static int anchor_of(int ns, int part, int s) {
  if (s == 0) return 0;
  if (ns == 2) return kAnchor[part] & 15;
  return s == 1 ? (kAnchor[part] >> 4) & 15 : (kAnchor[part] >> 8) & 15;
}

/* Bit reader over the 128-bit block. */
typedef struct {
  uint64_t lo, hi;
} Bits128;
static Bits128 load128(const uint8_t *p) {
  Bits128 b;
  memcpy(&b.lo, p, 8);
  memcpy(&b.hi, p + 8, 8);
  return b;
}
static uint32_t get(Bits128 b, int off, int n) {
  if (n == 0) return 0;
  uint64_t v;
  if (off >= 64) v = b.hi >> (off - 64);
  else if (off + n <= 64) v = b.lo >> off;
  else v = (b.lo >> off) | (b.hi << (64 - off));
  return (uint32_t)(v & ((1ull << n) - 1));
}

/* ---------------------------------------------------------------- BC7 */
typedef struct {
  int ns, pb, rb, isb, cb, ab, epb, spb, ib, ib2;
} Bc7Mode;
static const Bc7Mode kBc7[8] = {
    {3, 4, 0, 0, 4, 0, 1, 0, 3, 0}, {2, 6, 0, 0, 6, 0, 0, 1, 3, 0}, {3, 6, 0, 0, 5, 0, 0, 0, 2, 0},
    {2, 6, 0, 0, 7, 0, 1, 0, 2, 0}, {1, 0, 2, 1, 5, 6, 0, 0, 2, 3}, {1, 0, 2, 0, 7, 8, 0, 0, 2, 2},
    {1, 0, 0, 0, 7, 7, 1, 0, 4, 0}, {2, 6, 0, 0, 5, 5, 1, 0, 2, 0},
};

static int bc7_interp(int e0, int e1, int idx, int bits) {
  const int w = bits == 2 ? kW2[idx] : bits == 3 ? kW3[idx] : kW4[idx];
  return ((64 - w) * e0 + w * e1 + 32) >> 6;
}

/* Texel t of a BC7 block as B, G, R, A bytes. */
static void bc7_texel(Bits128 b, int t, uint8_t out[4]) {
  int m = 0;
  while (m < 8 && !((b.lo >> m) & 1)) m++;
  if (m == 8) {
    out[0] = out[1] = out[2] = out[3] = 0;
    return;
  }
  const Bc7Mode M = kBc7[m];
  int pos = m + 1;
  const int part = (int)get(b, pos, M.pb);
  pos += M.pb;
  const int rot = (int)get(b, pos, M.rb);
  pos += M.rb;
  const int sel = (int)get(b, pos, M.isb);
  pos += M.isb;
  const int ne = 2 * M.ns;
  int ep[6][4];  /* endpoint e, channel R, G, B, A */
  for (int c = 0; c < 3; c++)
    for (int e = 0; e < ne; e++, pos += M.cb) ep[e][c] = (int)get(b, pos, M.cb);
  for (int e = 0; e < ne; e++) {
    ep[e][3] = M.ab ? (int)get(b, pos, M.ab) : 255;
    pos += M.ab;
  }
  int pbit[6] = {0, 0, 0, 0, 0, 0};
  if (M.epb)
    for (int e = 0; e < ne; e++) pbit[e] = (int)get(b, pos++, 1);
  if (M.spb) {
    const int p0 = (int)get(b, pos, 1), p1 = (int)get(b, pos + 1, 1);
    pos += 2;
    pbit[0] = pbit[1] = p0;
    pbit[2] = pbit[3] = p1;
  }
  const int pbits = M.epb || M.spb;
  const int cbits = M.cb + pbits, abits = M.ab + M.epb;
  for (int e = 0; e < ne; e++) {
    for (int c = 0; c < 3; c++) {
      int v = pbits ? (ep[e][c] << 1 | pbit[e]) : ep[e][c];
      v = (v << (8 - cbits)) & 0xFF;
      ep[e][c] = v | (v >> cbits);
    }
    if (M.ab) {
      int v = M.epb ? (ep[e][3] << 1 | pbit[e]) : ep[e][3];
      v = (v << (8 - abits)) & 0xFF;
      ep[e][3] = v | (v >> abits);
    }
  }
  /* Index stream: texels in order, anchors one bit shorter; the secondary
   * stream (modes 4, 5) follows the primary one. */
  const int primStart = pos;
  int primLen = 0;
  for (int k = 0; k < 16; k++) primLen += M.ib - (k == anchor_of(M.ns, part, subset_of(M.ns, part, k)));
  const int secStart = primStart + primLen;
  int pOff = 0, sOff = 0, pIdx = 0, sIdx = 0;
  for (int k = 0; k <= t; k++) {
    const int anchor = k == anchor_of(M.ns, part, subset_of(M.ns, part, k));
    const int pn = M.ib - anchor, sn = M.ib2 ? M.ib2 - anchor : 0;
    if (k == t) {
      pIdx = (int)get(b, primStart + pOff, pn);
      sIdx = (int)get(b, secStart + sOff, sn);
    }
    pOff += pn;
    sOff += sn;
  }
  const int s = subset_of(M.ns, part, t);
  const int *e0 = ep[2 * s], *e1 = ep[2 * s + 1];
  /* Colour takes the secondary index when the selection bit is set, alpha
   * the secondary index when there is one and the bit is clear. */
  const int cIdx = (M.ib2 && sel) ? sIdx : pIdx, cBits = (M.ib2 && sel) ? M.ib2 : M.ib;
  const int aIdx = (M.ib2 && !sel) ? sIdx : pIdx, aBits = (M.ib2 && !sel) ? M.ib2 : M.ib;
  int rgba[4];
  for (int c = 0; c < 3; c++) rgba[c] = bc7_interp(e0[c], e1[c], cIdx, cBits);
  rgba[3] = bc7_interp(e0[3], e1[3], aIdx, aBits);
  if (rot) {
    const int tmp = rgba[3];
    rgba[3] = rgba[rot - 1];
    rgba[rot - 1] = tmp;
  }
  out[0] = (uint8_t)rgba[2];
  out[1] = (uint8_t)rgba[1];
  out[2] = (uint8_t)rgba[0];
  out[3] = (uint8_t)rgba[3];
}

/* ---------------------------------------------------------------- BC6H */
/* Header layouts of the 14 modes in the notation of the format tables:
 * fields in stream order, "r0[9:0]" = bits 9..0 of endpoint 0's red
 * (lowest first in the stream), "r0[10:15]" = bits 10..15 with bit 15
 * first, "d[4:0]" the partition index.  The mode bits come first. */
typedef struct {
  int mode, modeBits, regions, epBits, delta, dBits[3];
  const char *layout;
} Bc6Mode;
static const Bc6Mode kBc6[14] = {
    {0, 2, 2, 10, 1, {5, 5, 5},
     "g2[4] b2[4] b3[4] r0[9:0] g0[9:0] b0[9:0] r1[4:0] g3[4] g2[3:0] g1[4:0] b3[0] g3[3:0] b1[4:0] b3[1] "
     "b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3] d[4:0]"},
    {1, 2, 2, 7, 1, {6, 6, 6},
     "g2[5] g3[5:4] r0[6:0] b3[1:0] b2[4] g0[6:0] b2[5] b3[2] g2[4] b0[6:0] b3[3] b3[5] b3[4] r1[5:0] g2[3:0] "
     "g1[5:0] g3[3:0] b1[5:0] b2[3:0] r2[5:0] r3[5:0] d[4:0]"},
    {2, 5, 2, 11, 1, {5, 4, 4},
     "r0[9:0] g0[9:0] b0[9:0] r1[4:0] r0[10] g2[3:0] g1[3:0] g0[10] b3[0] g3[3:0] b1[3:0] b0[10] b3[1] b2[3:0] "
     "r2[4:0] b3[2] r3[4:0] b3[3] d[4:0]"},
    {6, 5, 2, 11, 1, {4, 5, 4},
     "r0[9:0] g0[9:0] b0[9:0] r1[3:0] r0[10] g3[4] g2[3:0] g1[4:0] g0[10] g3[3:0] b1[3:0] b0[10] b3[1] b2[3:0] "
     "r2[3:0] b3[0] b3[2] r3[3:0] g2[4] b3[3] d[4:0]"},
    {10, 5, 2, 11, 1, {4, 4, 5},
     "r0[9:0] g0[9:0] b0[9:0] r1[3:0] r0[10] b2[4] g2[3:0] g1[3:0] g0[10] b3[0] g3[3:0] b1[4:0] b0[10] b2[3:0] "
     "r2[3:0] b3[1] b3[2] r3[3:0] b3[4] b3[3] d[4:0]"},
    {14, 5, 2, 9, 1, {5, 5, 5},
     "r0[8:0] b2[4] g0[8:0] g2[4] b0[8:0] b3[4] r1[4:0] g3[4] g2[3:0] g1[4:0] b3[0] g3[3:0] b1[4:0] b3[1] "
     "b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3] d[4:0]"},
    {18, 5, 2, 8, 1, {6, 5, 5},
     "r0[7:0] g3[4] b2[4] g0[7:0] b3[2] g2[4] b0[7:0] b3[3] b3[4] r1[5:0] g2[3:0] g1[4:0] b3[0] g3[3:0] b1[4:0] "
     "b3[1] b2[3:0] r2[5:0] r3[5:0] d[4:0]"},
    {22, 5, 2, 8, 1, {5, 6, 5},
     "r0[7:0] b3[0] b2[4] g0[7:0] g2[5] g2[4] b0[7:0] g3[5] b3[4] r1[4:0] g3[4] g2[3:0] g1[5:0] g3[3:0] "
     "b1[4:0] b3[1] b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3] d[4:0]"},
    {26, 5, 2, 8, 1, {5, 5, 6},
     "r0[7:0] b3[1] b2[4] g0[7:0] b2[5] g2[4] b0[7:0] b3[5] b3[4] r1[4:0] g3[4] g2[3:0] g1[4:0] b3[0] g3[3:0] "
     "b1[5:0] b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3] d[4:0]"},
    {30, 5, 2, 6, 0, {6, 6, 6},
     "r0[5:0] g3[4] b3[0] b3[1] b2[4] g0[5:0] g2[5] b2[5] b3[2] g2[4] b0[5:0] g3[5] b3[3] b3[5] b3[4] r1[5:0] "
     "g2[3:0] g1[5:0] g3[3:0] b1[5:0] b2[3:0] r2[5:0] r3[5:0] d[4:0]"},
    {3, 5, 1, 10, 0, {10, 10, 10}, "r0[9:0] g0[9:0] b0[9:0] r1[9:0] g1[9:0] b1[9:0]"},
    {7, 5, 1, 11, 1, {9, 9, 9}, "r0[9:0] g0[9:0] b0[9:0] r1[8:0] r0[10] g1[8:0] g0[10] b1[8:0] b0[10]"},
    {11, 5, 1, 12, 1, {8, 8, 8}, "r0[9:0] g0[9:0] b0[9:0] r1[7:0] r0[10:11] g1[7:0] g0[10:11] b1[7:0] b0[10:11]"},
    {15, 5, 1, 16, 1, {4, 4, 4}, "r0[9:0] g0[9:0] b0[9:0] r1[3:0] r0[10:15] g1[3:0] g0[10:15] b1[3:0] b0[10:15]"},
};

static int parse_int(const char **s) {
  int v = 0;
  while (**s >= '0' && **s <= '9') v = 10 * v + (*(*s)++ - '0');
  return v;
}

/* Reads the header of `M` from the stream at bit `pos` into ep[4][3] and
 * the partition; returns the first index bit. */
static int bc6_header(const Bc6Mode *M, Bits128 b, int ep[4][3], int *part) {
  int pos = M->modeBits;
  memset(ep, 0, 4 * 3 * sizeof(int));
  *part = 0;
  const char *s = M->layout;
  while (*s) {
    while (*s == ' ') s++;
    if (!*s) break;
    const char kind = *s++;
    int e = -1;
    if (kind != 'd') e = *s++ - '0';
    s++; /* '[' */
    const int a = parse_int(&s);
    int z = a;
    if (*s == ':') {
      s++;
      z = parse_int(&s);
    }
    s++; /* ']' */
    const int n = (a > z ? a - z : z - a) + 1;
    for (int i = 0; i < n; i++) {
      /* stream bit i goes to bit z + i ([a:z], a >= z) or a - i... : the
       * lowest-numbered bit first unless the range is written ascending */
      const int bit = a >= z ? z + i : z - i;
      const int v = (int)get(b, pos++, 1);
      if (kind == 'd') *part |= v << bit;
      else ep[e][kind == 'r' ? 0 : kind == 'g' ? 1 : 2] |= v << bit;
    }
  }
  return pos;
}

static int sign_extend(int v, int bits) {
  const int m = 1 << (bits - 1);
  return (int16_t)(uint16_t)((v ^ m) - m);
}

static int bc6_unquantize(int v, int bits, int isSigned) {
  if (!isSigned) {
    if (bits >= 15 || v == 0) return v;
    if (v == (1 << bits) - 1) return 0xFFFF;
    return (int)((((uint32_t)v << 16) + 0x8000u) >> bits);
  }
  if (bits >= 16 || v == 0) return v;
  int16_t s = (int16_t)(uint16_t)v;
  const int neg = s < 0;
  int mag = neg ? -(int)s : (int)s;
  int r = mag >= (1 << (bits - 1)) - 1 ? 0x7FFF : ((mag << 15) + 0x4000) >> (bits - 1);
  return (uint16_t)(neg ? -r : r);
}

/* Texel t of a BC6H block as four half floats. */
static void bc6_texel(Bits128 b, int t, int isSigned, uint16_t out[4]) {
  out[3] = 0x3C00;
  int modeVal = (b.lo & 2) == 0 ? (int)(b.lo & 3) : (int)(b.lo & 31);
  const Bc6Mode *M = NULL;
  for (int i = 0; i < 14; i++)
    if (kBc6[i].mode == modeVal) M = &kBc6[i];
  if (!M) {
    out[0] = out[1] = out[2] = 0;
    return;
  }
  int ep[4][3], part;
  int pos = bc6_header(M, b, ep, &part);
  const int ne = 2 * M->regions;
  int bits[4][3];
  for (int e = 0; e < 4; e++)
    for (int c = 0; c < 3; c++) bits[e][c] = (e == 0 || !M->delta) ? M->epBits : M->dBits[c];
  /* sign extension: every endpoint of a signed format, the deltas of an
   * unsigned one */
  for (int e = 0; e < ne; e++)
    for (int c = 0; c < 3; c++)
      if (isSigned || (M->delta && e > 0)) ep[e][c] = (uint16_t)sign_extend(ep[e][c], bits[e][c]);
  if (M->delta)
    for (int e = 1; e < ne; e++)
      for (int c = 0; c < 3; c++) {
        ep[e][c] = (ep[0][c] + ep[e][c]) & ((1 << M->epBits) - 1);
        bits[e][c] = M->epBits;
        if (isSigned) ep[e][c] = (uint16_t)sign_extend(ep[e][c], M->epBits);
      }
  for (int e = 0; e < ne; e++)
    for (int c = 0; c < 3; c++) ep[e][c] = bc6_unquantize(ep[e][c], bits[e][c], isSigned);
  /* indices: 4 bits (one region) or 3 bits, anchors one bit shorter */
  const int ib = M->regions == 1 ? 4 : 3;
  int idx = 0;
  for (int k = 0; k <= t; k++) {
    const int anchor = k == 0 || (M->regions == 2 && k == (kAnchor[part] & 15));
    const int n = ib - anchor;
    if (k == t) idx = (int)get(b, pos, n);
    pos += n;
  }
  const int s = M->regions == 2 ? (kPart2[part] >> t) & 1 : 0;
  const int w = ib == 3 ? kW3[idx] : kW4[idx];
  for (int c = 0; c < 3; c++) {
    int a = ep[2 * s][c], z = ep[2 * s + 1][c];
    if (isSigned) {
      a = (int16_t)(uint16_t)a;
      z = (int16_t)(uint16_t)z;
    }
    uint32_t v = (uint32_t)((a * (64 - w) + z * w + 32) >> 6);
    if (isSigned) {
      v = (v & 0x80000000u) ? ((((~v + 1u) * 31u) >> 5) | 0x8000u) : ((v * 31u) >> 5);
      if (v == 0x8000u) v = 0;
    } else {
      v = (v * 31u) >> 6;
    }
    out[c] = (uint16_t)v;
  }
}

/* BC_Decoder::Decode for n = 6 (BC6H, isSigned selects the SF16 form) and
 * n = 7 (BC7).  Returns 1. */
int swo_decode_bptc(const uint8_t *src, uint8_t *dst, int w, int h, int dstPitch, int dstBpp, int n, int isSigned) {
  const int bw = (w + 3) / 4;
  for (int by = 0; by < (h + 3) / 4; by++)
    for (int bx = 0; bx < bw; bx++) {
      const Bits128 b = load128(src + 16 * (size_t)(by * bw + bx));
      for (int y = 0; y < 4 && 4 * by + y < h; y++)
        for (int x = 0; x < 4 && 4 * bx + x < w; x++) {
          uint8_t *p = dst + (size_t)(4 * by + y) * dstPitch + (size_t)(4 * bx + x) * dstBpp;
          if (n == 7) {
            bc7_texel(b, x + 4 * y, p);
          } else {
            uint16_t v[4];
            bc6_texel(b, x + 4 * y, isSigned, v);
            memcpy(p, v, 8);
          }
        }
    }
  return 1;
}
