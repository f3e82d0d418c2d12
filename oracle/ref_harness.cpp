// ref_harness.cpp -- C entry points into reference sources compiled as-is.
//
// TEST INFRASTRUCTURE ONLY (see oracle/Makefile, target `ref`).  This file
// is this project's own code; it is linked with unmodified reference
// translation units that compile standalone with g++ and the reference's
// own include directories:
//   src/Device/Clipper.cpp    (Sutherland-Hodgman clipping)
//   src/Pipeline/Constants.cpp (sample offsets and anisotropy tables)
//   src/System/Half.cpp       (needed by Constants.cpp)
//   src/Device/BC_Decoder.cpp (BC1-BC7 block decoding)
//   src/Device/ETC_Decoder.cpp (ETC2 / EAC block decoding)
//   src/System/Debug.cpp      (needed by BC_Decoder.cpp)
// The result, oracle/_ref/libswr_ref.so, lets the tests check the oracle's
// clipper bit-exactly and its constant tables value by value.  The rest of
// the draw path is JIT-generated through Reactor/Subzero and needs the
// reference's CMake build, so it is not built here.
#include <cstdint>
#include <cstring>
#include <new>
#include <xmmintrin.h>

#include "Device/BC_Decoder.hpp"
#include "Device/ETC_Decoder.hpp"
#include "Device/Clipper.hpp"
#include "Device/Polygon.hpp"
#include "Device/Renderer.hpp"
#include "Pipeline/Constants.hpp"

extern "C" {

// Clips triangle (v0, v1, v2) with sw::Clipper::Clip; writes the resulting
// polygon (n float4s) to `out`.  depthClipNegativeOneToOne is false, which
// is the only DrawCall field Clip() reads; the rest of the object is unused.
__attribute__((visibility("default"))) int ref_clip_triangle(const float *v0, const float *v1, const float *v2,
                                                              int flagsOr, float *out, int *outN) {
  unsigned csr = _mm_getcsr();
  _mm_setcsr(csr | 0x8040);  // FTZ | DAZ as on the reference's worker threads
  alignas(64) static unsigned char drawStorage[sizeof(sw::DrawCall)];
  memset(drawStorage, 0, sizeof(drawStorage));
  const sw::DrawCall &draw = *reinterpret_cast<const sw::DrawCall *>(drawStorage);
  sw::float4 a, b, c;
  memcpy(&a, v0, 16);
  memcpy(&b, v1, 16);
  memcpy(&c, v2, 16);
  sw::Polygon polygon(&a, &b, &c);
  bool visible = sw::Clipper::Clip(polygon, flagsOr, draw);
  *outN = polygon.n;
  for (int i = 0; i < polygon.n && i < 16; i++) memcpy(out + 4 * i, polygon.P[polygon.i][i], 16);
  _mm_setcsr(csr);
  return visible ? 1 : 0;
}

// Copies the on-path tables of sw::Constants.
__attribute__((visibility("default"))) void ref_constants(int *xf, int *yf, int *msOffsets, uint16_t *cweight,
                                                          float *uvweight, float *uvstart) {
  static sw::Constants *constants = new sw::Constants();
  for (int q = 0; q < 4; q++) {
    xf[q] = constants->Xf[q];
    yf[q] = constants->Yf[q];
  }
  msOffsets[0] = sw::yMinMultiSampleOffset;
  msOffsets[1] = sw::yMaxMultiSampleOffset;
  for (int n = 0; n < 17; n++) {
    cweight[n] = constants->cWeight[n][0];
    uvweight[n] = constants->uvWeight[n][0];
    uvstart[n] = constants->uvStart[n][0];
  }
}

// sw::Constants' sRGB decode table for 8-bit sampled texels.
__attribute__((visibility("default"))) void ref_srgb_lut(uint16_t *out) {
  static sw::Constants *constants = new sw::Constants();
  for (int i = 0; i < 256; i++) out[i] = constants->sRGBtoLinearFF_FF00[i];
}

// BC_Decoder::Decode as the reference's Image::decodeBC calls it.
__attribute__((visibility("default"))) int ref_decode_bc(const uint8_t *src, uint8_t *dst, int w, int h, int dstPitch,
                                                          int dstBpp, int n, int isNoAlphaU) {
  return BC_Decoder::Decode(src, dst, w, h, dstPitch, dstBpp, n, isNoAlphaU != 0) ? 1 : 0;
}

// ETC_Decoder::Decode; `type` is its InputType (0 R11 signed .. 6 RGBA).
__attribute__((visibility("default"))) int ref_decode_etc(const uint8_t *src, uint8_t *dst, int w, int h,
                                                           int dstPitch, int dstBpp, int type) {
  return ETC_Decoder::Decode(src, dst, w, h, dstPitch, dstBpp, static_cast<ETC_Decoder::InputType>(type)) ? 1 : 0;
}

}  // extern "C"
