// swr_batch.hip -- the tiled front end for a batch of draws.
//
// A batch is a list of draws rendered in one pass of each stage: the
// primitives of all draws share one record array (draw d owns slots
// [triBase_d, triBase_d + T_d)), and the screen tiles of all draws share
// one global tile index (draw d owns tiles [tileBase_d, tileBase_d + tiles_d)).
//
//   setup_batch_kernel    setup of every primitive of every draw
//   scan_*                exclusive prefix sum of the per-primitive tile
//                         counts (block scans, then each block adds the
//                         totals before it; no memset, fixed sizes)
//   emit_index_kernel     the first primitive slot of every 256-pair block
//                         (done by the scan's second launch when the pair
//                         capacity is known beforehand)
//   bin_emit_batch_kernel one thread per (tile, primitive) pair of the
//                         primitives' bounding boxes up to a fixed capacity,
//                         with the 8x8 blocks of the tile the primitive's
//                         edges may reach; pairs that reach none, and
//                         entries past the real count, get the key
//                         0xFFFFFFFF so that they sort last
//   radix_*               stable LSD radix sort of the pairs by tile, 8 bits
//                         per pass: per-block digit histograms, their block
//                         scans, and an in-order scatter that adds the block
//                         totals itself (waves rank their keys with
//                         ballots, so equal keys keep list order, which
//                         keeps every tile's primitives in API order; the
//                         block is staged in LDS sorted, then written in
//                         runs)
//   tile_range_batch_kernel  start / end of each tile's run of pairs
//
// Every size is fixed when the work is enqueued and nothing is cleared with
// a memset, so the whole front end can be captured into a HIP graph and
// replayed.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swr_params.hpp"

namespace swr {

constexpr int kScanItems = 1024;   // items per scan block (256 threads x 4)
constexpr int kSortItems = 4096;   // items per radix-sort block (4 waves x 16 steps x 64 lanes)
constexpr int kEmitBlock = 256;    // pairs per emission workgroup

// Largest d in [0, n) with base[d] <= v (base non-decreasing, base[0] <= v).
__device__ __forceinline__ int find_segment(const uint32_t *__restrict__ base, int n, uint32_t v) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (base[mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

// blockBase[d]: first 256-thread block of draw d (the polygon path as in
// setup_lanes).
__global__ __launch_bounds__(256) void setup_batch_kernel(const DrawParams *__restrict__ params,
                                                          const uint32_t *__restrict__ blockBase, int nDraws,
                                                          TriRecord *__restrict__ rec, int *__restrict__ poly,
                                                          uint32_t *__restrict__ tileCount) {
  __shared__ PolyWork work[4][kPolySlots];
  const int d = find_segment(blockBase, nDraws, blockIdx.x);
  const DrawParams &P = params[d];
  const int t = (int)(blockIdx.x - blockBase[d]) * 256 + (int)threadIdx.x;
  setup_lanes(P, t < P.triCount ? t : -1, (size_t)P.triBase + (size_t)t, d, rec, poly, tileCount,
              work[threadIdx.x >> 6]);
}

// ---- exclusive scan (u32) --------------------------------------------------
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = __shfl_up(v, off, 64);
    if (lane >= off) v += o;
  }
  return v;
}

// Block-local exclusive scan of 1024 items; the block's total to sums[b].
__global__ __launch_bounds__(256) void scan_partial_kernel(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                           uint32_t n, uint32_t *__restrict__ sums) {
  __shared__ uint32_t wsum[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t base = blockIdx.x * kScanItems + threadIdx.x * 4;
  uint32_t v[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    s += v[k];
  }
  const uint32_t incl = wave_incl_scan(s, lane);
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t prefix = 0;
  for (int w = 0; w < wave; w++) prefix += wsum[w];
  uint32_t run = prefix + incl - s;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) sums[blockIdx.x] = prefix + incl;
}

// In-place exclusive scan of the m block sums, by one workgroup of 1024.
__global__ __launch_bounds__(1024) void scan_sums_kernel(uint32_t *__restrict__ sums, uint32_t m) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t carry;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < m; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < m ? sums[i] : 0u;
    const uint32_t incl = wave_incl_scan(v, lane);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t prefix = carry;
    for (int w = 0; w < wave; w++) prefix += wsum[w];
    if (i < m) sums[i] = prefix + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry = prefix + incl;
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void scan_add_kernel(uint32_t *__restrict__ out, uint32_t n,
                                                       const uint32_t *__restrict__ sums) {
  const uint32_t add = sums[blockIdx.x];
  const uint32_t base = blockIdx.x * kScanItems + threadIdx.x * 4;
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (base + k < n) out[base + k] += add;
}

// Scans of at most this many blocks finish in one launch after
// scan_partial_kernel: each block sums the totals of the blocks before it
// itself (scan_finish_kernel), and the radix scatter scans the block totals
// of its histogram in LDS.
constexpr uint32_t kFusedScanBlocks = 1024;

// Emission index of the pairs of slot g, whose pairs are [lo, hi0) (see
// emit_index_kernel).
__device__ __forceinline__ void emit_slot(uint32_t g, uint32_t lo, uint32_t hi0, uint32_t capacity,
                                          uint32_t *__restrict__ first, uint32_t *__restrict__ split) {
  const uint32_t hi = hi0 < capacity ? hi0 : capacity;
  for (uint32_t b = (lo + kEmitBlock - 1) / kEmitBlock; b * kEmitBlock < hi; b++) first[b] = g;
  if (lo <= capacity && capacity < hi0) *split = g;
}

// Second launch of a scan of n = T + 1 counts `in` (in != out): block b adds
// the totals of blocks [0, b) to its items of out.  With `first`, it also
// does emit_index_kernel's work for slots [0, T) (slot g's pairs end at
// out[g] + in[g]); the last item (the total) writes *split = T when
// everything fits below `capacity`.
__global__ __launch_bounds__(256) void scan_finish_kernel(uint32_t *__restrict__ out, uint32_t n,
                                                          const uint32_t *__restrict__ sums,
                                                          const uint32_t *__restrict__ in, uint32_t capacity,
                                                          uint32_t *__restrict__ first, uint32_t *__restrict__ split) {
  __shared__ uint32_t wsum[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t p = 0;
  for (uint32_t i = threadIdx.x; i < blockIdx.x; i += 256) p += sums[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) p += __shfl_xor(p, o, 64);
  if (lane == 0) wsum[wave] = p;
  __syncthreads();
  const uint32_t add = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  const uint32_t base = blockIdx.x * kScanItems + threadIdx.x * 4;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t g = base + k;
    if (g >= n) break;
    const uint32_t v = out[g] + add;
    out[g] = v;
    if (first) {
      if (g + 1 < n) emit_slot(g, v, v + in[g], capacity, first, split);
      else if (v <= capacity) *split = g;
    }
  }
}

// ---- pair emission ---------------------------------------------------------
// off: exclusive prefix of the tile counts over the T record slots (off[T]
// = total pairs).  Pair k belongs to the slot g with off[g] <= k <
// off[g + 1].  first[b] is the slot of pair b * kEmitBlock (written for the
// blocks that start below the total): slot g writes it for the blocks whose
// first pair falls in its range.
//
// *split receives the first slot whose pairs do not all fit below
// `capacity` (T when everything fits): exactly one thread writes it, the
// slot whose non-empty pair range holds pair `capacity`, or slot T - 1 when
// the total fits.  The binned pass renders the pairs of slots below the
// split; the slots from the split on are rendered afterwards by the
// unbinned pass of the raster kernel, so a run whose geometry outgrew its
// plan still renders every primitive, in order.
__global__ __launch_bounds__(256) void emit_index_kernel(const uint32_t *__restrict__ off, uint32_t T,
                                                         uint32_t capacity, uint32_t *__restrict__ first,
                                                         uint32_t *__restrict__ split) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= T) return;
  const uint32_t hi0 = off[g + 1];
  emit_slot(g, off[g], hi0, capacity, first, split);
  if (g == T - 1 && hi0 <= capacity) *split = T;
}

// The 8x8 blocks of a tile that a triangle may cover, as bits 0-3 (block b
// of the tile in row-major order, tileW / 8 per row).  All sample points of
// a block's pixels lie within 6/16 px of their centres (28.4 coordinate
// 16 x + xo); a block is dropped when one edge of the triangle has the
// block's centres grown by half a pixel on every side strictly on its outer
// side.  Covered samples lie in the closed triangle (the span rule takes
// x >= ceil(Xleft / 16) and x < ceil(Xright / 16) of the spanning edges),
// so no covered sample is dropped.  Together with the bounding-box test of
// the list walk this is the separating-axis test of the triangle against
// the block.  Clipped polygons, lines and points (n != 3) and degenerate
// triangles keep every block.
__device__ __forceinline__ uint32_t tile_block_mask(const TriRecord &R, int tx, int ty, int tileW, int tileH) {
  const int bw = tileW >> 3, bh = tileH >> 3;
  const uint32_t all = (1u << (bw * bh)) - 1u;
  if ((R.nflags & 0xFF) != 3) return all;
  const long long X0 = R.X[0], X1 = R.X[1], X2 = R.X[2], Y0 = R.Y[0], Y1 = R.Y[1], Y2 = R.Y[2];
  const long long area = (X1 - X0) * (Y2 - Y0) - (Y1 - Y0) * (X2 - X0);
  if (area == 0) return all;
  const long long s = area > 0 ? 1 : -1;
  // Block (i, j) of the tile, grown: x in [bx + 128 i, bx + 128 i + 128]
  // (28.4), likewise y.
  const long long bx = 16ll * tx * tileW - 8, by = 16ll * ty * tileH - 8;
  uint32_t out = 0;
#pragma unroll
  for (int e = 0; e < 3; e++) {
    const long long xa = e == 0 ? X0 : (e == 1 ? X1 : X2), ya = e == 0 ? Y0 : (e == 1 ? Y1 : Y2);
    const long long xb = e == 0 ? X1 : (e == 1 ? X2 : X0), yb = e == 0 ? Y1 : (e == 1 ? Y2 : Y0);
    // s * f(x, y), f = (xb - xa)(y - ya) - (yb - ya)(x - xa), is positive
    // inside; its largest value over a block is at the corner picked by the
    // signs of its slopes, and moves by 128 slopes from block to block.
    const long long cx = -s * (yb - ya), cy = s * (xb - xa);
    const long long v = cx * ((cx > 0 ? bx + 128 : bx) - xa) + cy * ((cy > 0 ? by + 128 : by) - ya);
    const long long stepX = 128 * cx, stepY = 128 * cy;
    for (int j = 0; j < bh; j++)
      for (int i = 0; i < bw; i++)
        if (v + i * stepX + j * stepY < 0) out |= 1u << (j * bw + i);
  }
  return all & ~out;
}

// Pairs of the slots from *split on are not emitted (their keys, like those
// of the unused capacity, are 0xFFFFFFFF) and the total needed is reported
// in *overflow (left untouched when everything fits).  A workgroup emits
// kEmitBlock consecutive pairs; their slots lie in [first[b], first[b + 1]]
// (the last block searches up to T - 1), a range each thread searches,
// staged in LDS when it fits.  A pair's value is its slot with, in bits
// 28-31, the blocks of its tile the primitive may cover (tile_block_mask);
// a pair that covers none of them gets the key 0xFFFFFFFF and leaves the
// tile's list.
constexpr int kEmitSpan = 1024;

__global__ __launch_bounds__(kEmitBlock) void bin_emit_batch_kernel(const DrawParams *__restrict__ params,
                                                                    const uint32_t *__restrict__ tileBase,
                                                                    const TriRecord *__restrict__ rec,
                                                                    const uint32_t *__restrict__ off, uint32_t T,
                                                                    const uint32_t *__restrict__ first,
                                                                    uint32_t capacity, uint32_t *__restrict__ keys,
                                                                    uint32_t *__restrict__ vals,
                                                                    uint32_t *__restrict__ overflow,
                                                                    const uint32_t *__restrict__ split,
                                                                    uint32_t *__restrict__ emitted) {
  __shared__ uint32_t sOff[kEmitSpan + 1];
  const uint32_t k = blockIdx.x * kEmitBlock + threadIdx.x;
  const uint32_t needed = off[T];
  const uint32_t total = off[*split];  // <= capacity
  if (k == 0 && needed > capacity) *overflow = needed;
  if (k == 0) *emitted = total;
  const uint32_t k0 = blockIdx.x * kEmitBlock;
  const bool any = k0 < total && k0 < capacity;  // workgroup-uniform
  int t0 = 0, span = 0;
  if (any) {
    t0 = (int)first[blockIdx.x];
    const uint32_t next = k0 + kEmitBlock;
    const int t1 = (next < total && next < capacity) ? (int)first[blockIdx.x + 1] : (int)T - 1;
    span = t1 - t0 + 1;
    if (span <= kEmitSpan)
      for (int i = threadIdx.x; i <= span; i += kEmitBlock) sOff[i] = off[t0 + i];
  }
  __syncthreads();
  if (k >= capacity) return;
  if (k >= total) {
    keys[k] = 0xFFFFFFFFu;
    vals[k] = 0u;
    return;
  }
  // Largest slot g in [t0, t0 + span) with off[g] <= k: the slot whose
  // range holds k (empty slots share their prefix value with the next).
  int lo = 0, hi = span;
  if (span <= kEmitSpan) {
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (sOff[mid] <= k) lo = mid;
      else hi = mid;
    }
  } else {
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (off[t0 + mid] <= k) lo = mid;
      else hi = mid;
    }
  }
  const int g = t0 + lo;
  const TriRecord &R = rec[g];
  const DrawParams &P = params[R.draw];
  const uint32_t i = k - off[g];
  // Tiles are 8 or 16 pixels a side and the box is not negative (the
  // render area starts at x, y >= 0): shifts divide exactly.
  const int sw = __builtin_ctz((uint32_t)P.tileW), sh = __builtin_ctz((uint32_t)P.tileH);
  const int tx0 = R.xMin >> sw, tx1 = (R.xMax - 1) >> sw;
  const int ty0 = R.yMin >> sh;
  const uint32_t w = (uint32_t)(tx1 - tx0 + 1);
  const uint32_t row = i / w;
  const int tx = tx0 + (int)(i - row * w), ty = ty0 + (int)row;
  const uint32_t bm = tile_block_mask(R, tx, ty, P.tileW, P.tileH);
  keys[k] = bm ? tileBase[R.draw] + (uint32_t)(ty * P.tilesX + tx) : 0xFFFFFFFFu;
  vals[k] = (uint32_t)g | (bm << 28);
}

// ---- stable LSD radix sort, 8 bits per pass ---------------------------------
// The sort and the tile ranges cover the *count pairs that were emitted
// (bin_emit_batch_kernel's total, at most the capacity the grids are sized
// for): blocks past it write empty histograms and nothing else, so unused
// capacity costs no sorting.
// hist is digit-major: hist[digit * nb + block].
__global__ __launch_bounds__(256) void radix_hist_kernel(const uint32_t *__restrict__ keys,
                                                         const uint32_t *__restrict__ count, int shift,
                                                         uint32_t *__restrict__ hist, uint32_t nb) {
  __shared__ uint32_t h[256];
  const uint32_t n = *count;
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortItems;
#pragma unroll
  for (int k = 0; k < kSortItems / 256; k++) {
    const uint32_t i = base + k * 256 + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// One workgroup of 4 waves per block of kSortItems keys.  Each wave ranks
// its 1024 consecutive keys 64 at a time in order (a lane's rank among the
// lanes with its digit comes from eight ballots, so equal digits keep their
// input order); the waves' counts then give every key its place in the
// block sorted by digit, the block is staged in LDS in that order, and
// written out with consecutive threads on consecutive positions of each
// digit's run.
//
// histScan is the exclusive scan of the digit-major histogram; with
// chunkSums it is only scanned within blocks of kScanItems entries
// (scan_partial_kernel), chunkSums holding the m block totals, and each
// workgroup scans those totals itself (m <= kFusedScanBlocks).
__global__ __launch_bounds__(256) void radix_scatter_kernel(const uint32_t *__restrict__ keysIn,
                                                            const uint32_t *__restrict__ valsIn,
                                                            const uint32_t *__restrict__ count, int shift,
                                                            const uint32_t *__restrict__ histScan, uint32_t nb,
                                                            const uint32_t *__restrict__ chunkSums, uint32_t m,
                                                            uint32_t *__restrict__ keysOut,
                                                            uint32_t *__restrict__ valsOut) {
  constexpr int kSteps = kSortItems / 256;
  const uint32_t n = *count;
  if (blockIdx.x * (uint32_t)kSortItems >= n) return;  // workgroup-uniform
  __shared__ uint32_t cnt[4][256];  // per wave and digit: running count, then the wave's offset
  __shared__ uint32_t lstart[256];  // block-local start of each digit
  __shared__ uint32_t gstart[256];  // output start of each digit for this block
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t skey[kSortItems], sval[kSortItems];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 4 * 256; i += 256) (&cnt[0][0])[i] = 0u;
  {
    const uint32_t e = threadIdx.x * nb + blockIdx.x;
    uint32_t g = histScan[e];
    if (chunkSums) {
      // Exclusive scan of the block totals into skey (free until the keys
      // are staged), 256 at a time.
      uint32_t carry = 0;
      for (uint32_t c0 = 0; c0 < m; c0 += 256) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t v = i < m ? chunkSums[i] : 0u;
        const uint32_t incl = wave_incl_scan(v, lane);
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        uint32_t pre = carry;
        for (int w = 0; w < wave; w++) pre += wsum[w];
        if (i < m) skey[i] = pre + incl - v;
        carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
      }
      g += skey[e / kScanItems];
    }
    gstart[threadIdx.x] = g;
  }
  __syncthreads();

  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t base = blockIdx.x * kSortItems + wave * (kSortItems / 4);
  uint32_t keyv[kSteps], valv[kSteps], rankv[kSteps];
#pragma unroll
  for (int step = 0; step < kSteps; step++) {
    const uint32_t i = base + step * 64 + lane;
    keyv[step] = i < n ? keysIn[i] : 0u;
    valv[step] = i < n ? valsIn[i] : 0u;
  }
#pragma unroll
  for (int step = 0; step < kSteps; step++) {
    const bool valid = base + step * 64 + lane < n;
    const uint32_t digit = (keyv[step] >> shift) & 0xFFu;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const uint64_t m = __ballot((digit >> b) & 1u);
      peers &= ((digit >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__builtin_popcountll(peers & lt);
    const uint32_t c = (uint32_t)__builtin_popcountll(peers);
    rankv[step] = valid ? cnt[wave][digit] + rank : 0u;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (valid && rank + 1 == c) cnt[wave][digit] += c;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  }
  __syncthreads();

  // Thread d: the waves' offsets within digit d, and the block-local start
  // of digit d (exclusive scan of the digit totals).
  {
    const int d = threadIdx.x;
    const uint32_t c0 = cnt[0][d], c1 = cnt[1][d], c2 = cnt[2][d], c3 = cnt[3][d];
    const uint32_t tot = c0 + c1 + c2 + c3;
    uint32_t incl = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t pre = 0;
    for (int w = 0; w < wave; w++) pre += wsum[w];
    lstart[d] = pre + incl - tot;
    cnt[0][d] = 0u;
    cnt[1][d] = c0;
    cnt[2][d] = c0 + c1;
    cnt[3][d] = c0 + c1 + c2;
  }
  __syncthreads();
#pragma unroll
  for (int step = 0; step < kSteps; step++) {
    if (base + step * 64 + lane < n) {
      const uint32_t digit = (keyv[step] >> shift) & 0xFFu;
      const uint32_t pos = lstart[digit] + cnt[wave][digit] + rankv[step];
      skey[pos] = keyv[step];
      sval[pos] = valv[step];
    }
  }
  __syncthreads();
  const uint32_t b0 = blockIdx.x * kSortItems;
  const uint32_t valid = n - b0 < (uint32_t)kSortItems ? n - b0 : (uint32_t)kSortItems;
  for (uint32_t e = threadIdx.x; e < valid; e += 256) {
    const uint32_t key = skey[e];
    const uint32_t digit = (key >> shift) & 0xFFu;
    const uint32_t o = gstart[digit] + (e - lstart[digit]);
    keysOut[o] = key;
    valsOut[o] = sval[e];
  }
}

// Tile ranges over the *count sorted keys; keys >= tiles (dropped pairs) are
// skipped.
__global__ void tile_range_batch_kernel(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ count,
                                        uint32_t tiles, uint32_t *__restrict__ start, uint32_t *__restrict__ end) {
  const uint32_t n = *count;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  if (k >= tiles) return;
  if (i == 0 || keys[i - 1] != k) start[k] = i;
  if (i == n - 1 || keys[i + 1] != k) end[k] = i + 1;
}

}  // namespace swr
