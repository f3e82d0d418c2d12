// swr_batch.hip -- the tiled front end for a batch of draws.
//
// A batch is a list of draws rendered in one pass of each stage: the
// primitives of all draws share one record array (draw d owns slots
// [triBase_d, triBase_d + T_d)), and the screen tiles of all draws share
// one global tile index (draw d owns tiles [tileBase_d, tileBase_d + tiles_d)).
//
//   setup_batch_kernel    setup of every primitive of every draw
//   scan_*                exclusive prefix sum of the per-primitive tile
//                         counts (three passes, no memset, fixed sizes)
//   bin_emit_batch_kernel one thread per (tile, primitive) pair up to a
//                         fixed capacity; entries past the real count get
//                         the key 0xFFFFFFFF so that they sort last
//   radix_*               stable LSD radix sort of the pairs by tile, 8 bits
//                         per pass: per-block digit histograms, a scan, and
//                         an in-order scatter (one wave per block ranks its
//                         keys with ballots, so equal keys keep list order,
//                         which keeps every tile's primitives in API order)
//   tile_range_kernel     start / end of each tile's run of pairs
//
// Every size is fixed when the work is enqueued and nothing is cleared with
// a memset, so the whole front end can be captured into a HIP graph and
// replayed.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swr_params.hpp"

namespace swr {

constexpr int kScanItems = 1024;   // items per scan block (256 threads x 4)
constexpr int kSortItems = 1024;   // items per radix-sort block (64 lanes x 16)

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

// blockBase[d]: first 256-thread block of draw d.
__global__ __launch_bounds__(256) void setup_batch_kernel(const DrawParams *__restrict__ params,
                                                          const uint32_t *__restrict__ blockBase, int nDraws,
                                                          TriRecord *__restrict__ rec, int *__restrict__ poly,
                                                          uint32_t *__restrict__ tileCount) {
  const int d = find_segment(blockBase, nDraws, blockIdx.x);
  const DrawParams &P = params[d];
  const int t = (int)(blockIdx.x - blockBase[d]) * 256 + (int)threadIdx.x;
  if (t < P.triCount) setup_one(P, t, (size_t)P.triBase + (size_t)t, d, rec, poly, tileCount);
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

// ---- pair emission ---------------------------------------------------------
// off: exclusive prefix of the tile counts over the T record slots (off[T]
// = total pairs).  Pairs beyond `capacity` are dropped and the needed count
// is reported in *overflow (left untouched when everything fits).
__global__ __launch_bounds__(256) void bin_emit_batch_kernel(const DrawParams *__restrict__ params,
                                                             const uint32_t *__restrict__ tileBase,
                                                             const TriRecord *__restrict__ rec,
                                                             const uint32_t *__restrict__ off, uint32_t T,
                                                             uint32_t capacity, uint32_t *__restrict__ keys,
                                                             uint32_t *__restrict__ vals,
                                                             uint32_t *__restrict__ overflow) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t total = off[T];
  if (k == 0 && total > capacity) *overflow = total;
  if (k >= capacity) return;
  if (k >= total) {
    keys[k] = 0xFFFFFFFFu;
    vals[k] = 0u;
    return;
  }
  const int g = find_segment(off, (int)T, k);
  // Skip empty slots that share the prefix value (find the last slot <= k
  // whose range is non-empty: off[g] <= k < off[g + 1] holds for it).
  const TriRecord &R = rec[g];
  const DrawParams &P = params[R.draw];
  const uint32_t i = k - off[g];
  const int tx0 = R.xMin / P.tileW, tx1 = (R.xMax - 1) / P.tileW;
  const int ty0 = R.yMin / P.tileH;
  const uint32_t w = (uint32_t)(tx1 - tx0 + 1);
  keys[k] = tileBase[R.draw] + (uint32_t)((ty0 + (int)(i / w)) * P.tilesX + tx0 + (int)(i % w));
  vals[k] = (uint32_t)g;
}

// ---- stable LSD radix sort, 8 bits per pass ---------------------------------
// hist is digit-major: hist[digit * nb + block].
__global__ __launch_bounds__(256) void radix_hist_kernel(const uint32_t *__restrict__ keys, uint32_t n, int shift,
                                                         uint32_t *__restrict__ hist, uint32_t nb) {
  __shared__ uint32_t h[256];
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

// One wave per block of kSortItems keys, taken 64 at a time in order: each
// lane's rank among the lanes with its digit comes from eight ballots, so
// keys with equal digits keep their input order.
__global__ __launch_bounds__(64) void radix_scatter_kernel(const uint32_t *__restrict__ keysIn,
                                                           const uint32_t *__restrict__ valsIn, uint32_t n, int shift,
                                                           const uint32_t *__restrict__ histScan, uint32_t nb,
                                                           uint32_t *__restrict__ keysOut,
                                                           uint32_t *__restrict__ valsOut) {
  __shared__ uint32_t next[256];
  const int lane = threadIdx.x;
#pragma unroll
  for (int k = 0; k < 4; k++) next[lane + 64 * k] = histScan[(lane + 64 * k) * nb + blockIdx.x];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t base = blockIdx.x * kSortItems;
  for (int step = 0; step < kSortItems / 64; step++) {
    const uint32_t i = base + step * 64 + lane;
    const bool valid = i < n;
    const uint32_t key = valid ? keysIn[i] : 0u;
    const uint32_t val = valid ? valsIn[i] : 0u;
    const uint32_t digit = (key >> shift) & 0xFFu;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const uint64_t m = __ballot((digit >> b) & 1u);
      peers &= ((digit >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__builtin_popcountll(peers & lt);
    const uint32_t cnt = (uint32_t)__builtin_popcountll(peers);
    uint32_t pos = 0;
    if (valid) pos = next[digit] + rank;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (valid && rank + 1 == cnt) next[digit] += cnt;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (valid) {
      keysOut[pos] = key;
      valsOut[pos] = val;
    }
  }
}

// Tile ranges over sorted keys; keys >= tiles (unused capacity) are skipped.
__global__ void tile_range_batch_kernel(const uint32_t *__restrict__ keys, uint32_t n, uint32_t tiles,
                                        uint32_t *__restrict__ start, uint32_t *__restrict__ end) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  if (k >= tiles) return;
  if (i == 0 || keys[i - 1] != k) start[k] = i;
  if (i == n - 1 || keys[i + 1] != k) end[k] = i + 1;
}

}  // namespace swr
