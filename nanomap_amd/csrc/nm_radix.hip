// nm_radix.hip -- stable LSD radix sort of (u32 key, u32 value) pairs.
//
// Used to group hit events by voxel (in ascending pixel order) and to bin
// points by voxel in the voxel filter.  8-bit digits; each pass is
//   hist:    per-tile digit histogram (LDS atomics), digit-major layout;
//   scan:    one workgroup turns the histogram into exclusive offsets;
//   scatter: stable rank of every key inside its tile, computed with
//            wavefront ballots on the 8 digit bits (a 64-lane match), a
//            per-wave digit count in LDS and a running per-digit count
//            across the tile's strided iterations.
// Tiles are 2048 keys: 256 threads x 8 strided iterations, so iteration j
// holds keys [j*256, j*256+256) of the tile in lane order and the input
// order is (iteration, wave, lane) -- the ranking is therefore stable.
#include "nm_internal.h"

namespace nm {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = kRadixTile / kThreads;  // 8
constexpr int kBins = 256;
constexpr int kWaves = kThreads / 64;

__global__ __launch_bounds__(kThreads) void k_radix_hist(
    const uint32_t* __restrict__ keys, int n, int shift,
    uint32_t* __restrict__ hist, int nblocks) {
  __shared__ uint32_t h[kBins];
  const int tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const int base = blockIdx.x * kRadixTile;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int i = base + j * kThreads + tid;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xff], 1u);
  }
  __syncthreads();
  hist[tid * nblocks + blockIdx.x] = h[tid];
}

// Exclusive scan of hist[0, total) by one workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_radix_scan(uint32_t* hist,
                                                     int total) {
  __shared__ uint32_t wsum[16];
  const int tid = threadIdx.x;
  const int per = (total + 1023) / 1024;
  const int b0 = tid * per;
  const int b1 = min(total, b0 + per);
  uint32_t local = 0;
  for (int i = b0; i < b1; ++i) local += hist[i];
  // inclusive scan of `local` across the block
  const int lane = tid & 63, wave = tid >> 6;
  uint32_t x = local;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  if (tid < 16) {
    uint32_t w = wsum[tid];
    for (int o = 1; o < 16; o <<= 1) {
      const uint32_t y = __shfl_up(w, o, 16);
      if (tid >= o) w += y;
    }
    wsum[tid] = w;
  }
  __syncthreads();
  uint32_t run = x - local + (wave ? wsum[wave - 1] : 0u);
  for (int i = b0; i < b1; ++i) {
    const uint32_t v = hist[i];
    hist[i] = run;
    run += v;
  }
}

__global__ __launch_bounds__(kThreads) void k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int n,
    int shift, const uint32_t* __restrict__ hist, int nblocks) {
  __shared__ uint32_t running[kBins];
  __shared__ uint32_t offset[kBins];
  __shared__ uint32_t wcnt[kWaves][kBins];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  running[tid] = 0;
  offset[tid] = hist[tid * nblocks + blockIdx.x];
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int base = blockIdx.x * kRadixTile;
  for (int j = 0; j < kItems; ++j) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    const int i = base + j * kThreads + tid;
    const bool valid = i < n;
    uint32_t key = 0, val = 0, d = 0;
    if (valid) {
      key = kin[i];
      val = vin[i];
      d = (key >> shift) & 0xff;
    }
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t bb = __ballot(valid && bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = offset[d] + running[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) tot += wcnt[w][tid];
    running[tid] += tot;
  }
}

}  // namespace

int radix_max_blocks(int n) { return (n + kRadixTile - 1) / kRadixTile; }

hipError_t radix_sort_pairs(uint32_t* keys[2], uint32_t* vals[2], int n,
                            int bits, uint32_t* hist, int* out_buf,
                            hipStream_t s) {
  int cur = 0;
  *out_buf = 0;
  if (n <= 0) return hipSuccess;
  const int nblocks = radix_max_blocks(n);
  for (int shift = 0; shift < bits; shift += 8) {
    k_radix_hist<<<nblocks, kThreads, 0, s>>>(keys[cur], n, shift, hist,
                                              nblocks);
    k_radix_scan<<<1, 1024, 0, s>>>(hist, kBins * nblocks);
    k_radix_scatter<<<nblocks, kThreads, 0, s>>>(
        keys[cur], vals[cur], keys[cur ^ 1], vals[cur ^ 1], n, shift, hist,
        nblocks);
    cur ^= 1;
  }
  *out_buf = cur;
  return hipGetLastError();
}

}  // namespace nm
