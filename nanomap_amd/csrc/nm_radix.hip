// nm_radix.hip -- stable LSD radix sort of (u32 key, u32 value) pairs.
//
// Used to group hit events by voxel (in ascending pixel order) and to bin
// points by voxel in the voxel filter.  Digits of D = 8..11 bits, chosen so
// the pass count is minimal (27 key bits: 3 passes of 9); each pass is
//   hist:    per-tile digit histogram, digit-major layout.  In sorts of
//            more than 2^20 keys, lanes holding the same digit are matched
//            with D wavefront ballots and the first of them adds their
//            number (hit keys repeat a voxel's cell many times, and a plain
//            LDS atomic per key serializes on them); smaller sorts add one
//            per key;
//   rowscan: one workgroup per digit: exclusive prefix over the tiles and
//            the digit total (the scatter scans the 2^D totals in LDS);
//   scatter: stable rank of every key inside its tile, from the same
//            ballot match, a per-wave digit count in LDS and a running
//            per-digit count across the tile's strided iterations, both
//            updated by the first lane of each match only.
// Optionally the first pass drops keys equal to kInvalid (the hit sort's
// rays without a hit, a quarter to a third of all points): they are not
// counted or written, the first pass's scatter stores the number kept, and
// every later pass reads its key count from there.
// A tile is 256 threads x `items` strided iterations, so iteration j holds
// keys [j*256, j*256+256) of the tile in lane order and the input order is
// (iteration, wave, lane) -- the ranking is therefore stable.  Small sorts
// use 8 iterations (2048 keys) for enough tiles to fill the GPU; large
// ones up to 64, so the digit-major histogram (2^D words per tile, written
// and read one word per digit row) stays well below the size of the keys.
#include <algorithm>

#include "nm_internal.h"

namespace nm {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxItems = 64;
constexpr int kWaves = kThreads / 64;
constexpr int kMinItems = 8;
constexpr int kChunk = 8;  // iterations whose loads are issued together
static_assert(kMinItems % kChunk == 0, "items must be a multiple of kChunk");

// Lanes of the wavefront holding the same digit as this one (D ballots).
template <int D>
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < D; ++b) {
    const bool bit = (d >> b) & 1;
    const uint64_t bb = __ballot(valid && bit);
    peers &= bit ? bb : ~bb;
  }
  return peers;
}

template <int D>
__global__ __launch_bounds__(kThreads) void k_radix_hist(
    const uint32_t* __restrict__ keys, int n, int shift,
    uint32_t* __restrict__ hist, int nblocks, int items,
    const int* __restrict__ n_dev, bool drop, bool match) {
  constexpr int kBins = 1 << D;
  __shared__ uint32_t h[kBins];
  const int tid = threadIdx.x, lane = tid & 63;
  if (n_dev) n = *n_dev;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int b = tid; b < kBins; b += kThreads) h[b] = 0;
  __syncthreads();
  const int base = blockIdx.x * items * kThreads;
  // the keys of kChunk iterations are requested together (items is a
  // multiple of kChunk), so their loads overlap instead of each iteration
  // waiting for its own
  for (int j0 = 0; j0 < items; j0 += kChunk) {
    uint32_t kk[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      const int i = base + (j0 + u) * kThreads + tid;
      kk[u] = i < n ? keys[i] : kInvalid;
    }
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      const int i = base + (j0 + u) * kThreads + tid;
      const uint32_t key = kk[u];
      const bool valid = i < n && !(drop && key == kInvalid);
      const uint32_t d = valid ? (key >> shift) & (kBins - 1) : 0u;
      if (match) {
        const uint64_t peers = digit_peers<D>(d, valid);
        if (valid && (peers & lt) == 0)
          atomicAdd(&h[d], (uint32_t)__popcll(peers));
      } else if (valid) {
        atomicAdd(&h[d], 1u);
      }
    }
  }
  __syncthreads();
  for (int b = tid; b < kBins; b += kThreads) hist[b * nblocks + blockIdx.x] = h[b];
}

// Row scan: workgroup d turns row d of the digit-major histogram (the
// counts of digit d in every tile) into exclusive prefixes and writes the
// row total to totals[d].  Rows are short (one entry per 2048 keys), so
// each workgroup loops over its row in coalesced 256-wide chunks.
__global__ __launch_bounds__(256) void k_radix_rowscan(
    uint32_t* __restrict__ hist, int nblocks, uint32_t* __restrict__ totals) {
  __shared__ uint32_t ws[4];
  uint32_t* row = hist + (size_t)blockIdx.x * nblocks;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t run = 0;
  for (int c0 = 0; c0 < nblocks; c0 += 256) {
    const int c = c0 + tid;
    const uint32_t v = c < nblocks ? row[c] : 0u;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) ws[wave] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (int w = 0; w < wave; ++w) pre += ws[w];
    if (c < nblocks) row[c] = run + pre + x - v;
    run += ws[0] + ws[1] + ws[2] + ws[3];
    __syncthreads();
  }
  if (tid == 0) totals[blockIdx.x] = run;
}

template <int D>
__global__ __launch_bounds__(kThreads) void k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int n,
    int shift, const uint32_t* __restrict__ hist, int nblocks,
    const uint32_t* __restrict__ totals, int items,
    const int* __restrict__ n_dev, bool drop, int* __restrict__ n_out) {
  constexpr int kBins = 1 << D;
  __shared__ uint32_t running[kBins];
  __shared__ uint32_t offset[kBins];
  __shared__ uint32_t wcnt[kWaves][kBins];
  __shared__ uint32_t ws[kWaves];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  if (n_dev) n = *n_dev;
  // digit base = exclusive scan of the digit totals, in LDS
  {
    constexpr int kPer = kBins / kThreads > 0 ? kBins / kThreads : 1;
    uint32_t loc[kPer];
    uint32_t sum = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int b = tid * kPer + q;
      loc[q] = b < kBins ? totals[b] : 0u;
      sum += loc[q];
    }
    uint32_t x = sum;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) ws[wave] = x;
    __syncthreads();
    uint32_t pre = x - sum;
    for (int w = 0; w < wave; ++w) pre += ws[w];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int b = tid * kPer + q;
      if (b < kBins) {
        running[b] = 0;
        offset[b] = pre + hist[b * nblocks + blockIdx.x];
      }
      pre += loc[q];
    }
    // the last thread's running sum is the number of keys kept
    if (n_out && blockIdx.x == 0 && tid == kThreads - 1) *n_out = (int)pre;
  }
  for (int b = tid; b < kBins; b += kThreads) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wcnt[w][b] = 0;
  }
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int base = blockIdx.x * items * kThreads;
  // keys and values of kChunk iterations are requested together (items is
  // a multiple of kChunk); the iterations then run in order, so the ranks
  // stay stable
  for (int j0 = 0; j0 < items; j0 += kChunk) {
    uint32_t kk[kChunk], vv[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      const int i = base + (j0 + u) * kThreads + tid;
      kk[u] = i < n ? kin[i] : kInvalid;
      vv[u] = i < n ? vin[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      const int i = base + (j0 + u) * kThreads + tid;
      const uint32_t key = kk[u];
      const bool valid = i < n && !(drop && key == kInvalid);
      const uint32_t d = valid ? (key >> shift) & (kBins - 1) : 0u;
      const uint64_t peers = digit_peers<D>(d, valid);
      const uint32_t rank = (uint32_t)__popcll(peers & lt);
      const uint32_t cnt = (uint32_t)__popcll(peers);
      const bool leader = valid && rank == 0;
      if (leader) wcnt[wave][d] = cnt;
      __syncthreads();
      if (valid) {
        uint32_t pos = offset[d] + running[d] + rank;
        for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
        kout[pos] = key;
        vout[pos] = vv[u];
      }
      __syncthreads();
      // every lane has read: the leaders clear their wave's count and move
      // the digit's running count on (ordered before the next reads by the
      // next iteration's barrier)
      if (leader) {
        wcnt[wave][d] = 0;
        atomicAdd(&running[d], cnt);
      }
    }
  }
}

template <int D>
void radix_pass(uint32_t* kin, uint32_t* vin, uint32_t* kout, uint32_t* vout,
                int n, int shift, uint32_t* hist, int nblocks, int items,
                const int* n_dev, bool drop, int* n_out, hipStream_t s) {
  uint32_t* totals = hist + (size_t)(1 << D) * nblocks;
  // large sorts match equal digits before the LDS add (their tiles repeat
  // a voxel's key many times); small ones are faster with one add per key
  const bool match = n > (1 << 20);
  k_radix_hist<D><<<nblocks, kThreads, 0, s>>>(kin, n, shift, hist, nblocks,
                                               items, n_dev, drop, match);
  k_radix_rowscan<<<1 << D, 256, 0, s>>>(hist, nblocks, totals);
  k_radix_scatter<D><<<nblocks, kThreads, 0, s>>>(
      kin, vin, kout, vout, n, shift, hist, nblocks, totals, items, n_dev,
      drop, n_out);
}

}  // namespace

// Iterations per tile for n keys: the smallest power of two from 8 to 64
// that keeps the tiles at 2048 or fewer.  Small sorts use 8 (2048 keys a
// tile; smaller tiles measured slower: each tile's prologue reads the 2^D
// digit totals and its row of the histogram), large ones keep the
// digit-major histogram (2^D words per tile) well below the size of the
// keys.
int radix_items(int n) {
  int items = kMinItems;
  while (items < kMaxItems && (int64_t)n > (int64_t)items * kThreads * 2048)
    items *= 2;
  return items;
}
int radix_blocks(int n) {
  const int items = radix_items(n);
  return (int)(((int64_t)n + (int64_t)items * kThreads - 1) / ((int64_t)items * kThreads));
}
// The most tiles any sort of up to n keys uses: at most 2048 while the
// tiles can still grow, at most one per 64 iterations beyond.
int radix_max_blocks(int n) {
  const int64_t small = std::min<int64_t>(
      2048, ((int64_t)n + kMinItems * kThreads - 1) / (kMinItems * kThreads));
  const int64_t large = ((int64_t)n + (int64_t)kMaxItems * kThreads - 1) /
                        ((int64_t)kMaxItems * kThreads);
  return (int)std::max(small, large);
}
int radix_hist_words(int n) { return 2048 * (radix_max_blocks(n) + 2); }

hipError_t radix_sort_pairs(uint32_t* keys[2], uint32_t* vals[2], int n,
                            int bits, uint32_t* hist, int* out_buf,
                            hipStream_t s, int* n_kept) {
  int cur = 0;
  *out_buf = 0;
  if (n <= 0) return hipSuccess;
  const int items = radix_items(n);
  const int nblocks = radix_blocks(n);
  const int passes = (bits + 10) / 11;
  const int D = (bits + passes - 1) / passes < 8 ? 8 : (bits + passes - 1) / passes;
  for (int shift = 0; shift < bits; shift += D) {
    uint32_t *ki = keys[cur], *vi = vals[cur], *ko = keys[cur ^ 1],
             *vo = vals[cur ^ 1];
    // with n_kept: the first pass drops kInvalid keys and stores the count
    // kept, which the later passes read
    const bool first = shift == 0;
    const int* n_dev = (n_kept && !first) ? n_kept : nullptr;
    const bool drop = n_kept && first;
    int* n_out = first ? n_kept : nullptr;
    switch (D) {
      case 8: radix_pass<8>(ki, vi, ko, vo, n, shift, hist, nblocks, items, n_dev, drop, n_out, s); break;
      case 9: radix_pass<9>(ki, vi, ko, vo, n, shift, hist, nblocks, items, n_dev, drop, n_out, s); break;
      case 10: radix_pass<10>(ki, vi, ko, vo, n, shift, hist, nblocks, items, n_dev, drop, n_out, s); break;
      default: radix_pass<11>(ki, vi, ko, vo, n, shift, hist, nblocks, items, n_dev, drop, n_out, s); break;
    }
    cur ^= 1;
  }
  *out_buf = cur;
  return hipGetLastError();
}

}  // namespace nm
