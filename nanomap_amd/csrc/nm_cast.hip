// nm_cast.hip -- the hit/miss cast and the clamped log-odds map update.
//
// Replaces frustumCastCloud (src/nanomap/kernels/frustumCastCloud.cu:311)
// and the host fold into the map (Handler::voxelUpdateFromFrustumBuffer,
// src/nanomap/handler/Handler.cpp:134), with the semantics of the
// reference CPU path (Handler::populateTempGrid/integrateTempGrid,
// Handler.cpp:28-122).  Per frame:
//
//  1 discover   one lane per point: ray set-up + voxel DDA (nm_dda.h);
//               marks every 8^3 leaf the ray enters in a dense cube table
//               around the eye, records the cell of the ray's hit voxel.
//  2 compact    prefix sum over the cube table -> dense tile slots, in
//               ascending cube order (deterministic).
//  3 hit sort   (tile cell, pixel) pairs of all hits, stable radix sort:
//               every hit voxel's hits in ascending pixel order.
//  4 segments   per hit voxel: hit count into its tile, hit bit in the
//               tile's 512-bit mask, zeroed gap counters.
//  5 accumulate one lane per point again: integer miss counts into the
//               tiles (atomics, merged across the wavefront when every
//               lane hits the same voxel).  A miss in a voxel that also
//               holds hits increments the gap counter of the first later
//               hit (binary search in the sorted hits), so the order of
//               hits and misses in pixel order is known exactly.
//  6 update     one workgroup per tile: the temp value of each voxel is
//               the FP32 sum of its events in pixel order, rebuilt from
//               the counts with nm::fold (exact, any rounding mode); then
//               map += temp, clamp, activity thresholds; tiles cleared.
//
// Counts are integers, so the result does not depend on the order in
// which lanes run: the map is bit-identical to the sequential CPU fold.
#include "nm_dda.h"
#include "nm_internal.h"

namespace nm {
namespace {

constexpr int kCastThreads = 256;

__device__ __forceinline__ int local_of(int x, int y, int z) {
  return (z & 7) + ((y & 7) << 3) + ((x & 7) << 6);
}

// Cube index of the leaf holding voxel (x,y,z), or -1 outside the cube.
__device__ __forceinline__ int cube_index(const Cube& c, int x, int y,
                                          int z) {
  const int lx = (x >> 3) - c.cmin[0];
  const int ly = (y >> 3) - c.cmin[1];
  const int lz = (z >> 3) - c.cmin[2];
  if ((unsigned)lx >= (unsigned)c.side || (unsigned)ly >= (unsigned)c.side ||
      (unsigned)lz >= (unsigned)c.side)
    return -1;
  return (lx * c.side + ly) * c.side + lz;
}

template <int M>
__device__ __forceinline__ bool ray_of(const CastParams& p, int i,
                                       RayDir& r) {
  if (p.use_rays) {
    if (i >= *p.n_dev) return false;
    const float4 q = reinterpret_cast<const float4*>(p.rays4)[i];
    const double L = (double)q.w;
    if (!(L > 0.0 && L < (double)p.max_time)) return false;
    r.d[0] = q.x;
    r.d[1] = q.y;
    r.d[2] = q.z;
    r.L = L;
    return true;
  }
  const float* pt =
      reinterpret_cast<const float*>(p.cloud + (size_t)i * (size_t)p.step);
  return ray_from_point<M>(p.R, pt[0], pt[1], pt[2], p.res, p.max_time, r);
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void k_reset_counters(FrameCounters* c) {
  if (threadIdx.x == 0) *c = FrameCounters{};
}

template <int M>
__global__ __launch_bounds__(kCastThreads) void k_discover(
    CastParams p, int32_t* __restrict__ leaf_slot,
    uint32_t* __restrict__ ray_hit, FrameCounters* __restrict__ cnt) {
  const int i = blockIdx.x * kCastThreads + threadIdx.x;
  uint32_t hit_cell = kInvalid;
  uint32_t visits = 0;
  int valid = 0, errors = 0;
  RayDir r;
  if (i < p.n && ray_of<M>(p, i, r)) {
    valid = 1;
    Dda s;
    dda_init<M>(p.eye, r, s);
    int c = cube_index(p.cube, s.vx, s.vy, s.vz);
    if (c < 0)
      ++errors;
    else
      leaf_slot[c] = -2;
    int px = s.vx, py = s.vy, pz = s.vz;
    while (dda_step<M>(s, r.L)) {
      ++visits;  // miss on (px, py, pz)
      if (((s.vx ^ px) | (s.vy ^ py) | (s.vz ^ pz)) & ~7) {
        c = cube_index(p.cube, s.vx, s.vy, s.vz);
        if (c < 0)
          ++errors;
        else
          leaf_slot[c] = -2;
      }
      px = s.vx;
      py = s.vy;
      pz = s.vz;
    }
    if (s.t < (double)p.max_time) {
      ++visits;
      const int ce = cube_index(p.cube, px, py, pz);
      if (ce >= 0) hit_cell = (uint32_t)ce * 512u + (uint32_t)local_of(px, py, pz);
    }
  }
  if (i < p.n) ray_hit[i] = hit_cell;
  const uint64_t sv = wave_sum(visits);
  const uint64_t sr = wave_sum((uint64_t)valid);
  const uint64_t sh = wave_sum((uint64_t)(hit_cell != kInvalid));
  const uint64_t se = wave_sum((uint64_t)errors);
  if ((threadIdx.x & 63) == 0) {
    if (sv) atomicAdd((unsigned long long*)&cnt->visits, (unsigned long long)sv);
    if (sr) atomicAdd((unsigned long long*)&cnt->rays, (unsigned long long)sr);
    if (sh) atomicAdd(&cnt->n_hit_rays, (int)sh);
    if (se) atomicAdd(&cnt->errors, (int)se);
  }
}

// ---- leaf compaction: exclusive scan over the cube table --------------
__global__ __launch_bounds__(256) void k_scan_count(
    const int32_t* __restrict__ leaf_slot, int64_t vol,
    int32_t* __restrict__ block_counts) {
  const int64_t base = (int64_t)blockIdx.x * kScanChunk;
  int n = 0;
  for (int j = threadIdx.x; j < kScanChunk; j += 256) {
    const int64_t c = base + j;
    if (c < vol && leaf_slot[c] == -2) ++n;
  }
  n = (int)wave_sum((uint64_t)n);
  __shared__ int ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) block_counts[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// One workgroup: exclusive scan of block_counts; total -> n_leaves.
__global__ __launch_bounds__(1024) void k_scan_blocks(
    int32_t* __restrict__ block_counts, int nblocks,
    FrameCounters* __restrict__ cnt) {
  __shared__ int wsum[16];
  const int tid = threadIdx.x;
  const int per = (nblocks + 1023) / 1024;
  const int b0 = tid * per, b1 = min(nblocks, b0 + per);
  int local = 0;
  for (int b = b0; b < b1; ++b) local += block_counts[b];
  const int lane = tid & 63, wave = tid >> 6;
  int x = local;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  if (tid < 16) {
    int w = wsum[tid];
    for (int o = 1; o < 16; o <<= 1) {
      const int y = __shfl_up(w, o, 16);
      if (tid >= o) w += y;
    }
    wsum[tid] = w;
  }
  __syncthreads();
  int run = x - local + (wave ? wsum[wave - 1] : 0);
  for (int b = b0; b < b1; ++b) {
    const int v = block_counts[b];
    block_counts[b] = run;
    run += v;
  }
  if (tid == 1023) cnt->n_leaves = wsum[15];
}

__global__ __launch_bounds__(256) void k_scan_assign(
    int32_t* __restrict__ leaf_slot, int64_t vol,
    const int32_t* __restrict__ block_counts,
    int32_t* __restrict__ slot_leaf, int tile_cap,
    FrameCounters* __restrict__ cnt) {
  __shared__ int ws[4];
  const int64_t base = (int64_t)blockIdx.x * kScanChunk;
  int run = block_counts[blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (int j0 = 0; j0 < kScanChunk; j0 += 256) {
    const int64_t c = base + j0 + threadIdx.x;
    const bool f = c < vol && leaf_slot[c] == -2;
    const uint64_t b = __ballot(f);
    if (lane == 0) ws[wave] = __popcll(b);
    __syncthreads();
    int pre = run + __popcll(b & lt);
    for (int w = 0; w < wave; ++w) pre += ws[w];
    if (f) {
      if (pre < tile_cap) {
        leaf_slot[c] = pre;
        slot_leaf[pre] = (int32_t)c;
      } else {
        leaf_slot[c] = -1;  // dropped: the frame reports overflow
        atomicAdd(&cnt->overflow, 1);
      }
    }
    run += ws[0] + ws[1] + ws[2] + ws[3];
    __syncthreads();
  }
}

// Sort keys of the hit events: tile cell (slot * 512 + local), value =
// pixel index.  Rays without a hit get the largest key and sort last.
__global__ __launch_bounds__(256) void k_hit_keys(
    int n, const uint32_t* __restrict__ ray_hit,
    const int32_t* __restrict__ leaf_slot, uint32_t* __restrict__ keys,
    uint32_t* __restrict__ vals) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t h = ray_hit[i];
  uint32_t key = kInvalid;
  if (h != kInvalid) {
    const int slot = leaf_slot[h >> 9];
    if (slot >= 0) key = (uint32_t)slot * 512u + (h & 511u);
  }
  keys[i] = key;
  vals[i] = (uint32_t)i;
}

// First index e in [lo, hi) with (keys[e], vals[e]) >= (key, val).
__device__ __forceinline__ int lower_bound_pair(const uint32_t* keys,
                                                const uint32_t* vals, int lo,
                                                int hi, uint32_t key,
                                                uint32_t val) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const uint32_t k = keys[mid];
    if (k < key || (k == key && vals[mid] < val))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_hit_segments(
    const uint32_t* __restrict__ keys, const FrameCounters* __restrict__ cnt,
    uint32_t* __restrict__ tiles, uint32_t* __restrict__ hit_mask,
    uint32_t* __restrict__ gap) {
  const int nh = cnt->n_hit_rays;
  for (int e = blockIdx.x * 256 + threadIdx.x; e < nh;
       e += gridDim.x * 256) {
    gap[e] = 0;
    const uint32_t key = keys[e];
    if (key == kInvalid) continue;
    if (e > 0 && keys[e - 1] == key) continue;
    // segment start: find its end by binary search
    int lo = e + 1, hi = nh;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (keys[mid] <= key)
        lo = mid + 1;
      else
        hi = mid;
    }
    tiles[2 * (size_t)key + 1] = (uint32_t)(lo - e);
    const uint32_t slot = key >> 9, li = key & 511;
    atomicOr(&hit_mask[slot * kMaskWords + (li >> 5)], 1u << (li & 31));
  }
}

template <int M>
__global__ __launch_bounds__(kCastThreads) void k_accumulate(
    CastParams p, const int32_t* __restrict__ leaf_slot,
    uint32_t* __restrict__ tiles, const uint32_t* __restrict__ hit_mask,
    const uint32_t* __restrict__ skeys, const uint32_t* __restrict__ svals,
    uint32_t* __restrict__ gap, FrameCounters* __restrict__ cnt) {
  const int i = blockIdx.x * kCastThreads + threadIdx.x;
  RayDir r;
  int errors = 0;
  if (i < p.n && ray_of<M>(p, i, r)) {
    const int nh = cnt->n_hit_rays;
    Dda s;
    dda_init<M>(p.eye, r, s);
    int px = s.vx, py = s.vy, pz = s.vz;
    int lkey = 0x7fffffff;  // leaf of the cached slot (packed, any value)
    int slot = -1;
    bool first = true;
    while (dda_step<M>(s, r.L)) {
      // miss on (px, py, pz)
      const int c = cube_index(p.cube, px, py, pz);
      if (first || c != lkey) {
        first = false;
        lkey = c;
        slot = c >= 0 ? leaf_slot[c] : -1;
      }
      if (slot < 0) {
        ++errors;
      } else {
        const int li = local_of(px, py, pz);
        const uint32_t cell = (uint32_t)slot * 512u + (uint32_t)li;
        const uint32_t mw = hit_mask[slot * kMaskWords + (li >> 5)];
        if ((mw >> (li & 31)) & 1u) {
          const int pos =
              lower_bound_pair(skeys, svals, 0, nh, cell, (uint32_t)i);
          if (pos < nh && skeys[pos] == cell) atomicAdd(&gap[pos], 1u);
        }
        atomicAdd(&tiles[2 * (size_t)cell], 1u);
      }
      px = s.vx;
      py = s.vy;
      pz = s.vz;
    }
  }
  const uint64_t se = wave_sum((uint64_t)errors);
  if ((threadIdx.x & 63) == 0 && se) atomicAdd(&cnt->errors, (int)se);
}

// ---- map hash ----------------------------------------------------------
__device__ __forceinline__ unsigned long long pack_leaf(int x, int y, int z) {
  const unsigned long long m = (1ull << 21) - 1;
  return (((unsigned long long)((x >> 3) + (1 << 20)) & m) << 42) |
         (((unsigned long long)((y >> 3) + (1 << 20)) & m) << 21) |
         ((unsigned long long)((z >> 3) + (1 << 20)) & m);
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// Leaf index of the leaf with origin (x,y,z), inserting it if absent;
// -1 if the pool or the table is full.  Within one update kernel every
// leaf is handled by exactly one workgroup, so an inserted key's value is
// written before any later kernel reads it.
__device__ int map_get_or_insert(const MapView& m, int x, int y, int z,
                                 FrameCounters* cnt) {
  const unsigned long long key = pack_leaf(x, y, z);
  const unsigned long long mask = (unsigned long long)m.hash_cap - 1;
  unsigned long long h = mix64(key) & mask;
  for (int64_t probe = 0; probe < m.hash_cap; ++probe) {
    const unsigned long long k = m.hash_keys[h];
    if (k == key) return m.hash_vals[h];
    if (k == ~0ull) {
      const unsigned long long old = atomicCAS(&m.hash_keys[h], ~0ull, key);
      if (old == ~0ull) {
        const int idx = atomicAdd(m.leaf_count, 1);
        if (idx >= m.leaf_cap) {
          m.hash_vals[h] = -1;
          atomicAdd(&cnt->overflow, 1);
          return -1;
        }
        m.origins[3 * idx + 0] = x;
        m.origins[3 * idx + 1] = y;
        m.origins[3 * idx + 2] = z;
        m.hash_vals[h] = idx;
        return idx;
      }
      if (old == key) return m.hash_vals[h];
    }
    h = (h + 1) & mask;
  }
  atomicAdd(&cnt->overflow, 1);
  return -1;
}

// First sorted index of `cell` among keys[0, nh).
__device__ __forceinline__ int lower_bound_key(const uint32_t* keys, int nh,
                                               uint32_t cell) {
  int lo = 0, hi = nh;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < cell)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

template <int M>
__global__ __launch_bounds__(512) void k_update(
    Cube cube, int32_t* __restrict__ leaf_slot,
    const int32_t* __restrict__ slot_leaf, uint32_t* __restrict__ tiles,
    uint32_t* __restrict__ hit_mask, const uint32_t* __restrict__ skeys,
    const uint32_t* __restrict__ gap, MapView m, float hit, float miss,
    int tile_cap, FrameCounters* __restrict__ cnt) {
  __shared__ int s_idx;
  const int li = threadIdx.x;
  const int lane = li & 63, wave = li >> 6;
  const int nleaves = min(cnt->n_leaves, tile_cap);
  const int nh = cnt->n_hit_rays;
  uint64_t n_vox = 0, n_hit = 0, n_mixed = 0;
  for (int slot = blockIdx.x; slot < nleaves; slot += gridDim.x) {
    const int c = slot_leaf[slot];
    const int lz = c % cube.side;
    const int ly = (c / cube.side) % cube.side;
    const int lx = c / (cube.side * cube.side);
    const int ox = (lx + cube.cmin[0]) * 8;
    const int oy = (ly + cube.cmin[1]) * 8;
    const int oz = (lz + cube.cmin[2]) * 8;
    const size_t cell = (size_t)slot * 512 + li;
    const uint32_t k = tiles[2 * cell];      // misses
    const uint32_t j = tiles[2 * cell + 1];  // hits
    const bool touched = (k | j) != 0;
    float tv = 0.0f;
    if (touched) {
      if (j == 0) {
        tv = fold<M>(0.0f, miss, k);
      } else {
        const int s0 = lower_bound_key(skeys, nh, (uint32_t)cell);
        uint32_t used = 0;
        float acc = 0.0f;
        for (uint32_t g = 0; g < j; ++g) {
          const uint32_t gm = gap[s0 + g];
          acc = fold<M>(acc, miss, gm);
          used += gm;
          acc = add<M>(acc, hit);
        }
        tv = fold<M>(acc, miss, k - used);
      }
    }
    const bool nz = touched && tv != 0.0f;
    if (li == 0) s_idx = -1;
    __syncthreads();
    const int any = __syncthreads_or(nz);
    if (any) {
      if (li == 0) s_idx = map_get_or_insert(m, ox, oy, oz, cnt);
      __syncthreads();
      const int idx = s_idx;
      if (idx >= 0) {
        float* mv = m.values + (size_t)idx * 512;
        uint32_t* mm = m.masks + (size_t)idx * kMaskWords;
        bool act = (mm[li >> 5] >> (li & 31)) & 1u;
        if (nz) {
          float v = add<M>(mv[li], tv);
          if (v > m.clamp_max) {
            v = m.clamp_max;
          } else if (v < m.clamp_min) {
            v = m.clamp_min;
          }
          if (v > m.hit_thr) {
            act = true;
          } else if (v < m.miss_thr) {
            act = false;
          }
          mv[li] = v;
        }
        const uint64_t b = __ballot(act);
        if (lane == 0) {
          mm[2 * wave] = (uint32_t)b;
          mm[2 * wave + 1] = (uint32_t)(b >> 32);
        }
      }
    }
    if (touched) {
      tiles[2 * cell] = 0;
      tiles[2 * cell + 1] = 0;
    }
    if (li < kMaskWords) hit_mask[(size_t)slot * kMaskWords + li] = 0;
    if (li == 0) leaf_slot[c] = -1;
    n_vox += touched;
    n_hit += (j != 0);
    n_mixed += (j != 0 && k != 0);
    __syncthreads();
  }
  n_vox = wave_sum(n_vox);
  n_hit = wave_sum(n_hit);
  n_mixed = wave_sum(n_mixed);
  if (lane == 0) {
    if (n_vox) atomicAdd((unsigned long long*)&cnt->voxels, (unsigned long long)n_vox);
    if (n_hit) atomicAdd((unsigned long long*)&cnt->hit_voxels, (unsigned long long)n_hit);
    if (n_mixed) atomicAdd((unsigned long long*)&cnt->mixed_voxels, (unsigned long long)n_mixed);
  }
}

__global__ void k_map_clear(MapView m) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t h = t; h < m.hash_cap; h += stride) {
    m.hash_keys[h] = ~0ull;
    m.hash_vals[h] = -1;
  }
  for (int64_t v = t; v < (int64_t)m.leaf_cap * 512; v += stride)
    m.values[v] = 0.0f;
  for (int64_t w = t; w < (int64_t)m.leaf_cap * kMaskWords; w += stride)
    m.masks[w] = 0;
  if (t == 0) *m.leaf_count = 0;
}

}  // namespace

hipError_t launch_reset_counters(FrameCounters* c, hipStream_t s) {
  k_reset_counters<<<1, 64, 0, s>>>(c);
  return hipGetLastError();
}

hipError_t launch_discover(int rounding, const CastParams& p,
                           const CtxBuffers& b, hipStream_t s) {
  if (p.n <= 0) return hipSuccess;
  const int grid = (p.n + kCastThreads - 1) / kCastThreads;
  if (rounding == kDown)
    k_discover<kDown><<<grid, kCastThreads, 0, s>>>(p, b.leaf_slot, b.ray_hit,
                                                    b.counters);
  else
    k_discover<kNearest><<<grid, kCastThreads, 0, s>>>(p, b.leaf_slot,
                                                       b.ray_hit, b.counters);
  return hipGetLastError();
}

hipError_t launch_compact_leaves(const CtxBuffers& b, int64_t cube_vol,
                                 int tile_cap, hipStream_t s) {
  const int nblocks = (int)((cube_vol + kScanChunk - 1) / kScanChunk);
  k_scan_count<<<nblocks, 256, 0, s>>>(b.leaf_slot, cube_vol, b.scan_blocks);
  k_scan_blocks<<<1, 1024, 0, s>>>(b.scan_blocks, nblocks, b.counters);
  k_scan_assign<<<nblocks, 256, 0, s>>>(b.leaf_slot, cube_vol, b.scan_blocks,
                                        b.slot_leaf, tile_cap, b.counters);
  return hipGetLastError();
}

hipError_t launch_hit_keys(const CastParams& p, const CtxBuffers& b,
                           hipStream_t s) {
  if (p.n <= 0) return hipSuccess;
  k_hit_keys<<<(p.n + 255) / 256, 256, 0, s>>>(p.n, b.ray_hit, b.leaf_slot,
                                               b.keys[0], b.vals[0]);
  return hipGetLastError();
}

hipError_t launch_hit_segments(const CtxBuffers& b, int n, int sorted_buf,
                               hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int grid = min((n + 255) / 256, 2048);
  k_hit_segments<<<grid, 256, 0, s>>>(b.keys[sorted_buf], b.counters, b.tiles,
                                      b.hit_mask, b.gap);
  return hipGetLastError();
}

hipError_t launch_accumulate(int rounding, const CastParams& p,
                             const CtxBuffers& b, int sorted_buf,
                             hipStream_t s) {
  if (p.n <= 0) return hipSuccess;
  const int grid = (p.n + kCastThreads - 1) / kCastThreads;
  if (rounding == kDown)
    k_accumulate<kDown><<<grid, kCastThreads, 0, s>>>(
        p, b.leaf_slot, b.tiles, b.hit_mask, b.keys[sorted_buf],
        b.vals[sorted_buf], b.gap, b.counters);
  else
    k_accumulate<kNearest><<<grid, kCastThreads, 0, s>>>(
        p, b.leaf_slot, b.tiles, b.hit_mask, b.keys[sorted_buf],
        b.vals[sorted_buf], b.gap, b.counters);
  return hipGetLastError();
}

hipError_t launch_update(int rounding, const Cube& cube, const CtxBuffers& b,
                         const MapView& m, float hit, float miss,
                         int sorted_buf, int tile_cap, hipStream_t s) {
  const int grid = 2048;
  if (rounding == kDown)
    k_update<kDown><<<grid, 512, 0, s>>>(cube, b.leaf_slot, b.slot_leaf,
                                         b.tiles, b.hit_mask,
                                         b.keys[sorted_buf], b.gap, m, hit,
                                         miss, tile_cap, b.counters);
  else
    k_update<kNearest><<<grid, 512, 0, s>>>(cube, b.leaf_slot, b.slot_leaf,
                                            b.tiles, b.hit_mask,
                                            b.keys[sorted_buf], b.gap, m, hit,
                                            miss, tile_cap, b.counters);
  return hipGetLastError();
}

hipError_t launch_map_clear(const MapView& m, hipStream_t s) {
  k_map_clear<<<1024, 256, 0, s>>>(m);
  return hipGetLastError();
}

}  // namespace nm
