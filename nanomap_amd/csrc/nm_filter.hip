// nm_filter.hip -- deterministic voxel filter.
//
// Replaces voxelFilter + rayFromFilter (FP32 mode, PrecisionType 0,
// src/nanomap/kernels/filterCloud.cu:84-133, 279-318) and its lower-memory
// variants (PrecisionType 1-3, filterCloud.cu:134-277, 319-434): FP16
// sums and count (voxelFilterHalf2/rayFromFilterHalf2), and voxel-centre
// rays from an int count (Simple) or an 8-bit count that wraps
// (SimpleQuad: 256 points make a count of 0 and no ray).  The reference bins
// points with float atomicAdd into a dense buffer over the frustum's voxel
// bounding box (about 273 MB cleared and scanned per 640x480 frame at
// 0.05 m) and appends rays through an atomic counter, so both the sums
// and the ray order depend on thread timing.  Here:
//   keys    one lane per point: validity, r = R p, bin = floor(r/res +
//           voxelOriginOffset) relative to floor(eye), packed into a
//           32-bit key in (x, y, z) order; value = pixel index;
//   sort    stable radix sort of (key, pixel) -> each bin's points are
//           contiguous and in ascending pixel order;
//   bins    one lane per bin: FP32 (or FP16) sum of frac(r/res) in pixel
//           order, centroid (or the voxel centre), ray (dir, length) if
//           0 < length < max_time;
//   emit    prefix sum of the kept flags -> rays in ascending key order.
// Memory is O(points), independent of range and voxel size.
#include "nm_internal.h"

namespace nm {
namespace {

struct FilterKey {
  int bits;    // per axis
  int offset;  // added to each bin coordinate
};

// Validity (filterCloud.cu:98-112): no NaN coordinate, inside the
// sensor-frame box on every axis, 0 < |p| < maxRange; then the bin
// floor(r/res + voxelOriginOffset) and the fractions r/res - floor(r/res)
// of r = R p (filterCloud.cu:114-128).
__device__ __forceinline__ void point_bin(const CastParams& p,
                                          const FilterBounds& fb,
                                          const float* origin, int i,
                                          bool& valid, int b[3], float f[3]) {
  float px, py, pz;
  load_point(p, i, px, py, pz);
  valid = false;
  if (__builtin_isnan(px) || __builtin_isnan(py) || __builtin_isnan(pz))
    return;
  if (fb.use_box && (px < fb.lo[0] || px > fb.hi[0] || py < fb.lo[1] ||
                     py > fb.hi[1] || pz < fb.lo[2] || pz > fb.hi[2]))
    return;
  const float norm = __builtin_sqrtf(px * px + py * py + pz * pz);
  if (!(norm > 0.0f && norm < fb.max_range)) return;
  const float r[3] = {p.R[0] * px + p.R[1] * py + p.R[2] * pz,
                      p.R[3] * px + p.R[4] * py + p.R[5] * pz,
                      p.R[6] * px + p.R[7] * py + p.R[8] * pz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float q = r[a] / p.res;
    b[a] = (int)__builtin_floorf(q + origin[a]);
    f[a] = q - (float)(int)__builtin_floorf(q);
  }
  valid = true;
}

__global__ __launch_bounds__(256) void k_filter_keys(
    CastParams p, FilterBounds fb, float o0, float o1, float o2,
    FilterKey fk, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
    FrameCounters* __restrict__ cnt) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.n) return;
  const float origin[3] = {o0, o1, o2};
  bool valid;
  int b[3];
  float f[3];
  point_bin(p, fb, origin, i, valid, b, f);
  uint32_t key = kInvalid;
  if (valid) {
    const int lim = (1 << fk.bits) - 1;  // top value reserved
    const int x = b[0] + fk.offset, y = b[1] + fk.offset, z = b[2] + fk.offset;
    if (x >= 0 && y >= 0 && z >= 0 && x < lim && y < lim && z < lim) {
      key = ((uint32_t)x << (2 * fk.bits)) | ((uint32_t)y << fk.bits) |
            (uint32_t)z;
    } else {
      atomicAdd(&cnt->errors, 1);
    }
  }
  // the sort's first pass drops the kInvalid keys and counts the rest
  // into n_filter_points (a per-wavefront atomic on that one counter
  // serialised at the memory side: 48 us for a 640x480 frame)
  keys[i] = key;
  vals[i] = (uint32_t)i;
}

// precision: 0 FP32 centroid, 1 FP16 centroid, 2 voxel centre (int
// count), 3 voxel centre (8-bit count).
__global__ __launch_bounds__(256) void k_filter_bins(
    CastParams p, FilterBounds fb, float o0, float o1, float o2,
    FilterKey fk, int precision, const uint32_t* __restrict__ keys,
    const uint32_t* __restrict__ vals,
    const FrameCounters* __restrict__ cnt, float4* __restrict__ tmp,
    int32_t* __restrict__ flags) {
  const int nv = cnt->n_filter_points;
  const float origin[3] = {o0, o1, o2};
  for (int e = blockIdx.x * 256 + threadIdx.x; e < p.n;
       e += gridDim.x * 256) {
    int keep = 0;
    if (e < nv && (e == 0 || keys[e - 1] != keys[e])) {
      const uint32_t key = keys[e];
      const uint32_t m = (1u << fk.bits) - 1;
      const int bx = (int)((key >> (2 * fk.bits)) & m) - fk.offset;
      const int by = (int)((key >> fk.bits) & m) - fk.offset;
      const int bz = (int)(key & m) - fk.offset;
      // the bin is keys[e, end): a galloping search finds its end in
      // O(log n) loads instead of one dependent load per point
      int end;
      {
        int lo = e + 1, hi = nv, step = 1;
        for (;;) {
          const int probe = e + step;
          if (probe >= nv) break;
          if (keys[probe] != key) {
            hi = probe;
            break;
          }
          lo = probe + 1;
          step <<= 1;
        }
        while (lo < hi) {  // first index in [lo, hi) outside the bin
          const int mid = (lo + hi) >> 1;
          if (keys[mid] == key)
            lo = mid + 1;
          else
            hi = mid;
        }
        end = lo;
      }
      float pt[3];
      bool emit = true;
      if (precision <= 1) {
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        _Float16 hx = 0, hy = 0, hz = 0, hn = 0;
        const uint32_t n = (uint32_t)(end - e);
        // points are read eight at a time so their loads overlap; the sums
        // still take them one by one in ascending pixel order
        for (int q0 = e; q0 < end; q0 += 8) {
          float fx[8], fy[8], fz[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            if (q0 + u < end) {
              bool valid;
              int b[3];
              float f[3];
              point_bin(p, fb, origin, (int)vals[q0 + u], valid, b, f);
              fx[u] = f[0];
              fy[u] = f[1];
              fz[u] = f[2];
            }
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            if (q0 + u < end) {
              if (precision == 0) {
                sx += fx[u];
                sy += fy[u];
                sz += fz[u];
              } else {  // each term rounded to FP16, each sum in FP16
                hx = hx + (_Float16)fx[u];
                hy = hy + (_Float16)fy[u];
                hz = hz + (_Float16)fz[u];
                hn = hn + (_Float16)1.0f;
              }
            }
          }
        }
        if (precision == 1) {
          sx = (float)hx;
          sy = (float)hy;
          sz = (float)hz;
        }
        const float c = precision == 0 ? (float)n : (float)hn;
        pt[0] = sx / c + (float)bx;
        pt[1] = sy / c + (float)by;
        pt[2] = sz / c + (float)bz;
      } else {
        const uint32_t n = (uint32_t)(end - e);
        emit = precision == 2 || (n & 255u) != 0u;
        pt[0] = (float)bx + 0.5f;
        pt[1] = (float)by + 0.5f;
        pt[2] = (float)bz + 0.5f;
      }
      const float len =
          __builtin_sqrtf(pt[0] * pt[0] + pt[1] * pt[1] + pt[2] * pt[2]);
      if (emit && len > 0.0f && len < p.max_time) {
        tmp[e] = make_float4(pt[0] / len, pt[1] / len, pt[2] / len, len);
        keep = 1;
      }
    }
    flags[e] = keep;
  }
}

__device__ __forceinline__ int wave_sum_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_flag_count(
    const int32_t* __restrict__ flags, int n, int32_t* __restrict__ blocks) {
  const int base = blockIdx.x * kScanChunk;
  int c = 0;
  for (int j = threadIdx.x; j < kScanChunk; j += 256)
    if (base + j < n) c += flags[base + j];
  c = wave_sum_i(c);
  __shared__ int ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) blocks[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ __launch_bounds__(1024) void k_flag_scan(
    int32_t* __restrict__ blocks, int nblocks,
    FrameCounters* __restrict__ cnt) {
  __shared__ int wsum[16];
  const int tid = threadIdx.x;
  const int per = (nblocks + 1023) / 1024;
  const int b0 = tid * per, b1 = min(nblocks, b0 + per);
  int local = 0;
  for (int b = b0; b < b1; ++b) local += blocks[b];
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
    const int v = blocks[b];
    blocks[b] = run;
    run += v;
  }
  if (tid == 1023) cnt->n_filter_rays = wsum[15];
}

__global__ __launch_bounds__(256) void k_flag_emit(
    const int32_t* __restrict__ flags, int n,
    const int32_t* __restrict__ blocks, const float4* __restrict__ tmp,
    float4* __restrict__ rays) {
  __shared__ int ws[4];
  const int base = blockIdx.x * kScanChunk;
  int run = blocks[blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (int j0 = 0; j0 < kScanChunk; j0 += 256) {
    const int e = base + j0 + threadIdx.x;
    const bool f = e < n && flags[e];
    const uint64_t b = __ballot(f);
    if (lane == 0) ws[wave] = __popcll(b);
    __syncthreads();
    int pos = run + __popcll(b & lt);
    for (int w = 0; w < wave; ++w) pos += ws[w];
    if (f) rays[pos] = tmp[e];
    run += ws[0] + ws[1] + ws[2] + ws[3];
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_voxel_filter(const CastParams& p, const float origin[3],
                               const FilterBounds& fb, int precision,
                               const Workspace& b, hipStream_t s) {
  if (p.n <= 0) return hipSuccess;
  FilterKey fk;
  const int need = (int)__builtin_ceilf(p.max_time) + 3;
  fk.bits = 1;
  while ((1 << (fk.bits - 1)) < need) ++fk.bits;
  fk.offset = 1 << (fk.bits - 1);
  if (3 * fk.bits > 31) return hipErrorInvalidValue;
  const int grid = (p.n + 255) / 256;
  k_filter_keys<<<grid, 256, 0, s>>>(p, fb, origin[0], origin[1], origin[2],
                                     fk, b.keys[0], b.vals[0], b.counters);
  uint32_t* keys[2] = {b.keys[0], b.keys[1]};
  uint32_t* vals[2] = {b.vals[0], b.vals[1]};
  int out = 0;
  hipError_t e = radix_sort_pairs(keys, vals, p.n, 3 * fk.bits, b.radix_hist,
                                  &out, s, &b.counters->n_filter_points);
  if (e != hipSuccess) return e;
  k_filter_bins<<<min(grid, 4096), 256, 0, s>>>(
      p, fb, origin[0], origin[1], origin[2], fk, precision, b.keys[out],
      b.vals[out],
      b.counters, reinterpret_cast<float4*>(b.filter_tmp), b.flags);
  const int nblocks = (p.n + kScanChunk - 1) / kScanChunk;
  k_flag_count<<<nblocks, 256, 0, s>>>(b.flags, p.n, b.scan_blocks);
  k_flag_scan<<<1, 1024, 0, s>>>(b.scan_blocks, nblocks, b.counters);
  k_flag_emit<<<nblocks, 256, 0, s>>>(
      b.flags, p.n, b.scan_blocks, reinterpret_cast<const float4*>(b.filter_tmp),
      reinterpret_cast<float4*>(b.filter_rays));
  return hipGetLastError();
}

}  // namespace nm
