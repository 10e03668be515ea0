// nm_cast.hip -- the hit/miss cast and the clamped log-odds map update.
//
// Replaces frustumCastCloud (src/nanomap/kernels/frustumCastCloud.cu:311)
// and the host fold into the map (Handler::voxelUpdateFromFrustumBuffer,
// src/nanomap/handler/Handler.cpp:134), with the semantics of the
// reference CPU path (Handler::populateTempGrid/integrateTempGrid,
// Handler.cpp:28-122).  Every kernel reads its frames from a device table
// of jobs (one frame of one sensor each), so one sequence of launches
// serves one frame or a batch of many sensors with their own poses, ranges,
// voxel sizes and maps.  Per sequence:
//
//  1 discover   one lane per point: ray set-up (nm_dda.h), then the
//               leaf-level form of the voxel DDA (per-axis crossing times
//               merged a leaf at a time); marks every 8^3 leaf the ray
//               enters in the job's dense cube table around the eye,
//               records the cell of the ray's hit voxel.
//  2 compact    prefix sum over all jobs' cube tables -> dense tile slots,
//               in ascending (job, cube) order (deterministic).
//  3 hit sort   (tile cell, point) pairs of all hits, stable radix sort
//               (rays without a hit dropped by the first pass): every hit
//               voxel's hits in ascending pixel order.
//  4 segments   per hit voxel: end of its run of sorted hits into its
//               tile (the start into hit_start), hit bit in the
//               tile's 512-bit mask, zeroed gap counters.
//  5 accumulate one lane per point again, voxel DDA: integer miss counts
//               into the tiles (LDS-aggregated per work item).  A miss in
//               a voxel that also holds hits increments the gap counter of
//               the first later hit (binary search in the sorted hits), so
//               the order of hits and misses in pixel order is known
//               exactly.
//  6 temp       one lane per tile voxel: the temp value of each voxel is
//               the FP32 sum of its events in pixel order, rebuilt from
//               the counts with nm::fold (exact, any rounding mode); a
//               wavefront per voxel with many hits (k_heavy, on a forked
//               stream beside the per-voxel pass).
//  7 apply      one wavefront per tile: map += temp, clamp, activity
//               thresholds; tiles cleared.
//
// Counts are integers, so the result depends neither on the order in which
// lanes run nor on how the work items are shaped and scheduled: each job's
// map is bit-identical to the sequential CPU fold of that frame alone.
//
// Work items of the cast kernels are 256 rays: a pixel tile of an image
// cloud (layouts in nm_internal.h) or 256 consecutive points otherwise.
// They are taken one per workgroup, or drawn by persistent workgroups from
// a device queue (job order or costliest first).
#include <climits>

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
__device__ __forceinline__ bool ray_of(const Job& p, int i, RayDir& r) {
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
  float x, y, z;
  load_point(p, i, x, y, z);
  return ray_from_point<M>(p.R, x, y, z, p.res, p.max_time, r);
}

// Inclusive prefix sum over the 64 lanes of a fully active wavefront, in
// DPP moves: lanes read lane - 1, 2, 4, 8 inside their 16-lane row (0 past
// the row start), then lane 15 of each row is added to the next row (rows
// 1 and 3) and lane 31 to rows 2 and 3.
__device__ __forceinline__ uint32_t wave_incl_scan32(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
  return v;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__host__ __device__ inline void layout_dims(int layout, int& tw, int& th,
                                            int& ww) {
  switch (layout) {
    case kLayout32x8: tw = 32; th = 8; ww = 8; break;
    case kLayout64x4: tw = 64; th = 4; ww = 16; break;
    case kLayout8x32: tw = 8; th = 32; ww = 8; break;
    default: tw = 16; th = 16; ww = 8; break;
  }
}

// Point of thread `tid` in work item `t` of job J, or -1.  Image clouds
// are walked in pixel tiles so the rays of a wavefront stay close together:
// less divergence in the DDA, more shared voxels for the LDS aggregation.
__device__ __forceinline__ int point_of(const Job& J, int t, int tid) {
  if (J.tiled) {
    int tw, th, ww;
    layout_dims(J.layout, tw, th, ww);
    const int w = tid >> 6, lane = tid & 63;
    int sx, sy;
    if (J.layout == kLayout16x16) {
      sx = (w & 1) * 8;
      sy = (w >> 1) * 8;
    } else if (J.layout == kLayout8x32) {
      sx = 0;
      sy = w * 8;
    } else {
      sx = w * ww;
      sy = 0;
    }
    const int tiles_x = (J.width + tw - 1) / tw;
    const int x = (t % tiles_x) * tw + sx + lane % ww;
    const int y = (t / tiles_x) * th + sy + lane / ww;
    if (x >= J.width || y >= J.height) return -1;
    return y * J.width + x;
  }
  const int i = t * kCastThreads + tid;
  return i < J.n ? i : -1;
}

// Work items of a launch: one per workgroup (head == null) or drawn from
// the queue at *head by persistent workgroups, in `order` if given.
struct ItemQueue {
  const int32_t* item_job;  // job of each item (null: job 0)
  const int32_t* order;
  int32_t* head;
  int nitems;
};

// Next item of this workgroup, or -1.  `round` counts the calls.
__device__ __forceinline__ int take_item(const ItemQueue& q, int& round) {
  if (!q.head) {
    const int it = (int)blockIdx.x;
    return (round++ == 0 && it < q.nitems) ? it : -1;
  }
  __shared__ int s_item;
  __syncthreads();  // every lane is done with the previous item
  if (threadIdx.x == 0) s_item = atomicAdd(q.head, 1);
  __syncthreads();
  const int it = __builtin_amdgcn_readfirstlane(s_item);
  if (it >= q.nitems) return -1;
  return q.order ? __builtin_amdgcn_readfirstlane(q.order[it]) : it;
}

__device__ __forceinline__ int job_of_item(const ItemQueue& q, int item) {
  return q.item_job ? __builtin_amdgcn_readfirstlane(q.item_job[item]) : 0;
}

__global__ void k_begin(Job j0, int has_job0, Job* __restrict__ jobs,
                        FrameCounters* __restrict__ c) {
  if (threadIdx.x == 0) {
    *c = FrameCounters{};
    if (has_job0) jobs[0] = j0;
  }
}

// Per-item partial sums of the counters, written with plain stores and
// reduced later: global atomics on a handful of addresses serialize at the
// memory side and cost more than the traversal itself.
__device__ __forceinline__ void block_partials(uint64_t* __restrict__ out,
                                               const uint64_t* v, int n) {
  __shared__ uint64_t s_part[4][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int q = 0; q < n; ++q) {
    const uint64_t w = wave_sum(v[q]);
    if (lane == 0) s_part[wave][q] = w;
  }
  __syncthreads();
  if (threadIdx.x < n)
    out[threadIdx.x] = s_part[0][threadIdx.x] + s_part[1][threadIdx.x] +
                       s_part[2][threadIdx.x] + s_part[3][threadIdx.x];
}

// Leaf-level form of the voxel DDA for the discover pass.  Each axis's
// crossing times form their own sequence, c(k+1) = round(c(k) + delta),
// independent of the other axes: the DDA only merges them in time order
// (ties: z, then y, then x) and stops at the first crossing after L.  The
// leaf changes only at every eighth crossing of an axis, so the pass merges
// just those leaf-entering crossings and produces each axis's times eight
// at a time.  Inside one binade [2^e, 2^(e+1)) every time is a multiple of
// its ulp u and a rounded step adds the fixed D = round(delta / u) u, so
// eight steps that stay in the binade (and at or below L) are one exact
// jump; elsewhere the steps are taken one by one.  The visit count is the
// number of crossings at or below L, the end voxel is the start voxel moved
// by that many crossings per axis, and the exit time is the first crossing
// above L.
// FP64 rounding mode of the wave (MODE.FP_ROUND bits 3:2, shared by FP64
// and FP16) around the round-down voxel loop: toward -inf inside it, so
// each crossing time c + delta is one v_add_f64, and round-to-nearest
// again after it, for the emulated set-up of the next ray.  The loop's only
// FP operation is that add (the rest is integer and LDS work).  The DDA
// state passes through the asm statement that sets the mode, so the
// round-to-nearest code producing it completes before the change and the
// loop's adds cannot start ahead of it; the loop exits only through the
// statement that restores the mode.  The s_nop covers the mode write's
// latency before the first dependent VALU instruction.
__device__ __forceinline__ void fp64_round_down_begin(Dda& s, double& L) {
  asm volatile("s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 2\n\ts_nop 3"
               : "+v"(s.nx), "+v"(s.ny), "+v"(s.nz), "+v"(s.dx), "+v"(s.dy),
                 "+v"(s.dz), "+v"(L));
}

__device__ __forceinline__ void fp64_round_nearest() {
  asm volatile("s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0\n\ts_nop 3" ::
                   : "memory");
}

struct AxisRun {
  double t;     // time of the pending leaf-entering crossing
  double d;     // time between crossings
  double over;  // first crossing time above L (DBL_MAX until produced)
  int n;        // crossings at or below L produced so far
};

// kHwDown: the wave's FP64 rounding mode is toward -inf (the round-down
// sum is the hardware add).
template <int M, bool kHwDown>
__device__ __forceinline__ void axis_steps(AxisRun& a, int k, double L) {
  for (int i = 0; i < k; ++i) {
    a.t = (M == kDown && !kHwDown) ? add_down_pos(a.t, a.d) : a.t + a.d;
    if (a.t <= L)
      ++a.n;
    else if (a.over == 1.7976931348623157e308)
      a.over = a.t;
  }
}

// The next leaf-entering crossing: eight crossings on.  The in-binade jump
// uses only exact operations, so it is the same in either rounding mode.
template <int M, bool kHwDown>
__device__ __forceinline__ void axis_leaf(AxisRun& a, double L) {
  const uint32_t e = (uint32_t)(bits_of(a.t) >> 52);  // t >= 0: exponent
  if (e > 52u && e < 0x7feu) {
    const double u = double_of((uint64_t)(e - 52u) << 52);
    const double top = double_of((uint64_t)(e + 1u) << 52);
    const double q = a.d * double_of((uint64_t)(2098u - e) << 52);  // d / u
    const double fq = floor_(q);
    // to nearest, a tie's step depends on the parity of t: no jump
    if (M == kDown || q - fq != 0.5) {
      const double D = (M == kDown ? fq : rint_(q)) * u;  // exact
      // every partial sum exact below top; the eighth step's sum below top
      const double t7 = a.t + 7.0 * D;
      if (D > 0.0 && t7 < top && top - t7 > a.d && t7 + D <= L) {
        a.t = t7 + D;
        a.n += 8;
        return;
      }
    }
  }
  axis_steps<M, kHwDown>(a, 8, L);
}

// First leaf-entering crossing of an axis from its first crossing time c1:
// crossing k moves the voxel from v0 + s(k - 1) to v0 + s k.
template <int M, bool kHwDown>
__device__ __forceinline__ AxisRun axis_start(double c1, double d, int s,
                                              int v0, double L) {
  AxisRun a;
  a.d = d;
  a.over = 1.7976931348623157e308;
  a.n = 0;
  a.t = c1;
  if (s == 0) return a;  // never crosses (time DBL_MAX)
  if (c1 <= L)
    a.n = 1;
  else
    a.over = c1;
  const int k0 = s > 0 ? 8 - (v0 & 7) : (v0 & 7) + 1;
  axis_steps<M, kHwDown>(a, k0 - 1, L);
  return a;
}

// Moves one axis of a voxel-DDA state to time `lo`: every crossing at or
// below lo is applied (the voxel moves one step per crossing) and the next
// crossing becomes the first one above lo.  The crossing times are the
// axis's own rounded sequence c(k+1) = c(k) + d, as the voxel loop produces
// them, eight at a time where a binade allows (axis_leaf); the round-down
// adds are emulated here, outside the loop's hardware-mode region.
template <int M>
__device__ __forceinline__ void advance_axis(double& next, double d, int sgn,
                                             int& v, double lo) {
  if (sgn == 0 || !(next <= lo)) return;  // no crossing at or below lo
  AxisRun a = axis_start<M, false>(next, d, sgn, v, lo);
  while (a.over == 1.7976931348623157e308) axis_leaf<M, false>(a, lo);
  next = a.over;
  v += sgn * a.n;
}

// The state the voxel loop has once it has taken every step at or below
// time lo: the three axes' crossing sequences are independent, so each is
// advanced on its own.
template <int M>
__device__ __forceinline__ void dda_advance(Dda& s, double lo) {
  advance_axis<M>(s.nx, s.dx, s.sx, s.vx, lo);
  advance_axis<M>(s.ny, s.dy, s.sy, s.vy, lo);
  advance_axis<M>(s.nz, s.dz, s.sz, s.vz, lo);
}

// Largest value over the workgroup (256 threads).
__device__ __forceinline__ uint32_t block_max(uint32_t v) {
  __shared__ uint32_t s_max[4];
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = v;
  __syncthreads();
  return max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
}

// Sort key of an item's cost (its longest ray, in voxel visits): longer
// rays give smaller keys, so an ascending sort puts them first.
constexpr int kCostBits = 16;
__device__ __forceinline__ uint32_t cost_key(uint32_t visits) {
  const uint32_t top = (1u << kCostBits) - 1;
  return top - (visits < top ? visits : top);
}

// cost_keys/cost_vals (queue 3 only, else null): the item's cost key and
// index, the input of the item-order sort.
template <int M>
__global__ __launch_bounds__(kCastThreads) void k_discover(
    const Job* __restrict__ jobs, ItemQueue q, int32_t* __restrict__ leaf_slot,
    uint32_t* __restrict__ ray_hit, uint64_t* __restrict__ part,
    uint32_t* __restrict__ cost_keys, uint32_t* __restrict__ cost_vals,
    uint32_t* __restrict__ item_longest) {
  int round = 0;
  for (;;) {
    const int item = take_item(q, round);
    if (item < 0) break;
    const Job J = jobs[job_of_item(q, item)];
    int32_t* __restrict__ ls = leaf_slot + J.cube_off;
    const int i = point_of(J, item - J.item_off, threadIdx.x);
    uint32_t hit_cell = kInvalid;
    uint32_t visits = 0;
    int valid = 0, errors = 0;
    RayDir r;
    if (i >= 0 && ray_of<M>(J, i, r)) {
      valid = 1;
      Dda s;
      dda_init<M>(J.eye, r, s);
      double L = r.L;
      constexpr bool kHw = M == kDown;
      if (kHw) fp64_round_down_begin(s, L);
      AxisRun ax = axis_start<M, kHw>(s.nx, s.dx, s.sx, s.vx, L);
      AxisRun ay = axis_start<M, kHw>(s.ny, s.dy, s.sy, s.vy, L);
      AxisRun az = axis_start<M, kHw>(s.nz, s.dz, s.sz, s.vz, L);
      // a voxel of the current leaf per axis (only v >> 3 is used)
      int lx = s.vx, ly = s.vy, lz = s.vz;
      int c = cube_index(J.cube, lx, ly, lz);
      if (c < 0)
        ++errors;
      else
        ls[c] = -2;
      for (;;) {
        const double t = fmin_(ax.t, fmin_(ay.t, az.t));
        if (!(t <= L)) break;
        // the axis is chosen by selects, and one axis_leaf serves all lanes
        const bool bz = az.t == t, by = !bz && ay.t == t, bx = !bz && !by;
        AxisRun a;
        a.t = t;
        a.d = bz ? az.d : (by ? ay.d : ax.d);
        a.over = bz ? az.over : (by ? ay.over : ax.over);
        a.n = bz ? az.n : (by ? ay.n : ax.n);
        axis_leaf<M, kHw>(a, L);
        if (bx) ax = a;
        if (by) ay = a;
        if (bz) az = a;
        lx += bx ? 8 * s.sx : 0;
        ly += by ? 8 * s.sy : 0;
        lz += bz ? 8 * s.sz : 0;
        c = cube_index(J.cube, lx, ly, lz);
        if (c < 0)
          ++errors;
        else
          ls[c] = -2;
      }
      if (kHw) fp64_round_nearest();
      visits = (uint32_t)(ax.n + ay.n + az.n);
      if (fmin_(ax.over, fmin_(ay.over, az.over)) < (double)J.max_time) {
        ++visits;
        const int ex = s.vx + s.sx * ax.n, ey = s.vy + s.sy * ay.n,
                  ez = s.vz + s.sz * az.n;
        const int ce = cube_index(J.cube, ex, ey, ez);
        if (ce >= 0)
          hit_cell = (uint32_t)ce * 512u + (uint32_t)local_of(ex, ey, ez);
      }
    }
    if (i >= 0) ray_hit[(size_t)J.ray_off + i] = hit_cell;
    const uint64_t v[4] = {visits, (uint64_t)valid,
                           (uint64_t)(hit_cell != kInvalid), (uint64_t)errors};
    block_partials(part + (size_t)item * 4, v, 4);
    if (cost_keys || item_longest) {
      const uint32_t longest = block_max(visits);
      if (threadIdx.x == 0) {
        if (cost_keys) {
          cost_keys[item] = cost_key(longest);
          cost_vals[item] = (uint32_t)item;
        }
        if (item_longest) item_longest[item] = longest;
      }
    }
  }
}

// ---- leaf compaction: exclusive scan over the cube tables --------------
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

// One workgroup: exclusive scan of block_counts; total -> n_leaves.  Also
// reduces the discover partials of all items (visits, rays, hit rays,
// errors) into the sequence counters.
// Sliced accumulate: the visits per time slice V are picked per frame
// from this ladder -- the first whose extra slice items fit in the extra
// work items the accumulate launch has -- so a frame whose ray blocks
// already fill the GPU gets no extra items and is not sliced.  A ray block
// whose longest ray has v visits is walked in min(nmax, max(1, v / V))
// time slices.
constexpr int kSliceLadder = 12;
__device__ __forceinline__ constexpr uint32_t slice_ladder(int q) {
  constexpr uint32_t kV[kSliceLadder] = {40,  48,  56,  64,  80,  96,
                                         128, 160, 200, 256, 320, 400};
  return kV[q];
}

__device__ __forceinline__ uint32_t block_slices(uint32_t longest,
                                                 uint32_t per, int nmax) {
  return min((uint32_t)nmax, max(1u, longest / per));
}

// Slice plan, run by one 1024-thread workgroup after the discover pass:
// picks V and lists the extra work items (every slice but the first of
// each ray block, as block << 3 | slice, in block order).
__device__ void plan_slices(const uint32_t* __restrict__ longest, int nblk,
                            int nmax, int extra,
                            FrameCounters* __restrict__ cnt,
                            uint32_t* __restrict__ items) {
  __shared__ uint32_t s_ex[kSliceLadder][16];
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_pick;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int chunk = (nblk + 1023) / 1024;
  const int b0 = min(tid * chunk, nblk), b1 = min(b0 + chunk, nblk);
  uint32_t ex[kSliceLadder];
#pragma unroll
  for (int q = 0; q < kSliceLadder; ++q) ex[q] = 0;
  for (int b = b0; b < b1; ++b) {
    const uint32_t v = longest[b];
#pragma unroll
    for (int q = 0; q < kSliceLadder; ++q)
      ex[q] += block_slices(v, slice_ladder(q), nmax) - 1u;
  }
#pragma unroll
  for (int q = 0; q < kSliceLadder; ++q) {
    const uint32_t w = wave_incl_scan32(ex[q]);
    if (lane == 63) s_ex[q][wave] = w;
  }
  __syncthreads();
  if (wave == 0) {
    uint32_t tot = 0;
    if (lane < kSliceLadder)
      for (int w = 0; w < 16; ++w) tot += s_ex[lane][w];
    const uint64_t fits =
        __ballot(lane < kSliceLadder && tot <= (uint32_t)extra);
    if (lane == 0) {
      uint32_t pick = 0;  // no ladder entry fits: no slicing
      if (fits) {
        const int q = __builtin_ctzll(fits);
#pragma unroll
        for (int k = 0; k < kSliceLadder; ++k)
          if (k == q) pick = slice_ladder(k);
      }
      s_pick = pick;
      cnt->slice_visits = pick;
    }
  }
  __syncthreads();
  const uint32_t per = s_pick;
  if (per == 0u) return;  // slice_extra stays 0
  uint32_t mine = 0;
  for (int b = b0; b < b1; ++b)
    mine += block_slices(longest[b], per, nmax) - 1u;
  const uint32_t x = wave_incl_scan32(mine);
  if (lane == 63) s_w[wave] = x;
  __syncthreads();
  uint32_t pos = x - mine;
  for (int w = 0; w < wave; ++w) pos += s_w[w];
  for (int b = b0; b < b1; ++b) {
    const uint32_t n = block_slices(longest[b], per, nmax);
    for (uint32_t sl = 1; sl < n; ++sl) items[pos++] = ((uint32_t)b << 3) | sl;
  }
  if (tid == 1023) cnt->slice_extra = (int32_t)(pos);
}

__global__ __launch_bounds__(1024) void k_scan_blocks(
    int32_t* __restrict__ block_counts, int nblocks,
    const uint64_t* __restrict__ part, int npart,
    FrameCounters* __restrict__ cnt, const uint32_t* __restrict__ longest,
    int nmax, int extra, uint32_t* __restrict__ slice_items) {
  __shared__ int wsum[16];
  __shared__ unsigned long long psum[4];
  const int tid = threadIdx.x;
  if (tid < 4) psum[tid] = 0;
  __syncthreads();
  {
    uint64_t acc[4] = {0, 0, 0, 0};
    for (int b = tid; b < npart; b += 1024)
      for (int q = 0; q < 4; ++q) acc[q] += part[(size_t)b * 4 + q];
    for (int q = 0; q < 4; ++q) {
      const uint64_t w = wave_sum(acc[q]);
      if ((tid & 63) == 0 && w) atomicAdd(&psum[q], (unsigned long long)w);
    }
  }
  __syncthreads();
  if (tid == 0) {
    cnt->visits += psum[0];
    cnt->rays += psum[1];
    cnt->n_hit_rays += (int32_t)psum[2];
    cnt->errors += (int32_t)psum[3];
  }
  const int per = (nblocks + 1023) / 1024;
  const int b0 = tid * per, b1 = min(nblocks, b0 + per);
  int local = 0;
  for (int b = b0; b < b1; ++b) local += block_counts[b];
  const int lane = tid & 63, wave = tid >> 6;
  const int x = (int)wave_incl_scan32((uint32_t)local);
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
  if (longest) plan_slices(longest, npart, nmax, extra, cnt, slice_items);
}

__global__ __launch_bounds__(256) void k_scan_assign(
    int32_t* __restrict__ leaf_slot, int64_t vol,
    const int32_t* __restrict__ block_counts,
    const int32_t* __restrict__ scan_job, int32_t* __restrict__ slot_leaf,
    int32_t* __restrict__ slot_job, int tile_cap,
    FrameCounters* __restrict__ cnt) {
  __shared__ int ws[4];
  const int64_t base = (int64_t)blockIdx.x * kScanChunk;
  int run = block_counts[blockIdx.x];
  const int job = scan_job[blockIdx.x];
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
        slot_job[pre] = job;
      } else {
        leaf_slot[c] = -1;  // dropped: the sequence reports overflow
        atomicAdd(&cnt->overflow, 1);
      }
    }
    run += ws[0] + ws[1] + ws[2] + ws[3];
    __syncthreads();
  }
}

// Sort keys of the hit events: tile cell (slot * 512 + local), value =
// global point index.  Rays without a hit get the largest key, sort last.
__global__ __launch_bounds__(256) void k_hit_keys(
    const Job* __restrict__ jobs, ItemQueue q,
    const uint32_t* __restrict__ ray_hit,
    const int32_t* __restrict__ leaf_slot, uint32_t* __restrict__ keys,
    uint32_t* __restrict__ vals) {
  const int item = blockIdx.x;
  if (item >= q.nitems) return;
  const Job J = jobs[job_of_item(q, item)];
  const int i = point_of(J, item - J.item_off, threadIdx.x);
  if (i < 0) return;
  const size_t g = (size_t)J.ray_off + i;
  const uint32_t h = ray_hit[g];
  uint32_t key = kInvalid;
  if (h != kInvalid) {
    const int slot = leaf_slot[J.cube_off + (h >> 9)];
    if (slot >= 0) key = (uint32_t)slot * 512u + (h & 511u);
  }
  keys[g] = key;
  vals[g] = (uint32_t)g;
}

__global__ __launch_bounds__(256) void k_hit_segments(
    const uint32_t* __restrict__ keys, const FrameCounters* __restrict__ cnt,
    uint32_t* __restrict__ tiles, uint32_t* __restrict__ hit_mask,
    uint32_t* __restrict__ hit_start, uint32_t* __restrict__ gap,
    const int32_t* __restrict__ slot_leaf, int32_t* __restrict__ leaf_slot) {
  const int nh = cnt->n_hit_keys;
  for (int e = blockIdx.x * 256 + threadIdx.x; e < nh;
       e += gridDim.x * 256) {
    gap[e] = 0;
    const uint32_t key = keys[e];
    if (key == kInvalid) continue;
    // the last hit of a voxel stores the end of its run, the first one the
    // start (the hit count is end - start)
    if (e + 1 == nh || keys[e + 1] != key)
      tiles[2 * (size_t)key + 1] = (uint32_t)(e + 1);
    const uint32_t prev = e > 0 ? keys[e - 1] : kInvalid;
    if (prev == key) continue;
    hit_start[key] = (uint32_t)e;
    const uint32_t slot = key >> 9, li = key & 511;
    atomicOr(&hit_mask[slot * kMaskWords + (li >> 5)], 1u << (li & 31));
    // the leaf's first hit voxel flags the leaf in the cube table (sole
    // writer: keys are sorted leaf-major), so the accumulate pass reads
    // its hit mask only inside leaves that hold hits
    if (e == 0 || (prev >> 9) != slot)
      leaf_slot[slot_leaf[slot]] = (int32_t)slot | kSlotHasHits;
  }
}

// LDS aggregation of miss counts: an open-addressing table of (key,
// count) per workgroup.  Counts are 16-bit halves of 32-bit words (slot h
// in half h & 1 of word h >> 1): a key gains at most one per ray of the
// 256-ray item, and the smaller table fits one more workgroup per CU.  A key is a voxel cell (slot * 512 + local index,
// below 2^31), whose count goes to its miss counter, or kGapKey | i, whose
// count goes to gap[i] (misses of a hit voxel before its i-th sorted hit).  A 16x16 pixel tile touches a few thousand distinct
// voxels for tens of thousands of visits, so the table turns almost every
// visit into an LDS atomic and leaves one global atomic per distinct voxel
// per work item, at flush.  A visit that finds no free slot within kProbe
// probes falls back to a global atomic.  Two unrolled probes measured
// fastest: longer probe loops diverge across the wavefront and cost more
// than the rare global fallback.  For one frame 4096 slots beat 2048 (fewer
// collisions) and 8192 (occupancy); a batch, with many more work items in
// flight, can prefer 2048 slots for the extra resident workgroups -- the
// table size is a template parameter (S = log2 slots).
constexpr int kProbe = 2;
constexpr uint32_t kGapKey = 0x80000000u;

// The global counter of a table key.
__device__ __forceinline__ uint32_t* key_counter(uint32_t* tiles,
                                                 uint32_t* gap, uint32_t key) {
  return (key & kGapKey) ? gap + (key & ~kGapKey) : tiles + 2 * (size_t)key;
}

template <int S>
__device__ __forceinline__ void lds_count(uint32_t* s_key, uint32_t* s_cnt,
                                          uint32_t* tiles, uint32_t* gap,
                                          uint32_t cell, uint32_t inc) {
  // multiplicative hash of the key's low 24 bits (a full-rate multiply):
  // two of the few thousand cells of one work item rarely agree in those
  // bits, and when they do they only share a probe sequence (the table
  // compares whole keys)
  uint32_t h = (uint32_t)__umul24(cell, 0x9e3779u) >> (32 - S);
  int found = -1;
#pragma unroll
  for (int q = 0; q < kProbe; ++q) {
    if (found < 0) {
      const uint32_t k = s_key[h];
      if (k == cell) {
        found = (int)h;
      } else if (k == kInvalid) {
        const uint32_t old = atomicCAS(&s_key[h], kInvalid, cell);
        if (old == kInvalid || old == cell) found = (int)h;
      }
      h = (h + 1) & ((1u << S) - 1);
    }
  }
  // The two adds differ in scope on purpose: the LDS one is a workgroup
  // atomic (ds_add), the fallback a device one; with equal scopes the
  // compiler merges them into one flat atomic, which also waits on the
  // global memory counter.
  if (found >= 0)
    __hip_atomic_fetch_add(&s_cnt[found >> 1], inc << ((found & 1) * 16),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else
    atomicAdd(key_counter(tiles, gap, cell), inc);
}

// lds_count for the active lanes.  Neighbouring rays at the same step
// mostly stand in the same voxel (an 8x8 pixel wavefront covers a few
// voxels at any depth), so lanes are merged before the table is touched:
// when they all name the same voxel (the voxels next to the eye) one lane
// adds the lane count; otherwise each run of equal cells over consecutive
// active lanes of a 16-lane row (two 8-pixel rows of an 8x8 part, one of a
// 16x4 part; lane - 1 read with a DPP shift) is added once, by its first
// lane, with the run's length.  prev_bit is 1 << (lane - 1), or 0 for the
// first lane of a 16-lane row, which has no left neighbour.
template <int S>
__device__ __forceinline__ void wave_lds_count(uint32_t* s_key,
                                               uint32_t* s_cnt,
                                               uint32_t* tiles,
                                               uint32_t* gap,
                                               uint32_t cell,
                                               uint64_t prev_bit) {
  const uint64_t act = __ballot(1);
  const int lane = threadIdx.x & 63;
  const uint32_t first = __builtin_amdgcn_readfirstlane(cell);
  const uint64_t same = __builtin_amdgcn_ballot_w64(cell == first);
  if (same == act) {
    if (lane == (int)__builtin_ctzll(act))
      lds_count<S>(s_key, s_cnt, tiles, gap, cell, (uint32_t)__popcll(act));
    return;
  }
  // row_shr:1 -- lane l reads lane l - 1 of its 16-lane row
  const uint32_t left = (uint32_t)__builtin_amdgcn_update_dpp(
      (int)kInvalid, (int)cell, 0x111, 0xf, 0xf, false);
  const bool head = (act & prev_bit) == 0ull || left != cell;
  const uint64_t stop =
      __builtin_amdgcn_ballot_w64(head) | ~act | 0x0001000100010001ull;
  if (head) {
    // a run ends at the next stop bit above the lane, which is at most 16
    // lanes up (every row start is a stop), so 32 bits of the shifted mask
    // suffice; past lane 63 there is none, which min(., 63 - lane) covers
    // (lane 63's shift by 0 included)
    const uint32_t after = (uint32_t)(stop >> ((lane + 1) & 63));
    const uint32_t len =
        min((uint32_t)__builtin_ctzg(after, 32), (uint32_t)(63 - lane)) + 1u;
    lds_count<S>(s_key, s_cnt, tiles, gap, cell, len);
  }
}

// Cube-table entry of the leaf holding (x,y,z): slot | kSlotHasHits, or -1.
__device__ __forceinline__ int slot_of(const Cube& cube,
                                       const int32_t* leaf_slot, int x, int y,
                                       int z) {
  const int c = cube_index(cube, x, y, z);
  return c >= 0 ? leaf_slot[c] : -1;
}

template <int M, int S>
__global__ __launch_bounds__(kCastThreads) void k_accumulate(
    const Job* __restrict__ jobs, ItemQueue q,
    const int32_t* __restrict__ leaf_slot, uint32_t* __restrict__ tiles,
    const uint32_t* __restrict__ hit_mask,
    const uint32_t* __restrict__ hit_start,
    const uint32_t* __restrict__ svals, uint32_t* __restrict__ gap,
    FrameCounters* __restrict__ cnt,
    const uint32_t* __restrict__ item_longest,
    const uint32_t* __restrict__ slice_items) {
  constexpr int kSlots = 1 << S;
  __shared__ uint32_t s_key[kSlots];
  __shared__ uint32_t s_cnt[kSlots / 2];
  for (int k = threadIdx.x; k < kSlots; k += kCastThreads) {
    s_key[k] = kInvalid;
    if (k < kSlots / 2) s_cnt[k] = 0;
  }
  int round = 0;
  int errors = 0;
  for (;;) {
    const int item = take_item(q, round);
    if (item < 0) break;
    __syncthreads();  // the table is clear
    const Job J = jobs[job_of_item(q, item)];
    const int32_t* __restrict__ ls = leaf_slot + J.cube_off;
    // a sliced job (single-job launches) walks a ray block in nsl time
    // slices: slice k of a ray takes the crossings in (T_k, T_k+1],
    // T_k = L k / nsl (T_0 below every crossing, T_nsl = L), so the slices
    // of a ray count exactly the misses of its whole walk.  Items
    // [0, nitems) are the blocks' first slices; the extra items that
    // k_slice_plan lists are their other slices.
    const int nmax = J.slices > 1 ? J.slices : 1;
    int blk = item - J.item_off, slice = 0;
    if (blk >= J.nitems) {
      const int k = blk - J.nitems;
      if (k >= cnt->slice_extra) continue;
      const uint32_t e = slice_items[k];
      blk = (int)(e >> 3);
      slice = (int)(e & 7u);
    }
    int nsl = 1;
    if (nmax > 1) {
      const uint32_t per = cnt->slice_visits;
      if (per)
        nsl = (int)min((uint32_t)nmax,
                       max(1u, item_longest[J.item_off + blk] / per));
    }
    const int i = point_of(J, blk, threadIdx.x);
    const uint32_t gi = (uint32_t)(J.ray_off + i);
    RayDir r;
    if (i >= 0 && ray_of<M>(J, i, r)) {
      Dda s;
      dda_init<M>(J.eye, r, s);
      double L = r.L;
      if (nsl > 1) {
        if (slice > 0) dda_advance<M>(s, r.L * slice / nsl);
        if (slice < nsl - 1) L = r.L * (slice + 1) / nsl;
      }
      // Walk state beside the DDA, kept so that a step that stays in its
      // leaf costs two adds:
      //   cell  slot * 512 + local index of the current voxel; a step along
      //         an axis adds ci (+-64, +-8, +-1) unless it leaves the leaf;
      //   U     per axis (4-bit fields x, y, z), the voxel's place in its
      //         leaf counted along the axis's step direction (0..7): a step
      //         adds 1 to the field, and a field reaching 8 has left the
      //         leaf;
      //   c     the cube-table index of the current leaf, moved by +-side^2,
      //         +-side or +-1 as the walk enters the next leaf.
      // A ray stays inside its frame's cube (the cube reaches a leaf and
      // two voxels beyond max_time > L around the eye, and the walk stops
      // at L), so c always names the leaf that holds the voxel; a c outside
      // the table reads no entry (and counts errors, as an absent leaf
      // does).
      const int side = J.cube.side;
      const int lx = (s.vx >> 3) - J.cube.cmin[0];
      const int ly = (s.vy >> 3) - J.cube.cmin[1];
      const int lz = (s.vz >> 3) - J.cube.cmin[2];
      const uint32_t vol = (uint32_t)(side * side * side);
      int c = (lx * side + ly) * side + lz;
      if ((unsigned)lx >= (unsigned)side || (unsigned)ly >= (unsigned)side ||
          (unsigned)lz >= (unsigned)side)
        c = -1;
      // place of v among 0..7 counted along step direction sgn (0 for an
      // axis that never steps)
      const auto place = [](int sgn, int v) {
        return (uint32_t)(sgn > 0 ? (v & 7) : (sgn < 0 ? 7 - (v & 7) : 0));
      };
      uint32_t U = (place(s.sx, s.vx) << 8) | (place(s.sy, s.vy) << 4) |
                   place(s.sz, s.vz);
      const int cix = s.sx * 64, ciy = s.sy * 8, ciz = s.sz;
      const int csx = s.sx * side * side, csy = s.sy * side, csz = s.sz;
      // entries are read at a 32-bit byte offset (vol < 2^30, so 4 c fits)
      const char* __restrict__ lsb = reinterpret_cast<const char*>(ls);
      const auto entry_at = [lsb](uint32_t k) {
        return *reinterpret_cast<const int32_t*>(lsb + k * 4u);
      };
      int entry = -1;
      if ((uint32_t)c < vol) entry = entry_at((uint32_t)c);
      uint32_t cell = ((uint32_t)(entry & ~kSlotHasHits) << 9) |
                      (uint32_t)local_of(s.vx, s.vy, s.vz);
      const int lane = threadIdx.x & 63;
      const uint64_t prev_bit = (lane & 15) ? 1ull << (lane - 1) : 0ull;
      if (M == kDown) fp64_round_down_begin(s, L);
      // The loop is software-pipelined around its two memory reads: the
      // hit-mask word of the current voxel and the next leaf's entry are
      // requested before the table work for the current voxel and used
      // after it, so their latency overlaps that work.
      // the DDA step's axis (dda_step, nested-select form): x only if
      // strictly smallest, else y if y < z, else z; t its crossing time
      bool bx, by;
      double t;
      {
        const double nx = s.nx, ny = s.ny, nz = s.nz;
        bx = (nx < ny) && (nx < nz);
        by = !bx && (ny < nz);
        t = bx ? nx : (by ? ny : nz);
      }
      while (t <= L) {
        // a miss on the current voxel (its exit time t <= L)
        // in a leaf holding hits, also the voxel's run of sorted hits
        // [s0, end), which the gap count below needs if the voxel has hits
        uint32_t mw = 0, s0 = 0, end = 0;
        if (entry >= 0 && (entry & kSlotHasHits)) {
          mw = hit_mask[cell >> 5];  // kMaskWords words per slot
          s0 = hit_start[cell];
          end = tiles[2 * (size_t)cell + 1];
        }
        // the step: next crossing time of the axis (round-down: the wave's
        // FP64 mode is toward -inf here, so this is the rounded-down sum)
        const double dt = bx ? s.dx : (by ? s.dy : s.dz);
        const double nt = t + dt;
        s.nx = bx ? nt : s.nx;
        s.ny = by ? nt : s.ny;
        s.nz = (!bx && !by) ? nt : s.nz;
        // the voxel it moves to; leaving the leaf, the axis's place starts
        // again at 0 (the one field at 8 drops to 0) and the next leaf's
        // entry is read
        const int ci = bx ? cix : (by ? ciy : ciz);
        const uint32_t ui = bx ? 0x100u : (by ? 0x10u : 0x1u);
        U += ui;
        const bool cross = (U & 0x888u) != 0u;
        int next_entry = entry;
        if (cross) {
          U &= ~0x888u;
          c += bx ? csx : (by ? csy : csz);
          next_entry = -1;
          if ((uint32_t)c < vol) next_entry = entry_at((uint32_t)c);
        }
        if (entry < 0) {
          ++errors;
        } else {
          wave_lds_count<S>(s_key, s_cnt, tiles, gap, cell, prev_bit);
          if ((mw >> (cell & 31)) & 1u) {
            // a voxel holding hits also counts the miss in the gap before
            // its first later hit (the voxel's hits are sorted[s0, end) in
            // ascending point order); neighbouring rays mostly share the
            // gap, so these counts are merged in the table as well
            uint32_t lo = s0, hi = end;
            while (lo < hi) {
              const uint32_t mid = (lo + hi) >> 1;
              if (svals[mid] < gi)
                lo = mid + 1;
              else
                hi = mid;
            }
            if (lo < end)
              wave_lds_count<S>(s_key, s_cnt, tiles, gap, kGapKey | lo,
                                prev_bit);
          }
        }
        // the local index wraps to the far side of the next leaf (li + ci -
        // 8 ci) when the step leaves the leaf
        if (cross)
          cell = ((uint32_t)(next_entry & ~kSlotHasHits) << 9) |
                 ((cell & 511u) - 7u * (uint32_t)ci);
        else
          cell += (uint32_t)ci;
        entry = next_entry;
        {
          const double nx = s.nx, ny = s.ny, nz = s.nz;
          bx = (nx < ny) && (nx < nz);
          by = !bx && (ny < nz);
          t = bx ? nx : (by ? ny : nz);
        }
      }
      if (M == kDown) fp64_round_nearest();
    }
    __syncthreads();
    // flush the item's counts and leave the table empty for the next one
    for (int w = threadIdx.x; w < kSlots / 2; w += kCastThreads) {
      const uint32_t c2 = s_cnt[w];
      if (c2 != 0u) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const uint32_t key = s_key[2 * w + u];
          if (key != kInvalid) {
            atomicAdd(key_counter(tiles, gap, key), (c2 >> (16 * u)) & 0xffffu);
            s_key[2 * w + u] = kInvalid;
          }
        }
        s_cnt[w] = 0;
      } else {
        s_key[2 * w] = kInvalid;
        s_key[2 * w + 1] = kInvalid;
      }
    }
  }
  const uint64_t se = wave_sum((uint64_t)errors);
  if ((threadIdx.x & 63) == 0 && se) atomicAdd(&cnt->errors, (int)se);
}

// ---- map hash (pack_leaf, mix64: nm_internal.h) --------------------------
// Leaf origin (index space) of entry c of job J's cube table.
__device__ __forceinline__ void slot_origin(const Job& J, int c, int& ox,
                                            int& oy, int& oz) {
  const int side = J.cube.side;
  const int lz = c % side;
  const int ly = (c / side) % side;
  const int lx = c / (side * side);
  ox = (lx + J.cube.cmin[0]) * 8;
  oy = (ly + J.cube.cmin[1]) * 8;
  oz = (lz + J.cube.cmin[2]) * 8;
}

// Map leaf of every discovered tile, one lane per tile so the probe
// latencies overlap: map_idx = leaf index, or -1 if the leaf is new to the
// map.  A new leaf gets a pool index reserved here, in slot order (one
// atomic per wavefront and map), and its hash insertion point; the update
// inserts it only if one of its voxels changes (Handler.cpp:116-118
// touches the map only for nonzero temp values), otherwise the reserved
// index is a hole.
__global__ __launch_bounds__(256) void k_map_lookup(
    const Job* __restrict__ jobs, const int32_t* __restrict__ slot_job,
    const int32_t* __restrict__ slot_leaf, const MapView* __restrict__ maps,
    MapView single, int tile_cap, FrameCounters* __restrict__ cnt,
    int32_t* __restrict__ map_idx, int32_t* __restrict__ map_new) {
  const int nleaves = min(cnt->n_leaves, tile_cap);
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (int base = blockIdx.x * 256; base < nleaves; base += gridDim.x * 256) {
    const int slot = base + threadIdx.x;
    int idx = -1;
    bool unknown = false;
    int mapi = -1;
    MapView m = single;
    if (slot < nleaves) {
      const Job& J = jobs[slot_job[slot]];
      mapi = J.map;
      if (maps) m = maps[mapi];
      int ox, oy, oz;
      slot_origin(J, slot_leaf[slot] - (int)J.cube_off, ox, oy, oz);
      const unsigned long long key = pack_leaf(ox, oy, oz);
      const unsigned long long mask = (unsigned long long)m.hash_cap - 1;
      unsigned long long h = mix64(key) & mask;
      unknown = true;
      for (int64_t probe = 0; probe < m.hash_cap; ++probe) {
        const unsigned long long k = m.hash_keys[h];
        if (k == key) {
          idx = m.hash_vals[h];
          unknown = false;
          break;
        }
        if (k == ~0ull) break;
        h = (h + 1) & mask;
      }
    }
    // reserve pool indices for the unknown leaves, one atomic per map
    // present in the wavefront
    uint64_t want = __ballot(unknown);
    int first = 0;
    uint64_t mine = 0;
    while (want) {
      const int leader = (int)__builtin_ctzll(want);
      const int lm = __shfl(mapi, leader, 64);
      const uint64_t peers = want & __ballot(mapi == lm);
      int f = 0;
      if (lane == leader) f = atomicAdd(m.leaf_count, (int)__popcll(peers));
      f = __shfl(f, leader, 64);
      if (mapi == lm) {
        first = f;
        mine = peers;
      }
      want &= ~peers;
    }
    if (slot < nleaves) {
      map_idx[slot] = idx;
      int reserved = -1;
      if (unknown) {
        reserved = first + (int)__popcll(mine & lt);
        if (reserved >= m.leaf_cap) {
          atomicAdd(&cnt->overflow, 1);
          reserved = -1;
        }
      }
      map_new[slot] = reserved;
    }
  }
}

// Insert leaf `idx` (reserved by k_map_lookup) under the key of (x,y,z);
// false if the table is full.
__device__ bool map_insert(const MapView& m, int x, int y, int z, int idx) {
  const unsigned long long key = pack_leaf(x, y, z);
  const unsigned long long mask = (unsigned long long)m.hash_cap - 1;
  unsigned long long h = mix64(key) & mask;
  for (int64_t probe = 0; probe < m.hash_cap; ++probe) {
    const unsigned long long k = m.hash_keys[h];
    if (k == ~0ull) {
      const unsigned long long old = atomicCAS(&m.hash_keys[h], ~0ull, key);
      if (old == ~0ull) {
        m.hash_vals[h] = idx;
        m.origins[3 * idx + 0] = x;
        m.origins[3 * idx + 1] = y;
        m.origins[3 * idx + 2] = z;
        return true;
      }
    }
    h = (h + 1) & mask;
  }
  return false;
}

template <int M>
__device__ __noinline__ float fold_call(float v, float a, uint32_t k) {
  return fold<M, true>(v, a, k);
}

// fold<M> with its common short case (a few steps) inline and the rest out
// of line, for call sites that are unrolled.
template <int M>
__device__ __forceinline__ float fold_small(float v, float a, uint32_t k) {
  if (k <= 2) {
    if (k > 0) v = add<M, true>(v, a);
    if (k > 1) v = add<M, true>(v, a);
    return v;
  }
  return fold_call<M>(v, a, k);
}

// Pixel-order sum of one voxel's hit sequence, computed by a whole
// wavefront: gap[g] misses precede hit g (g < j); *used = total of the
// gaps.  Lane i takes event i of a chunk of 64 (its gap run, then its
// hit).  While the running sum v stays in one binade [2^e, 2^(e+1)) -- in
// magnitude, same sign -- every value is a multiple of the ulp u and each
// rounded step adds a fixed multiple of u: round(h / u) for a hit and
// round(m / u) for a miss (rounded toward -inf, or to nearest when no tie
// arises).  One exact integer prefix sum over the lanes then gives v after
// every event, and the events are taken up to the first one whose run
// leaves the binade's interior; that event is summed one step at a time
// (fold), and the jump is tried again from there (near zero, where a
// binade holds few events, events are summed one at a time).  Near zero, a
// chunk with few nonzero gaps is summed as runs of hits instead: one fold
// per run.  The result equals the sequential sum bit for bit.

// Increment of one event in units of the ulp (q = a / u exactly), or
// false when it cannot be taken as a fixed step: larger than the binade,
// or a round-to-nearest tie (whose result depends on the parity of v).
template <int M>
__device__ __forceinline__ bool unit_step(double q, long long* A) {
  if (!(abs_(q) < 16777216.0)) return false;
  if (M == kDown) {
    *A = (long long)floor_(q);
  } else {
    if (q - floor_(q) == 0.5) return false;
    *A = (long long)rint_(q);
  }
  return true;
}

// T (in ulps) lies strictly inside the binade: every value between T and
// a neighbour inside it rounds on the same grid.
__device__ __forceinline__ bool in_binade(long long T, bool neg) {
  const long long m = neg ? -T : T;
  return m >= (1ll << 23) + 1 && m <= (1ll << 24) - 1;
}

// Near zero a voxel's running sum is stepped in fixed point.  With U the
// smaller ulp of the two increments, every sum the sequence reaches is a
// multiple of U: the increments are, a rounded sum is a multiple of its own
// ulp (at least U once it is as large as the increments), and smaller sums
// are exact.  So v = V U for an integer V, and one rounded step is an
// integer add and a rounding of the sum X to the grid 2^(lead(|X|) - 23) of
// its binade: a mask toward -inf, or ties-to-even to nearest.  An exact
// cancellation gives -0 rounding toward -inf and +0 to nearest.
struct FixedSteps {
  int uexp;   // U = 2^uexp
  int Ah;     // hit increment in units of U
  int Am;     // miss increment in units of U
  int bound;  // fixed point while |V| < bound (|v| < small)
  bool ok;    // the increments and the bound fit the integer range
};

__device__ __forceinline__ FixedSteps fixed_steps(float hit, float miss,
                                                  float small) {
  FixedSteps f{};
  f.ok = false;
  const uint32_t eh = (bits_of(hit) >> 23) & 0xffu;
  const uint32_t em = (bits_of(miss) >> 23) & 0xffu;
  if (eh == 0u || em == 0u || eh == 0xffu || em == 0xffu) return f;
  const int e = (int)(eh < em ? eh : em) - 150;
  if (e < -126) return f;
  const double inv = double_of((uint64_t)(1023 - e) << 52);  // 2^-e
  const double ah = (double)hit * inv, am = (double)miss * inv;
  const double b = (double)small * inv;  // all three exact
  if (!(abs_(ah) < 134217728.0 && abs_(am) < 134217728.0 && b < 536870912.0))
    return f;
  f.uexp = e;
  f.Ah = (int)ah;
  f.Am = (int)am;
  f.bound = (int)-floor_(-b);  // ceil
  f.ok = true;
  return f;
}

// One rounded fixed-point step: X = V + A rounded to its binade's grid;
// *z takes the signed zero of an exact cancellation.
template <int M>
__device__ __forceinline__ int fixed_step(int V, int A, float* z) {
  const int X = V + A;
  if (X == 0) {
    *z = M == kDown ? -0.0f : 0.0f;
    return 0;
  }
  const uint32_t m = (uint32_t)(X < 0 ? -X : X);
  const int k = 8 - (int)__builtin_clz(m);  // lead(m) - 23
  if (k <= 0) return X;
  const int low = (1 << k) - 1;
  if (M == kDown) return X & ~low;
  return (X + (low >> 1) + ((X >> k) & 1)) & ~low;
}

// Events i0.. of a chunk stepped in fixed point while |v| < small (v must
// lie on the U grid, else nothing is done).  The event whose misses leave
// the region is finished in floating point.  Returns the next event.
template <int M>
__device__ uint32_t fixed_run(const FixedSteps& F, uint32_t g, uint32_t i0,
                              uint32_t n, float hit, float miss, float& v,
                              uint32_t& used) {
  const double inv_u = double_of((uint64_t)(1023 - F.uexp) << 52);
  const double Vd = (double)v * inv_u;  // exact
  if (Vd != floor_(Vd)) return i0;
  const float U = float_of((uint32_t)(F.uexp + 127) << 23);
  int V = __builtin_amdgcn_readfirstlane((int)Vd);
  float z = v;
  while (i0 < n) {
    uint32_t gi = (uint32_t)__builtin_amdgcn_readlane((int)g, (int)i0);
    used += gi;
    for (; gi > 0 && abs(V) < F.bound; --gi) V = fixed_step<M>(V, F.Am, &z);
    if (abs(V) >= F.bound) {
      v = (float)V * U;  // exact: V is on its binade's grid
      v = fold<M, true>(v, miss, gi);
      v = add<M, true>(v, hit);
      return i0 + 1;
    }
    V = fixed_step<M>(V, F.Ah, &z);
    ++i0;
    if (abs(V) >= F.bound) break;
  }
  v = V == 0 ? z : (float)V * U;
  return i0;
}

template <int M>
__device__ float wave_hit_sequence(const uint32_t* __restrict__ gap,
                                   uint32_t j, float miss, float hit,
                                   uint32_t* used_out) {
  const uint32_t lane = threadIdx.x & 63;
  // below this magnitude a binade holds only a few events: single steps
  const float small = 8.0f * fmaxf(fabsf(hit), fabsf(miss));
  float v = 0.0f;
  uint32_t used = 0;
  const FixedSteps F = fixed_steps(hit, miss, small);
  // Inside a binade the running sum is kept as S ulps (binade exponent ex,
  // steps Ah and Am ulps); v is brought up to date when the binade is left.
  bool in = false;
  uint32_t ex = 0;
  long long S = 0, Ah = 0, Am = 0;
  auto leave = [&]() {
    if (in) {
      const double u = double_of((uint64_t)(ex + 873u) << 52);
      v = (float)((double)S * u);  // exact
      in = false;
    }
  };
  // gaps are loaded kGroup chunks at a time, so one memory latency is
  // spread over kGroup * 64 events
  constexpr uint32_t kGroup = 8;
  uint32_t grp[kGroup];
  for (uint32_t base = 0; base < j; base += 64) {
    const uint32_t c = (base >> 6) & (kGroup - 1);
    if (c == 0) {
#pragma unroll
      for (uint32_t t = 0; t < kGroup; ++t) {
        const uint32_t e = base + t * 64 + lane;
        grp[t] = e < j ? gap[e] : 0u;
      }
    }
    const uint32_t n = j - base < 64u ? j - base : 64u;
    const uint32_t g = grp[c];
    uint64_t nzm = __ballot(g != 0u && lane < n);
    if (!in && !F.ok && __popcll(nzm) <= 8 && fabsf(v) < small) {
      // near zero without fixed-point steps (where the prefix sums below
      // step one event at a time) and few misses in this chunk: runs of
      // hits between them, one fold each
      uint32_t pos = 0;
      while (nzm) {
        const uint32_t p = (uint32_t)__builtin_ctzll(nzm);
        nzm &= nzm - 1;
        const uint32_t gp = (uint32_t)__builtin_amdgcn_readlane((int)g, (int)p);
        v = fold<M, true>(v, hit, p - pos);
        v = fold<M, true>(v, miss, gp);
        v = add<M, true>(v, hit);
        used += gp;
        pos = p + 1;
      }
      v = fold<M, true>(v, hit, n - pos);
      continue;
    }
    uint32_t i0 = 0;
    while (i0 < n) {
      if (!in && F.ok && fabsf(v) < small) {
        const uint32_t i1 = fixed_run<M>(F, g, i0, n, hit, miss, v, used);
        if (i1 != i0) {
          i0 = i1;
          continue;
        }
      }
      if (!in) {
        ex = (bits_of(v) >> 23) & 0xffu;
        if (ex != 0u && ex != 0xffu && !(fabsf(v) < small)) {
          const double inv_u = double_of((uint64_t)(1173u - ex) << 52);  // 2^-e'
          if (unit_step<M>((double)hit * inv_u, &Ah) &&
              unit_step<M>((double)miss * inv_u, &Am)) {
            S = (long long)((double)v * inv_u);  // exact
            in = true;
          }
        }
      }
      uint32_t f = 0;
      if (in) {
        const bool neg = S < 0;
        const bool act = lane >= i0 && lane < n;
        // the sum of (g Am + Ah) over events i0..lane, from the prefix
        // sum G of the gaps (below 2^31: they count one voxel's misses)
        const uint32_t G = wave_incl_scan32(act ? g : 0u);
        const long long P =
            (long long)G * Am + (long long)((int)lane - (int)i0 + 1) * Ah;
        const bool ok = act && in_binade(S + P, neg) &&
                        (g == 0u || in_binade(S + P - Ah, neg));
        const uint64_t bad = ~__ballot(ok) & (~0ull << i0);
        f = (bad ? (uint32_t)__builtin_ctzll(bad) : 64u) - i0;
        if (f) {
          const int last = (int)(i0 + f - 1);
          const uint32_t Gf = (uint32_t)__builtin_amdgcn_readlane((int)G, last);
          S += (long long)Gf * Am + (long long)f * Ah;
          used += Gf;
          i0 += f;
        }
      }
      if (i0 < n) {
        // event i0 leaves the binade (or no jump applies): one step
        leave();
        const uint32_t gi = (uint32_t)__builtin_amdgcn_readlane((int)g, (int)i0);
        v = fold<M, true>(v, miss, gi);
        v = add<M, true>(v, hit);
        used += gi;
        ++i0;
      }
    }
  }
  leave();
  *used_out = used;
  return v;
}

// Bits of a tile's hit word after k_temp: the voxel was touched, and it
// had misses (its miss word then holds the FP32 temp value).
constexpr uint32_t kTouchedBit = 1u << 30;
constexpr uint32_t kMixedBit = 1u << 31;
constexpr uint32_t kHitsMask = kTouchedBit - 1;

// Temp value of every touched voxel: the FP32 pixel-order sum of its
// events, rebuilt from the counts (Handler.cpp:71-82).  One lane per voxel
// over all tiles, so the serial sums of many voxels overlap.  Voxels with
// many hits (a surface close to the sensor, often a whole plane of a leaf)
// are listed for k_heavy, which gives each its own wavefront.  The result
// replaces the voxel's miss count in the tile.
// miss_runs[job][k] = fold(0, miss, k): the sum of a voxel with k misses
// and no hit.  A job whose table already holds its (miss, rounding) --
// runs_key, set by k_runs_mark after the table is built -- is skipped, so a
// batch replayed with unchanged log-odds rebuilds nothing.
__device__ __forceinline__ uint64_t runs_key_of(float miss, int M) {
  return ((uint64_t)bits_of(miss) << 8) | (uint64_t)M;
}

template <int M>
__global__ __launch_bounds__(256) void k_miss_runs(
    const Job* __restrict__ jobs, float* __restrict__ runs,
    const uint64_t* __restrict__ runs_key) {
  const float miss = jobs[blockIdx.y].miss;
  if (runs_key[blockIdx.y] == runs_key_of(miss, M)) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < kMissRuns)
    runs[(size_t)blockIdx.y * kMissRuns + k] = fold<M, true>(0.0f, miss, k);
}

template <int M>
__global__ void k_runs_mark(const Job* __restrict__ jobs, int njobs,
                            uint64_t* __restrict__ runs_key) {
  for (int j = threadIdx.x; j < njobs; j += blockDim.x)
    runs_key[j] = runs_key_of(jobs[j].miss, M);
}

// Pixel-order temp value of a voxel with k misses and the j sorted hits
// gap[s0, s0 + j) (gap g: the misses before hit g): the gap runs and hits
// folded one by one, then the misses after the last hit.
template <int M>
__device__ __forceinline__ float light_temp(const uint32_t* __restrict__ gap,
                                            uint32_t s0, uint32_t j,
                                            uint32_t k, float hit,
                                            float miss) {
  // gaps are loaded 8 at a time so the loads overlap
  uint32_t used = 0;
  float acc = 0.0f;
  for (uint32_t g = 0; g < j; g += 8) {
    uint32_t gm[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) gm[u] = g + u < j ? gap[s0 + g + u] : 0u;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (g + u < j) {
        if (gm[u] != 0u) {
          acc = fold_small<M>(acc, miss, gm[u]);
          used += gm[u];
        }
        acc = add<M, true>(acc, hit);
      }
    }
  }
  // (used <= k: every gap counts a miss that k counts too)
  return fold<M, true>(acc, miss, k >= used ? k - used : 0u);
}

// Temp values of every voxel with at most heavy_hits hits, in one launch
// of two parts that run side by side:
//  - blocks [0, lblocks): the voxels with 1..heavy_hits hits, one lane per
//    voxel, found where its run starts in the sorted hit keys, so the lanes
//    of a wavefront are all hit voxels and none idles through another's
//    loop (a run longer than heavy_hits is k_heavy's, which may already
//    have rewritten the word: j then exceeds heavy_hits as well);
//  - the other blocks: the voxels without hits, one lane per voxel of every
//    tile, the miss-only sum fold(0, miss, k) from the job's table (a voxel
//    holding hits has a nonzero hit word).
template <int M>
__global__ __launch_bounds__(256) void k_temp(
    const Job* __restrict__ jobs, const int32_t* __restrict__ slot_job,
    uint32_t* __restrict__ tiles, const uint32_t* __restrict__ hit_start,
    const uint32_t* __restrict__ gap, const FrameCounters* __restrict__ cnt,
    const uint32_t* __restrict__ keys, uint32_t heavy_hits, int tile_cap,
    const float* __restrict__ miss_runs, int lblocks) {
  if ((int)blockIdx.x < lblocks) {
    const int nh = cnt->n_hit_keys;
    for (int e = blockIdx.x * 256 + threadIdx.x; e < nh; e += lblocks * 256) {
      const uint32_t key = keys[e];
      if (key == kInvalid || (e > 0 && keys[e - 1] == key)) continue;
      const size_t cell = key;
      const uint32_t s0 = hit_start[cell];
      const uint32_t end = tiles[2 * cell + 1];
      const uint32_t j = end - s0;
      if (j == 0u || j > heavy_hits) continue;
      const Job& J = jobs[slot_job[cell >> 9]];
      const uint32_t k = tiles[2 * cell];
      const float tv = light_temp<M>(gap, s0, j, k, J.hit, J.miss);
      tiles[2 * cell] = bits_of(tv);
      tiles[2 * cell + 1] = j | kTouchedBit | (k != 0 ? kMixedBit : 0u);
    }
    return;
  }
  const int lane = threadIdx.x & 63;
  const int64_t nvox = (int64_t)min(cnt->n_leaves, tile_cap) * 512;
  const int64_t stride = (int64_t)(gridDim.x - lblocks) * blockDim.x;
  for (int64_t base = (int64_t)(blockIdx.x - lblocks) * blockDim.x +
                      (threadIdx.x & ~63);
       base < nvox; base += stride) {
    // the wavefront's 64 voxels lie in one tile
    const int slot = (int)(base >> 9);
    const int job = __builtin_amdgcn_readfirstlane(slot_job[slot]);
    const float miss = jobs[job].miss;
    const size_t cell = (size_t)base + lane;
    const uint32_t k = tiles[2 * cell];
    const uint32_t end = tiles[2 * cell + 1];
    if (k == 0u || end != 0u) continue;  // untouched, or holding hits
    const float tv = k < (uint32_t)kMissRuns
                         ? miss_runs[(size_t)job * kMissRuns + k]
                         : fold<M, true>(0.0f, miss, k);
    tiles[2 * cell] = bits_of(tv);
    tiles[2 * cell + 1] = kTouchedBit | kMixedBit;
  }
}

// Temp values of the listed many-hit voxels, one wavefront per voxel.
template <int M>
__global__ __launch_bounds__(256) void k_heavy(
    const Job* __restrict__ jobs, const int32_t* __restrict__ slot_job,
    uint32_t* __restrict__ tiles, const uint32_t* __restrict__ hit_start,
    const uint32_t* __restrict__ gap, const FrameCounters* __restrict__ cnt,
    const uint32_t* __restrict__ keys, uint32_t heavy_hits) {
  const int lane = threadIdx.x & 63;
  const int nh = cnt->n_hit_keys;
  const int nchunks = (nh + 63) >> 6;
  const int waves = gridDim.x * (blockDim.x >> 6);
  // Each wavefront scans chunks of 64 sorted hits for the first hit of a
  // voxel whose run is longer than heavy_hits (the run reaches
  // e + heavy_hits), and sums each such voxel's sequence in turn.
  for (int c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       c < nchunks; c += waves) {
    const int e = c * 64 + lane;
    const uint32_t key = e < nh ? keys[e] : kInvalid;
    const uint32_t prev = (e > 0 && e <= nh) ? keys[e - 1] : kInvalid;
    const bool heavy = key != kInvalid && key != prev &&
                       (int64_t)e + heavy_hits < (int64_t)nh &&
                       keys[e + heavy_hits] == key;
    uint64_t m = __ballot(heavy);
    while (m) {
      const int l = (int)__builtin_ctzll(m);
      m &= m - 1;
      const size_t cell = (uint32_t)__builtin_amdgcn_readlane((int)key, l);
      const int slot = (int)(cell >> 9);
      const Job& J = jobs[__builtin_amdgcn_readfirstlane(slot_job[slot])];
      const uint32_t k = tiles[2 * cell];
      const uint32_t s0 = hit_start[cell];
      const uint32_t j = tiles[2 * cell + 1] - s0;
      uint32_t used = 0;
      const float acc =
          wave_hit_sequence<M>(gap + s0, j, J.miss, J.hit, &used);
      const float tv = fold<M, true>(acc, J.miss, k >= used ? k - used : 0u);
      if (lane == 0) {
        tiles[2 * cell] = bits_of(tv);
        tiles[2 * cell + 1] = j | kTouchedBit | (k != 0 ? kMixedBit : 0u);
      }
    }
  }
}

// Folds the temp values into each tile's map leaf (Handler.cpp:88-122):
// map += temp for nonzero temp values, clamp, activity thresholds, and
// clears the tile for the next frame.  One wavefront per tile, lane l
// taking voxels l + 64 u (u < 8): the tile's eight rows and the map leaf's
// are loaded together, and the leaf-wide decisions are ballots, so a
// wavefront waits on one chain of memory latencies per tile.
template <int M>
__global__ __launch_bounds__(256) void k_apply(
    const Job* __restrict__ jobs, const int32_t* __restrict__ slot_job,
    int32_t* __restrict__ leaf_slot, const int32_t* __restrict__ slot_leaf,
    uint32_t* __restrict__ tiles, uint32_t* __restrict__ hit_mask,
    const int32_t* __restrict__ map_idx, const int32_t* __restrict__ map_new,
    const MapView* __restrict__ maps, MapView single, int tile_cap,
    FrameCounters* __restrict__ cnt, uint32_t* __restrict__ slot_stats) {
  const int lane = threadIdx.x & 63;
  const int nleaves = min(cnt->n_leaves, tile_cap);
  const int nwaves = gridDim.x * (blockDim.x >> 6);
  uint2* __restrict__ cells = reinterpret_cast<uint2*>(tiles);
  // the next tile's rows and slot entries are requested before this
  // tile's are folded
  int slot = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  uint2 tn[8];
  int jbn = 0, cn = 0, knownn = -1;
  if (slot < nleaves) {
#pragma unroll
    for (int u = 0; u < 8; ++u)
      tn[u] = cells[(size_t)slot * 512 + u * 64 + lane];
    jbn = slot_job[slot];
    cn = slot_leaf[slot];
    knownn = map_idx[slot];
  }
  for (; slot < nleaves; slot += nwaves) {
    uint2 t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = tn[u];
    const int jb = __builtin_amdgcn_readfirstlane(jbn);
    const int c = __builtin_amdgcn_readfirstlane(cn);
    const int known = __builtin_amdgcn_readfirstlane(knownn);
    const int ns = slot + nwaves;
    if (ns < nleaves) {
#pragma unroll
      for (int u = 0; u < 8; ++u)
        tn[u] = cells[(size_t)ns * 512 + u * 64 + lane];
      jbn = slot_job[ns];
      cn = slot_leaf[ns];
      knownn = map_idx[ns];
    }
    const Job& J = jobs[jb];
    const MapView m = maps ? maps[J.map] : single;
    float old_v[8];
    uint32_t old_w[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      old_v[u] = 0.0f;
      old_w[u] = 0u;
    }
    if (known >= 0) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        old_v[u] = m.values[(size_t)known * 512 + u * 64 + lane];
        old_w[u] = m.masks[(size_t)known * kMaskWords + 2 * u + (lane >> 5)];
      }
    }
    bool any_nz = false;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      any_nz |= (t[u].y & kTouchedBit) != 0 && float_of(t[u].x) != 0.0f;
    const bool any = __ballot(any_nz) != 0;
    int idx = known;
    if (known < 0) {  // a leaf new to the map: its pool leaf is fresh (zero)
      int got = -1;
      if (lane == 0) {
        const int r = map_new[slot];
        bool ok = false;
        if (r >= 0) {
          int ox, oy, oz;
          slot_origin(J, c - (int)J.cube_off, ox, oy, oz);
          ok = any && map_insert(m, ox, oy, oz, r);
          if (any && !ok) atomicAdd(&cnt->overflow, 1);
          if (!ok) {  // the reserved pool leaf stays a (zero) hole
            m.origins[3 * r] = INT_MIN;
            atomicAdd(m.leaf_count + 1, 1);
          }
        }
        got = ok ? r : -1;
      }
      idx = __builtin_amdgcn_readfirstlane(got);
    }
    uint32_t n_touched = 0, n_hit = 0, n_mixed = 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t t1 = t[u].y;
      const bool touched = (t1 & kTouchedBit) != 0;
      const uint32_t nhits = t1 & kHitsMask;
      const float tv = touched ? float_of(t[u].x) : 0.0f;
      const bool nz = touched && tv != 0.0f;
      if (any && idx >= 0) {
        bool act = (old_w[u] >> (lane & 31)) & 1u;
        if (nz) {
          float v = add<M, true>(old_v[u], tv);
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
          m.values[(size_t)idx * 512 + u * 64 + lane] = v;
        }
        const uint64_t bits = __ballot(act);
        if (lane < 2)
          m.masks[(size_t)idx * kMaskWords + 2 * u + lane] =
              (uint32_t)(lane == 0 ? bits : bits >> 32);
      }
      if (touched) cells[(size_t)slot * 512 + u * 64 + lane] = make_uint2(0u, 0u);
      // per-slot statistics: voxels touched, with hits, with hits and misses
      n_touched += (uint32_t)__popcll(__ballot(touched));
      n_hit += (uint32_t)__popcll(__ballot(nhits != 0));
      n_mixed += (uint32_t)__popcll(__ballot(nhits != 0 && (t1 & kMixedBit)));
    }
    if (lane < kMaskWords) hit_mask[(size_t)slot * kMaskWords + lane] = 0;
    if (lane == 0) {
      leaf_slot[c] = -1;
      slot_stats[(size_t)slot * 4 + 0] = n_touched;
      slot_stats[(size_t)slot * 4 + 1] = n_hit;
      slot_stats[(size_t)slot * 4 + 2] = n_mixed;
    }
  }
}

// Without a map: clear the per-frame state of the last cast.
__global__ __launch_bounds__(512) void k_clear_slots(
    int32_t* __restrict__ leaf_slot, const int32_t* __restrict__ slot_leaf,
    uint32_t* __restrict__ tiles, uint32_t* __restrict__ hit_mask,
    const FrameCounters* __restrict__ cnt, int tile_cap) {
  const int nleaves = min(cnt->n_leaves, tile_cap);
  for (int slot = blockIdx.x; slot < nleaves; slot += gridDim.x) {
    const size_t cell = (size_t)slot * 512 + threadIdx.x;
    tiles[2 * cell] = 0;
    tiles[2 * cell + 1] = 0;
    if (threadIdx.x < kMaskWords)
      hit_mask[(size_t)slot * kMaskWords + threadIdx.x] = 0;
    if (threadIdx.x == 0) leaf_slot[slot_leaf[slot]] = -1;
  }
}

// One workgroup per job: the job's statistics from the per-item discover
// partials and the per-slot update statistics.
__global__ __launch_bounds__(256) void k_job_stats(
    const Job* __restrict__ jobs, int njobs,
    const uint64_t* __restrict__ part, const int32_t* __restrict__ scan_blocks,
    const uint32_t* __restrict__ slot_stats,
    const FrameCounters* __restrict__ cnt, int tile_cap,
    JobStats* __restrict__ out) {
  __shared__ unsigned long long s[7];
  const int jb = blockIdx.x;
  const Job& J = jobs[jb];
  const int total = min(cnt->n_leaves, tile_cap);
  int b0 = scan_blocks[J.cube_off / kScanChunk];
  int b1 = jb + 1 < njobs ? scan_blocks[jobs[jb + 1].cube_off / kScanChunk]
                          : cnt->n_leaves;
  b0 = min(b0, total);
  b1 = min(b1, total);
  if (threadIdx.x < 7) s[threadIdx.x] = 0;
  __syncthreads();
  uint64_t a[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int it = threadIdx.x; it < J.nitems; it += 256)
    for (int q = 0; q < 4; ++q) a[q] += part[(size_t)(J.item_off + it) * 4 + q];
  for (int sl = b0 + threadIdx.x; sl < b1; sl += 256)
    for (int q = 0; q < 3; ++q) a[4 + q] += slot_stats[(size_t)sl * 4 + q];
  for (int q = 0; q < 7; ++q) {
    const uint64_t w = wave_sum(a[q]);
    if ((threadIdx.x & 63) == 0 && w) atomicAdd(&s[q], (unsigned long long)w);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    JobStats r;
    r.visits = s[0];
    r.rays = s[1];
    r.hit_rays = s[2];
    r.errors = s[3];
    r.voxels = s[4];
    r.hit_voxels = s[5];
    r.mixed_voxels = s[6];
    r.slot_begin = b0;
    r.slot_end = b1;
    out[jb] = r;
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
  if (t == 0) {
    m.leaf_count[0] = 0;
    m.leaf_count[1] = 0;
  }
}

ItemQueue cast_queue(const Pipeline& p, const Workspace& b, int which,
                     int* grid) {
  ItemQueue q;
  q.item_job = p.njobs > 1 ? b.item_job : nullptr;
  q.order = nullptr;
  q.head = nullptr;
  const int n = which == 1 && p.accum_items > 0 ? p.accum_items : p.nitems;
  q.nitems = n;
  *grid = n;
  if (p.tuning.queue != 0) {
    q.head = &b.counters->queue[which];
    if (p.tuning.queue >= 2) q.order = b.item_order;
    if (p.tuning.queue == 3 && which == 1 && p.accum_order)
      q.order = p.accum_order;
    const int g = p.cu_count * (p.tuning.grid_mult > 0 ? p.tuning.grid_mult : 1);
    *grid = n < g ? n : g;
  }
  return q;
}

}  // namespace

int job_items(const Job& j, int layout) {
  if (j.n <= 0) return 0;
  if (j.tiled) {
    int tw, th, ww;
    layout_dims(layout, tw, th, ww);
    return ((j.width + tw - 1) / tw) * ((j.height + th - 1) / th);
  }
  return (j.n + kCastThreads - 1) / kCastThreads;
}

hipError_t launch_begin(const Workspace& b, const Job* job0, hipStream_t s) {
  Job j{};
  if (job0) j = *job0;
  k_begin<<<1, 64, 0, s>>>(j, job0 ? 1 : 0, b.jobs, b.counters);
  return hipGetLastError();
}

hipError_t launch_discover(const Pipeline& p, const Workspace& b,
                           hipStream_t s) {
  if (p.nitems <= 0) return hipSuccess;
  int grid = 0;
  const ItemQueue q = cast_queue(p, b, 0, &grid);
  uint32_t* ck = p.tuning.queue == 3 ? b.cost_keys[0] : nullptr;
  uint32_t* cv = p.tuning.queue == 3 ? b.cost_vals[0] : nullptr;
  uint32_t* il = p.accum_items > p.nitems ? b.item_longest : nullptr;
  if (p.rounding == kDown)
    k_discover<kDown><<<grid, kCastThreads, 0, s>>>(
        b.jobs, q, b.leaf_slot, b.ray_hit, b.part_discover, ck, cv, il);
  else
    k_discover<kNearest><<<grid, kCastThreads, 0, s>>>(
        b.jobs, q, b.leaf_slot, b.ray_hit, b.part_discover, ck, cv, il);
  return hipGetLastError();
}


int accumulate_resident(int cu_count) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void*>(&k_accumulate<kDown, 12>),
          kCastThreads, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 6;
  return per_cu * cu_count;
}

hipError_t launch_item_order(Pipeline& p, const Workspace& b, hipStream_t s) {
  p.accum_order = nullptr;
  if (p.tuning.queue != 3 || p.nitems <= 0) return hipSuccess;
  uint32_t* keys[2] = {b.cost_keys[0], b.cost_keys[1]};
  uint32_t* vals[2] = {b.cost_vals[0], b.cost_vals[1]};
  int out = 0;
  // stable: items of equal cost keep the discover pass's order
  const hipError_t e = radix_sort_pairs(keys, vals, p.nitems, kCostBits,
                                        b.radix_hist, &out, s);
  if (e != hipSuccess) return e;
  p.accum_order = reinterpret_cast<const int32_t*>(b.cost_vals[out]);
  return hipSuccess;
}

hipError_t launch_compact_leaves(const Pipeline& p, const Workspace& b,
                                 hipStream_t s) {
  const int nblocks = (int)((p.cube_vol + kScanChunk - 1) / kScanChunk);
  if (nblocks <= 0) return hipSuccess;
  k_scan_count<<<nblocks, 256, 0, s>>>(b.leaf_slot, p.cube_vol, b.scan_blocks);
  // a sliced single-job launch also plans its accumulate slices here
  const bool sliced = p.njobs == 1 && p.accum_items > p.nitems;
  k_scan_blocks<<<1, 1024, 0, s>>>(
      b.scan_blocks, nblocks, b.part_discover, p.nitems, b.counters,
      sliced ? b.item_longest : nullptr,
      sliced ? p.host_jobs[0].slices : 1, p.accum_items - p.nitems,
      b.slice_items);
  k_scan_assign<<<nblocks, 256, 0, s>>>(b.leaf_slot, p.cube_vol, b.scan_blocks,
                                        b.scan_job, b.slot_leaf, b.slot_job,
                                        p.tile_cap, b.counters);
  return hipGetLastError();
}

hipError_t launch_hit_keys(const Pipeline& p, const Workspace& b,
                           hipStream_t s) {
  if (p.nitems <= 0) return hipSuccess;
  ItemQueue q;
  q.item_job = p.njobs > 1 ? b.item_job : nullptr;
  q.order = nullptr;
  q.head = nullptr;
  q.nitems = p.nitems;
  k_hit_keys<<<p.nitems, 256, 0, s>>>(b.jobs, q, b.ray_hit, b.leaf_slot,
                                      b.keys[0], b.vals[0]);
  return hipGetLastError();
}

hipError_t launch_hit_segments(const Pipeline& p, const Workspace& b,
                               int sorted_buf, hipStream_t s) {
  if (p.n_rays <= 0) return hipSuccess;
  const int64_t g = (p.n_rays + 255) / 256;
  const int grid = g < 2048 ? (int)g : 2048;
  k_hit_segments<<<grid, 256, 0, s>>>(b.keys[sorted_buf], b.counters, b.tiles,
                                      b.hit_mask, b.hit_start, b.gap,
                                      b.slot_leaf, b.leaf_slot);
  return hipGetLastError();
}

template <int M, int S>
void accumulate_as(const Pipeline& p, const Workspace& b, int sorted_buf,
                   const ItemQueue& q, int grid, hipStream_t s) {
  k_accumulate<M, S><<<grid, kCastThreads, 0, s>>>(
      b.jobs, q, b.leaf_slot, b.tiles, b.hit_mask, b.hit_start,
      b.vals[sorted_buf], b.gap, b.counters, b.item_longest, b.slice_items);
}

hipError_t launch_accumulate(const Pipeline& p, const Workspace& b,
                             int sorted_buf, hipStream_t s) {
  if (p.nitems <= 0) return hipSuccess;
  int grid = 0;
  const ItemQueue q = cast_queue(p, b, 1, &grid);
  const bool small = p.tuning.lds_slots == 2048;
  if (p.rounding == kDown) {
    if (small)
      accumulate_as<kDown, 11>(p, b, sorted_buf, q, grid, s);
    else
      accumulate_as<kDown, 12>(p, b, sorted_buf, q, grid, s);
  } else {
    if (small)
      accumulate_as<kNearest, 11>(p, b, sorted_buf, q, grid, s);
    else
      accumulate_as<kNearest, 12>(p, b, sorted_buf, q, grid, s);
  }
  return hipGetLastError();
}

namespace {

// The map lookup and, unless already valid, the miss-run tables.
void map_lookup_and_runs(const Pipeline& p, const Workspace& b,
                         const MapView* map_table, const MapView& single,
                         hipStream_t s) {
  k_map_lookup<<<256, 256, 0, s>>>(b.jobs, b.slot_job, b.slot_leaf, map_table,
                                   single, p.tile_cap, b.counters, b.map_idx,
                                   b.map_new);
  const dim3 rgrid(kMissRuns / 256, p.njobs);
  if (!p.miss_runs_ready) {
    if (p.rounding == kDown) {
      k_miss_runs<kDown><<<rgrid, 256, 0, s>>>(b.jobs, b.miss_runs,
                                               b.runs_key);
      k_runs_mark<kDown><<<1, 64, 0, s>>>(b.jobs, p.njobs, b.runs_key);
    } else {
      k_miss_runs<kNearest><<<rgrid, 256, 0, s>>>(b.jobs, b.miss_runs,
                                                  b.runs_key);
      k_runs_mark<kNearest><<<1, 64, 0, s>>>(b.jobs, p.njobs, b.runs_key);
    }
  }
}

}  // namespace

hipError_t launch_map_prepare(Pipeline& p, const Workspace& b,
                              const MapView* map_table, const MapView& single,
                              hipStream_t s) {
  p.map_prepared = false;
  if (!p.side || !p.early_fork || !p.prepared) return hipErrorInvalidValue;
  hipError_t e = hipEventRecord(p.early_fork, s);
  if (e == hipSuccess) e = hipStreamWaitEvent(p.side, p.early_fork, 0);
  if (e != hipSuccess) return e;
  map_lookup_and_runs(p, b, map_table, single, p.side);
  e = hipEventRecord(p.prepared, p.side);
  if (e != hipSuccess) return e;
  p.map_prepared = true;
  return hipGetLastError();
}

hipError_t launch_update(const Pipeline& p, const Workspace& b,
                         int sorted_buf, const MapView* map_table,
                         const MapView& single, hipStream_t s) {
  const bool down = p.rounding == kDown;
  // eight wavefronts per SIMD to hide the latency of the heavy voxels'
  // serial chains
  const int hgrid = p.cu_count * 8;
  const int tgrid = p.cu_count * 8;
  const int agrid = p.cu_count * 8;
  // With a side stream (a branch of the batch graph) the heavy voxels'
  // chains, which need only the cast's results, run beside the map lookup
  // and the per-voxel sums; both join before the apply pass.
  hipStream_t hs = s;
  if (p.side) {
    hipError_t e = hipEventRecord(p.fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(p.side, p.fork, 0);
    if (e != hipSuccess) return e;
    hs = p.side;
  }
  if (down)
    k_heavy<kDown><<<hgrid, 256, 0, hs>>>(b.jobs, b.slot_job, b.tiles,
                                          b.hit_start, b.gap, b.counters,
                                          b.keys[sorted_buf], p.heavy_hits);
  else
    k_heavy<kNearest><<<hgrid, 256, 0, hs>>>(b.jobs, b.slot_job, b.tiles,
                                             b.hit_start, b.gap, b.counters,
                                             b.keys[sorted_buf],
                                             p.heavy_hits);
  if (p.side) {
    const hipError_t e = hipEventRecord(p.join, p.side);
    if (e != hipSuccess) return e;
  }
  if (p.map_prepared) {
    // the lookup and the miss-run tables ran beside the cast
    const hipError_t e = hipStreamWaitEvent(s, p.prepared, 0);
    if (e != hipSuccess) return e;
  } else {
    map_lookup_and_runs(p, b, map_table, single, s);
  }
  const int lgrid = p.cu_count * 4;
  if (down)
    k_temp<kDown><<<lgrid + tgrid, 256, 0, s>>>(
        b.jobs, b.slot_job, b.tiles, b.hit_start, b.gap, b.counters,
        b.keys[sorted_buf], p.heavy_hits, p.tile_cap, b.miss_runs, lgrid);
  else
    k_temp<kNearest><<<lgrid + tgrid, 256, 0, s>>>(
        b.jobs, b.slot_job, b.tiles, b.hit_start, b.gap, b.counters,
        b.keys[sorted_buf], p.heavy_hits, p.tile_cap, b.miss_runs, lgrid);
  if (p.side) {
    const hipError_t e = hipStreamWaitEvent(s, p.join, 0);
    if (e != hipSuccess) return e;
  }
  if (down)
    k_apply<kDown><<<agrid, 256, 0, s>>>(
        b.jobs, b.slot_job, b.leaf_slot, b.slot_leaf, b.tiles, b.hit_mask,
        b.map_idx, b.map_new, map_table, single, p.tile_cap, b.counters,
        b.slot_stats);
  else
    k_apply<kNearest><<<agrid, 256, 0, s>>>(
        b.jobs, b.slot_job, b.leaf_slot, b.slot_leaf, b.tiles, b.hit_mask,
        b.map_idx, b.map_new, map_table, single, p.tile_cap, b.counters,
        b.slot_stats);
  return hipGetLastError();
}

hipError_t launch_clear_slots(const Pipeline& p, const Workspace& b,
                              hipStream_t s) {
  k_clear_slots<<<kUpdateGrid, 512, 0, s>>>(b.leaf_slot, b.slot_leaf, b.tiles,
                                            b.hit_mask, b.counters,
                                            p.tile_cap);
  return hipGetLastError();
}

hipError_t launch_job_stats(const Pipeline& p, const Workspace& b,
                            hipStream_t s) {
  if (p.njobs <= 0) return hipSuccess;
  k_job_stats<<<p.njobs, 256, 0, s>>>(b.jobs, p.njobs, b.part_discover,
                                      b.scan_blocks, b.slot_stats, b.counters,
                                      p.tile_cap, b.job_stats);
  return hipGetLastError();
}

hipError_t launch_map_clear(const MapView& m, hipStream_t s) {
  k_map_clear<<<1024, 256, 0, s>>>(m);
  return hipGetLastError();
}

}  // namespace nm
