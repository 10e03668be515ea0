// nm_internal.h -- device data layout and launcher declarations shared by
// the HIP translation units of libnanomap_amd.
#ifndef NM_INTERNAL_H
#define NM_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <string>

#include "../../include/nanomap_amd.h"

namespace nm {

constexpr uint32_t kInvalid = 0xffffffffu;
constexpr int kLeafVoxels = 512;  // 8^3
constexpr int kMaskWords = 16;    // 512 bits

// Device-side counters of one launch sequence (one frame, or one batch of
// frames).  Zeroed by the first kernel of the sequence (no hipMemsetAsync,
// so the sequence can be graph-captured).
struct FrameCounters {
  int32_t n_leaves;    // leaves discovered (tile slots in use), all jobs
  int32_t n_hit_rays;  // rays whose end voxel received a hit, all jobs
  int32_t n_hit_keys;  // sorted hit events (hits inside allocated leaves)
  int32_t errors;      // consistency errors (expected 0)
  int32_t overflow;    // tile or map capacity exceeded
  uint64_t rays;       // valid rays
  uint64_t visits;     // voxel updates
  int32_t n_filter_rays;    // rays emitted by the voxel filter
  int32_t n_filter_points;  // points binned by the voxel filter
  int32_t queue[4];         // work-queue heads of the persistent kernels
  uint32_t slice_visits;    // sliced accumulate: visits per time slice
                            // (k_slice_plan; 0 = no slicing)
  int32_t slice_extra;      // sliced accumulate: entries of slice_items
};

// Statistics of one job, reduced on demand from per-item and per-slot
// partial sums (k_job_stats).
struct JobStats {
  unsigned long long rays, visits, hit_rays, errors;
  unsigned long long voxels, hit_voxels, mixed_voxels;
  int32_t slot_begin, slot_end;  // tile slots of this job
};

// The sparse map on the device: an open-addressing hash of packed leaf
// coordinates -> leaf index, and a leaf pool of 512 FP32 log-odds plus a
// 512-bit active mask per leaf.  Stands in for the openvdb::FloatGrid of
// map::Map (include/nanomap/map/OccupancyMap.h:44-51, 69).
struct MapView {
  unsigned long long* hash_keys;  // [hash_cap], ~0 = empty
  int32_t* hash_vals;             // [hash_cap]
  float* values;                  // [leaf_cap * 512]
  uint32_t* masks;                // [leaf_cap * 16]
  int32_t* origins;               // [leaf_cap * 3]
  int32_t* leaf_count;            // device counters: [0] pool slots used,
                                  // [1] holes (slots whose origin x is
                                  // INT_MIN: reserved, never inserted)
  int64_t hash_cap;               // power of two
  int32_t leaf_cap;
  float clamp_min, clamp_max, hit_thr, miss_thr;
};

// Map hash key of the leaf holding voxel (x, y, z): leaf coordinates
// (x >> 3 ...) biased by 2^20, 21 bits each; and the hash of a key.
__host__ __device__ __forceinline__ unsigned long long pack_leaf(int x, int y,
                                                                  int z) {
  const unsigned long long m = (1ull << 21) - 1;
  return (((unsigned long long)((x >> 3) + (1 << 20)) & m) << 42) |
         (((unsigned long long)((y >> 3) + (1 << 20)) & m) << 21) |
         ((unsigned long long)((z >> 3) + (1 << 20)) & m);
}

__host__ __device__ __forceinline__ unsigned long long mix64(
    unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// Pool index of the leaf holding voxel (x, y, z), or -1 (open addressing,
// linear probing; an empty slot ends the probe).
__device__ __forceinline__ int map_find(const MapView& m, int x, int y,
                                        int z) {
  const unsigned long long key = pack_leaf(x, y, z);
  const unsigned long long mask = (unsigned long long)m.hash_cap - 1;
  unsigned long long h = mix64(key) & mask;
  for (int64_t probe = 0; probe < m.hash_cap; ++probe) {
    const unsigned long long k = m.hash_keys[h];
    if (k == key) return m.hash_vals[h];
    if (k == ~0ull) return -1;
    h = (h + 1) & mask;
  }
  return -1;
}

// Cube of candidate leaves around the eye: leaf coordinates
// [cmin, cmin + side) per axis.  Every voxel a ray can reach lies inside,
// because rays are shorter than max_time voxels.
struct Cube {
  int cmin[3];
  int side;
};

// Pixel-tile layouts of the cast work items (256 rays each, four
// wavefronts): width x height of the item and of each wavefront's part.
enum Layout : int {
  kLayout16x16 = 0,  // 2x2 wavefronts of 8x8
  kLayout32x8 = 1,   // 4x1 wavefronts of 8x8
  kLayout64x4 = 2,   // 4x1 wavefronts of 16x4
  kLayout8x32 = 3,   // 1x4 wavefronts of 8x8
  kLayoutCount = 4
};

// One frame of one sensor ("job").  The cast kernels read jobs from a
// device table, so one launch serves any number of sensors; a plain frame
// is a table of one job.
struct Job {
  float R[9];
  double eye[3];
  float res;
  float max_time;
  float hit;
  float miss;
  Cube cube;
  const unsigned char* cloud;  // xyz f32 points with byte stride `step`,
                               // or (depth) a width x height f32 depth image
  const float* depth_ax;       // depth image: x factor of each column
  const float* depth_ay;       // depth image: y factor of each row
  const float* rays4;          // or pre-filtered rays (dir, length)
  const int32_t* n_dev;        // device count of rays4 entries
  int n;                       // points (upper bound on rays)
  int step;
  int width;
  int height;
  int tiled;                   // 1: points form a width x height image
  int use_rays;                // 1: cast rays4 instead of points
  int layout;                  // Layout of the work items (tiled only)
  int map;                     // index into the map table
  int depth;                   // 1: `cloud` is a depth image
  int slices;                  // accumulate work per ray block: up to
                               // this many time slices (<= kMaxSlices), as
                               // the slice plan decides; single-job
                               // launches only
  // placement in the shared workspace
  int64_t cube_off;  // first entry of the job's cube table (kScanChunk-aligned)
  int ray_off;       // global index of the job's first point
  int item_off;      // first cast work item of the job
  int nitems;        // cast work items of the job
};
using CastParams = Job;

// Sensor-frame point i of a job: the i-th xyz point of the cloud, or from
// a depth image the pinhole back-projection (d ax[u], d ay[v], d) of pixel
// (u, v) = (i % width, i / width) with z-depth d (FP32 products; NaN
// depth gives a NaN point).
__device__ __forceinline__ void load_point(const Job& p, int i, float& x,
                                           float& y, float& z) {
  if (p.depth) {
    const float d = reinterpret_cast<const float*>(p.cloud)[i];
    const int v = i / p.width;
    const int u = i - v * p.width;
    x = d * p.depth_ax[u];
    y = d * p.depth_ay[v];
    z = d;
    return;
  }
  const float* pt =
      reinterpret_cast<const float*>(p.cloud + (size_t)i * (size_t)p.step);
  x = pt[0];
  y = pt[1];
  z = pt[2];
}

// How the cast kernels walk their work items (chosen by the autotuner for
// batches; plain frames use the default).
struct Tuning {
  int layout = kLayout16x16;
  int queue = 0;      // 0: one workgroup per item; 1: persistent workgroups
                      // draining a queue in job order; 2: same, longest-
                      // range jobs first; 3: as 2 for the discover pass,
                      // and the accumulate pass takes the items in order
                      // of their longest ray, longest first (measured by
                      // the discover pass)
  int grid_mult = 4;  // persistent workgroups per compute unit
  int lds_slots = 4096;  // LDS aggregation table of the accumulate pass:
                         // 4096 or 2048 entries
};

// Device workspace shared by all jobs of a launch sequence.
struct Workspace {
  int32_t* leaf_slot;   // [cube_cap]  -1 free, -2 discovered, >= 0 slot
                        //             (| kSlotHasHits once hits are sorted)
  int32_t* slot_leaf;   // [tile_cap]  workspace cube index of each slot
  int32_t* slot_job;    // [tile_cap]  job of each slot
  int32_t* map_idx;     // [tile_cap]  map leaf of each slot, -1 if new
  int32_t* map_new;     // [tile_cap]  pool index reserved for a new leaf
  uint32_t* slot_stats; // [tile_cap * 4]  voxels, hit voxels, mixed voxels
  uint32_t* tiles;      // [tile_cap * 512 * 2]  (miss, hit) per voxel
  uint32_t* hit_mask;   // [tile_cap * 16]  voxels holding a hit
  uint32_t* hit_start;  // [tile_cap * 512] first sorted hit of a hit voxel
  uint32_t* ray_hit;    // [ray_cap]  job-cube cell of each ray's hit
  uint32_t* keys[2];    // [ray_cap]  radix sort buffers
  uint32_t* vals[2];
  uint32_t* gap;        // [ray_cap]  misses before each sorted hit
  uint32_t* radix_hist; // radix_hist_words(ray_cap)
  int32_t* scan_blocks; // [cube_cap / kScanChunk + 2]
  int32_t* scan_job;    // [cube_cap / kScanChunk + 2]  job of each chunk
  FrameCounters* counters;
  JobStats* job_stats;  // [max_jobs]
  Job* jobs;            // [max_jobs]  device job table
  MapView* maps;        // [max_jobs]  device map table
  uint64_t* part_discover;  // [item_cap * 4] per-item counter sums
  int32_t* item_job;    // [item_cap]  job of each work item
  int32_t* item_order;  // [item_cap]  queue order of the work items
  uint32_t* item_longest;  // [item_cap]  visits of each item's longest ray
                           //   (sliced accumulate)
  uint32_t* slice_items;   // [item_cap * (kMaxSlices - 1)]  the extra
                           //   accumulate items: block << 3 | slice
  uint32_t* cost_keys[2];  // [item_cap]  item cost keys (queue 3) and the
  uint32_t* cost_vals[2];  //             item indices sorted by them
  unsigned char* cloud; // staging for host clouds
  float* filter_rays;   // [ray_cap * 4]
  float* filter_tmp;    // [ray_cap * 4]
  int32_t* flags;       // [ray_cap]
  float* miss_runs;     // [max_jobs * kMissRuns]  fold(0, miss, k) per job
  uint64_t* runs_key;   // [max_jobs]  (miss bits, rounding) each table holds
};

// Sizes of a workspace.
struct WorkspaceDims {
  int64_t cube_cap = 0;  // entries of all jobs' cube tables
  int tile_cap = 0;      // tile slots
  int64_t ray_cap = 0;   // points of all jobs
  int item_cap = 0;      // cast work items of all jobs
  int max_jobs = 1;
  bool host_clouds = false;  // staging for host clouds + voxel filter
};

// Allocates a workspace (cube tables free, tiles zero); synchronous.
hipError_t workspace_alloc(const WorkspaceDims& d, Workspace* out);
void workspace_free(Workspace* b);

constexpr int kScanChunk = 4096;
// A voxel of a frame that only misses holds fold(0, miss, k): the update
// reads it from a per-job table for k < kMissRuns.
constexpr int kMissRuns = 4096;
// Bit of a cube-table entry marking a leaf that holds hit voxels (set after
// the hit sort, cleared with the entry by the update).
constexpr int32_t kSlotHasHits = 1 << 30;
constexpr int kUpdateGrid = 2048;
// Voxels with more hits than Pipeline::heavy_hits in one launch sequence
// are summed by a whole wavefront each, so that no lane sums a long
// sequence alone (measured best from 16 hits on, for a single frame and for
// a batch).
constexpr uint32_t kFrameHeavyHits = 16;
constexpr uint32_t kBatchHeavyHits = 16;

// ---- launchers (nm_cast.hip, nm_radix.hip, nm_filter.hip) -------------
// The frame sequence over a device job table `b.jobs[0, njobs)`:
//   begin -> discover -> compact -> hit keys -> radix sort -> segments ->
//   accumulate -> map lookup -> update (or clear, without a map).
// `host_jobs` is the host copy of the table (grid sizes are taken from
// it); `job0` non-null makes the begin kernel store that job as entry 0
// (plain frames, no separate upload).
struct Pipeline {
  const Job* host_jobs;
  int njobs;
  int64_t cube_vol;   // workspace cube entries in use
  int64_t n_rays;     // points of all jobs
  int nitems;         // cast work items of all jobs
  int accum_items = 0;  // work items of the accumulate pass (0: nitems; a
                        // sliced job adds the slice plan's extra items)
  int tile_cap;
  int sort_bits;
  uint32_t heavy_hits = kFrameHeavyHits;  // see kFrameHeavyHits
  bool miss_runs_ready = false;  // miss_runs already hold these jobs' table
  // optional second stream for the heavy voxels (fork/join events)
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr;
  hipEvent_t join = nullptr;
  // launch_map_prepare ran the map lookup and the miss-run tables on the
  // side stream (early_fork, then prepared recorded there): launch_update
  // then only waits for `prepared` before the per-voxel sums
  hipEvent_t early_fork = nullptr;
  hipEvent_t prepared = nullptr;
  bool map_prepared = false;
  int rounding;
  int cu_count;
  Tuning tuning;
  // queue 3: the accumulate pass's item order (set by launch_item_costs)
  const int32_t* accum_order = nullptr;
};

hipError_t launch_begin(const Workspace& b, const Job* job0, hipStream_t s);
hipError_t launch_discover(const Pipeline& p, const Workspace& b,
                           hipStream_t s);
hipError_t launch_compact_leaves(const Pipeline& p, const Workspace& b,
                                 hipStream_t s);
hipError_t launch_hit_keys(const Pipeline& p, const Workspace& b,
                           hipStream_t s);
hipError_t launch_hit_segments(const Pipeline& p, const Workspace& b,
                               int sorted_buf, hipStream_t s);
hipError_t launch_accumulate(const Pipeline& p, const Workspace& b,
                             int sorted_buf, hipStream_t s);
// Sliced accumulate (single-job launches with accum_items > nitems): the
// compact pass also picks the visits per time slice so that the extra
// slice items fit in the accum_items - nitems extra work items, and lists
// them (block, slice).
constexpr int kMaxSlices = 8;  // time slices per ray block, at most
// Accumulate workgroups resident on the current device at once.
int accumulate_resident(int cu_count);
// Queue 3: sorts the items by the cost keys the discover pass wrote and
// points p.accum_order at the result.
hipError_t launch_item_order(Pipeline& p, const Workspace& b, hipStream_t s);
// The update's steps that need only the compacted leaves and the job
// table -- the map lookup (leaf indices, new-leaf reservation) and the
// miss-run tables -- on the side stream, forked from `s` after the leaf
// compaction so they overlap the hit sort and the voxel pass.  Sets
// p.map_prepared.  Needs p.side, p.early_fork and p.prepared.
hipError_t launch_map_prepare(Pipeline& p, const Workspace& b,
                              const MapView* map_table, const MapView& single,
                              hipStream_t s);
// map_table: device table indexed by Job::map, or null to use `single`.
hipError_t launch_update(const Pipeline& p, const Workspace& b,
                         int sorted_buf, const MapView* map_table,
                         const MapView& single, hipStream_t s);
// Clears the per-frame state of the last cast without touching a map.
hipError_t launch_clear_slots(const Pipeline& p, const Workspace& b,
                              hipStream_t s);
hipError_t launch_job_stats(const Pipeline& p, const Workspace& b,
                            hipStream_t s);
hipError_t launch_map_clear(const MapView& m, hipStream_t s);

// Cast work items of a job under a layout.
int job_items(const Job& j, int layout);

// Stable LSD radix sort of (keys, vals) over the low `bits` key bits.
// Sorts buffer 0 -> result in buffer returned via *out_buf (0 or 1).  With
// n_kept (a device address), keys equal to kInvalid are dropped and the
// number of keys kept, which lead the result, is stored there.
hipError_t radix_sort_pairs(uint32_t* keys[2], uint32_t* vals[2], int n,
                            int bits, uint32_t* hist, int* out_buf,
                            hipStream_t s, int* n_kept = nullptr);
int radix_max_blocks(int n);
int radix_hist_words(int n);  // histogram buffer size for n keys

// Point bounds of the voxel filter (filterCloud.cu:100-112): a point is
// binned only if it has no NaN coordinate, lies inside [lo, hi] on every
// axis (when use_box), and 0 < |p| < max_range.
struct FilterBounds {
  float lo[3], hi[3];
  float max_range;
  int use_box;
};

// Voxel filter: bins points by voxel, writes one ray per bin into
// b.filter_rays (count -> counters->n_filter_rays).  origin: fractional
// part of the eye, eye - floor(eye), as FP32 (the reference's
// voxelOriginOffset, include/nanomap/sensor/SensorData.h:55).
hipError_t launch_voxel_filter(const CastParams& p, const float origin[3],
                               const FilterBounds& fb, int precision,
                               const Workspace& b, hipStream_t s);

void set_error(const std::string& msg);

// Log-odds increments and clamps are bounded so that every temp sum (at
// most 2^32 terms) and every map value stays a finite float, which the
// round-down sums of the update rely on.
constexpr float kMaxLogOdds = 1e27f;
inline bool logodds_ok(float v) { return std::fabs(v) <= kMaxLogOdds; }

// Bits needed for keys below max_key.
inline int sort_bits_for(int64_t max_key) {
  int bits = 1;
  while (bits < 32 && ((int64_t)1 << bits) <= max_key) ++bits;
  return bits;
}

// Cube of candidate leaves of a frame: every voxel within max_time of the
// eye, with one leaf of margin.
inline Cube frame_cube(const nm_frame& f) {
  Cube c;
  const int r = (int)ceil(((double)f.max_time + 2.0) / 8.0) + 1;
  c.side = 2 * r + 1;
  for (int a = 0; a < 3; ++a) {
    const int v = (int)floor(f.eye[a]);
    c.cmin[a] = (v >> 3) - r;
  }
  return c;
}

}  // namespace nm

// The device-resident map behind the opaque nm_map handle.
struct nm_map {
  int device = 0;
  float grid_res = 0;
  nm::MapView v{};
  hipStream_t stream = nullptr;
  bool auto_grow = false;
  // staging of host-array point queries (nm_map_query), grown on demand
  void* query_buf = nullptr;
  int64_t query_cap = 0;  // points
};

namespace nm {
// Grows m's pool and hash to at least `capacity` leaves (nm_map.hip).
int map_reserve(nm_map* m, int64_t capacity);
}  // namespace nm

#endif  // NM_INTERNAL_H
