// nm_internal.h -- device data layout and launcher declarations shared by
// the HIP translation units of libnanomap_amd.
#ifndef NM_INTERNAL_H
#define NM_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/nanomap_amd.h"

namespace nm {

constexpr uint32_t kInvalid = 0xffffffffu;
constexpr int kLeafVoxels = 512;  // 8^3
constexpr int kMaskWords = 16;    // 512 bits

// Device-side counters of one frame.  Zeroed by a fill kernel at the start
// of each frame (no hipMemsetAsync, so the sequence can be graph-captured).
struct FrameCounters {
  int32_t n_leaves;    // leaves discovered (tile slots in use)
  int32_t n_hit_rays;  // rays whose end voxel received a hit
  int32_t errors;      // consistency errors (expected 0)
  int32_t overflow;    // tile or map capacity exceeded
  uint64_t rays;       // valid rays
  uint64_t visits;     // voxel updates
  uint64_t voxels;     // unique voxels touched
  uint64_t hit_voxels;
  uint64_t mixed_voxels;
  int32_t n_filter_rays;    // rays emitted by the voxel filter
  int32_t n_filter_points;  // points binned by the voxel filter
};

// The sparse map on the device: an open-addressing hash of packed leaf
// coordinates -> leaf index, and a leaf pool of 512 FP32 log-odds plus a
// 512-bit active mask per leaf.  Stands in for the openvdb::FloatGrid of
// map::Map (include/nanomap/map/OccupancyMap.h:132).
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

// Cube of candidate leaves around the eye: leaf coordinates
// [cmin, cmin + side) per axis.  Every voxel a ray can reach lies inside,
// because rays are shorter than max_time voxels.
struct Cube {
  int cmin[3];
  int side;
};

// Per-frame parameters of the cast kernels (passed by value).
struct CastParams {
  float R[9];
  double eye[3];
  float res;
  float max_time;
  float hit;
  float miss;
  Cube cube;
  const unsigned char* cloud;  // xyz f32 points with byte stride `step`
  const float* rays4;          // or pre-filtered rays (dir, length)
  const int32_t* n_dev;        // device count of rays4 entries
  int n;                       // points (upper bound on rays)
  int step;
  int width;
  int height;
  int tiled;                   // 1: points form a width x height image
  int use_rays;                // 1: cast rays4 instead of points
};

// Device workspace of one context (one sensor stream).
struct CtxBuffers {
  int32_t* leaf_slot;   // [cube_cap]  -1 free, -2 discovered, >= 0 slot
                        //             (| kSlotHasHits once hits are sorted)
  int32_t* slot_leaf;   // [tile_cap]  cube index of each slot
  int32_t* map_idx;     // [tile_cap]  map leaf of each slot, -1 if new
  int32_t* map_new;     // [tile_cap]  pool index reserved for a new leaf
  uint32_t* tiles;      // [tile_cap * 512 * 2]  (miss, hit) per voxel
  uint32_t* hit_mask;   // [tile_cap * 16]  voxels holding a hit
  uint32_t* hit_start;  // [tile_cap * 512] first sorted hit of a hit voxel
  uint32_t* ray_hit;    // [max_points]  cube cell of each ray's hit
  uint32_t* keys[2];    // [max_points]  radix sort buffers
  uint32_t* vals[2];
  uint32_t* gap;        // [max_points]  misses before each sorted hit
  uint32_t* radix_hist; // [256 * max_radix_blocks]
  int32_t* scan_blocks; // [cube_cap / kScanChunk + 1]
  FrameCounters* counters;
  uint64_t* part_discover;  // [discover blocks * 4] per-block counter sums
  uint64_t* part_update;    // [kUpdateGrid * 4] voxels, hit, mixed
  unsigned char* cloud; // [max_points * 16] staging for host clouds
  float* filter_rays;   // [max_points * 4]
  float* filter_tmp;    // [max_points * 4]
  int32_t* flags;       // [max_points]
};

constexpr int kScanChunk = 4096;
// Bit of a cube-table entry marking a leaf that holds hit voxels (set after
// the hit sort, cleared with the entry by the update).
constexpr int32_t kSlotHasHits = 1 << 30;
constexpr int kUpdateGrid = 2048;
constexpr int kRadixTile = 2048;
// Voxels with more hits than this in one frame are summed by a wavefront.
constexpr uint32_t kHeavyHits = 256;

// ---- launchers (nm_cast.hip, nm_radix.hip, nm_filter.hip) -------------
hipError_t launch_reset_counters(FrameCounters* c, hipStream_t s);
hipError_t launch_discover(int rounding, const CastParams& p,
                           const CtxBuffers& b, hipStream_t s);
hipError_t launch_compact_leaves(const CastParams& p, const CtxBuffers& b,
                                 int64_t cube_vol, int tile_cap,
                                 hipStream_t s);
hipError_t launch_hit_keys(const CastParams& p, const CtxBuffers& b,
                           hipStream_t s);
hipError_t launch_hit_segments(const CtxBuffers& b, int n, int sorted_buf,
                               hipStream_t s);
hipError_t launch_accumulate(int rounding, const CastParams& p,
                             const CtxBuffers& b, int sorted_buf,
                             hipStream_t s);
hipError_t launch_update(int rounding, const Cube& cube, const CtxBuffers& b,
                         const MapView& m, float hit, float miss,
                         int sorted_buf, int tile_cap, hipStream_t s);
hipError_t launch_map_clear(const MapView& m, hipStream_t s);

// Stable LSD radix sort of (keys, vals) over the low `bits` key bits.
// Sorts buffer 0 -> result in buffer returned via *out_buf (0 or 1).
hipError_t radix_sort_pairs(uint32_t* keys[2], uint32_t* vals[2], int n,
                            int bits, uint32_t* hist, int* out_buf,
                            hipStream_t s);
int radix_max_blocks(int n);
int radix_hist_words(int n);  // histogram buffer size for n keys

// Voxel filter: bins points by voxel, writes one ray per bin into
// b.filter_rays (count -> counters->n_filter_rays).  origin: fractional
// part of the eye, eye - floor(eye), as FP32 (the reference's
// voxelOriginOffset, include/nanomap/sensor/SensorData.h:55).
hipError_t launch_voxel_filter(const CastParams& p, const float origin[3],
                               const CtxBuffers& b, hipStream_t s);

void set_error(const std::string& msg);

}  // namespace nm

#endif  // NM_INTERNAL_H
