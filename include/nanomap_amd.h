/* nanomap_amd.h -- C ABI of the MI355X-native NanoMap frustum-sensor path.
 *
 * This is the drop-in boundary.  The reference drives its GPU path through
 * three extern "C" launchers that take C++ objects by reference:
 *
 *   void filterCloud(nanomap::gpu::SensorBucket&, cudaStream_t)
 *        -- src/nanomap/kernels/filterCloud.cu:476, declared
 *           include/nanomap/handler/Handler.h:46
 *   void frustumCastCloud(nanomap::gpu::SensorBucket&, cudaStream_t,
 *                         cudaStream_t)
 *        -- src/nanomap/kernels/frustumCastCloud.cu:311, Handler.h:49
 *   void generateCloud(nanomap::gpu::SensorBucket&,
 *                      nanovdb::GridHandle<CudaDeviceBuffer>&, cudaStream_t)
 *        -- src/nanomap/kernels/generateCloud.cu:84, SimHandler.h:48
 *
 * and folds the result into the map on the host
 * (Handler::voxelUpdateFromFrustumBuffer, src/nanomap/handler/Handler.cpp:134).
 * Here every entry point takes plain pointers, sizes and POD structs, and
 * returns an int status.  The map stays resident in device memory.
 *
 * Conventions
 *  - Every function returns NM_OK (0) or an NM_E* code; nm_last_error()
 *    returns a message for the calling thread's last failure.
 *  - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *    Calls enqueue work and return; nothing blocks unless documented.
 *  - A context owns its device buffers and serves one stream at a time;
 *    distinct contexts may be used from distinct threads.
 *  - Coordinates are voxel indices (index space = world metres / voxel
 *    size).  Leaves are 8^3 voxels with origins at multiples of 8; the
 *    voxel order inside a leaf is z + 8*y + 64*x (frustumCastCloud.cu:215).
 *  - rounding: NM_ROUND_DOWN reproduces the reference host path, which runs
 *    with FE_DOWNWARD left set (include/nanomap/sensor/SensorData.h:42);
 *    NM_ROUND_NEAREST is IEEE default rounding.
 */
#ifndef NANOMAP_AMD_H
#define NANOMAP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NM_OK 0
#define NM_EINVAL 1     /* bad argument                               */
#define NM_ENOMEM 2     /* allocation failed                          */
#define NM_EOVERFLOW 3  /* a capacity was exceeded; result incomplete */
#define NM_EHIP 4       /* HIP runtime error                          */
#define NM_ENODEV 5     /* no usable gfx950 device                    */
#define NM_ESTATE 6     /* call out of order (e.g. update w/o cast)   */

#define NM_ROUND_NEAREST 0
#define NM_ROUND_DOWN 1

#define NM_FILTER_NONE 0  /* every point is a ray: Handler::populateTempGrid */
#define NM_FILTER_VOXEL 1 /* voxel-centroid filter: filterCloud.cu:84,279   */
/* The voxel filter's lower-memory variants (PrecisionType 1-3,
 * filterCloud.cu:134-277, 319-434), with the sums taken in ascending pixel
 * order instead of by atomics: */
#define NM_FILTER_VOXEL_HALF 2   /* centroid from FP16 sums and FP16 count  */
#define NM_FILTER_VOXEL_COUNT 3  /* voxel centre, int count                 */
#define NM_FILTER_VOXEL_COUNT8 4 /* voxel centre, 8-bit count: a voxel whose
                                    count is a multiple of 256 wraps to 0
                                    and emits no ray                        */

typedef struct nm_ctx nm_ctx; /* per-sensor device workspace             */
typedef struct nm_map nm_map; /* device-resident sparse log-odds map     */

/* Context sizing.  Replaces SensorAllocator's buffer sizing
 * (include/nanomap/allocator/SensorAllocator.h:33-108). */
typedef struct nm_ctx_config {
  int device;          /* HIP device ordinal                              */
  int max_points;      /* largest cloud (pixels) per frame                */
  float max_time;      /* largest max_range / voxel size, in voxels       */
  int tile_capacity;   /* leaf tiles per frame; 0 = the worst case        */
} nm_ctx_config;

/* Per-frame parameters, precomputed on the host.  Replaces the device copy
 * of gpu::Sensor + gpu::GridData (include/nanomap/gpu/Sensor.h:189-211,
 * GridData.h:38-62; uploaded at SensorAllocator.h:120-121). */
typedef struct nm_frame {
  float rotation[9];    /* row-major R = normalized(q) * F (Handler.cpp:41) */
  double eye[3];        /* sensor origin in index space (Handler.cpp:34)    */
  float grid_res;       /* voxel size, metres                               */
  float max_time;       /* max range / grid_res (Handler.cpp:39)            */
  float logodds_hit;    /* added on the end voxel of a ray (|.| <= 1e27)    */
  float logodds_miss;   /* added on every voxel a ray passes (|.| <= 1e27)  */
  int rounding;         /* NM_ROUND_*                                       */
  int filter;           /* NM_FILTER_*                                      */
  int width, height;    /* cloud layout (pcl_width, pcl_height)             */
  /* Voxel filter only (NM_FILTER_VOXEL*; the unfiltered cast follows the
   * reference's CPU path, which has neither check).  Sensor-frame bounds
   * in metres: a point with a coordinate below box_min or above box_max
   * is dropped (filterCloud.cu:100-111); the reference's bounds are the
   * frustum's worldMin/worldMax (FrustumData::calculateSensorBounds,
   * include/nanomap/sensor/FrustumData.h:50-83), see nm_frustum_box.
   * All six zero: no bound.  max_range: a point is binned only if
   * 0 < |p| < max_range (metres, the sensor's maxRange, filterCloud.cu:112);
   * 0 takes max_time * grid_res. */
  float box_min[3], box_max[3];
  float max_range;
} nm_frame;

/* Counters of the most recent frame of a context. */
typedef struct nm_frame_stats {
  int64_t points;        /* input points                                   */
  int64_t rays;          /* rays cast (valid points or filter bins)        */
  int64_t visits;        /* voxel updates (misses + hits)                  */
  int64_t voxels;        /* unique voxels touched                          */
  int64_t leaves;        /* 8^3 leaves touched                             */
  int64_t hit_voxels;    /* voxels with at least one hit                   */
  int64_t mixed_voxels;  /* voxels with both hits and misses               */
  int64_t map_leaves;    /* leaves in the map after the update             */
  int32_t errors;        /* internal consistency errors (0 expected)       */
  int32_t overflow;      /* nonzero if a capacity was exceeded             */
} nm_frame_stats;

const char* nm_last_error(void);
int nm_version(void);

/* ---- context ---------------------------------------------------------- */
int nm_ctx_create(const nm_ctx_config* cfg, nm_ctx** out);
int nm_ctx_destroy(nm_ctx* ctx);

/* ---- map: map::Map (include/nanomap/map/OccupancyMap.h:35-61) ---------- */
/* clamp_min/clamp_max: emptyClampingThreshold/occupiedClampingThreshold
 * (finite, |.| <= 1e27); hit_thr/miss_thr: logOddsHitThreshold/
 * logOddsMissThreshold. */
int nm_map_create(int device, float grid_res, float clamp_min,
                  float clamp_max, float hit_thr, float miss_thr,
                  int leaf_capacity, nm_map** out);
int nm_map_destroy(nm_map* map);
int nm_map_clear(nm_map* map, void* stream);
/* Number of allocated leaves (synchronizes the map's last stream). */
int nm_map_leaf_count(nm_map* map, int64_t* count);
/* Copy all leaves to host arrays sized leaf_count: origins (3 ints/leaf),
 * values (512 floats/leaf), masks (16 words/leaf, bit v = voxel v active).
 * Leaves are in allocation order.  Synchronous. */
int nm_map_download(nm_map* map, int32_t* origins, float* values,
                    uint32_t* masks);
/* Device pointers of the map arrays (for kernels and zero-copy readers).
 * The leaf pool may hold holes: slots reserved for a leaf that was never
 * inserted (no voxel of it changed); a hole's origin x is INT32_MIN and its
 * values and mask are zero.  Holes are skipped by nm_map_leaf_count and
 * nm_map_download. */
int nm_map_device_arrays(nm_map* map, void** values, void** masks,
                         void** origins, void** hash_keys, void** hash_vals,
                         int64_t* hash_capacity);

/* Pool capacity in leaves. */
int nm_map_capacity(nm_map* map, int64_t* leaf_capacity);
/* The parameters the map was created with (any pointer may be NULL). */
int nm_map_params(nm_map* map, float* grid_res, float* clamp_min,
                  float* clamp_max, float* hit_thr, float* miss_thr);
/* Grows the map to hold at least leaf_capacity leaves (pool and hash),
 * keeping every leaf, value and mask.  Synchronous; a no-op when the map
 * already holds that many.  The reference's OpenVDB grid grows per leaf
 * (OccupancyMap.h:44, 69); this pool grows on request or by auto-grow. */
int nm_map_reserve(nm_map* map, int64_t leaf_capacity);
/* Auto-grow (off by default): nm_update_map first waits for the cast and,
 * when the map's free leaves are fewer than the leaves the cast touched,
 * grows the map (at least doubling) before folding the cast in, so no
 * update is lost to a full map.  Off, the update stays asynchronous and a
 * full map is reported as overflow in the frame statistics. */
int nm_map_set_auto_grow(nm_map* map, int enable);

/* Batched point queries on the device: for n voxels (ijk, 3 int32 each)
 * the log-odds value (0 where the leaf is absent: the grid background) and
 * the active state (1/0).  Replaces reads through map::Map::occupiedGrid()
 * / occAccessor() (include/nanomap/map/OccupancyMap.h:54,60).
 * on_device = 1: ijk, values and active are device pointers and the call
 * is asynchronous on `stream`; 0: host arrays, staged through the map's
 * query buffer, and the call returns when the results are in place. */
int nm_map_query(nm_map* map, const int32_t* ijk, int64_t n, float* values,
                 uint8_t* active, int on_device, void* stream);
/* Dense copy of a box of the map into device memory: out[(i * ny + j) * nz
 * + k] = log-odds of voxel (lo[0] + i, lo[1] + j, lo[2] + k), `fill` where
 * the voxel's leaf is absent; asynchronous on `stream`.  No counterpart in
 * the reference: the dense grid the soft march reads (nanomap_amd_soft.h),
 * built from the device map without a host round trip. */
int nm_map_dense(nm_map* map, const int32_t lo[3], int nx, int ny, int nz,
                 float fill, float* out, void* stream);

/* Snapshot and restore.  nm_map_save writes the map to `path`: a 64-byte
 * header ("NMAPLEAF", format version 1, leaf count, voxel size, clamps,
 * thresholds), then per leaf its origin (3 int32), 512 FP32 log-odds
 * (voxel z + 8y + 64x) and 16 active-mask words, little-endian, leaves in
 * nm_map_download order.  nm_map_load creates a map from such a file on
 * `device` with room for max(leaf_capacity, leaves in the file) leaves, or
 * with leaf_capacity 0 for max(2 x those leaves, 2^18); the restored map
 * continues exactly where the saved one stood.  The
 * reference never saves its map; it only loads .vdb grids for simulation,
 * SimManager.h:171-213. */
int nm_map_save(nm_map* map, const char* path);
int nm_map_load(int device, const char* path, int64_t leaf_capacity,
                nm_map** out);

/* NanoVDB export and import.  The reference hands its map to the GPU as a
 * NanoVDB grid (nanovdb::openToNanoVDB, src/nanomap/handler/
 * SimHandler.cpp:22-23) and its simulator reads grids from files
 * (SimManager.h:196-210).  These write a map's leaves as a NanoVDB 32.4
 * FloatGrid (breadth-first root/upper/lower/leaf nodes, every voxel in a
 * leaf, background 0, no checksum) and read FloatGrids back.
 * nm_nvdb_build: the grid buffer of n host leaves (nm_map_download
 *   layout) into `buffer` (*bytes in: its size, out: the grid's size; a
 *   null buffer only returns the size).  Host only.
 * nm_nvdb_write: the same grid as an uncompressed .nvdb file (16-byte
 *   header, 176-byte grid record, name, grid).  Host only.
 * nm_nvdb_read: the leaves of the FloatGrid named grid_name (the first
 *   FloatGrid for "" or NULL) of an uncompressed .nvdb file: leaf nodes as
 *   they are, active tiles expanded into leaves (at most 2^22 of them);
 *   with null arrays *n_leaves returns the count, else the arrays hold
 *   *n_leaves entries.  Host only.
 * nm_map_write_nvdb: nm_map_download, then nm_nvdb_write.
 * nm_map_load_nvdb: a map from nm_nvdb_read's leaves, with the grid's
 *   voxel size and the given clamps and thresholds (a grid file carries
 *   none); leaf_capacity as for nm_map_load. */
int nm_nvdb_build(const int32_t* origins, const float* values,
                  const uint32_t* masks, int64_t n_leaves, double voxel_size,
                  const char* grid_name, void* buffer, uint64_t* bytes);
int nm_nvdb_write(const char* path, const int32_t* origins,
                  const float* values, const uint32_t* masks, int64_t n_leaves,
                  double voxel_size, const char* grid_name);
int nm_nvdb_read(const char* path, const char* grid_name, int64_t* n_leaves,
                 int32_t* origins, float* values, uint32_t* masks,
                 double* voxel_size);
int nm_map_write_nvdb(nm_map* map, const char* path, const char* grid_name);
int nm_map_load_nvdb(int device, const char* path, const char* grid_name,
                     float clamp_min, float clamp_max, float hit_thr,
                     float miss_thr, int64_t leaf_capacity, nm_map** out);

/* ---- plain frame path --------------------------------------------------
 * nm_cast: filterCloud + frustumCastCloud (filterCloud.cu:476,
 * frustumCastCloud.cu:311).  `cloud` holds n_points xyz-f32 points with
 * byte stride `step` (PointCloud::updateData, include/nanomap/gpu/
 * PointCloud.h:121-134).  cloud_on_device = 0 means a host pointer (copied
 * with one H2D inside the call).  Produces per-voxel (miss, hit) counts in
 * the context's leaf tiles. */
int nm_cast(nm_ctx* ctx, const nm_frame* frame, const void* cloud,
            int n_points, int step, int cloud_on_device, void* stream);
/* Handler::voxelUpdateFromFrustumBuffer / integrateTempGrid
 * (Handler.cpp:134-182, 88-122) on the device: folds the counts of the
 * last nm_cast into `map` and clears the tiles. */
int nm_update_map(nm_ctx* ctx, nm_map* map, void* stream);
/* Manager::insertPointCloud's device work (Handler.cpp:237-284):
 * nm_cast then nm_update_map. */
int nm_frame_submit(nm_ctx* ctx, nm_map* map, const nm_frame* frame,
                    const void* cloud, int n_points, int step,
                    int cloud_on_device, void* stream);

/* ---- depth-image ingestion ---------------------------------------------
 * A pinhole depth camera: pixel (u, v) of a width x height z-depth image
 * (metres, row-major, NaN = no return) back-projects to the sensor-frame
 * point (d ax[u], d ay[v], d), where ax[u] = ((u + 0.5) - cx) / fx and
 * ay[v] = -((v + 0.5) - cy) / fy are each rounded once to FP32 and each
 * coordinate is one FP32 product (optical frame: x right, y up, z
 * forward, as the sample sensor's FrameTransform expects,
 * config/sensors/640x480.txt:22-25).  The reference takes xyz clouds only
 * (PointCloud::updateData, include/nanomap/gpu/PointCloud.h:121-134); a
 * depth image is a third of the bytes to upload, and the conversion runs
 * inside the cast's ray set-up. */
typedef struct nm_depth_camera {
  int width, height;
  float fx, fy, cx, cy;
} nm_depth_camera;
/* fx = (W/2) / tan(hfov/2), fy = (H/2) / tan(vfov/2), cx = W/2, cy = H/2. */
int nm_depth_camera_from_fov(int width, int height, float hfov_deg,
                             float vfov_deg, nm_depth_camera* out);
/* Host back-projection of a depth image into xyz (3 floats per pixel):
 * the points the device computes, bit for bit. */
int nm_depth_to_cloud(const nm_depth_camera* cam, const float* depth,
                      float* xyz);
/* nm_cast / nm_frame_submit of a depth image of cam->width x cam->height
 * floats: a device pointer when depth_on_device, else host memory copied
 * with one H2D inside the call. */
int nm_cast_depth(nm_ctx* ctx, const nm_frame* frame,
                  const nm_depth_camera* cam, const float* depth,
                  int depth_on_device, void* stream);
int nm_frame_submit_depth(nm_ctx* ctx, nm_map* map, const nm_frame* frame,
                          const nm_depth_camera* cam, const float* depth,
                          int depth_on_device, void* stream);

/* Stats of the last frame.  Synchronizes the context's stream. */
int nm_frame_stats_get(nm_ctx* ctx, nm_frame_stats* out);

/* Counts of the last nm_cast (call before nm_update_map).  Synchronous.
 * *n_leaves receives the leaf count; arrays sized max_leaves: origins
 * (3 ints/leaf) and miss/hit (512 per leaf), leaves in tile order. */
int nm_counts_download(nm_ctx* ctx, int64_t max_leaves, int64_t* n_leaves,
                       int32_t* origins, uint32_t* miss, uint32_t* hit);

/* Rays produced by the voxel filter of the last nm_cast with a voxel
 * filter (NM_FILTER_VOXEL*): (dir xyz, length) per ray, index units.
 * Synchronous. */
int nm_filter_rays_download(nm_ctx* ctx, int max_rays, int* n_rays,
                            float* rays4);

/* ---- stage timing -------------------------------------------------------
 * When enabled, nm_cast/nm_update_map record a HIP event pair around each
 * stage on the call's stream.  nm_ctx_stage_times synchronizes and returns,
 * per stage, the summed milliseconds and the number of timed launches since
 * the last reset.  Stage ids: NM_STAGE_*. */
#define NM_STAGE_FILTER 0     /* voxel filter (bin sort + centroids)       */
#define NM_STAGE_DISCOVER 1   /* ray set-up + DDA, leaf discovery          */
#define NM_STAGE_COMPACT 2    /* leaf table prefix sum                     */
#define NM_STAGE_HITSORT 3    /* hit keys, radix sort, hit segments        */
#define NM_STAGE_ACCUMULATE 4 /* DDA again, miss counts into leaf tiles    */
#define NM_STAGE_UPDATE 5     /* temp values, clamped log-odds map update  */
#define NM_STAGE_COUNT 6
int nm_ctx_set_timing(nm_ctx* ctx, int enable);
/* Times only the stages whose bit (1 << NM_STAGE_*) is set; 0 turns timing
 * off.  Fewer event pairs disturb the timed work less.  No counterpart in
 * the reference (which prints host timers, Manager.h:65). */
int nm_ctx_set_timing_mask(nm_ctx* ctx, unsigned mask);
int nm_ctx_stage_times(nm_ctx* ctx, float* ms, int64_t* launches);
int nm_ctx_reset_timing(nm_ctx* ctx);

/* ---- device buffers ------------------------------------------------------
 * For host code that holds no HIP headers: device memory for clouds the
 * library fills and reads back (e.g. simulated clouds fed to
 * nm_frame_submit with cloud_on_device = 1). */
int nm_device_alloc(int device, int64_t bytes, void** out);
int nm_device_free(int device, void* ptr);
/* Synchronous device-to-host copy. */
int nm_device_download(int device, void* dst, const void* src, int64_t bytes);

/* ---- host-side helpers (no device work) ------------------------------- */
/* fold(v, a, k): v <- round(v + a), k times, in `rounding` (FP32). */
float nm_fold_host(float v, float a, uint32_t k, int rounding);

#ifdef __cplusplus
}
#endif
#endif /* NANOMAP_AMD_H */
