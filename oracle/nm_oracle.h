/* nm_oracle.h -- CPU restatement of NanoMap's frustum-sensor mapping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP path in nanomap_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every function restates a reference function; the citation is given at
 * the definition in nm_oracle.cpp.  Rounding mode is an explicit argument:
 * 0 = round-to-nearest-even, 1 = round toward -inf (what the reference
 * executes, because SensorData::updatePose leaves FE_DOWNWARD set,
 * include/nanomap/sensor/SensorData.h:42).
 */
#ifndef NM_ORACLE_H
#define NM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_temp or_temp; /* temporary grid: per-frame accumulation */
typedef struct or_map or_map;   /* persistent log-odds map                */

/* ---- temp grid (Handler::populateTempGrid) ---------------------------- */
or_temp* or_temp_create(void);
void or_temp_destroy(or_temp* t);
void or_temp_clear(or_temp* t);

/* Cast every point of a strided xyz-f32 cloud into the temp grid.
 * R: row-major 3x3 float rotation (sensor optical frame -> world),
 * eye: ray origin in index space (double), res: voxel size (m),
 * max_time: max range / res (float).  Returns the number of valid rays. */
int or_cast_cloud(or_temp* t, const unsigned char* cloud, int n_points,
                  int step, const float R[9], const double eye[3], float res,
                  float max_time, float logodds_hit, float logodds_miss,
                  int rounding);

/* Cast pre-filtered rays: each ray is (dir xyz, length) in index units,
 * relative to eye (the output of or_voxel_filter). */
int or_cast_rays(or_temp* t, const float* rays4, int n_rays,
                 const double eye[3], float max_time, float logodds_hit,
                 float logodds_miss, int rounding);

int64_t or_temp_size(const or_temp* t);
/* Export touched voxels, leaf by leaf (leaves sorted by origin, voxels in
 * z + 8y + 64x order).  Any pointer may be NULL. */
void or_temp_export(const or_temp* t, int32_t* ijk, uint32_t* miss,
                    uint32_t* hit, float* value);
uint64_t or_temp_visits(const or_temp* t);

/* ---- map (Handler::integrateTempGrid, map::Map) ----------------------- */
or_map* or_map_create(float clamp_min, float clamp_max, float hit_thr,
                      float miss_thr);
void or_map_destroy(or_map* m);
void or_map_integrate(or_map* m, const or_temp* t, int rounding);
int64_t or_map_leaf_count(const or_map* m);
/* Export every allocated leaf sorted by origin: origin (3 ints per leaf),
 * 512 values per leaf (z fastest, then y, then x), 16 mask words per leaf
 * (bit v of word v/32 is voxel v's active state). */
void or_map_export(const or_map* m, int32_t* origins, float* values,
                   uint32_t* masks);
/* Look up one voxel: returns 1 if its leaf exists. */
int or_map_get(const or_map* m, int i, int j, int k, float* value,
               int* active);

/* ---- single-ray DDA trace (for known-answer tests) -------------------- */
/* d: rotated point in metres (float).  Writes up to max_out voxels (3 ints
 * each) in visit order; *n_out = number of voxel updates, *hit = 1 if the
 * last written voxel received the hit.  Returns 0 if the ray is skipped. */
int or_dda_trace(const double eye[3], const float d[3], float res,
                 float max_time, int rounding, int32_t* out_ijk, int max_out,
                 int* n_out, int* hit);

/* ---- sequential FP32 fold: v_{i+1} = round(v_i + a), k times ---------- */
float or_fold(float v0, float a, uint32_t k, int rounding);

/* ---- voxel filter (filterCloud.cu voxelFilter + rayFromFilter) ------- */
/* Deterministic restatement: drops points with a NaN coordinate, outside
 * the sensor-frame box (box = min xyz, max xyz in metres; NULL = no box) or
 * without 0 < |p| < max_range (metres), bins the rest by voxel, averages in
 * pixel order, and emits one ray (dir xyz, length) per occupied bin in
 * ascending key order.  Returns the number of rays (written up to
 * max_rays). */
int or_voxel_filter(const unsigned char* cloud, int n_points, int step,
                    const float R[9], const double eye[3], float res,
                    float max_time, float max_range, const float* box,
                    float* rays4, int max_rays, int precision);

/* ---- sensor-simulation cast (generateCloud.cu:20-82, Sensor.h:119-156) - */
typedef struct or_simgrid or_simgrid;
/* Occupancy grid: voxels with value > threshold of the given leaves
 * (origins: 3 ints/leaf, values: 512 floats/leaf).  Topology: a leaf exists
 * iff it holds an active voxel; internal nodes of NanoVDB's 5-4-3 tree
 * (16^3 leaves per lower node, 32^3 lower nodes per upper node) exist iff
 * they hold a leaf. */
or_simgrid* or_simgrid_build(const int32_t* origins, const float* values,
                             int64_t n_leaves, float threshold);
void or_simgrid_destroy(or_simgrid* g);
int64_t or_simgrid_active_count(const or_simgrid* g);
/* Cast a W x H frustum (type 0), a W x H-beam spherical scanner (type 1)
 * or W explicit ray directions (type 2, rays: W x 3 floats, H = 1).  Per
 * pixel (row-major): first active voxel ijk (INT32_MIN if
 * none in range), hit distance t*res in metres (-1 if none), and the
 * simulated sensor-frame point (NaN if no return).  hierarchical:
 * 1 = NanoVDB HDDA with node dims 1/8/128/4096, 0 = plain dim-1 DDA.
 * Returns the number of returns. */
int or_sim_cast(const or_simgrid* g, const float pos[3], const float R[9],
                float res, int width, int height, float hfov_deg,
                float vfov_deg, float min_range, float max_range, int type,
                const float* rays, int hierarchical, int32_t* ijk,
                float* dist, float* xyz);

/* ---- soft expected-depth march (include/nanomap_amd_soft.h), FP64 ------
 * grid: dense [nx][ny][nz]; cams: n_cams x 12 doubles (row-major optical ->
 * world rotation, then position).  Outputs accumulate (+=). */
void or_soft_forward(const double* grid, int nx, int ny, int nz,
                     const double origin[3], double res, double background,
                     const double* cams, int n_cams, int width, int height,
                     double hfov_deg, double vfov_deg, double step,
                     int n_samples, double* depth);
void or_soft_backward(const double* grid, int nx, int ny, int nz,
                      const double origin[3], double res, double background,
                      const double* cams, int n_cams, int width, int height,
                      double hfov_deg, double vfov_deg, double step,
                      int n_samples, const double* grad_depth,
                      double* grad_pose, double* grad_logodds);

#ifdef __cplusplus
}
#endif
#endif
