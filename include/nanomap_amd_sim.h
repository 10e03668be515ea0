/* nanomap_amd_sim.h -- sensor-simulation cast of the MI355X NanoMap path.
 *
 * Replaces
 *   extern "C" void generateCloud(nanomap::gpu::SensorBucket&,
 *       nanovdb::GridHandle<nanovdb::CudaDeviceBuffer>&, cudaStream_t)
 *   -- src/nanomap/kernels/generateCloud.cu:84, declared
 *      include/nanomap/handler/SimHandler.h:48,
 * whose grid is the agent map converted to NanoVDB (SimHandler.cpp:22-23)
 * and whose per-pixel ray is gpu::Sensor::getRay (include/nanomap/gpu/
 * Sensor.h:119-156: the frustum branch 121-133 and the spherical laser
 * branch 134-155, which the reference only runs on its CPU path).
 */
#ifndef NANOMAP_AMD_SIM_H
#define NANOMAP_AMD_SIM_H

#include <stdint.h>

#include "nanomap_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nm_simgrid nm_simgrid; /* device occupancy grid (5-4-3 tree) */

#define NM_SENSOR_FRUSTUM 0 /* pinhole depth camera (SharedParameters _type 0) */
#define NM_SENSOR_LASER 1   /* spherical scanner (_type 1)                   */
#define NM_SENSOR_RAYS 2    /* explicit ray directions (an agent's
                               observation sphere, Agent.h:186-195)       */

/* One simulated sensor.  rotation = normalized(q) * F (row-major).
 * Frustum: width x height pixels; hfov/vfov are independent (the reference
 * derives hfov from the pixel aspect, include/nanomap/sensor/
 * FrustumData.h:24 -- pass 2*atan(aspect*tan(vfov/2)) to reproduce it);
 * the range limits are z-depths (Sensor.h:129-130).
 * Laser: width x height beams over hfov x vfov degrees (LaserData.h:23-26:
 * width = (int)(hfov / angular step + 1)); beam (w, h) has yaw
 * u = -hfov/2 + w hfov/width and pitch v = -vfov/2 + h vfov/height and
 * direction (sin u cos v, sin v, cos u cos v) in the sensor frame
 * (Sensor.h:134-155); the range limits are distances.
 * Rays: `width` directions (3 floats each, host memory, read during the
 * call) in the sensor frame, height 1; the range limits are distances.
 * Laser and ray directions are normalised, rotated by `rotation` and
 * normalised again, as the laser branch of getRay does. */
typedef struct nm_sim_sensor {
  float position[3];
  float rotation[9];
  float grid_res;
  int width, height;
  float hfov_deg, vfov_deg;
  float min_range, max_range;
  int type;           /* NM_SENSOR_* */
  const float* rays;  /* NM_SENSOR_RAYS: width x 3 floats */
} nm_sim_sensor;

/* Occupancy grid of the voxels whose value exceeds `threshold`, from host
 * arrays (origins: 3 ints/leaf, values: 512 floats/leaf) or from a map. */
int nm_simgrid_from_leaves(int device, const int32_t* origins,
                           const float* values, int64_t n_leaves,
                           float threshold, nm_simgrid** out);
int nm_simgrid_from_map(nm_map* map, float threshold, nm_simgrid** out);
int nm_simgrid_destroy(nm_simgrid* grid);
/* active voxel count, bbox (min xyz, max xyz), node counts (upper, lower,
 * leaf); any pointer may be NULL. */
int nm_simgrid_info(nm_simgrid* grid, int64_t* active, int32_t bbox[6],
                    int32_t counts[3]);

/* generateCloud for n_sensors sensors in one launch.  Outputs are device
 * arrays indexed by pixel (sensor-major, row-major within a sensor):
 * ijk (3 int32: first active voxel, INT32_MIN if no return), dist (metres
 * along the ray, -1 if none), xyz (sensor-frame point default_dir * dist,
 * NaN if none).  Any output may be NULL.  *returns receives the number of
 * returns.  Synchronizes `stream`.  The sensors' descriptors live in a
 * device buffer owned by the grid, allocated when a call needs more room
 * than any before it (never per call). */
int nm_sim_cast(nm_simgrid* grid, const nm_sim_sensor* sensors,
                int n_sensors, int32_t* ijk, float* dist, float* xyz,
                int* returns, void* stream);

/* nm_sim_cast without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The sensor descriptors and the return counter go to a fixed device
 * buffer of NM_SIM_STATIC_BYTES per GPU, held until the process ends;
 * sensor sets that do not fit (about 300 frustum sensors, fewer lasers:
 * each laser needs 8 bytes per beam column and row) are rejected with
 * NM_EINVAL.  Calls on one GPU share that buffer: calls that overlap on
 * one GPU (other threads, other streams) race.  Same results as
 * nm_sim_cast. */
#define NM_SIM_STATIC_BYTES 65536
int nm_sim_cast_static(nm_simgrid* grid, const nm_sim_sensor* sensors,
                       int n_sensors, int32_t* ijk, float* dist, float* xyz,
                       int* returns, void* stream);


#ifdef __cplusplus
}
#endif
#endif /* NANOMAP_AMD_SIM_H */
