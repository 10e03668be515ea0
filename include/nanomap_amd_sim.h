/* nanomap_amd_sim.h -- sensor-simulation cast of the MI355X NanoMap path.
 *
 * Replaces
 *   extern "C" void generateCloud(nanomap::gpu::SensorBucket&,
 *       nanovdb::GridHandle<nanovdb::CudaDeviceBuffer>&, cudaStream_t)
 *   -- src/nanomap/kernels/generateCloud.cu:84, declared
 *      include/nanomap/handler/SimHandler.h:48,
 * whose grid is the agent map converted to NanoVDB (SimHandler.cpp:22-23)
 * and whose per-pixel ray is gpu::Sensor::getRay (include/nanomap/gpu/
 * Sensor.h:213-227).
 */
#ifndef NANOMAP_AMD_SIM_H
#define NANOMAP_AMD_SIM_H

#include <stdint.h>

#include "nanomap_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nm_simgrid nm_simgrid; /* device occupancy grid (5-4-3 tree) */

/* One simulated frustum sensor.  rotation = normalized(q) * F (row-major);
 * hfov/vfov are independent (the reference derives hfov from the pixel
 * aspect, include/nanomap/sensor/FrustumData.h:24 -- pass
 * 2*atan(aspect*tan(vfov/2)) to reproduce it). */
typedef struct nm_sim_sensor {
  float position[3];
  float rotation[9];
  float grid_res;
  int width, height;
  float hfov_deg, vfov_deg;
  float min_range, max_range;
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
 * returns.  Synchronizes `stream`. */
int nm_sim_cast(nm_simgrid* grid, const nm_sim_sensor* sensors,
                int n_sensors, int32_t* ijk, float* dist, float* xyz,
                int* returns, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* NANOMAP_AMD_SIM_H */
