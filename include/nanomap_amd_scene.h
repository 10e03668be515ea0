/* nanomap_amd_scene.h -- seeded synthetic workloads and host-side frame
 * set-up for the MI355X NanoMap path (benchmarks, tests, examples).
 *
 * No device work.  These functions have no counterpart in the reference
 * (which is driven by external ROS nodes, README.md:59-67), except
 * nm_frame_from_pose, which restates the reference's per-frame host
 * set-up: R = normalized(q) * F (src/nanomap/handler/Handler.cpp:41),
 * eye = worldToIndex(position) (Handler.cpp:34) and
 * max_time = maxRange / gridRes (Handler.cpp:39).
 */
#ifndef NANOMAP_AMD_SCENE_H
#define NANOMAP_AMD_SCENE_H

#include <stdint.h>

#include "nanomap_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nm_scene nm_scene;

/* Pinhole depth camera.  Optical frame: x right, y up, z forward (the
 * frame the reference's FrameTransform maps from, config/sensors/
 * 640x480.txt:22-25).  Pixel (u, v) looks along
 * ((u + 0.5 - W/2)/fx, -(v + 0.5 - H/2)/fy, 1), fx = (W/2)/tan(hfov/2). */
typedef struct nm_camera {
  int width, height;
  float hfov_deg, vfov_deg;
  float max_range;   /* z-depth cap, metres                      */
  float noise_k;     /* depth noise sigma = noise_k * depth^2     */
  float dropout;     /* probability that a pixel becomes NaN      */
  float frame[9];    /* row-major F: optical frame -> body frame  */
} nm_camera;

/* Closed room [0,world] with n_boxes axis-aligned boxes (centres uniform,
 * sides uniform in [box_min, box_max] per axis) and n_cylinders vertical
 * cylinders of radius uniform in [cyl_rmin, cyl_rmax]. */
int nm_scene_create(uint64_t seed, const double world[3], int n_boxes,
                    double box_min, double box_max, int n_cylinders,
                    double cyl_rmin, double cyl_rmax, nm_scene** out);
int nm_scene_destroy(nm_scene* s);

/* Render one depth frame at pose (position, quaternion w,x,y,z) into
 * width*height xyz FP32 points (stride 12 bytes), row-major pixels. */
int nm_scene_render(const nm_scene* s, const nm_camera* cam,
                    const float pos[3], const float quat[4], uint64_t seed,
                    float* xyz);

/* A pose uniform in free space (clearance metres from every primitive),
 * yaw uniform in [-180, 180) deg, pitch uniform in +-max_pitch_deg. */
int nm_scene_sample_pose(const nm_scene* s, uint64_t seed,
                         float max_pitch_deg, float clearance, float pos[3],
                         float quat[4]);

/* Smooth random walk of n poses: step_m metres along the heading per
 * frame, heading and pitch perturbed by up to step_deg per frame. */
int nm_scene_walk(const nm_scene* s, uint64_t seed, int n, float step_m,
                  float step_deg, float max_pitch_deg, float clearance,
                  float* pos_out, float* quat_out);

/* Per-frame host set-up in the given rounding mode (see above); also sets
 * out->max_range = max_range and a zero (unbounded) filter box. */
int nm_frame_from_pose(const float pos[3], const float quat[4],
                       const float F[9], float grid_res, float max_range,
                       float logodds_hit, float logodds_miss, int rounding,
                       nm_frame* out);

/* Sensor-frame bounds of a frustum sensor, the voxel filter's point box
 * (nm_frame.box_min/box_max): FrustumData::calculateSensorBounds
 * (include/nanomap/sensor/FrustumData.h:50-83) -- aspect = width/height,
 * half height h = range * tan(vfov/2), half width = aspect * h (the
 * horizontal FOV is derived from the aspect, FrustumData.h:24), at
 * min_range and max_range, plus the origin; box = per-axis min and max of
 * those nine points.  Evaluated in `rounding` (the reference constructs
 * its sensors with FE_DOWNWARD left set by an earlier one). */
int nm_frustum_box(int width, int height, float vfov_deg, float min_range,
                   float max_range, int rounding, float box_min[3],
                   float box_max[3]);

#ifdef __cplusplus
}
#endif
#endif /* NANOMAP_AMD_SCENE_H */
