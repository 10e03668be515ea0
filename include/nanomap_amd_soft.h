/* nanomap_amd_soft.h -- differentiable expected-depth ray march.
 *
 * No counterpart in the reference: NanoMap's casts are hard (first-hit
 * voxel DDA, src/nanomap/kernels/generateCloud.cu:20-82, and the hit/miss
 * cast).  This is the soft form of the sensor model, for gradient-based
 * scan-to-map pose refinement and map optimisation:
 *
 *   sample i of a ray at distance d_i = (i + 0.5) * step, i < n_samples;
 *   l_i = quadratic B-spline interpolation of the dense log-odds grid at
 *         the sample (C1 in position, so the depth is C1 in the pose);
 *   p_i = sigmoid(l_i),  w_i = p_i * prod_{j<i} (1 - p_j),
 *   D   = sum_i w_i d_i        (expected termination depth, metres).
 *
 * The adjoint recomputes each ray's march from its far end (no stored
 * traces; the forward pass keeps one number per ray, the total
 * log-transmittance) and returns dL/dpose and dL/dlog-odds for upstream
 * gradients dL/dD.  The pose
 * gradient is taken with respect to a world-frame perturbation
 * xi = (rho, omega) of each camera: sample x = exp(omega^) (x0 - eye) +
 * eye + rho, i.e. translation rho and a rotation omega about the camera
 * centre, both in world axes, at xi = 0.
 *
 * Every array is device memory; nothing here allocates or synchronizes.
 * The _f64 entry points run the same march in FP64 (for finite-difference
 * checks).
 */
#ifndef NANOMAP_AMD_SOFT_H
#define NANOMAP_AMD_SOFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Dense log-odds grid: value of voxel (i, j, k) at [(i * ny + j) * nz + k];
 * voxel (i, j, k) covers origin + res * [i, i+1) x [j, j+1) x [k, k+1). */
typedef struct nm_soft_grid {
  int nx, ny, nz;
  double origin[3]; /* world position (m) of the corner of voxel (0,0,0) */
  float res;        /* voxel size (m) */
  float background; /* log-odds outside the grid */
} nm_soft_grid;

/* Pinhole camera: pixel (u, v) looks along R * normalize(x, y, 1) with
 * x = (2 (u + 0.5) / width - 1) tan(hfov/2), y = (2 (v + 0.5) / height - 1)
 * tan(vfov/2); R is optical -> world (row-major). */
typedef struct nm_soft_camera {
  double rotation[9];
  double position[3];
} nm_soft_camera;

typedef struct nm_soft_params {
  int width, height;
  float hfov_deg, vfov_deg;
  float step;    /* metres between samples */
  int n_samples; /* samples per ray */
} nm_soft_params;

/* depth[c * height * width + v * width + u] for camera c, and (optional,
 * same layout) the ray's total log-transmittance sum_i log(1 - p_i), which
 * the adjoint needs. */
int nm_soft_forward(const nm_soft_grid* grid, const float* logodds,
                    const nm_soft_camera* cams, int n_cams,
                    const nm_soft_params* params, float* depth,
                    double* log_transmittance, void* stream);

/* Adjoint of nm_soft_forward for upstream gradients grad_depth (same
 * layout as depth), given the forward's log_transmittance.  Each output is
 * optional (null = not computed) and is accumulated into (+=):
 *   grad_pose     [n_cams * 6]  (rho, omega) per camera, FP64;
 *   grad_ray_pose [rays * 6]    per-ray (rho, omega) of dD/dxi (ignores
 *                               grad_depth), for Gauss-Newton;
 *   grad_logodds  [nx * ny * nz] dL/dlog-odds. */
int nm_soft_backward(const nm_soft_grid* grid, const float* logodds,
                     const nm_soft_camera* cams, int n_cams,
                     const nm_soft_params* params,
                     const double* log_transmittance,
                     const float* grad_depth, double* grad_pose,
                     float* grad_ray_pose, float* grad_logodds, void* stream);

/* The same march in FP64 (grid values in FP64). */
int nm_soft_forward_f64(const nm_soft_grid* grid, const double* logodds,
                        const nm_soft_camera* cams, int n_cams,
                        const nm_soft_params* params, double* depth,
                        double* log_transmittance, void* stream);
int nm_soft_backward_f64(const nm_soft_grid* grid, const double* logodds,
                         const nm_soft_camera* cams, int n_cams,
                         const nm_soft_params* params,
                         const double* log_transmittance,
                         const double* grad_depth, double* grad_pose,
                         double* grad_logodds, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NANOMAP_AMD_SOFT_H */
