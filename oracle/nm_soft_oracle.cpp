// nm_soft_oracle.cpp -- CPU restatement (FP64) of the soft expected-depth
// ray march of include/nanomap_amd_soft.h.
//
// TEST INFRASTRUCTURE ONLY (see nm_oracle.h).  The soft march has no
// counterpart in the reference; this restates its model independently of
// the HIP kernels: samples at (i + 1/2) step along each pixel ray, log-odds
// from a quadratic B-spline of the dense grid, p = sigmoid(l),
// w_i = p_i prod_{j<i} (1 - p_j), D = sum_i w_i d_i.  The adjoint here uses
// explicit suffix sums over stored per-sample arrays (the kernels use the
// D - D_<=k identity and no stored arrays), so the two formulations check
// each other.
#include <cmath>
#include <vector>

#include "nm_oracle.h"

namespace {

struct Grid {
  const double* g;
  long nx, ny, nz;
  double o[3];
  double res;
  double bg;
  double at(long i, long j, long k) const {
    if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return bg;
    return g[(i * ny + j) * nz + k];
  }
};

// Quadratic B-spline of the grid at voxel coordinates q: value, gradient
// (per voxel), and the tap base / weights for the log-odds adjoint.
double spline(const Grid& G, const double q[3], double grad[3], long base[3],
              double w[3][3]) {
  double dw[3][3];
  for (int a = 0; a < 3; ++a) {
    const double f = std::floor(q[a]);
    base[a] = (long)f;
    const double s = q[a] - f - 0.5;
    w[a][0] = 0.5 * (0.5 - s) * (0.5 - s);
    w[a][1] = 0.75 - s * s;
    w[a][2] = 0.5 * (0.5 + s) * (0.5 + s);
    dw[a][0] = -(0.5 - s);
    dw[a][1] = -2.0 * s;
    dw[a][2] = 0.5 + s;
  }
  double v = 0, gx = 0, gy = 0, gz = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        const double C = G.at(base[0] + a - 1, base[1] + b - 1, base[2] + c - 1);
        v += w[0][a] * w[1][b] * w[2][c] * C;
        gx += dw[0][a] * w[1][b] * w[2][c] * C;
        gy += w[0][a] * dw[1][b] * w[2][c] * C;
        gz += w[0][a] * w[1][b] * dw[2][c] * C;
      }
  grad[0] = gx;
  grad[1] = gy;
  grad[2] = gz;
  return v;
}

void pixel_dir(const double* cam, int u, int v, int W, int H, double tx,
               double ty, double dir[3]) {
  const double x = (2.0 * (u + 0.5) / W - 1.0) * tx;
  const double y = (2.0 * (v + 0.5) / H - 1.0) * ty;
  const double n = std::sqrt(x * x + y * y + 1.0);
  const double l[3] = {x / n, y / n, 1.0 / n};
  for (int a = 0; a < 3; ++a)
    dir[a] = cam[3 * a] * l[0] + cam[3 * a + 1] * l[1] + cam[3 * a + 2] * l[2];
}

struct Sample {
  double d, l, p, g[3], w[3][3];
  long base[3];
};

// March one ray; fills the per-sample arrays, returns D.
double march(const Grid& G, const double* cam, const double dir[3],
             double step, int ns, std::vector<Sample>& s) {
  s.resize(ns);
  double T = 1.0, D = 0.0;
  for (int i = 0; i < ns; ++i) {
    Sample& k = s[i];
    k.d = (i + 0.5) * step;
    double q[3];
    for (int a = 0; a < 3; ++a)
      q[a] = (cam[9 + a] + dir[a] * k.d - G.o[a]) / G.res;
    k.l = spline(G, q, k.g, k.base, k.w);
    k.p = 1.0 / (1.0 + std::exp(-k.l));
    D += k.p * T * k.d;
    T *= 1.0 - k.p;
  }
  return D;
}

}  // namespace

extern "C" {

void or_soft_forward(const double* grid, int nx, int ny, int nz,
                     const double origin[3], double res, double background,
                     const double* cams, int n_cams, int width, int height,
                     double hfov_deg, double vfov_deg, double step,
                     int n_samples, double* depth) {
  const Grid G{grid, nx, ny, nz, {origin[0], origin[1], origin[2]}, res,
               background};
  const double pi = 3.14159265358979323846;
  const double tx = std::tan(hfov_deg * pi / 360.0);
  const double ty = std::tan(vfov_deg * pi / 360.0);
  std::vector<Sample> s;
  for (int c = 0; c < n_cams; ++c)
    for (int v = 0; v < height; ++v)
      for (int u = 0; u < width; ++u) {
        double dir[3];
        pixel_dir(cams + 12 * c, u, v, width, height, tx, ty, dir);
        depth[((long)c * height + v) * width + u] =
            march(G, cams + 12 * c, dir, step, n_samples, s);
      }
}

void or_soft_backward(const double* grid, int nx, int ny, int nz,
                      const double origin[3], double res, double background,
                      const double* cams, int n_cams, int width, int height,
                      double hfov_deg, double vfov_deg, double step,
                      int n_samples, const double* grad_depth,
                      double* grad_pose, double* grad_logodds) {
  const Grid G{grid, nx, ny, nz, {origin[0], origin[1], origin[2]}, res,
               background};
  const double pi = 3.14159265358979323846;
  const double tx = std::tan(hfov_deg * pi / 360.0);
  const double ty = std::tan(vfov_deg * pi / 360.0);
  std::vector<Sample> s;
  std::vector<double> T(n_samples + 1), S(n_samples + 1);
  for (int c = 0; c < n_cams; ++c)
    for (int v = 0; v < height; ++v)
      for (int u = 0; u < width; ++u) {
        const double go = grad_depth[((long)c * height + v) * width + u];
        if (go == 0.0) continue;
        double dir[3];
        const double* cam = cams + 12 * c;
        pixel_dir(cam, u, v, width, height, tx, ty, dir);
        march(G, cam, dir, step, n_samples, s);
        // T[k]: transmittance before sample k; S[k]: sum_{i>k} w_i d_i
        T[0] = 1.0;
        for (int k = 0; k < n_samples; ++k) T[k + 1] = T[k] * (1.0 - s[k].p);
        S[n_samples - 1] = 0.0;
        for (int k = n_samples - 2; k >= 0; --k)
          S[k] = S[k + 1] + s[k + 1].p * T[k + 1] * s[k + 1].d;
        for (int k = 0; k < n_samples; ++k) {
          const double p = s[k].p;
          const double dDdp = T[k] * s[k].d - S[k] / (1.0 - p);
          const double dl = go * dDdp * p * (1.0 - p);
          const double g[3] = {s[k].g[0] / res, s[k].g[1] / res,
                               s[k].g[2] / res};
          const double r[3] = {dir[0] * s[k].d, dir[1] * s[k].d,
                               dir[2] * s[k].d};
          if (grad_pose) {
            double* gp = grad_pose + 6 * c;
            gp[0] += dl * g[0];
            gp[1] += dl * g[1];
            gp[2] += dl * g[2];
            gp[3] += dl * (r[1] * g[2] - r[2] * g[1]);
            gp[4] += dl * (r[2] * g[0] - r[0] * g[2]);
            gp[5] += dl * (r[0] * g[1] - r[1] * g[0]);
          }
          if (grad_logodds)
            for (int a = 0; a < 3; ++a)
              for (int b = 0; b < 3; ++b)
                for (int cc = 0; cc < 3; ++cc) {
                  const long i = s[k].base[0] + a - 1, j = s[k].base[1] + b - 1,
                             kk = s[k].base[2] + cc - 1;
                  if (i < 0 || j < 0 || kk < 0 || i >= nx || j >= ny || kk >= nz)
                    continue;
                  grad_logodds[(i * ny + j) * nz + kk] +=
                      dl * s[k].w[0][a] * s[k].w[1][b] * s[k].w[2][cc];
                }
        }
      }
}

}  // extern "C"
