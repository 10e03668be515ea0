// nm_scene.cpp -- seeded synthetic workloads and host-side frame set-up.
//
// Scenes: a closed room [0,X]x[0,Y]x[0,Z] holding axis-aligned boxes
// (centres uniform in the room, side lengths uniform per axis) and
// vertical cylinders spanning the room height.  A pinhole depth camera
// renders z-depth by exact ray/primitive intersection in FP64, adds
// Gaussian noise of sigma = k*d^2, drops pixels to NaN with probability p,
// caps missing or out-of-range returns at the range, and emits the
// sensor-frame xyz FP32 cloud that Manager::insertPointCloud consumes
// (include/nanomap/manager/Manager.h:52).  Randomness is splitmix64,
// seeded per pixel, so rendering is deterministic under threading.
//
// nm_frame_from_pose restates the host-side pose set-up of the reference:
// R = normalized(q).toRotationMatrix() * F (Handler.cpp:41, Sensor.h:87-88),
// eye = worldToIndex(position) (Handler.cpp:34), max_time = maxRange /
// gridRes (Handler.cpp:39), evaluated in the frame's rounding mode.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cfenv>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/nanomap_amd.h"
#include "../../include/nanomap_amd_scene.h"

namespace {

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  double normal() {
    double u1 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    const double u2 = uniform();
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
};

struct Box {
  double lo[3], hi[3];
};
struct Cyl {
  double cx, cy, r;
};

// Quaternion (w, x, y, z) to a row-major 3x3 rotation, in double.
void quat_to_mat(const double q[4], double m[9]) {
  const double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  const double w = q[0] / n, x = q[1] / n, y = q[2] / n, z = q[3] / n;
  m[0] = 1 - 2 * (y * y + z * z);
  m[1] = 2 * (x * y - w * z);
  m[2] = 2 * (x * z + w * y);
  m[3] = 2 * (x * y + w * z);
  m[4] = 1 - 2 * (x * x + z * z);
  m[5] = 2 * (y * z - w * x);
  m[6] = 2 * (x * z - w * y);
  m[7] = 2 * (y * z + w * x);
  m[8] = 1 - 2 * (x * x + y * y);
}

void yaw_pitch_quat(double yaw, double pitch, float q[4]) {
  // q = qz(yaw) * qy(pitch)
  const double cy = cos(yaw / 2), sy = sin(yaw / 2);
  const double cp = cos(pitch / 2), sp = sin(pitch / 2);
  q[0] = (float)(cy * cp);
  q[1] = (float)(-sy * sp);
  q[2] = (float)(cy * sp);
  q[3] = (float)(sy * cp);
}

}  // namespace

struct nm_scene {
  double world[3];
  std::vector<Box> boxes;
  std::vector<Cyl> cyls;

  // Nearest positive hit distance along o + t d (d not necessarily unit).
  double trace(const double o[3], const double d[3]) const {
    double best = INFINITY;
    // room walls, floor and ceiling (o is inside the room)
    for (int a = 0; a < 3; ++a) {
      if (d[a] > 0) best = std::min(best, (world[a] - o[a]) / d[a]);
      if (d[a] < 0) best = std::min(best, (0.0 - o[a]) / d[a]);
    }
    for (const Box& b : boxes) {
      double t0 = 0.0, t1 = best;
      bool hit = true;
      for (int a = 0; a < 3 && hit; ++a) {
        if (d[a] == 0.0) {
          if (o[a] < b.lo[a] || o[a] > b.hi[a]) hit = false;
          continue;
        }
        double ta = (b.lo[a] - o[a]) / d[a], tb = (b.hi[a] - o[a]) / d[a];
        if (ta > tb) std::swap(ta, tb);
        t0 = std::max(t0, ta);
        t1 = std::min(t1, tb);
        if (t0 > t1) hit = false;
      }
      if (hit && t0 > 0.0 && t0 < best) best = t0;
    }
    const double a2 = d[0] * d[0] + d[1] * d[1];
    if (a2 > 0) {
      for (const Cyl& c : cyls) {
        const double ox = o[0] - c.cx, oy = o[1] - c.cy;
        const double b = ox * d[0] + oy * d[1];
        const double cc = ox * ox + oy * oy - c.r * c.r;
        const double disc = b * b - a2 * cc;
        if (disc < 0) continue;
        const double t = (-b - sqrt(disc)) / a2;
        if (t > 0 && t < best) best = t;
      }
    }
    return best;
  }

  bool free_at(const double p[3], double clearance) const {
    for (int a = 0; a < 3; ++a)
      if (p[a] < clearance || p[a] > world[a] - clearance) return false;
    for (const Box& b : boxes) {
      bool inside = true;
      for (int a = 0; a < 3; ++a)
        if (p[a] < b.lo[a] - clearance || p[a] > b.hi[a] + clearance)
          inside = false;
      if (inside) return false;
    }
    for (const Cyl& c : cyls) {
      const double dx = p[0] - c.cx, dy = p[1] - c.cy;
      if (dx * dx + dy * dy < (c.r + clearance) * (c.r + clearance))
        return false;
    }
    return true;
  }
};

extern "C" {

int nm_scene_create(uint64_t seed, const double world[3], int n_boxes,
                    double box_min, double box_max, int n_cylinders,
                    double cyl_rmin, double cyl_rmax, nm_scene** out) {
  if (!out || !world || n_boxes < 0 || n_cylinders < 0) return NM_EINVAL;
  nm_scene* s = new nm_scene();
  memcpy(s->world, world, sizeof(s->world));
  SplitMix rng(seed);
  for (int i = 0; i < n_boxes; ++i) {
    Box b;
    for (int a = 0; a < 3; ++a) {
      const double c = rng.uniform(0.0, world[a]);
      const double h = 0.5 * rng.uniform(box_min, box_max);
      b.lo[a] = c - h;
      b.hi[a] = c + h;
    }
    s->boxes.push_back(b);
  }
  for (int i = 0; i < n_cylinders; ++i) {
    Cyl c;
    c.cx = rng.uniform(0.0, world[0]);
    c.cy = rng.uniform(0.0, world[1]);
    c.r = rng.uniform(cyl_rmin, cyl_rmax);
    s->cyls.push_back(c);
  }
  *out = s;
  return NM_OK;
}

int nm_scene_destroy(nm_scene* s) {
  delete s;
  return NM_OK;
}

int nm_scene_render(const nm_scene* s, const nm_camera* cam,
                    const float pos[3], const float quat[4], uint64_t seed,
                    float* xyz) {
  if (!s || !cam || !pos || !quat || !xyz || cam->width <= 0 ||
      cam->height <= 0)
    return NM_EINVAL;
  const int W = cam->width, H = cam->height;
  const double fx = 0.5 * W / tan(cam->hfov_deg * M_PI / 360.0);
  const double fy = 0.5 * H / tan(cam->vfov_deg * M_PI / 360.0);
  const double q[4] = {quat[0], quat[1], quat[2], quat[3]};
  double Rq[9], RF[9];
  quat_to_mat(q, Rq);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0;
      for (int k = 0; k < 3; ++k) acc += Rq[3 * i + k] * (double)cam->frame[3 * k + j];
      RF[3 * i + j] = acc;
    }
  const double o[3] = {pos[0], pos[1], pos[2]};
  const double range = cam->max_range;
  auto work = [&](int row0, int row1) {
    for (int v = row0; v < row1; ++v) {
      for (int u = 0; u < W; ++u) {
        const int pix = v * W + u;
        SplitMix rng(seed ^ (0x9e3779b97f4a7c15ull * (uint64_t)(pix + 1)));
        const double oc[3] = {(u + 0.5 - 0.5 * W) / fx, -(v + 0.5 - 0.5 * H) / fy,
                              1.0};
        double d[3];
        for (int i = 0; i < 3; ++i)
          d[i] = RF[3 * i] * oc[0] + RF[3 * i + 1] * oc[1] + RF[3 * i + 2] * oc[2];
        double depth = s->trace(o, d);  // z-depth: oc has unit z
        if (depth < range) {
          depth += rng.normal() * cam->noise_k * depth * depth;
          if (!(depth < range)) depth = range;
          if (depth < 0) depth = 0;
        } else {
          depth = range;
        }
        float* p = xyz + 3 * (size_t)pix;
        if (cam->dropout > 0 && rng.uniform() < cam->dropout) {
          p[0] = p[1] = p[2] = NAN;
        } else {
          p[0] = (float)(oc[0] * depth);
          p[1] = (float)(oc[1] * depth);
          p[2] = (float)(oc[2] * depth);
        }
      }
    }
  };
  const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  std::vector<std::thread> ts;
  const int rows = (H + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const int r0 = t * rows, r1 = std::min(H, r0 + rows);
    if (r0 < r1) ts.emplace_back(work, r0, r1);
  }
  for (auto& t : ts) t.join();
  return NM_OK;
}

int nm_scene_sample_pose(const nm_scene* s, uint64_t seed,
                         float max_pitch_deg, float clearance, float pos[3],
                         float quat[4]) {
  if (!s || !pos || !quat) return NM_EINVAL;
  SplitMix rng(seed);
  for (int tries = 0; tries < 100000; ++tries) {
    double p[3];
    for (int a = 0; a < 3; ++a) p[a] = rng.uniform(0.0, s->world[a]);
    if (!s->free_at(p, clearance)) continue;
    const double yaw = rng.uniform(-M_PI, M_PI);
    const double pitch = rng.uniform(-1.0, 1.0) * max_pitch_deg * M_PI / 180.0;
    for (int a = 0; a < 3; ++a) pos[a] = (float)p[a];
    yaw_pitch_quat(yaw, pitch, quat);
    return NM_OK;
  }
  return NM_EOVERFLOW;
}

int nm_scene_walk(const nm_scene* s, uint64_t seed, int n, float step_m,
                  float step_deg, float max_pitch_deg, float clearance,
                  float* pos_out, float* quat_out) {
  if (!s || n <= 0 || !pos_out || !quat_out) return NM_EINVAL;
  SplitMix rng(seed);
  double p[3];
  {
    float p0[3], q0[4];
    int st = nm_scene_sample_pose(s, rng.next(), 0.0f, clearance, p0, q0);
    if (st != NM_OK) return st;
    for (int a = 0; a < 3; ++a) p[a] = p0[a];
  }
  double yaw = rng.uniform(-M_PI, M_PI), pitch = 0.0;
  const double sd = step_deg * M_PI / 180.0, mp = max_pitch_deg * M_PI / 180.0;
  for (int k = 0; k < n; ++k) {
    for (int a = 0; a < 3; ++a) pos_out[3 * k + a] = (float)p[a];
    yaw_pitch_quat(yaw, pitch, quat_out + 4 * k);
    yaw += rng.uniform(-sd, sd);
    pitch = std::max(-mp, std::min(mp, pitch + rng.uniform(-0.5 * sd, 0.5 * sd)));
    for (int tries = 0; tries < 8; ++tries) {
      const double q[3] = {p[0] + step_m * cos(yaw), p[1] + step_m * sin(yaw), p[2]};
      if (s->free_at(q, clearance)) {
        p[0] = q[0];
        p[1] = q[1];
        break;
      }
      yaw += 0.5 * M_PI;
    }
  }
  return NM_OK;
}

int nm_frame_from_pose(const float pos[3], const float quat[4],
                       const float F[9], float grid_res, float max_range,
                       float logodds_hit, float logodds_miss, int rounding,
                       nm_frame* out) {
  if (!pos || !quat || !F || !out || !(grid_res > 0) || !(max_range > 0))
    return NM_EINVAL;
  const int saved = fegetround();
  fesetround(rounding == NM_ROUND_DOWN ? FE_DOWNWARD : FE_TONEAREST);
  // Eigen::Quaternionf::normalized(): coefficients divided by the norm.
  volatile float w = quat[0], x = quat[1], y = quat[2], z = quat[3];
  const float n = sqrtf(x * x + y * y + z * z + w * w);
  w = w / n;
  x = x / n;
  y = y / n;
  z = z / n;
  // QuaternionBase::toRotationMatrix (Eigen 3.3), FP32.
  const float tx = 2.0f * x, ty = 2.0f * y, tz = 2.0f * z;
  const float twx = tx * w, twy = ty * w, twz = tz * w;
  const float txx = tx * x, txy = ty * x, txz = tz * x;
  const float tyy = ty * y, tyz = tz * y, tzz = tz * z;
  const float q[9] = {1.0f - (tyy + tzz), txy - twz, txz + twy,
                      txy + twz, 1.0f - (txx + tzz), tyz - twx,
                      txz - twy, tyz + twx, 1.0f - (txx + tyy)};
  memset(out, 0, sizeof(*out));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      volatile float acc = q[3 * i] * F[j];
      acc = acc + q[3 * i + 1] * F[3 + j];
      acc = acc + q[3 * i + 2] * F[6 + j];
      out->rotation[3 * i + j] = acc;
    }
  // openvdb UniformScaleMap::applyInverseMap: xyz * (1 / voxelSize).
  volatile double inv = 1.0 / (double)grid_res;
  for (int a = 0; a < 3; ++a) out->eye[a] = (double)pos[a] * inv;
  volatile float mt = max_range / grid_res;
  out->max_time = mt;
  fesetround(saved);
  out->grid_res = grid_res;
  out->logodds_hit = logodds_hit;
  out->logodds_miss = logodds_miss;
  out->rounding = rounding;
  out->filter = NM_FILTER_NONE;
  out->max_range = max_range;
  return NM_OK;
}

int nm_frustum_box(int width, int height, float vfov_deg, float min_range,
                   float max_range, int rounding, float box_min[3],
                   float box_max[3]) {
  if (!box_min || !box_max || width <= 0 || height <= 0 ||
      !(vfov_deg > 0.0f && vfov_deg < 180.0f) || !(min_range >= 0.0f) ||
      !(max_range >= min_range) || !std::isfinite(max_range) ||
      (rounding != NM_ROUND_NEAREST && rounding != NM_ROUND_DOWN))
    return NM_EINVAL;
  const int saved = fegetround();
  fesetround(rounding == NM_ROUND_DOWN ? FE_DOWNWARD : FE_TONEAREST);
  // FrustumData.h:24, 69-72: FP32 aspect; the half extents are double
  // products rounded to float; the widths FP32 products.
  volatile float aspect = (float)width / (float)height;
  volatile double t = tan(vfov_deg * M_PI / 360);
  volatile float hh = (float)(max_range * t);
  volatile float hw = aspect * hh;
  volatile float ch = (float)(min_range * t);
  volatile float cw = aspect * ch;
  fesetround(saved);
  const float x[9] = {-cw, -cw, cw, cw, -hw, -hw, hw, hw, 0.0f};
  const float y[9] = {-ch, ch, ch, -ch, -hh, hh, hh, -hh, 0.0f};
  const float z[9] = {min_range, min_range, min_range, min_range,
                      max_range, max_range, max_range, max_range, 0.0f};
  const float* c[3] = {x, y, z};
  for (int a = 0; a < 3; ++a) {
    box_min[a] = *std::min_element(c[a], c[a] + 9);
    box_max[a] = *std::max_element(c[a], c[a] + 9);
  }
  return NM_OK;
}

}  // extern "C"
