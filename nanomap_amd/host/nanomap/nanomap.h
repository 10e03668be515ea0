// nanomap/nanomap.h -- common types of the host API (header-only, over the
// C ABI in include/nanomap_amd.h).
//
// Mirrors include/nanomap/nanomap.h of the reference: nanomap::Pose with a
// position and an orientation quaternion.  The reference uses Eigen types;
// here small structs with the same member names and accessors stand in for
// them (Eigen is not a dependency of this library).
#ifndef NANOMAP_AMD_HOST_NANOMAP_H
#define NANOMAP_AMD_HOST_NANOMAP_H

#include <cmath>
#include <stdexcept>
#include <string>

#include "../../../include/nanomap_amd.h"
#include "../../../include/nanomap_amd_scene.h"  // nm_frame_from_pose

namespace nanomap {

struct Vector3f {
  float v[3] = {0.0f, 0.0f, 0.0f};
  Vector3f() = default;
  Vector3f(float x, float y, float z) : v{x, y, z} {}
  float& operator()(int i) { return v[i]; }
  float operator()(int i) const { return v[i]; }
  float x() const { return v[0]; }
  float y() const { return v[1]; }
  float z() const { return v[2]; }
};

// Quaternion with Eigen's constructor argument order (w, x, y, z).
struct Quaternionf {
  float qw = 1.0f, qx = 0.0f, qy = 0.0f, qz = 0.0f;
  Quaternionf() = default;
  Quaternionf(float w_, float x_, float y_, float z_)
      : qw(w_), qx(x_), qy(y_), qz(z_) {}
  // From a rotation matrix m[row][col] (Eigen's Quaternion(Matrix3)
  // conversion: the trace form, or the largest diagonal element's form).
  explicit Quaternionf(const float m[3][3]) {
    const float t = m[0][0] + m[1][1] + m[2][2];
    if (t > 0.0f) {
      const float s = std::sqrt(t + 1.0f);
      qw = 0.5f * s;
      const float r = 0.5f / s;
      qx = (m[2][1] - m[1][2]) * r;
      qy = (m[0][2] - m[2][0]) * r;
      qz = (m[1][0] - m[0][1]) * r;
      return;
    }
    int i = 0;
    if (m[1][1] > m[0][0]) i = 1;
    if (m[2][2] > m[i][i]) i = 2;
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    const float s = std::sqrt(m[i][i] - m[j][j] - m[k][k] + 1.0f);
    float q[3];
    q[i] = 0.5f * s;
    const float r = 0.5f / s;
    qw = (m[k][j] - m[j][k]) * r;
    q[j] = (m[j][i] + m[i][j]) * r;
    q[k] = (m[k][i] + m[i][k]) * r;
    qx = q[0];
    qy = q[1];
    qz = q[2];
  }
  float w() const { return qw; }
  float x() const { return qx; }
  float y() const { return qy; }
  float z() const { return qz; }
  // Hamilton product (Eigen's operator*): this rotation after o.
  Quaternionf operator*(const Quaternionf& o) const {
    return Quaternionf(qw * o.qw - qx * o.qx - qy * o.qy - qz * o.qz,
                       qw * o.qx + qx * o.qw + qy * o.qz - qz * o.qy,
                       qw * o.qy + qy * o.qw + qz * o.qx - qx * o.qz,
                       qw * o.qz + qz * o.qw + qx * o.qy - qy * o.qx);
  }
};

// include/nanomap/nanomap.h:20-28.  (The reference's default orientation
// Quaternionf(0,0,0,1) is w = 0, z = 1: a rotation by pi about z; kept.)
struct Pose {
  Vector3f position;
  Quaternionf orientation;
  Pose() : position(0.0f, 0.0f, 0.0f), orientation(0.0f, 0.0f, 0.0f, 1.0f) {}
};

// Status check for the C ABI: throws with the library's message.
inline void check(int status) {
  if (status != NM_OK) throw std::runtime_error(nm_last_error());
}

}  // namespace nanomap

#endif
