// nm_dda.h -- per-ray set-up and voxel DDA of the hit/miss cast.
//
// Semantics (restated from the reference host path, Handler.cpp:58-85,
// which uses openvdb::math::DDA<Ray<double>,0> and Ray::setDir):
//   d   = R p                      FP32, ((r0 x + r1 y) + r2 z), no FMA
//   len = sqrt((dx dx + dy dy) + dz dz)      FP64
//   L   = len / res; skip unless L < max_time (NaN skips)
//   d  *= 1 / len   unless len <= 1e-7       (Vec3::normalize)
//   inv = 1 / d; pos = eye + d * 0; voxel = floor(pos)
//   per axis: d == 0 -> step 0, next = DBL_MAX
//             inv > 0 -> step +1, next = 0 + ((v + 1) - pos) inv, delta = inv
//             else    -> step -1, next = 0 + (v - pos) inv,       delta = -inv
//   step(): a = argmin(next) (x only if strictly smallest, else y if
//           y < z, else z); t = next[a]; next[a] += delta[a]; v[a] += step;
//           continue while t <= L
//   every voxel stepped out of gets a miss; the end voxel gets a hit iff
//   its exit time t < max_time.
// Every operation rounds in mode M (nm_fp.h).
#ifndef NM_DDA_H
#define NM_DDA_H

#include "nm_fp.h"

namespace nm {

struct RayDir {
  double d[3];
  double L;
};

template <int M>
NM_HD bool ray_from_point(const float* R, float px, float py, float pz,
                          float res, float max_time, RayDir& r) {
  const float fx = add<M>(add<M>(mul<M>(R[0], px), mul<M>(R[1], py)),
                          mul<M>(R[2], pz));
  const float fy = add<M>(add<M>(mul<M>(R[3], px), mul<M>(R[4], py)),
                          mul<M>(R[5], pz));
  const float fz = add<M>(add<M>(mul<M>(R[6], px), mul<M>(R[7], py)),
                          mul<M>(R[8], pz));
  double x = fx, y = fy, z = fz;
  const double len =
      sqrt<M>(add<M>(add<M>(mul<M>(x, x), mul<M>(y, y)), mul<M>(z, z)));
  const double L = div<M>(len, (double)res);
  if (!(L < (double)max_time)) return false;
  if (!(len <= 1.0e-7)) {
    const double s = div<M>(1.0, len);
    x = mul<M>(x, s);
    y = mul<M>(y, s);
    z = mul<M>(z, s);
  }
  r.d[0] = x;
  r.d[1] = y;
  r.d[2] = z;
  r.L = L;
  return true;
}

NM_HD double fmin_(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fmin(a, b);
#else
  return std::fmin(a, b);
#endif
}

// Voxel DDA state.  Axes are kept in named scalars so the per-step
// argmin compiles to selects rather than indexed register access.
struct Dda {
  double nx, ny, nz;  // next crossing time per axis
  double dx, dy, dz;  // time between crossings per axis
  int sx, sy, sz;     // step per axis (-1, 0, +1)
  int vx, vy, vz;     // current voxel
  double t;           // time of the last crossing
};

template <int M>
NM_HD void dda_axis(double eye, double d, int& v, int& s, double& next,
                    double& delta) {
  const double inv = div<M>(1.0, d);
  const double pos = add<M>(eye, mul<M>(d, 0.0));
  v = (int)floor_(pos);
  if (d == 0.0) {
    s = 0;
    next = 1.7976931348623157e308;
    delta = 1.7976931348623157e308;
  } else if (inv > 0) {
    s = 1;
    next = add<M>(0.0, mul<M>(sub<M>((double)(v + 1), pos), inv));
    delta = inv;
  } else {
    s = -1;
    next = add<M>(0.0, mul<M>(sub<M>((double)v, pos), inv));
    delta = -inv;
  }
}

template <int M>
NM_HD void dda_init(const double* eye, const RayDir& r, Dda& s) {
  dda_axis<M>(eye[0], r.d[0], s.vx, s.sx, s.nx, s.dx);
  dda_axis<M>(eye[1], r.d[1], s.vy, s.sy, s.ny, s.dy);
  dda_axis<M>(eye[2], r.d[2], s.vz, s.sz, s.nz, s.dz);
  s.t = 0.0;
}

// One DDA::step(); returns t <= L.  Branch-free: one rounded add.  The
// state is read into scalars and written back unconditionally, so the
// axis choice is a select between register values (a select between two
// struct fields can otherwise become a load from a selected address, which
// pins the state to scratch memory).
//
// Three equivalent forms of the axis choice, picked per kernel by
// measurement:
//   kMin = 1:  t = min(nx, ny, nz); the axis is z if nz is minimal, else y
//              if ny is, else x (+0 and -0 compare equal here as in the
//              rule);
//   kMin = 0:  x only if strictly smallest, else y if y < z, else z, with t
//              and the increment selected by bit masks;
//   kMin = 2:  the same rule with t and the increment chosen by nested
//              selects (two v_cndmask per 32-bit half instead of three and
//              an or).
//
// kHwDown: the caller has set the wave's FP64 rounding mode to toward -inf,
// so the round-down sum is the hardware add (one instruction instead of the
// emulated add_down_pos).
template <int M, int kMin = 0, bool kHwDown = false>
NM_HD bool dda_step(Dda& s, double L) {
  const double nx = s.nx, ny = s.ny, nz = s.nz;
  const double ddx = s.dx, ddy = s.dy, ddz = s.dz;
  bool bx, by, bz;
  double t, dt;
  if (kMin == 2) {
    bx = (nx < ny) && (nx < nz);
    by = !bx && (ny < nz);
    bz = !bx && !by;
    t = bx ? nx : (by ? ny : nz);
    dt = bx ? ddx : (by ? ddy : ddz);
  } else if (kMin == 1) {
    t = fmin_(nx, fmin_(ny, nz));
    bz = nz == t;
    by = !bz && ny == t;
    bx = !bz && !by;
    dt = bz ? ddz : (by ? ddy : ddx);
  } else {
    bx = (nx < ny) && (nx < nz);
    by = !bx && (ny < nz);
    bz = !bx && !by;
    const uint64_t mx = 0ull - (uint64_t)bx, my = 0ull - (uint64_t)by;
    const uint64_t mz = ~(mx | my);
    t = double_of((bits_of(nx) & mx) | (bits_of(ny) & my) | (bits_of(nz) & mz));
    dt = double_of((bits_of(ddx) & mx) | (bits_of(ddy) & my) |
                   (bits_of(ddz) & mz));
  }
  // t >= 0 and dt > 0 (next times start at 0 + a non-negative product and
  // grow by |1/d|), so the round-down sum needs no sign or zero cases
  const double nt = (M == kDown && !kHwDown) ? add_down_pos(t, dt) : t + dt;
  s.nx = bx ? nt : nx;
  s.ny = by ? nt : ny;
  s.nz = bz ? nt : nz;
  s.vx += bx ? s.sx : 0;
  s.vy += by ? s.sy : 0;
  s.vz += bz ? s.sz : 0;
  s.t = t;
  return t <= L;
}

}  // namespace nm

#endif  // NM_DDA_H
