// nanomap/sensor/SensorData.h -- sensor description of the host API.
//
// Mirrors include/nanomap/sensor/SensorData.h, SharedParameters.h and
// FrustumData.h:20-83 / LaserData.h:20-48: the per-sensor constants read
// from a sensor config file (resolution, field of view, ranges, hit/miss
// probabilities, frame transform), the frustum's bounds in voxel units and
// in metres (the voxel filter's point box worldMin/worldMax), and the pose
// of the last updatePose() with the offsets and bounding boxes derived from
// it (SensorData.h:40-146).  The leaf-buffer sizing of the reference (a
// rotation sweep, FrustumData.h:85-131) is not needed: the device path
// sizes its leaf table from the range.
#ifndef NANOMAP_AMD_HOST_SENSORDATA_H
#define NANOMAP_AMD_HOST_SENSORDATA_H

#include <algorithm>
#include <cfenv>
#include <cmath>
#include <string>

#include "../nanomap.h"

namespace nanomap {
namespace sensor {

// Integer box [min, max] per axis (nanovdb::CoordBBox).
struct CoordBBox {
  int min[3] = {0, 0, 0};
  int max[3] = {-1, -1, -1};
  int dim(int a) const { return max[a] - min[a] + 1; }
};

struct SharedParameters {
  int _type = 0;  // 0 frustum (depth camera), 1 laser
  int _rate = 0;
  float _gridRes = 0;
  int _hRes = 0;
  int _vRes = 0;
  int _pclSize = 0;
  float _aspect = 0;
  float _vfov = 0;
  float _hfov = 0;
  nanomap::Pose _pose;
  float _frameTransform[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};  // [row][col]
  float _probHit = 0;
  float _probMiss = 0;
  float _maxRange = 0;
  float _minRange = 0;
  float _minVoxelRange = 0;
  float _maxVoxelRange = 0;
  int _leafEdge = 8;
  // updatePose (SensorData.h:40-83)
  float _voxelPosition[3] = {0, 0, 0};
  int _leafOffset[3] = {0, 0, 0};
  int _voxelOffset[3] = {0, 0, 0};
  float _leafOriginOffset[3] = {0, 0, 0};
  float _voxelOriginOffset[3] = {0, 0, 0};
  // frustum corners [axis][corner]: near plane 0-3, far plane 4-7, origin 8
  // (FrustumData::calculateSensorBounds), in voxel units and in metres
  float _sensorBounds[3][9] = {};
  float _worldBounds[3][9] = {};
  float _worldMin[3] = {0, 0, 0};
  float _worldMax[3] = {0, 0, 0};
  // rotateView (SensorData.h:85-146)
  float _transformedBounds[3][9] = {};
  float _voxelTransformedBounds[3][9] = {};
  float _leafTransformedBounds[3][9] = {};
  CoordBBox _leafBounds;
  CoordBBox _voxelBounds;
  int _leafBufferSize = 0;
};

class SensorData {
 public:
  SensorData() = default;
  SensorData(std::string name, int id, int type, const float frameTransform[3][3],
             float gridRes, int leafEdge, int hRes, int vRes, float vfov,
             float hfov, int rate, float maxRange, float minRange,
             float probHit, float probMiss)
      : _sensorName(std::move(name)), _sensorId(id) {
    _shared._type = type;
    _shared._gridRes = gridRes;
    _shared._leafEdge = leafEdge;
    _shared._hRes = hRes;
    _shared._vRes = vRes;
    _shared._pclSize = hRes * vRes;
    _shared._aspect = vRes > 0 ? (float)hRes / (float)vRes : 0.0f;
    _shared._vfov = vfov;
    _shared._hfov = hfov;
    _shared._rate = rate;
    _shared._maxRange = maxRange;
    _shared._minRange = minRange;
    _shared._minVoxelRange = minRange / gridRes;
    _shared._maxVoxelRange = maxRange / gridRes;
    _shared._probHit = probHit;
    _shared._probMiss = probMiss;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) _shared._frameTransform[r][c] = frameTransform[r][c];
    if (type == 0 && hRes > 0 && vRes > 0 && vfov > 0 && vfov < 180)
      calculateSensorBounds();
  }

  // FrustumData::calculateSensorBounds (FrustumData.h:50-83): the corners
  // of the near (minRange) and far (maxRange) planes, half height
  // range * tan(vfov/2), half width aspect * half height, and the origin;
  // worldMin/worldMax are their per-axis extremes in metres.  Evaluated
  // rounding toward -inf, the mode the reference runs in once its first
  // sensor has been built (FrustumData.h:87 leaves FE_DOWNWARD set).
  void calculateSensorBounds() {
    const SharedParameters& s = _shared;
    const int saved = std::fegetround();
    std::fesetround(FE_DOWNWARD);
    const double t = std::tan(s._vfov * M_PI / 360);
    const float aspect = (float)s._hRes / (float)s._vRes;
    const float rv[2] = {s._minVoxelRange, s._maxVoxelRange};
    const float rw[2] = {s._minRange, s._maxRange};
    float (*out[2])[9] = {_shared._sensorBounds, _shared._worldBounds};
    for (int u = 0; u < 2; ++u) {
      const float* r = u == 0 ? rv : rw;
      const float ch = (float)(r[0] * t), cw = aspect * ch;
      const float hh = (float)(r[1] * t), hw = aspect * hh;
      const float x[9] = {-cw, -cw, cw, cw, -hw, -hw, hw, hw, 0.0f};
      const float y[9] = {-ch, ch, ch, -ch, -hh, hh, hh, -hh, 0.0f};
      const float z[9] = {r[0], r[0], r[0], r[0], r[1], r[1], r[1], r[1], 0.0f};
      for (int c = 0; c < 9; ++c) {
        out[u][0][c] = x[c];
        out[u][1][c] = y[c];
        out[u][2][c] = z[c];
      }
    }
    std::fesetround(saved);
    // the per-axis extremes of the metre corners, by the library's
    // computation of the same bounds (identical arithmetic)
    check(nm_frustum_box(s._hRes, s._vRes, s._vfov, s._minRange, s._maxRange,
                         NM_ROUND_DOWN, _shared._worldMin, _shared._worldMax));
  }

  // SensorData.h:40-83, rounding toward -inf: per axis (zero if the
  // position is exactly 0.0) voxelPosition = p / gridRes, voxelOffset =
  // nearbyint(voxelPosition) (a floor in that mode), leafOffset =
  // floor(voxelOffset / leafEdge) * leafEdge, and the two origin offsets.
  // The reference leaves the thread's rounding mode at FE_DOWNWARD; this
  // restores the caller's (the device frames carry their own mode).
  void updatePose(nanomap::Pose pose) {
    const int saved = std::fegetround();
    std::fesetround(FE_DOWNWARD);
    _shared._pose = pose;
    for (int a = 0; a < 3; ++a) {
      const float p = pose.position(a);
      if (p == 0.0f) {
        _shared._voxelPosition[a] = 0.0f;
        _shared._leafOffset[a] = 0;
        _shared._leafOriginOffset[a] = 0.0f;
        _shared._voxelOffset[a] = 0;
        _shared._voxelOriginOffset[a] = 0.0f;
        continue;
      }
      volatile float vp = p / _shared._gridRes;
      volatile float vo = std::nearbyint((float)vp);
      volatile float lo =
          std::floor(vo / (float)_shared._leafEdge) * (float)_shared._leafEdge;
      _shared._voxelPosition[a] = vp;
      _shared._leafOffset[a] = (int)lo;
      _shared._leafOriginOffset[a] = vp - (float)_shared._leafOffset[a];
      _shared._voxelOffset[a] = (int)vo;
      _shared._voxelOriginOffset[a] = vp - (float)_shared._voxelOffset[a];
    }
    std::fesetround(saved);
  }

  // SensorData.h:85-146: R = toRotationMatrix(q) * F (q as given, not
  // normalised), the corners rotated and shifted by the origin offsets, and
  // the leaf and voxel bounding boxes of the result (floor(nearbyint(.) /
  // leafEdge) and nearbyint(.), rounding toward -inf).
  void rotateView() {
    const int saved = std::fegetround();
    std::fesetround(FE_DOWNWARD);
    const auto& q = _shared._pose.orientation;
    const float w = q.w(), x = q.x(), y = q.y(), z = q.z();
    const float tx = 2.0f * x, ty = 2.0f * y, tz = 2.0f * z;
    const float twx = tx * w, twy = ty * w, twz = tz * w;
    const float txx = tx * x, txy = ty * x, txz = tz * x;
    const float tyy = ty * y, tyz = tz * y, tzz = tz * z;
    const float Q[3][3] = {{1.0f - (tyy + tzz), txy - twz, txz + twy},
                           {txy + twz, 1.0f - (txx + tzz), tyz - twx},
                           {txz - twy, tyz + twx, 1.0f - (txx + tyy)}};
    float R[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        volatile float acc = Q[i][0] * _shared._frameTransform[0][j];
        acc = acc + Q[i][1] * _shared._frameTransform[1][j];
        acc = acc + Q[i][2] * _shared._frameTransform[2][j];
        R[i][j] = acc;
      }
    for (int c = 0; c < 9; ++c)
      for (int i = 0; i < 3; ++i) {
        volatile float acc = R[i][0] * _shared._sensorBounds[0][c];
        acc = acc + R[i][1] * _shared._sensorBounds[1][c];
        acc = acc + R[i][2] * _shared._sensorBounds[2][c];
        _shared._transformedBounds[i][c] = acc;
        _shared._voxelTransformedBounds[i][c] = acc + _shared._voxelOriginOffset[i];
        _shared._leafTransformedBounds[i][c] = acc + _shared._leafOriginOffset[i];
      }
    const float e = (float)_shared._leafEdge;
    for (int a = 0; a < 3; ++a) {
      const float* lb = _shared._leafTransformedBounds[a];
      const float* vb = _shared._voxelTransformedBounds[a];
      _shared._leafBounds.min[a] =
          (int)std::floor(std::nearbyint(*std::min_element(lb, lb + 9)) / e);
      _shared._leafBounds.max[a] =
          (int)std::floor(std::nearbyint(*std::max_element(lb, lb + 9)) / e);
      _shared._voxelBounds.min[a] = (int)std::nearbyint(*std::min_element(vb, vb + 9));
      _shared._voxelBounds.max[a] = (int)std::nearbyint(*std::max_element(vb, vb + 9));
    }
    std::fesetround(saved);
    _shared._leafBufferSize = _shared._leafBounds.dim(0) *
                              _shared._leafBounds.dim(1) *
                              _shared._leafBounds.dim(2);
  }

  SharedParameters& sharedParameters() { return _shared; }
  const SharedParameters& sharedParameters() const { return _shared; }
  std::string sensorName() const { return _sensorName; }
  std::string getName() const { return _sensorName; }
  int getId() const { return _sensorId; }

  // log(p) - log(1 - p) of the hit / miss probabilities (Sensor.h:89-90;
  // evaluated in FP64 and rounded to FP32).
  float logOddsHit() const {
    return (float)(std::log((double)_shared._probHit) -
                   std::log(1.0 - _shared._probHit));
  }
  float logOddsMiss() const {
    return (float)(std::log((double)_shared._probMiss) -
                   std::log(1.0 - _shared._probMiss));
  }

 private:
  std::string _sensorName;
  int _sensorId = 0;
  SharedParameters _shared;
};

}  // namespace sensor
}  // namespace nanomap

#endif
