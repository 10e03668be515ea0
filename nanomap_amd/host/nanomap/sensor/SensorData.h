// nanomap/sensor/SensorData.h -- sensor description of the host API.
//
// Mirrors include/nanomap/sensor/SensorData.h, SharedParameters.h and
// FrustumData.h:20-60 / LaserData.h: the per-sensor constants read from a
// sensor config file (resolution, field of view, ranges, hit/miss
// probabilities, frame transform) and the pose of the last updatePose().
// The frustum bounds and leaf-buffer sizing of the reference are not
// needed: the device path sizes its leaf table from the range.
#ifndef NANOMAP_AMD_HOST_SENSORDATA_H
#define NANOMAP_AMD_HOST_SENSORDATA_H

#include <cmath>
#include <string>

#include "../nanomap.h"

namespace nanomap {
namespace sensor {

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
  }

  // The reference sets FE_DOWNWARD here and derives voxel offsets
  // (SensorData.h:40-83); on this path the rounding mode is part of each
  // frame (NM_ROUND_DOWN) and the offsets are derived on the device.
  void updatePose(nanomap::Pose pose) { _shared._pose = pose; }
  void rotateView() {}

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
