// nanomap/agent/Agent.h -- a simulated agent of the host API.
//
// Mirrors include/nanomap/agent/Agent.h:43-230: an agent read from an
// "#agentconfig" file -- its spawn (fixed position, or random), its
// observation sphere (NumObs rays, LenObs and CollisionDist in metres) and
// its sensors (name, position and orientation relative to the body) -- with
// its own map.  updatePose() places every sensor at the body position with
// the body orientation times the sensor's own (Agent.h:132-141).  The
// observation rays are the reference's spiral sphere (normalisedSphere,
// Agent.h:150-168, x = 0.1 + 1.2 n), with norms (CollisionDist / GridRes,
// LenObs / MappingRes) (Agent.h:185-195).
#ifndef NANOMAP_AMD_HOST_AGENT_H
#define NANOMAP_AMD_HOST_AGENT_H

#include <cmath>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "../map/OccupancyMap.h"
#include "../nanomap.h"

namespace nanomap {
namespace agent {

class Agent {
 public:
  Agent(const std::string& agentFile, float mappingRes, float gridRes,
        float probHitThres, float probMissThres, int device = 0)
      : _mappingRes(mappingRes), _gridRes(gridRes) {
    _map = std::make_shared<map::Map>(mappingRes, probHitThres, probMissThres,
                                      0.0f, 0.0f, 0, 1 << 16, device);
    std::ifstream in(agentFile);
    std::string line;
    if (!(in >> line) || line != "#agentconfig")
      throw std::runtime_error("agent config " + agentFile +
                               ": first line must be #agentconfig");
    float numObs = 0, lenObs = 0, collisionDist = 0;
    while (in >> line) {
      if (line == "Spawn:") {
        in >> line >> line;  // "Type:" and the type
        _spawnRandom = line == "Random";
        if (!_spawnRandom) {
          float q[4];
          in >> line >> _pose.position(0) >> _pose.position(1) >>
              _pose.position(2);
          in >> line >> q[0] >> q[1] >> q[2] >> q[3];  // read, not used
        }
        _pose.orientation = bodyOrientation();
      } else if (line == "Observations:") {
        in >> line >> numObs >> line >> lenObs >> line >> collisionDist;
      } else if (line == "Sensor:") {
        std::string name;
        Pose p;
        float w, x, y, z;
        in >> line >> name;
        in >> line >> p.position(0) >> p.position(1) >> p.position(2);
        in >> line >> w >> x >> y >> z;
        p.orientation = Quaternionf(w, x, y, z);
        _sensorNames.push_back(name);
        _sensorOrigins.push_back(p);
      } else if (line == "#endconfig") {
        break;
      }
    }
    _sensorPoses = _sensorOrigins;
    _numObs = (int)numObs;
    _lenObs = lenObs / _mappingRes;
    _collisionDist = collisionDist / _gridRes;
    generateObservationSphere();
    updatePose(_pose);
  }

  // The body orientation of a spawn: columns (0,-1,0), (1,0,0), (0,0,1)
  // (Agent.h:67-71, 80-84; SimManager.h:218-224).
  static Quaternionf bodyOrientation() {
    const float m[3][3] = {{0, 1, 0}, {-1, 0, 0}, {0, 0, 1}};
    return Quaternionf(m);
  }

  void updatePose(const Pose& pose) {
    _pose = pose;
    for (size_t i = 0; i < _sensorOrigins.size(); ++i) {
      _sensorPoses[i].position = _pose.position;
      _sensorPoses[i].orientation = _pose.orientation * _sensorOrigins[i].orientation;
    }
  }

  Pose pose() const { return _pose; }
  Pose sensorPose(int index) const { return _sensorPoses[index]; }
  std::vector<Pose> sensorOrigins() const { return _sensorOrigins; }
  std::vector<Pose> sensorPoses() const { return _sensorPoses; }
  const std::vector<std::string>& sensorNames() const { return _sensorNames; }
  std::shared_ptr<map::Map> map() { return _map; }
  bool spawnRandom() const { return _spawnRandom; }
  // observationRays: numObs unit directions (3 floats each, world axes);
  // observationNorms: (collision distance, observation length) per ray in
  // voxel units.
  const std::vector<float>& observationRays() const { return _obsRays; }
  const std::vector<float>& observationNorms() const { return _obsNorms; }
  int numObservations() const { return _numObs; }

 private:
  // normalisedSphere(n, x): s runs from -1 + 1/(n-1) in steps of
  // (2 - 2/(n-1))/(n-1); yaw j = s x, pitch k = (pi/2) sign(s)
  // (1 - sqrt(1 - |s|)); point (cos j cos k, sin j cos k, sin k), normalised.
  void generateObservationSphere() {
    _obsRays.clear();
    _obsNorms.clear();
    const int n = _numObs;
    if (n <= 1) return;
    const float x = 0.1f + 1.2f * (float)n;
    const float start = -1.0f + 1.0f / ((float)n - 1.0f);
    const float inc = (2.0f - 2.0f / ((float)n - 1.0f)) / ((float)n - 1.0f);
    for (int i = 0; i < n; ++i) {
      const float s = start + (float)i * inc;
      const float j = s * x;
      const float k = (float)(M_PI / 2.0) * std::copysign(1.0f, s) *
                      (1.0f - std::sqrt(1.0f - std::fabs(s)));
      float p[3] = {std::cos(j) * std::cos(k), std::sin(j) * std::cos(k),
                    std::sin(k)};
      const float len = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
      for (int a = 0; a < 3; ++a) _obsRays.push_back(p[a] / len);
      _obsNorms.push_back(_collisionDist);
      _obsNorms.push_back(_lenObs);
    }
  }

  float _mappingRes, _gridRes;
  Pose _pose;
  bool _spawnRandom = false;
  int _numObs = 0;
  float _lenObs = 0, _collisionDist = 0;
  std::vector<float> _obsRays, _obsNorms;
  std::vector<std::string> _sensorNames;
  std::vector<Pose> _sensorOrigins, _sensorPoses;
  std::shared_ptr<map::Map> _map;
};

}  // namespace agent
}  // namespace nanomap

#endif
