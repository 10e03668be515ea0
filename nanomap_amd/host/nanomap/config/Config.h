// nanomap/config/Config.h -- configuration of the host API.
//
// Mirrors include/nanomap/config/Config.h and src/nanomap/config/
// Config.cpp: the same text formats ("#config" ... "#endconfig" main file
// with one "Sensor:" line per sensor file; "#frustumconfig" /
// "#laserconfig" sensor files, FrameTransform read column by column:
// column i = file line i, Config.cpp:249-256), the same accessors.  As in
// the reference, sensors take the MappingRes voxel size (Config.cpp:167).
#ifndef NANOMAP_AMD_HOST_CONFIG_H
#define NANOMAP_AMD_HOST_CONFIG_H

#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "../sensor/SensorData.h"

namespace nanomap {
namespace config {

class Config {
 public:
  Config() = default;
  explicit Config(std::string configFile) : _config(std::move(configFile)) {
    _configCheck = true;
    init();
  }

  void init() {
    if (!_configCheck) return;
    loadConfig();
    loadSensors();
    _leafVolume = _leafEdge * _leafEdge * _leafEdge;
    _maxPCLSize = _frustumPCLSize = _laserPCLSize = 0;
    for (auto& s : _sensorData) {
      const auto& sh = s->sharedParameters();
      if (sh._pclSize > _maxPCLSize) _maxPCLSize = sh._pclSize;
      if (sh._type == 0 && sh._pclSize > _frustumPCLSize) _frustumPCLSize = sh._pclSize;
      if (sh._type == 1 && sh._pclSize > _laserPCLSize) _laserPCLSize = sh._pclSize;
    }
  }

  void loadConfig() {
    std::ifstream in(_config);
    std::string line;
    if (!(in >> line) || line != "#config")
      throw std::runtime_error("config " + _config + ": first line must be #config");
    while (in >> line) {
      if (line == "MappingRes:") in >> _mappingRes;
      else if (line == "ProbHitThres:") in >> _probHitThres;
      else if (line == "ProbMissThres:") in >> _probMissThres;
      else if (line == "GridRes:") in >> _gridRes;
      else if (line == "NodeEdge:") in >> _leafEdge;
      else if (line == "FrustumAllocationFactor:") in >> _frustumAllocationFactor;
      else if (line == "FilterType:") in >> _filterType;
      else if (line == "UpdateType:") in >> _updateType;
      else if (line == "PrecisionType:") in >> _precisionType;
      else if (line == "PublishSensor:" || line == "PublishType:") in >> _publishSensor;
      else if (line == "#endconfig") break;
    }
  }

  void loadSensors() {
    std::ifstream in(_config);
    std::string line, file;
    std::vector<std::string> files;
    in >> line;
    while (in >> line) {
      if (line == "Sensor:") {
        in >> file;
        files.push_back(file);
      } else if (line == "Agent:") {  // Config.cpp:129-131
        in >> file;
        _agentFiles.push_back(file);
      } else if (line == "SimGrid:") {  // SimManager.h:183-185
        in >> _simGrid;
      } else if (line == "#endconfig") {
        break;
      }
    }
    for (const auto& f : files)
      _sensorData.push_back(loadSensorData(_mappingRes, _leafEdge, f));
  }

  std::shared_ptr<sensor::SensorData> loadSensorData(float gridRes, int leafEdge,
                                                     const std::string& file) {
    std::ifstream in(file);
    std::string line, name;
    if (!(in >> line))
      throw std::runtime_error("sensor config " + file + " not readable");
    const bool laser = line == "#laserconfig";
    if (!laser && line != "#frustumconfig")
      throw std::runtime_error("sensor config " + file + ": unknown type " + line);
    int id = 0, rate = 0, hRes = 0, vRes = 0;
    float aHRes = 0, aVRes = 0, vfov = 0, hfov = 0, maxRange = 0, minRange = 0;
    float probHit = 0, probMiss = 0;
    float F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    while (in >> line) {
      if (line == "Name:") in >> name;
      else if (line == "Id:") in >> id;
      else if (line == "HRes:") in >> hRes;
      else if (line == "VRes:") in >> vRes;
      else if (line == "AngularHRes:") in >> aHRes;
      else if (line == "AngularVRes:") in >> aVRes;
      else if (line == "HFOV:") in >> hfov;
      else if (line == "VFOV:") in >> vfov;
      else if (line == "Rate:") in >> rate;
      else if (line == "MaxRange:") in >> maxRange;
      else if (line == "MinRange:") in >> minRange;
      else if (line == "ProbHit:") in >> probHit;
      else if (line == "ProbMiss:") in >> probMiss;
      else if (line == "FrameTransform:") {
        for (int col = 0; col < 3; ++col)
          for (int row = 0; row < 3; ++row) in >> F[row][col];
      } else if (line == "#endconfig") {
        break;
      }
    }
    if (laser) {
      // beams from the angular resolution, as LaserData.h:23-24 truncates
      hRes = aHRes > 0 ? (int)(hfov / aHRes + 1) : 0;
      vRes = aVRes > 0 ? (int)(vfov / aVRes + 1) : 0;
    }
    return std::make_shared<sensor::SensorData>(
        name, id, laser ? 1 : 0, F, gridRes, leafEdge, hRes, vRes, vfov, hfov,
        rate, maxRange, minRange, probHit, probMiss);
  }

  std::vector<std::shared_ptr<sensor::SensorData>>& sensorData() { return _sensorData; }
  // Agent config files ("Agent:" lines) and the simulation grid ("SimGrid:",
  // here a map snapshot written by nm_map_save).
  const std::vector<std::string>& agentFiles() const { return _agentFiles; }
  const std::string& simGrid() const { return _simGrid; }
  std::shared_ptr<sensor::SensorData> sensorData(int index, bool searchById = false) {
    if (searchById) {
      for (auto& s : _sensorData)
        if (s->getId() == index) return s;
      return nullptr;
    }
    return index >= 0 && index < (int)_sensorData.size() ? _sensorData[index] : nullptr;
  }
  std::shared_ptr<sensor::SensorData> sensorData(const std::string& name) {
    for (auto& s : _sensorData)
      if (s->getName() == name) return s;
    return nullptr;
  }

  int& filterType() { return _filterType; }
  int& updateType() { return _updateType; }
  int& precisionType() { return _precisionType; }
  int& publishSensor() { return _publishSensor; }
  float& gridRes() { return _gridRes; }
  float& mappingRes() { return _mappingRes; }
  int& leafEdge() { return _leafEdge; }
  int& leafVolume() { return _leafVolume; }
  float& frustumAllocationFactor() { return _frustumAllocationFactor; }
  int& maxPCLSize() { return _maxPCLSize; }
  int& frustumPCLSize() { return _frustumPCLSize; }
  int& laserPCLSize() { return _laserPCLSize; }
  float& probHitThres() { return _probHitThres; }
  float& probMissThres() { return _probMissThres; }
  std::string& config() { return _config; }

 private:
  bool _configCheck = false;
  std::string _config;
  float _gridRes = 0.1f;
  float _mappingRes = 0.1f;
  int _leafEdge = 8;
  int _leafVolume = 512;
  int _filterType = 0;
  int _updateType = 0;
  int _precisionType = 0;
  int _publishSensor = 0;
  float _probHitThres = 0.5f;
  float _probMissThres = 0.5f;
  float _frustumAllocationFactor = 0.0f;
  int _maxPCLSize = 0;
  int _frustumPCLSize = 0;
  int _laserPCLSize = 0;
  std::vector<std::shared_ptr<sensor::SensorData>> _sensorData;
  std::vector<std::string> _agentFiles;
  std::string _simGrid;
};

}  // namespace config
}  // namespace nanomap

#endif
