// nanomap/manager/SimManager.h -- multi-agent sensor simulation of the host
// API.
//
// Mirrors include/nanomap/manager/SimManager.h:64-290: a SimManager built
// from a Config holds a simulation grid and the config's agents (each with
// its own map); updateAgentPoses() moves them, updateAgentViews() simulates
// every sensor of every agent on the grid and folds the simulated clouds
// into the agents' maps, getMap() returns an agent's map,
// getAgentSimCollision() tests a move against the grid along the agent's
// observation rays, getRandomSpawn() draws a free spawn pose.
//
// Differences from the reference, each on purpose:
//  - the grid is a map snapshot (nm_map_save; "SimGrid:" in the config), a
//    NanoVDB FloatGrid file (.nvdb, nm_map_load_nvdb) or a map in memory,
//    not an OpenVDB .vdb file (OpenVDB is not a dependency); its voxels
//    with log-odds above the map's hit threshold are the occupied ones (the
//    reference's active voxels);
//  - updateAgentViews() simulates every agent's sensors in one launch
//    (nm_sim_cast) and, for clouds cast without the voxel filter, folds
//    each sensor round of all agents into their maps in one batch
//    (nm_batch_submit); the maps equal those of a loop of Manager calls on
//    the same clouds, bit for bit;
//  - getAgentSimCollision() reports a collision when an active voxel lies
//    within the collision distance along an observation ray (first active
//    voxel, nm_sim_cast); the reference compares the exit time of the first
//    active span of leaf nodes (VolumeRayIntersector::march, SimManager.h:
//    152-158), so a thick obstacle can pass unnoticed there;
//  - getRandomSpawn() is seeded (the reference seeds rand() with the time,
//    SimManager.h:230) and draws voxels in the grid's active bounding box
//    until one is observed free and collision-free for the observation
//    sphere.
#ifndef NANOMAP_AMD_HOST_SIMMANAGER_H
#define NANOMAP_AMD_HOST_SIMMANAGER_H

#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../../../include/nanomap_amd_batch.h"
#include "../../../../include/nanomap_amd_sim.h"
#include "../agent/Agent.h"
#include "Manager.h"

namespace nanomap {
namespace manager {

class SimManager {
 public:
  // The grid from the config's "SimGrid:" file: a map snapshot, or a
  // NanoVDB FloatGrid (a .nvdb file).
  explicit SimManager(std::shared_ptr<config::Config> config, int device = 0)
      : _config(std::move(config)), _device(device), _manager(_config, device) {
    if (_config->simGrid().empty())
      throw std::runtime_error("SimManager: the config names no SimGrid");
    const std::string& g = _config->simGrid();
    if (g.size() > 5 && g.compare(g.size() - 5, 5, ".nvdb") == 0) {
      // a NanoVDB FloatGrid (the first in the file), read with the map's
      // clamps and thresholds
      _simMapRef = std::make_shared<map::Map>(g, "", _config->probHitThres(),
                                              _config->probMissThres(), device);
      _simMap = std::shared_ptr<nm_map>(_simMapRef->handle(), [](nm_map*) {});
    } else {
      nm_map* m = nullptr;
      check(nm_map_load(device, g.c_str(), 0, &m));
      _simMap.reset(m, nm_map_destroy);
    }
    init();
  }
  // The grid from a map in memory.
  SimManager(std::shared_ptr<config::Config> config,
             std::shared_ptr<map::Map> simMap, int device = 0)
      : _config(std::move(config)), _device(device), _manager(_config, device),
        _simMapRef(std::move(simMap)) {
    _simMap = std::shared_ptr<nm_map>(_simMapRef->handle(), [](nm_map*) {});
    init();
  }
  ~SimManager() { closeHandler(); }
  SimManager(const SimManager&) = delete;
  SimManager& operator=(const SimManager&) = delete;

  // SimManager.h:85-91
  void updateAgentPoses(std::vector<Pose> poses) {
    for (size_t i = 0; i < poses.size() && i < _agents.size(); ++i)
      _agents[i]->updatePose(poses[i]);
  }
  std::vector<Pose> agentPoses() const {
    std::vector<Pose> out;
    for (const auto& a : _agents) out.push_back(a->pose());
    return out;
  }

  // SimManager.h:100-113: every sensor of every agent into its map.
  void updateAgentViews() { views(nullptr); }
  // SimManager.h:115-129: every view into one map, agent after agent.
  void updateAgentViews(std::shared_ptr<map::Map> map_ptr) { views(map_ptr); }

  std::shared_ptr<map::Map> getMap(int index) { return _agents.at(index)->map(); }
  std::vector<std::shared_ptr<agent::Agent>>& agents() { return _agents; }

  // The last simulated cloud of agent `index`'s sensor `j` (host copy of the
  // sensor-frame xyz points, NaN where no return; the reference publishes
  // it with PublishSensor 1, SimHandler.cpp:248-250).
  std::vector<float> lastCloud(int index, int j) const {
    for (const auto& v : _last)
      if (v.agent == index && v.sensor == j) {
        std::vector<float> out((size_t)v.n * 3);
        check(nm_device_download(_device, out.data(), _xyz + v.off * 3,
                                 (int64_t)out.size() * 4));
        return out;
      }
    throw std::runtime_error("SimManager::lastCloud: no such view");
  }

  // SimManager.h:141-167: true when the move by positionDelta (metres) is
  // free, false when an observation ray meets an active voxel within the
  // agent's collision distance.
  bool getAgentSimCollision(int index, Vector3f positionDelta) {
    const auto& a = _agents.at(index);
    if (a->numObservations() <= 0) return true;
    float eye[3];
    for (int k = 0; k < 3; ++k) eye[k] = a->pose().position(k) + positionDelta(k);
    return !collides(eye, *a);
  }

  // SimManager.h:214-290, seeded: a voxel of the grid's active bounding
  // box, observed free (below the miss threshold), whose observation sphere
  // (of agent 0, or six axis rays of 5 voxels without agents) has no active
  // voxel within the collision distance; orientation as the reference's
  // spawn.
  Pose getRandomSpawn(uint64_t seed = 0x5eed) {
    int64_t active = 0;
    int32_t bbox[6];
    check(nm_simgrid_info(_grid, &active, bbox, nullptr));
    if (active <= 0) throw std::runtime_error("getRandomSpawn: empty grid");
    uint64_t st = seed;
    auto next = [&]() {
      uint64_t z = (st += 0x9e3779b97f4a7c15ull);
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      return z ^ (z >> 31);
    };
    for (int tries = 0; tries < 20000; ++tries) {
      int32_t v[3];
      for (int k = 0; k < 3; ++k) {
        const uint64_t span = (uint64_t)(bbox[3 + k] - bbox[k] + 1);
        v[k] = bbox[k] + (int32_t)(next() % span);
      }
      float value = 0.0f;
      uint8_t act = 0;
      check(nm_map_query(_simMap.get(), v, 1, &value, &act, 0, nullptr));
      if (act || !(value < _missThr)) continue;  // observed free only
      float eye[3];
      for (int k = 0; k < 3; ++k) eye[k] = ((float)v[k] + 0.5f) * _simRes;
      if (!_agents.empty() ? collides(eye, *_agents[0]) : collidesAxes(eye))
        continue;
      Pose p;
      p.position = Vector3f(eye[0], eye[1], eye[2]);
      p.orientation = agent::Agent::bodyOrientation();
      return p;
    }
    throw std::runtime_error("getRandomSpawn: no free voxel found");
  }

  void closeHandler() {
    if (_batch) nm_batch_destroy(_batch);
    _batch = nullptr;
    if (_grid) nm_simgrid_destroy(_grid);
    _grid = nullptr;
    if (_xyz) nm_device_free(_device, _xyz);
    _xyz = nullptr;
    _xyz_bytes = 0;
    _manager.closeHandler();
  }

  Manager& manager() { return _manager; }

 private:
  struct View {
    int agent, sensor;
    sensor::SensorData* s;
    nm_frame f;
    size_t off;  // first point in the simulated-cloud buffer
    int n;
  };

  void init() {
    check(nm_map_params(_simMap.get(), &_simRes, nullptr, nullptr, &_hitThr,
                        &_missThr));
    check(nm_simgrid_from_map(_simMap.get(), _hitThr, &_grid));
    for (const auto& file : _config->agentFiles())
      _agents.push_back(std::make_shared<agent::Agent>(
          file, _config->mappingRes(), _config->gridRes(),
          _config->probHitThres(), _config->probMissThres(), _device));
    // one seed per agent, so random spawns differ from agent to agent
    for (size_t i = 0; i < _agents.size(); ++i)
      if (_agents[i]->spawnRandom())
        _agents[i]->updatePose(getRandomSpawn(0x5eed + 7919 * (uint64_t)i));
  }

  // The simulated sensor of a view: the frame's rotation and position, the
  // grid's voxel size; a frustum's horizontal FOV derived from the pixel
  // aspect (FrustumData.h:24).
  nm_sim_sensor simSensor(const sensor::SharedParameters& sh,
                          const Pose& pose, const nm_frame& f) const {
    nm_sim_sensor q{};
    for (int k = 0; k < 3; ++k) q.position[k] = pose.position(k);
    for (int k = 0; k < 9; ++k) q.rotation[k] = f.rotation[k];
    q.grid_res = _simRes;
    q.width = sh._hRes;
    q.height = sh._vRes;
    q.vfov_deg = sh._vfov;
    q.hfov_deg = sh._type == 0
                     ? (float)(2.0 * std::atan(sh._aspect *
                                               std::tan(sh._vfov * M_PI / 360)) *
                               180.0 / M_PI)
                     : sh._hfov;
    q.min_range = sh._minRange;
    q.max_range = sh._maxRange;
    q.type = sh._type == 1 ? NM_SENSOR_LASER : NM_SENSOR_FRUSTUM;
    q.rays = nullptr;
    return q;
  }

  void views(std::shared_ptr<map::Map> shared) {
    std::vector<View> views;
    std::vector<nm_sim_sensor> sims;
    size_t total = 0;
    for (size_t i = 0; i < _agents.size(); ++i) {
      const auto& names = _agents[i]->sensorNames();
      for (size_t j = 0; j < names.size(); ++j) {
        auto s = _config->sensorData(names[j]);
        if (!s) throw std::runtime_error("SimManager: no sensor " + names[j]);
        const auto& sh = s->sharedParameters();
        const Pose p = _agents[i]->sensorPose((int)j);
        View v{(int)i, (int)j, s.get(),
               _manager.frameFor(*s, p, sh._hRes, sh._vRes), total,
               sh._hRes * sh._vRes};
        sims.push_back(simSensor(sh, p, v.f));
        views.push_back(v);
        total += (size_t)v.n;
      }
    }
    _last = views;
    if (views.empty()) return;
    if (total * 12 > _xyz_bytes) {
      if (_xyz) check(nm_device_free(_device, _xyz));
      _xyz = nullptr;
      _xyz_bytes = 0;
      void* p = nullptr;
      check(nm_device_alloc(_device, (int64_t)total * 12, &p));
      _xyz = static_cast<float*>(p);
      _xyz_bytes = total * 12;
    }
    int returns = 0;
    check(nm_sim_cast(_grid, sims.data(), (int)sims.size(), nullptr, nullptr,
                      _xyz, &returns, nullptr));
    if (shared) {
      // one map: agent after agent, as the reference's loop
      for (const auto& v : views)
        _manager.insertDeviceCloud(v.s->getName(), v.f.width, v.f.height, 12,
                                   _xyz + v.off * 3, _agents[v.agent]->sensorPose(v.sensor),
                                   shared);
      return;
    }
    // each agent's map: the unfiltered views of one sensor round in one
    // batch, the filtered ones frame by frame
    size_t rounds = 0;
    for (const auto& a : _agents) rounds = std::max(rounds, a->sensorNames().size());
    for (size_t j = 0; j < rounds; ++j) {
      std::vector<nm_agent_frame> batch;
      for (const auto& v : views) {
        if (v.sensor != (int)j) continue;
        auto map = _agents[v.agent]->map();
        if (v.f.filter != NM_FILTER_NONE) {
          _manager.insertDeviceCloud(v.s->getName(), v.f.width, v.f.height, 12,
                                     _xyz + v.off * 3,
                                     _agents[v.agent]->sensorPose(v.sensor), map);
          continue;
        }
        reserveFor(map->handle(), v.f);
        nm_agent_frame a{};
        a.frame = v.f;
        a.cloud = _xyz + v.off * 3;
        a.n_points = v.n;
        a.step = 12;
        a.map = map->handle();
        batch.push_back(a);
      }
      if (!batch.empty()) submitBatch(batch);
    }
  }

  // Leaf cube of a frame (every leaf its rays can touch).
  static int64_t cubeLeaves(const nm_frame& f) {
    const int64_t r = (int64_t)std::ceil(((double)f.max_time + 2.0) / 8.0) + 1;
    const int64_t side = 2 * r + 1;
    return side * side * side;
  }

  // A batch does not grow maps: make room for every leaf the frame can add.
  static void reserveFor(nm_map* m, const nm_frame& f) {
    int64_t n = 0, cap = 0;
    check(nm_map_leaf_count(m, &n));
    check(nm_map_capacity(m, &cap));
    const int64_t need = n + cubeLeaves(f);
    if (need > cap) check(nm_map_reserve(m, std::max<int64_t>(need, 2 * cap)));
  }

  void submitBatch(const std::vector<nm_agent_frame>& frames) {
    int64_t points = 0, tiles = 0;
    float max_time = 0.0f;
    for (const auto& a : frames) {
      points += a.n_points;
      tiles += cubeLeaves(a.frame);
      max_time = std::max(max_time, a.frame.max_time);
    }
    tiles = std::min<int64_t>(tiles, (int64_t)1 << 22);
    if (!_batch || (int)frames.size() > _batchCfg.max_frames ||
        points > _batchCfg.max_points || max_time + 1.0f > _batchCfg.max_time ||
        tiles > _batchCfg.tile_capacity) {
      if (_batch) nm_batch_destroy(_batch);
      _batch = nullptr;
      _batchCfg.device = _device;
      _batchCfg.max_frames = (int)frames.size();
      _batchCfg.max_points = points;
      _batchCfg.max_time = max_time + 1.0f;
      _batchCfg.tile_capacity = tiles;
      check(nm_batch_create(&_batchCfg, &_batch));
    }
    check(nm_batch_submit(_batch, frames.data(), (int)frames.size(), nullptr));
    std::vector<nm_frame_stats> st(frames.size());
    check(nm_batch_stats(_batch, st.data(), (int)frames.size()));
    for (const auto& s : st)
      if (s.errors != 0 || s.overflow != 0)
        throw std::runtime_error("updateAgentViews: a view overflowed a capacity");
  }

  bool collides(const float eye[3], const agent::Agent& a) {
    nm_sim_sensor q{};
    for (int k = 0; k < 3; ++k) q.position[k] = eye[k];
    q.rotation[0] = q.rotation[4] = q.rotation[8] = 1.0f;
    q.grid_res = _simRes;
    q.width = a.numObservations();
    q.height = 1;
    q.min_range = 0.0f;
    q.max_range = a.observationNorms()[0] * _simRes;  // collision distance
    q.type = NM_SENSOR_RAYS;
    q.rays = a.observationRays().data();
    return castRays(q) > 0;
  }

  bool collidesAxes(const float eye[3]) {
    static const float axes[18] = {1, 0, 0, -1, 0, 0, 0, 1, 0,
                                   0, -1, 0, 0, 0, 1, 0, 0, -1};
    nm_sim_sensor q{};
    for (int k = 0; k < 3; ++k) q.position[k] = eye[k];
    q.rotation[0] = q.rotation[4] = q.rotation[8] = 1.0f;
    q.grid_res = _simRes;
    q.width = 6;
    q.height = 1;
    q.max_range = 5.0f * _simRes;
    q.type = NM_SENSOR_RAYS;
    q.rays = axes;
    return castRays(q) > 0;
  }

  int castRays(const nm_sim_sensor& q) {
    int returns = 0;
    check(nm_sim_cast(_grid, &q, 1, nullptr, nullptr, nullptr, &returns,
                      nullptr));
    return returns;
  }

  std::shared_ptr<config::Config> _config;
  int _device;
  Manager _manager;
  std::shared_ptr<map::Map> _simMapRef;
  std::shared_ptr<nm_map> _simMap;
  float _simRes = 0.0f, _hitThr = 0.0f, _missThr = 0.0f;
  nm_simgrid* _grid = nullptr;
  std::vector<std::shared_ptr<agent::Agent>> _agents;
  std::vector<View> _last;
  float* _xyz = nullptr;
  size_t _xyz_bytes = 0;
  nm_batch* _batch = nullptr;
  nm_batch_config _batchCfg{};
};

}  // namespace manager
}  // namespace nanomap

#endif
