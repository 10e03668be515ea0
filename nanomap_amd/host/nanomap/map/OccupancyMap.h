// nanomap/map/OccupancyMap.h -- map::Map of the host API.
//
// Mirrors include/nanomap/map/OccupancyMap.h:35-61: the constructor takes
// the voxel size and the hit/miss probability thresholds and derives the
// log-odds thresholds and the clamping values logit(0.97) / logit(0.12).
// The grid lives in device memory (nm_map): getValue()/query() and the
// occAccessor() read voxels through device point queries in place of the
// reference's OpenVDB accessor, leaves() and occupiedGrid() copy the whole
// grid to the host, save() and the snapshot constructor store and restore
// it.
#ifndef NANOMAP_AMD_HOST_MAP_H
#define NANOMAP_AMD_HOST_MAP_H

#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../nanomap.h"

namespace nanomap {
namespace map {

// Voxel reads with the method names of the OpenVDB accessor that the
// reference's occAccessor() returns (OccupancyMap.h:60): each call is one
// device point query on the map of record.  Coordinates are index space.
class Accessor {
 public:
  explicit Accessor(nm_map* m) : _m(m) {}
  // log-odds of voxel (i, j, k); 0 (the background) where its leaf is absent
  float getValue(int i, int j, int k) const {
    float v;
    bool a;
    read(i, j, k, v, a);
    return v;
  }
  bool isValueOn(int i, int j, int k) const {
    float v;
    bool a;
    read(i, j, k, v, a);
    return a;
  }
  // value into `value`, returns the active state
  bool probeValue(int i, int j, int k, float& value) const {
    bool a;
    read(i, j, k, value, a);
    return a;
  }

 private:
  void read(int i, int j, int k, float& v, bool& a) const {
    const int32_t ijk[3] = {i, j, k};
    uint8_t act = 0;
    check(nm_map_query(_m, ijk, 1, &v, &act, 0, nullptr));
    a = act != 0;
  }
  nm_map* _m;
};

// Host snapshot of every leaf, the stand-in for the OpenVDB grid that the
// reference's occupiedGrid() returns (OccupancyMap.h:54): origins (index
// space, multiples of 8), 512 log-odds per leaf (voxel z + 8 y + 64 x) and
// 16 active-mask words per leaf.
struct LeafGrid {
  float voxelSize = 0.0f;
  std::vector<int32_t> origins;
  std::vector<float> values;
  std::vector<uint32_t> masks;
  int64_t leafCount() const { return (int64_t)(origins.size() / 3); }
};

class Map {
 public:
  Map(float gridRes, float probThresHit, float probThresMiss,
      float mappingRes = 0, float plannerRes = 0, int nodeEdge = 0,
      int leafCapacity = 1 << 18, int device = 0)
      : _gridRes(gridRes) {
    (void)mappingRes;
    (void)plannerRes;
    (void)nodeEdge;
    _logOddsMissThreshold =
        (float)(std::log((double)probThresMiss) - std::log(1.0 - probThresMiss));
    _logOddsHitThreshold =
        (float)(std::log((double)probThresHit) - std::log(1.0 - probThresHit));
    _occupiedClampingThreshold = (float)(std::log(0.97) - std::log(1 - 0.97));
    _emptyClampingThreshold = (float)(std::log(0.12) - std::log(1 - 0.12));
    check(nm_map_create(device, gridRes, _emptyClampingThreshold,
                        _occupiedClampingThreshold, _logOddsHitThreshold,
                        _logOddsMissThreshold, leafCapacity, &_map));
    // like the reference's OpenVDB grid, the map never runs out of leaves
    check(nm_map_set_auto_grow(_map, 1));
  }
  // A map restored from an nm_map_save snapshot (voxel size, clamps and
  // thresholds come from the file).
  explicit Map(const std::string& snapshot, int device = 0) {
    check(nm_map_load(device, snapshot.c_str(), 0, &_map));
    check(nm_map_set_auto_grow(_map, 1));
    check(nm_map_params(_map, &_gridRes, &_emptyClampingThreshold,
                        &_occupiedClampingThreshold, &_logOddsHitThreshold,
                        &_logOddsMissThreshold));
  }
  // A map from the FloatGrid `gridName` ("" for the first) of a NanoVDB
  // file, with the clamps and thresholds of Map(gridRes, probThresHit,
  // probThresMiss); the voxel size comes from the grid.
  Map(const std::string& nvdbFile, const std::string& gridName,
      float probThresHit, float probThresMiss, int device = 0) {
    _logOddsMissThreshold =
        (float)(std::log((double)probThresMiss) - std::log(1.0 - probThresMiss));
    _logOddsHitThreshold =
        (float)(std::log((double)probThresHit) - std::log(1.0 - probThresHit));
    _occupiedClampingThreshold = (float)(std::log(0.97) - std::log(1 - 0.97));
    _emptyClampingThreshold = (float)(std::log(0.12) - std::log(1 - 0.12));
    check(nm_map_load_nvdb(device, nvdbFile.c_str(), gridName.c_str(),
                           _emptyClampingThreshold, _occupiedClampingThreshold,
                           _logOddsHitThreshold, _logOddsMissThreshold, 0,
                           &_map));
    check(nm_map_set_auto_grow(_map, 1));
    check(nm_map_params(_map, &_gridRes, nullptr, nullptr, nullptr, nullptr));
  }
  ~Map() { nm_map_destroy(_map); }

  void save(const std::string& path) const { check(nm_map_save(_map, path.c_str())); }
  // The map as a NanoVDB FloatGrid file (the grid form the reference hands
  // to its GPU kernels, SimHandler.cpp:22-23).
  void writeNanoVDB(const std::string& path,
                    const std::string& gridName = "occupancy") const {
    check(nm_map_write_nvdb(_map, path.c_str(), gridName.c_str()));
  }
  Map(const Map&) = delete;
  Map& operator=(const Map&) = delete;

  float occupiedClampingThreshold() const { return _occupiedClampingThreshold; }
  float emptyClampingThreshold() const { return _emptyClampingThreshold; }
  float logOddsHitThreshold() const { return _logOddsHitThreshold; }
  float logOddsMissThreshold() const { return _logOddsMissThreshold; }
  float gridRes() const { return _gridRes; }

  int64_t leafCount() const {
    int64_t n = 0;
    check(nm_map_leaf_count(_map, &n));
    return n;
  }

  // Host copy of every leaf: origins (index space, multiples of 8), 512
  // log-odds per leaf (voxel z + 8 y + 64 x), 16 active-mask words.
  void leaves(std::vector<int32_t>& origins, std::vector<float>& values,
              std::vector<uint32_t>& masks) const {
    const int64_t n = leafCount();
    origins.resize((size_t)n * 3);
    values.resize((size_t)n * 512);
    masks.resize((size_t)n * 16);
    if (n) check(nm_map_download(_map, origins.data(), values.data(), masks.data()));
  }

  // Value and activity of voxel (i, j, k), an accessor read
  // (occAccessor()->probeValue, OccupancyMap.h:60): one device point query.
  // A voxel of an absent leaf reads as the background, 0 and inactive.
  // Returns the active state.
  bool getValue(int i, int j, int k, float& value, bool& active) const {
    const int32_t ijk[3] = {i, j, k};
    uint8_t a = 0;
    check(nm_map_query(_map, ijk, 1, &value, &a, 0, nullptr));
    active = a != 0;
    return active;
  }

  // Batched reads: values and active states (1/0) of n voxels (ijk: 3 ints
  // each), from host arrays; device arrays go to nm_map_query directly.
  void query(const int32_t* ijk, int64_t n, float* values,
             uint8_t* active) const {
    check(nm_map_query(_map, ijk, n, values, active, 0, nullptr));
  }

  // OccupancyMap.h:60: a voxel accessor (device point queries)
  std::shared_ptr<Accessor> occAccessor() const {
    return std::make_shared<Accessor>(_map);
  }
  // OccupancyMap.h:54: the grid, as a host copy of every leaf
  std::shared_ptr<LeafGrid> occupiedGrid() const {
    auto g = std::make_shared<LeafGrid>();
    g->voxelSize = _gridRes;
    leaves(g->origins, g->values, g->masks);
    return g;
  }

  nm_map* handle() { return _map; }

 private:
  float _gridRes;
  float _logOddsMissThreshold;
  float _logOddsHitThreshold;
  float _occupiedClampingThreshold;
  float _emptyClampingThreshold;
  nm_map* _map = nullptr;
};

}  // namespace map
}  // namespace nanomap

#endif
