// nanomap/map/OccupancyMap.h -- map::Map of the host API.
//
// Mirrors include/nanomap/map/OccupancyMap.h:35-61: the constructor takes
// the voxel size and the hit/miss probability thresholds and derives the
// log-odds thresholds and the clamping values logit(0.97) / logit(0.12).
// The grid lives in device memory (nm_map); leaves() copies it to the host
// in place of the reference's OpenVDB grid accessor.
#ifndef NANOMAP_AMD_HOST_MAP_H
#define NANOMAP_AMD_HOST_MAP_H

#include <cmath>
#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

#include "../nanomap.h"

namespace nanomap {
namespace map {

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
  }
  ~Map() { nm_map_destroy(_map); }
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

  // Value and activity of voxel (i, j, k) from a host copy (false if the
  // voxel's leaf is not allocated: background 0, inactive).
  bool getValue(int i, int j, int k, float& value, bool& active) const {
    std::vector<int32_t> o;
    std::vector<float> v;
    std::vector<uint32_t> m;
    leaves(o, v, m);
    const int li = (k & 7) + ((j & 7) << 3) + ((i & 7) << 6);
    for (size_t l = 0; l < o.size() / 3; ++l)
      if (o[3 * l] == (i & ~7) && o[3 * l + 1] == (j & ~7) &&
          o[3 * l + 2] == (k & ~7)) {
        value = v[l * 512 + li];
        active = (m[l * 16 + (li >> 5)] >> (li & 31)) & 1u;
        return true;
      }
    value = 0.0f;
    active = false;
    return false;
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
