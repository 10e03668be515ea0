// nanomap/manager/Manager.h -- the mapping entry point of the host API.
//
// Mirrors include/nanomap/manager/Manager.h:47-79: a Manager built from a
// Config; insertPointCloud(sensorName, width, height, step, cloud, pose,
// map) updates the sensor's pose and folds one point cloud into the map.
// In the reference that call runs Handler::processPointCloud (filterCloud
// + frustumCastCloud on the GPU, then a host loop into the OpenVDB grid);
// here it is one nm_frame_submit into the device-resident map, with the
// reference CPU path's semantics (bit-exact to it, including its
// round-toward-minus-infinity arithmetic).  Like the reference call it
// returns once the map is updated; a frame that overflows a capacity throws
// (the map grows instead of overflowing: map::Map turns auto-grow on).
#ifndef NANOMAP_AMD_HOST_MANAGER_H
#define NANOMAP_AMD_HOST_MANAGER_H

#include <chrono>
#include <cstdio>
#include <map>
#include <memory>
#include <string>

#include "../config/Config.h"
#include "../map/OccupancyMap.h"

namespace nanomap {
namespace manager {

using Map = nanomap::map::Map;
using SensorData = nanomap::sensor::SensorData;

class Manager {
 public:
  explicit Manager(std::shared_ptr<nanomap::config::Config> config, int device = 0)
      : _config(std::move(config)), _device(device) {}
  ~Manager() { closeHandler(); }
  Manager(const Manager&) = delete;
  Manager& operator=(const Manager&) = delete;

  // Handler.cpp:250-278 for every sensor named sensorName (the reference
  // loops over all sensors and processes each match, Manager.h:52-60).
  void insertPointCloud(std::string sensorName, int pcl_width, int pcl_height,
                        int pcl_step, unsigned char* cloud,
                        const nanomap::Pose& pose,
                        std::shared_ptr<nanomap::map::Map> map) {
    insert(sensorName, pcl_width, pcl_height, pcl_step, cloud, 0, pose, map);
  }

  // insertPointCloud for a cloud already in device memory (a simulated
  // cloud, a camera driver's device buffer); no counterpart in the
  // reference, whose clouds always come from the host.
  void insertDeviceCloud(std::string sensorName, int pcl_width,
                         int pcl_height, int pcl_step, const void* device_cloud,
                         const nanomap::Pose& pose,
                         std::shared_ptr<nanomap::map::Map> map) {
    insert(sensorName, pcl_width, pcl_height, pcl_step, device_cloud, 1, pose,
           map);
  }

  // The frame of one cloud of `s` at `pose`: R = normalized(q) F, the eye,
  // max_time, the sensor's log-odds, round-down (the reference's host
  // mode), the cast's filter and, for the voxel filter, the frustum's
  // sensor-frame box.  Updates the sensor's pose (SensorData::updatePose,
  // rotateView) as the reference does.
  nm_frame frameFor(SensorData& s, const nanomap::Pose& pose, int width,
                    int height) const {
    s.updatePose(pose);
    s.rotateView();
    const auto& sh = s.sharedParameters();
    const float pos[3] = {pose.position(0), pose.position(1), pose.position(2)};
    const float quat[4] = {pose.orientation.w(), pose.orientation.x(),
                           pose.orientation.y(), pose.orientation.z()};
    float F[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) F[3 * r + c] = sh._frameTransform[r][c];
    nm_frame f;
    check(nm_frame_from_pose(pos, quat, F, sh._gridRes, sh._maxRange,
                             s.logOddsHit(), s.logOddsMiss(), NM_ROUND_DOWN,
                             &f));
    f.width = width;
    f.height = height;
    f.filter = filterFor(sh);
    if (f.filter != NM_FILTER_NONE) {
      // the voxel filter drops points outside the frustum's sensor-frame
      // box (filterCloud.cu:100-111)
      for (int a = 0; a < 3; ++a) {
        f.box_min[a] = sh._worldMin[a];
        f.box_max[a] = sh._worldMax[a];
      }
    }
    return f;
  }

  void printUpdateTime(int count) {
    std::printf("average frame time over %d frames: %.3f ms\n", count,
                count > 0 ? 1e3 * _seconds / count : 0.0);
  }

  void closeHandler() {
    for (auto& kv : _contexts) nm_ctx_destroy(kv.second);
    _contexts.clear();
  }

  std::shared_ptr<nanomap::config::Config> config() { return _config; }
  int64_t frames() const { return _frames; }

  // The device workspace of sensor `s` for clouds of n points.
  nm_ctx* context(const SensorData& s, int n, float max_time) {
    auto it = _contexts.find(s.getName());
    if (it != _contexts.end()) return it->second;
    nm_ctx_config c;
    c.device = _device;
    c.max_points = std::max(n, s.sharedParameters()._pclSize);
    c.max_time = max_time + 1.0f;
    c.tile_capacity = 0;
    nm_ctx* ctx = nullptr;
    check(nm_ctx_create(&c, &ctx));
    _contexts[s.getName()] = ctx;
    return ctx;
  }

  // The cast of a sensor's clouds: frustum sensors with FilterType 1 take
  // the voxel filter of their PrecisionType (filterCloud.cu:476-553);
  // laser sensors and FilterType 0 cast every point, as the reference's
  // CPU path does (Handler.cpp:247-249, 28-86).
  int filterFor(const sensor::SharedParameters& sh) const {
    if (sh._type != 0 || _config->filterType() != 1) return NM_FILTER_NONE;
    const int prec = _config->precisionType();
    if (prec < 0 || prec > 3)
      throw std::runtime_error("FilterType 1 needs PrecisionType 0-3, got " +
                               std::to_string(prec));
    return NM_FILTER_VOXEL + prec;
  }

 private:
  void insert(const std::string& sensorName, int pcl_width, int pcl_height,
              int pcl_step, const void* cloud, int on_device,
              const nanomap::Pose& pose, std::shared_ptr<map::Map> map) {
    for (auto& s : _config->sensorData()) {
      if (s->sensorName() != sensorName) continue;
      const auto t0 = std::chrono::steady_clock::now();
      nm_frame f = frameFor(*s, pose, pcl_width, pcl_height);
      const int n = pcl_width * pcl_height;
      nm_ctx* ctx = context(*s, n, f.max_time);
      check(nm_frame_submit(ctx, map->handle(), &f, cloud, n, pcl_step,
                            on_device, nullptr));
      // waits for the update, as the reference call does; a frame that
      // did not fit is an error, not a silently partial map
      nm_frame_stats st;
      check(nm_frame_stats_get(ctx, &st));
      if (st.errors != 0 || st.overflow != 0)
        throw std::runtime_error(
            "insertPointCloud: frame incomplete (" +
            std::to_string(st.overflow) + " capacity overflows, " +
            std::to_string(st.errors) + " errors)");
      _seconds += std::chrono::duration<double>(
                      std::chrono::steady_clock::now() - t0).count();
      ++_frames;
    }
  }

  std::shared_ptr<nanomap::config::Config> _config;
  int _device;
  std::map<std::string, nm_ctx*> _contexts;
  double _seconds = 0.0;
  int64_t _frames = 0;
};

}  // namespace manager
}  // namespace nanomap

#endif
