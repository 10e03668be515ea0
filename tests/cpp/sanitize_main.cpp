// sanitize_main -- the host-side C and C++ code under AddressSanitizer and
// UndefinedBehaviorSanitizer (built by nanomap_amd.build.build_sanitize_test
// with g++ -fsanitize=address,undefined; no GPU code is linked).
//
// Drives, on small inputs:
//   - csrc/nm_scene.cpp: scene generation, rendering, walks, pose sampling,
//     nm_frame_from_pose and nm_frustum_box in both rounding modes;
//   - the oracle (oracle/nm_oracle.cpp, nm_soft_oracle.cpp): point and ray
//     casts, the voxel filter in its four precision modes with and without
//     the frustum box, map integration and export, single-ray traces, folds,
//     the sensor-simulation grid and its frustum, laser and explicit-ray
//     casts (hierarchical and plain), and the soft march forward and adjoint;
//   - the header-only host parser (host/nanomap/config/Config.h and
//     sensor/SensorData.h): a config with a frustum and a laser sensor,
//     pose updates and view rotation.
// Exits 0 when every check holds; any sanitizer report aborts the run.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "../../include/nanomap_amd.h"
#include "../../include/nanomap_amd_scene.h"

#include "../../oracle/nm_oracle.h"
#include "nanomap/config/Config.h"

// nanomap::check() reports failures through the library's last-error
// string; this binary links only nm_scene.cpp, whose calls return codes.
extern "C" const char* nm_last_error(void) { return "nm_scene call failed"; }
// nm_nvdb.cpp reports errors through the library's message slot, which
// lives in the HIP part of the library
namespace nm {
void set_error(const std::string&) {}
}

static int failures = 0;
#define EXPECT(c)                                                   \
  do {                                                              \
    if (!(c)) {                                                     \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__,   \
                   __LINE__, #c);                                   \
      ++failures;                                                   \
    }                                                               \
  } while (0)

static const float kDefaultF[9] = {0, 0, 1, -1, 0, 0, 0, -1, 0};

static void scene_and_oracle() {
  const double world[3] = {12.0, 12.0, 4.0};
  nm_scene* scene = nullptr;
  EXPECT(nm_scene_create(3, world, 30, 0.3, 1.5, 20, 0.1, 0.4, &scene) == NM_OK);
  nm_camera cam{};
  cam.width = 64;
  cam.height = 48;
  cam.hfov_deg = 87.0f;
  cam.vfov_deg = 58.0f;
  cam.max_range = 6.0f;
  cam.noise_k = 0.002f;
  cam.dropout = 0.05f;
  for (int k = 0; k < 9; ++k) cam.frame[k] = kDefaultF[k];
  float pos[3 * 4], quat[4 * 4];
  EXPECT(nm_scene_walk(scene, 9, 4, 0.1f, 2.0f, 15.0f, 0.3f, pos, quat) == NM_OK);
  float sp[3], sq[4];
  EXPECT(nm_scene_sample_pose(scene, 17, 15.0f, 0.3f, sp, sq) == NM_OK);
  const int P = cam.width * cam.height;
  std::vector<float> cloud((size_t)P * 3);

  or_map* maps[2] = {or_map_create(-2.0f, 3.5f, 0.0f, 0.0f),
                     or_map_create(-2.0f, 3.5f, 0.0f, 0.0f)};
  for (int f = 0; f < 4; ++f) {
    EXPECT(nm_scene_render(scene, &cam, pos + 3 * f, quat + 4 * f, 100 + f,
                           cloud.data()) == NM_OK);
    for (int rounding = 0; rounding < 2; ++rounding) {
      nm_frame fr{};
      EXPECT(nm_frame_from_pose(pos + 3 * f, quat + 4 * f, kDefaultF, 0.1f,
                                5.0f, 0.85f, -0.4f, rounding, &fr) == NM_OK);
      float bmin[3], bmax[3];
      EXPECT(nm_frustum_box(cam.width, cam.height, cam.vfov_deg, 0.1f, 5.0f,
                            rounding, bmin, bmax) == NM_OK);
      EXPECT(bmin[2] <= 0.1f && bmax[2] >= 5.0f - 1e-5f);
      or_temp* t = or_temp_create();
      const int n = or_cast_cloud(t, reinterpret_cast<unsigned char*>(cloud.data()),
                                  P, 12, fr.rotation, fr.eye, fr.grid_res,
                                  fr.max_time, 0.85f, -0.4f, rounding);
      EXPECT(n > 0 && n <= P);
      const int64_t nv = or_temp_size(t);
      std::vector<int32_t> ijk((size_t)nv * 3);
      std::vector<uint32_t> miss(nv), hit(nv);
      std::vector<float> val(nv);
      or_temp_export(t, ijk.data(), miss.data(), hit.data(), val.data());
      uint64_t sum = 0;
      for (int64_t v = 0; v < nv; ++v) sum += miss[v] + hit[v];
      EXPECT(sum == or_temp_visits(t));
      or_map_integrate(maps[rounding], t, rounding);
      // voxel filter, four precision modes, with and without the box
      std::vector<float> rays((size_t)P * 4);
      for (int precision = 0; precision < 4; ++precision) {
        for (int boxed = 0; boxed < 2; ++boxed) {
          float box[6] = {bmin[0], bmin[1], bmin[2], bmax[0], bmax[1], bmax[2]};
          const int nr = or_voxel_filter(
              reinterpret_cast<unsigned char*>(cloud.data()), P, 12,
              fr.rotation, fr.eye, fr.grid_res, fr.max_time, 5.0f,
              boxed ? box : nullptr, rays.data(), P, precision);
          EXPECT(nr >= 0 && nr <= P);
          or_temp* tr = or_temp_create();
          EXPECT(or_cast_rays(tr, rays.data(), nr, fr.eye, fr.max_time, 0.85f,
                              -0.4f, rounding) >= 0);
          or_temp_destroy(tr);
        }
      }
      or_temp_clear(t);
      EXPECT(or_temp_size(t) == 0);
      or_temp_destroy(t);
      // one traced ray
      int32_t out[3 * 256];
      int n_out = 0, hitf = 0;
      const float d[3] = {1.3f, -0.7f, 2.9f};
      if (or_dda_trace(fr.eye, d, 0.1f, 50.0f, rounding, out, 256, &n_out, &hitf))
        EXPECT(n_out > 0 && n_out <= 256);
    }
  }
  // fold against sequential adds
  for (int rounding = 0; rounding < 2; ++rounding) {
    const float a = or_fold(1.0f, -1e-20f, 30, rounding);
    EXPECT(std::isfinite(a));
  }
  // the simulation grid from the round-down map
  const int64_t L = or_map_leaf_count(maps[1]);
  EXPECT(L > 0);
  std::vector<int32_t> org((size_t)L * 3);
  std::vector<float> values((size_t)L * 512);
  std::vector<uint32_t> masks((size_t)L * 16);
  or_map_export(maps[1], org.data(), values.data(), masks.data());
  float v0 = 0.0f;
  int act = 0;
  EXPECT(or_map_get(maps[1], org[0], org[1], org[2], &v0, &act) == 1);
  or_simgrid* g = or_simgrid_build(org.data(), values.data(), L, 0.0f);
  EXPECT(or_simgrid_active_count(g) > 0);
  const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const float eye[3] = {pos[0], pos[1], pos[2]};
  std::vector<int32_t> sijk((size_t)P * 3);
  std::vector<float> dist(P), sxyz((size_t)P * 3);
  const float dirs[6 * 3] = {1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1};
  for (int hier = 0; hier < 2; ++hier) {
    EXPECT(or_sim_cast(g, eye, I, 0.1f, cam.width, cam.height, 87.0f, 58.0f,
                       0.1f, 6.0f, 0, nullptr, hier, sijk.data(), dist.data(),
                       sxyz.data()) >= 0);
    EXPECT(or_sim_cast(g, eye, I, 0.1f, 90, 8, 360.0f, 30.0f, 0.1f, 6.0f, 1,
                       nullptr, hier, sijk.data(), dist.data(), sxyz.data()) >= 0);
    EXPECT(or_sim_cast(g, eye, I, 0.1f, 6, 1, 0.0f, 0.0f, 0.0f, 3.0f, 2, dirs,
                       hier, sijk.data(), dist.data(), sxyz.data()) >= 0);
  }
  or_simgrid_destroy(g);
  or_map_destroy(maps[0]);
  or_map_destroy(maps[1]);
  nm_scene_destroy(scene);
}

static void soft_march() {
  const int nx = 12, ny = 10, nz = 8;
  std::vector<double> grid((size_t)nx * ny * nz);
  for (size_t k = 0; k < grid.size(); ++k) grid[k] = std::sin(0.37 * (double)k) * 2.0;
  const double origin[3] = {0, 0, 0};
  double cams[2 * 12];
  for (int c = 0; c < 2; ++c) {
    const double R[9] = {0, 0, 1, -1, 0, 0, 0, -1, 0};
    for (int k = 0; k < 9; ++k) cams[12 * c + k] = R[k];
    cams[12 * c + 9] = 0.05;
    cams[12 * c + 10] = 0.5 + 0.1 * c;
    cams[12 * c + 11] = 0.4;
  }
  const int W = 16, H = 12, ns = 24;
  std::vector<double> depth((size_t)2 * W * H, 0.0), gd((size_t)2 * W * H, 1.0);
  std::vector<double> gpose(2 * 6, 0.0), glog(grid.size(), 0.0);
  or_soft_forward(grid.data(), nx, ny, nz, origin, 0.1, -2.0, cams, 2, W, H,
                  60.0, 45.0, 0.05, ns, depth.data());
  or_soft_backward(grid.data(), nx, ny, nz, origin, 0.1, -2.0, cams, 2, W, H,
                   60.0, 45.0, 0.05, ns, gd.data(), gpose.data(), glog.data());
  for (double d : depth) EXPECT(std::isfinite(d));
  for (double d : gpose) EXPECT(std::isfinite(d));
}

static void config_files(const std::string& dir) {
  const std::string cam = dir + "/sanitize_cam.txt", lidar = dir + "/sanitize_lidar.txt",
                    cfg = dir + "/sanitize_config.txt";
  std::ofstream(cam) << "#frustumconfig\nName:\ncam\nId:\n1\nHRes:\n160\nVRes:\n120\n"
                        "VFOV:\n45\nRate:\n30\nMaxRange:\n5\nMinRange:\n0.1\n"
                        "ProbHit:\n0.7\nProbMiss:\n0.4\nFrameTransform:\n"
                        "0.0 -1.0 0.0\n0.0 0.0 1.0\n1.0 0.0 0.0\n#endconfig\n";
  std::ofstream(lidar) << "#laserconfig\nName:\nlidar\nId:\n2\nAngularHRes:\n0.4\n"
                          "AngularVRes:\n2.0\nHFOV:\n360\nVFOV:\n30\nRate:\n10\n"
                          "MaxRange:\n30\nMinRange:\n0.5\nProbHit:\n0.8\n"
                          "ProbMiss:\n0.35\nFrameTransform:\n1.0 0.0 0.0\n"
                          "0.0 1.0 0.0\n0.0 0.0 1.0\n#endconfig\n";
  std::ofstream(cfg) << "#config\nMappingRes:\n0.1\nProbHitThres:\n0.5\n"
                        "ProbMissThres:\n0.5\nGridRes:\n0.1\nNodeEdge:\n8\n"
                        "FrustumAllocationFactor:\n0.4\nFilterType:\n1\n"
                        "UpdateType:\n0\nPrecisionType:\n0\nPublishSensor:\n1\n"
                        "Sensor:\n" << cam << "\nSensor:\n" << lidar << "\n#endconfig\n";
  nanomap::config::Config config(cfg);
  auto c = config.sensorData(std::string("cam"));
  auto l = config.sensorData(std::string("lidar"));
  EXPECT(c && l);
  if (!c || !l) return;
  EXPECT(c->sharedParameters()._hRes == 160 && c->sharedParameters()._vRes == 120);
  EXPECT(l->sharedParameters()._type == 1);
  nanomap::Pose p;
  p.position = nanomap::Vector3f(1.25f, -0.5f, 0.75f);
  p.orientation = nanomap::Quaternionf(0.9f, 0.1f, -0.3f, 0.2f);
  for (auto s : {c, l}) {
    s->updatePose(p);
    s->rotateView();
  }
  EXPECT(c->sharedParameters()._leafBufferSize > 0);
}

// NanoVDB export and import (csrc/nm_nvdb.cpp): a grid of leaves in
// several root, upper and lower nodes written and read back, then damaged
// files (truncated, child offsets out of the buffer) that must be refused.
static void nanovdb_files(const std::string& dir) {
  const int n = 5;
  const int32_t orgs[3 * n] = {0, 0, 0, 8, 0, 0, 128, -8, 0, -4096, 4096, 64, 16, 16, 16};
  std::vector<float> vals((size_t)n * 512);
  std::vector<uint32_t> masks((size_t)n * 16);
  for (size_t k = 0; k < vals.size(); ++k) vals[k] = std::sin(0.1 * (double)k);
  for (size_t k = 0; k < masks.size(); ++k) masks[k] = (uint32_t)(k * 2654435761u);
  uint64_t bytes = 0;
  EXPECT(nm_nvdb_build(orgs, vals.data(), masks.data(), n, 0.1, "g", nullptr,
                       &bytes) == NM_OK);
  std::vector<unsigned char> buf(bytes);
  EXPECT(nm_nvdb_build(orgs, vals.data(), masks.data(), n, 0.1, "g", buf.data(),
                       &bytes) == NM_OK);
  const std::string path = dir + "/sanitize.nvdb";
  EXPECT(nm_nvdb_write(path.c_str(), orgs, vals.data(), masks.data(), n, 0.1,
                       "g") == NM_OK);
  int64_t got = 0;
  double dx = 0;
  EXPECT(nm_nvdb_read(path.c_str(), "g", &got, nullptr, nullptr, nullptr, &dx) == NM_OK);
  EXPECT(got == n && dx == 0.1);
  std::vector<int32_t> o((size_t)got * 3);
  std::vector<float> v((size_t)got * 512);
  std::vector<uint32_t> m((size_t)got * 16);
  EXPECT(nm_nvdb_read(path.c_str(), "", &got, o.data(), v.data(), m.data(), &dx) == NM_OK);
  std::ifstream in(path, std::ios::binary);
  std::vector<char> file((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  const std::string cut = dir + "/sanitize_cut.nvdb";
  std::ofstream(cut, std::ios::binary).write(file.data(), (std::streamsize)file.size() / 2);
  EXPECT(nm_nvdb_read(cut.c_str(), "", &got, nullptr, nullptr, nullptr, &dx) != NM_OK);
  // every 8-byte word of the root table and upper node header set to a
  // large offset, one at a time
  const size_t grid_at = 16 + 176 + 2;
  for (size_t w = 0; w < 64; ++w) {
    std::vector<char> bad = file;
    const size_t at = grid_at + 736 + 8 * w;
    if (at + 8 > bad.size()) break;
    const int64_t big = (int64_t)1 << 40;
    std::memcpy(bad.data() + at, &big, 8);
    const std::string bp = dir + "/sanitize_bad.nvdb";
    std::ofstream(bp, std::ios::binary).write(bad.data(), (std::streamsize)bad.size());
    int64_t nb = 0;
    (void)nm_nvdb_read(bp.c_str(), "", &nb, nullptr, nullptr, nullptr, &dx);
  }
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  scene_and_oracle();
  soft_march();
  config_files(dir);
  nanovdb_files(dir);
  std::printf("sanitize: %d failed checks\n", failures);
  return failures == 0 ? 0 : 1;
}
