// test_host.cpp -- drives the host API (nanomap/manager/Manager.h etc.) for
// tests/test_host_api.py.
//
//   test_host parse <config>
//       parse a config with nanomap::config::Config and print its values;
//   test_host map <config> <sensor> <width> <height> <out>
//                 [--capacity <leaves>] <frame>...
//       each <frame> is "<cloud.bin>,px,py,pz,qw,qx,qy,qz": xyz float32
//       points (width x height, stride 12), pose position and orientation;
//       inserts the frames into one nanomap::map::Map (created with room
//       for <leaves> leaves, default 2^18; it grows as needed) with
//       Manager::insertPointCloud and writes the map's leaves to <out>
//       (int64 count, then origins int32 [n,3], values float32 [n,512],
//       masks uint32 [n,16]).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include "nanomap/manager/Manager.h"

static int parse(const char* file) {
  auto c = std::make_shared<nanomap::config::Config>(file);
  std::printf("MappingRes %.9g\nGridRes %.9g\nProbHitThres %.9g\n"
              "ProbMissThres %.9g\nNodeEdge %d\nFilterType %d\nUpdateType %d\n"
              "maxPCLSize %d\nsensors %zu\n",
              c->mappingRes(), c->gridRes(), c->probHitThres(),
              c->probMissThres(), c->leafEdge(), c->filterType(),
              c->updateType(), c->maxPCLSize(), c->sensorData().size());
  for (auto& s : c->sensorData()) {
    const auto& sh = s->sharedParameters();
    std::printf("sensor %s id %d type %d res %dx%d vfov %.9g range %.9g %.9g "
                "prob %.9g %.9g grid %.9g logodds %.9g %.9g F",
                s->getName().c_str(), s->getId(), sh._type, sh._hRes, sh._vRes,
                sh._vfov, sh._minRange, sh._maxRange, sh._probHit,
                sh._probMiss, sh._gridRes, s->logOddsHit(), s->logOddsMiss());
    for (int r = 0; r < 3; ++r)
      for (int col = 0; col < 3; ++col) std::printf(" %.9g", sh._frameTransform[r][col]);
    std::printf("\n");
  }
  auto byname = c->sensorData(c->sensorData().empty() ? "" : c->sensorData()[0]->getName());
  std::printf("lookup %d\n", byname ? 1 : 0);
  return 0;
}

static int map(int argc, char** argv) {
  auto cfg = std::make_shared<nanomap::config::Config>(argv[2]);
  const std::string sensor = argv[3];
  const int w = std::atoi(argv[4]), h = std::atoi(argv[5]);
  int first = 7, capacity = 1 << 18;
  if (argc > 8 && std::string(argv[7]) == "--capacity") {
    capacity = std::atoi(argv[8]);
    first = 9;
  }
  nanomap::manager::Manager manager(cfg);
  auto m = std::make_shared<nanomap::map::Map>(
      cfg->mappingRes(), cfg->probHitThres(), cfg->probMissThres(), 0.0f,
      0.0f, 0, capacity);
  std::vector<float> cloud((size_t)w * h * 3);
  for (int a = first; a < argc; ++a) {
    std::stringstream ss(argv[a]);
    std::string path, tok;
    std::getline(ss, path, ',');
    float v[7];
    for (int k = 0; k < 7; ++k) {
      std::getline(ss, tok, ',');
      v[k] = std::strtof(tok.c_str(), nullptr);
    }
    std::ifstream in(path, std::ios::binary);
    in.read(reinterpret_cast<char*>(cloud.data()), cloud.size() * 4);
    if (!in) {
      std::fprintf(stderr, "cannot read %s\n", path.c_str());
      return 2;
    }
    nanomap::Pose pose;
    pose.position = nanomap::Vector3f(v[0], v[1], v[2]);
    pose.orientation = nanomap::Quaternionf(v[3], v[4], v[5], v[6]);
    manager.insertPointCloud(sensor, w, h, 12,
                             reinterpret_cast<unsigned char*>(cloud.data()),
                             pose, m);
  }
  std::vector<int32_t> o;
  std::vector<float> val;
  std::vector<uint32_t> mask;
  m->leaves(o, val, mask);
  const int64_t n = (int64_t)(o.size() / 3);
  std::ofstream out(argv[6], std::ios::binary);
  out.write(reinterpret_cast<const char*>(&n), 8);
  out.write(reinterpret_cast<const char*>(o.data()), o.size() * 4);
  out.write(reinterpret_cast<const char*>(val.data()), val.size() * 4);
  out.write(reinterpret_cast<const char*>(mask.data()), mask.size() * 4);
  manager.printUpdateTime((int)manager.frames());
  manager.closeHandler();
  return out ? 0 : 3;
}

int main(int argc, char** argv) {
  try {
    if (argc >= 3 && std::string(argv[1]) == "parse") return parse(argv[2]);
    if (argc >= 7 && std::string(argv[1]) == "map") return map(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "usage: test_host parse <config> | map <config> "
                       "<sensor> <w> <h> <out> <cloud,px,py,pz,qw,qx,qy,qz>...\n");
  return 2;
}
