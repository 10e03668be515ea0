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
//   test_host sim <config> <rounds>
//       SimManager over the config's agents against a loop of Manager
//       calls on the same simulated clouds (see sim() below).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>

#include "nanomap/manager/Manager.h"
#include "nanomap/manager/SimManager.h"

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

static bool same_maps(const std::shared_ptr<nanomap::map::Map>& a,
                      const std::shared_ptr<nanomap::map::Map>& b);

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
  // the reference's read paths: occupiedGrid() holds the same leaves, and
  // occAccessor() reads the same values and states voxel by voxel
  auto grid = m->occupiedGrid();
  if (grid->leafCount() != n || grid->values != val || grid->masks != mask) {
    std::fprintf(stderr, "occupiedGrid differs from leaves()\n");
    return 5;
  }
  auto acc = m->occAccessor();
  for (int64_t l = 0; l < n; l += std::max<int64_t>(1, n / 16)) {
    for (int v = 0; v < 512; v += 37) {
      const int i = o[3 * l] + (v >> 6), j = o[3 * l + 1] + ((v >> 3) & 7),
                k = o[3 * l + 2] + (v & 7);
      float pv = 0.0f;
      const bool on = acc->probeValue(i, j, k, pv);
      const bool want_on = (mask[16 * l + (v >> 5)] >> (v & 31)) & 1u;
      if (pv != val[512 * l + v] || acc->getValue(i, j, k) != pv ||
          on != want_on || acc->isValueOn(i, j, k) != on) {
        std::fprintf(stderr, "occAccessor differs at leaf %lld voxel %d\n",
                     (long long)l, v);
        return 5;
      }
    }
  }
  // the map as a NanoVDB grid file (<out>.nvdb), and a map read back from
  // it holds the same leaves
  const std::string nv = std::string(argv[6]) + ".nvdb";
  m->writeNanoVDB(nv, "occupancy");
  auto back = std::make_shared<nanomap::map::Map>(
      nv, "occupancy", cfg->probHitThres(), cfg->probMissThres());
  if (back->gridRes() != m->gridRes() || !same_maps(m, back)) {
    std::fprintf(stderr, "NanoVDB round trip differs\n");
    return 6;
  }
  std::ofstream out(argv[6], std::ios::binary);
  out.write(reinterpret_cast<const char*>(&n), 8);
  out.write(reinterpret_cast<const char*>(o.data()), o.size() * 4);
  out.write(reinterpret_cast<const char*>(val.data()), val.size() * 4);
  out.write(reinterpret_cast<const char*>(mask.data()), mask.size() * 4);
  manager.printUpdateTime((int)manager.frames());
  manager.closeHandler();
  return out ? 0 : 3;
}

static bool same_maps(const std::shared_ptr<nanomap::map::Map>& a,
                      const std::shared_ptr<nanomap::map::Map>& b) {
  std::vector<int32_t> oa, ob;
  std::vector<float> va, vb;
  std::vector<uint32_t> ma, mb;
  a->leaves(oa, va, ma);
  b->leaves(ob, vb, mb);
  if (oa.size() != ob.size()) return false;
  // leaf order may differ: compare leaf by leaf through the origins
  std::map<std::tuple<int, int, int>, size_t> idx;
  for (size_t l = 0; l < ob.size() / 3; ++l)
    idx[std::make_tuple(ob[3 * l], ob[3 * l + 1], ob[3 * l + 2])] = l;
  for (size_t l = 0; l < oa.size() / 3; ++l) {
    auto it = idx.find(std::make_tuple(oa[3 * l], oa[3 * l + 1], oa[3 * l + 2]));
    if (it == idx.end()) return false;
    if (std::memcmp(&va[512 * l], &vb[512 * it->second], 2048) != 0) return false;
    if (std::memcmp(&ma[16 * l], &mb[16 * it->second], 64) != 0) return false;
  }
  return true;
}

// test_host sim <config> <rounds>: SimManager over the config's agents for
// <rounds> rounds (every agent moved 0.1 m along x between rounds), then
// the same simulated clouds through a loop of Manager::insertPointCloud
// into fresh maps; prints the per-agent leaf counts and whether every map
// is identical, then collision and spawn queries.
static int sim(int argc, char** argv) {
  auto cfg = std::make_shared<nanomap::config::Config>(argv[2]);
  const int rounds = std::atoi(argv[3]);
  nanomap::manager::SimManager sm(cfg);
  nanomap::manager::Manager loop(cfg);
  const size_t n = sm.agents().size();
  std::vector<std::shared_ptr<nanomap::map::Map>> ref;
  for (size_t i = 0; i < n; ++i)
    ref.push_back(std::make_shared<nanomap::map::Map>(
        cfg->mappingRes(), cfg->probHitThres(), cfg->probMissThres()));
  int64_t returns = 0;
  for (int r = 0; r < rounds; ++r) {
    if (r > 0) {
      std::vector<nanomap::Pose> poses = sm.agentPoses();
      for (auto& p : poses) p.position(0) += 0.1f;
      sm.updateAgentPoses(poses);
    }
    sm.updateAgentViews();
    for (size_t i = 0; i < n; ++i) {
      auto& a = sm.agents()[i];
      for (size_t j = 0; j < a->sensorNames().size(); ++j) {
        auto s = cfg->sensorData(a->sensorNames()[j]);
        const auto& sh = s->sharedParameters();
        std::vector<float> cloud = sm.lastCloud((int)i, (int)j);
        for (size_t k = 0; k < cloud.size(); k += 3)
          returns += cloud[k] == cloud[k] ? 1 : 0;
        loop.insertPointCloud(a->sensorNames()[j], sh._hRes, sh._vRes, 12,
                              reinterpret_cast<unsigned char*>(cloud.data()),
                              a->sensorPose((int)j), ref[i]);
      }
    }
  }
  bool all = true;
  int64_t leaves = 0;
  for (size_t i = 0; i < n; ++i) {
    const bool same = same_maps(sm.getMap((int)i), ref[i]);
    all = all && same;
    leaves += sm.getMap((int)i)->leafCount();
    if (!same) std::printf("agent %zu differs\n", i);
  }
  std::printf("agents %zu rounds %d returns %lld leaves %lld identical %d\n", n,
              rounds, (long long)returns, (long long)leaves, all ? 1 : 0);
  // collisions: agent 0's observation rays (sphere directions), its pose,
  // and the first of 200 moves of 0.05 k m along +x that collides (-1 if
  // none); then a seeded spawn
  auto& a0 = *sm.agents()[0];
  std::printf("rays %d %a\n", a0.numObservations(),
              a0.observationNorms()[0]);
  for (int k = 0; k < a0.numObservations(); ++k)
    std::printf("ray %a %a %a\n", a0.observationRays()[3 * k],
                a0.observationRays()[3 * k + 1], a0.observationRays()[3 * k + 2]);
  std::printf("pose0 %a %a %a\n", a0.pose().position(0), a0.pose().position(1),
              a0.pose().position(2));
  int first = -1;
  for (int k = 0; k < 200 && first < 0; ++k)
    if (!sm.getAgentSimCollision(0, nanomap::Vector3f(0.05f * k, 0, 0)))
      first = k;
  nanomap::Pose sp = sm.getRandomSpawn(7);
  std::printf("collision_step %d\nspawn %a %a %a\n", first, sp.position(0),
              sp.position(1), sp.position(2));
  sm.closeHandler();
  loop.closeHandler();
  return all ? 0 : 4;
}

int main(int argc, char** argv) {
  try {
    if (argc >= 3 && std::string(argv[1]) == "parse") return parse(argv[2]);
    if (argc >= 7 && std::string(argv[1]) == "map") return map(argc, argv);
    if (argc >= 4 && std::string(argv[1]) == "sim") return sim(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "usage: test_host parse <config> | map <config> "
                       "<sensor> <w> <h> <out> <cloud,px,py,pz,qw,qx,qy,qz>...\n");
  return 2;
}
