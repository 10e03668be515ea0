// nm_nvdb.cpp -- NanoVDB export and import of a map's leaves (host code).
//
// The reference hands its map to the GPU as a NanoVDB grid
// (nanovdb::openToNanoVDB, src/nanomap/handler/SimHandler.cpp:22-23) and
// its simulator reads grids from files (SimManager.h:196-210).  This file
// writes the leaves of a map as a NanoVDB FloatGrid, in memory or as an
// uncompressed .nvdb file, and reads such grids back.  NanoVDB itself is not
// part of this build; the layout below is restated from the published
// format of NanoVDB 32.4 (OpenVDB 10, the `nanovdb/util/` era the reference
// builds against), with its sizes, offsets and lookup rules:
//
//   grid buffer (32-byte aligned parts, breadth first):
//     GridData      672 B  magic "NanoVDB0", checksum, version, flags,
//                          grid index/count/size, name[256], Map (264 B),
//                          world bbox, voxel size, class, type, blind data
//     TreeData       64 B  byte offsets (from the tree) of the first leaf,
//                          lower, upper and root node; node and tile counts;
//                          active voxel count
//     RootData       64 B  index bbox, table size, background, min, max,
//                          average, std. deviation; then one 32 B tile per
//                          upper node: key, child offset (from the root),
//                          state, value
//     upper nodes 270400 B 32^3 children of 128^3 voxels: bbox, flags,
//                          value mask, child mask, stats, 8 B table entries
//     lower nodes  33856 B 16^3 children of 8^3 voxels, same layout
//     leaves        2144 B bbox min, bbox extent, flags, value mask (8
//                          words), stats, 512 FP32 values (x*64 + y*8 + z)
//   root key = (x >> 12) << 42 | (y >> 12) << 21 | (z >> 12) on the
//   coordinates as unsigned 32-bit values; a node's child slot is
//   ((x & m) >> c) << 2d | ((y & m) >> c) << d | ((z & m) >> c) for a node
//   of 2^d children per axis, each 2^c voxels wide, m the node's extent - 1;
//   a child's offset counts bytes from its parent node.
//   file: 16 B header (magic, version, grid count, codec), then per grid a
//   176 B record (sizes, name hash, voxel count, type, class, bboxes, voxel
//   size, name length, node and tile counts, codec, version), the name with
//   its terminating zero, and the grid buffer.
//
// The writer stores no tiles (every voxel lives in a leaf, as in the map),
// background 0 (the log-odds of an unknown voxel), grid class Unknown, and
// no checksum (the all-ones value NanoVDB uses for "not computed").
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "../../include/nanomap_amd.h"

namespace nm {
void set_error(const std::string& msg);
}

namespace {

constexpr uint64_t kMagic = 0x304244566f6e614eull;  // "NanoVDB0"
constexpr uint32_t kVersion = (32u << 21) | (4u << 10) | 2u;  // 32.4.2
constexpr uint32_t kTypeFloat = 1;
constexpr uint32_t kFlagBBox = 1u << 1, kFlagMinMax = 1u << 2,
                   kFlagBreadthFirst = 1u << 5;
constexpr uint64_t kNoChecksum = ~0ull;

constexpr size_t kGridSize = 672, kTreeSize = 64, kRootSize = 64,
                 kTileSize = 32, kUpperSize = 270400, kLowerSize = 33856,
                 kLeafSize = 2144;
// offsets inside the nodes
constexpr size_t kUpperTable = 8256, kLowerTable = 1088, kLeafValues = 96;
constexpr int kUpperLog2 = 5, kLowerLog2 = 4;

int fail(int code, const std::string& msg) {
  nm::set_error(msg);
  return code;
}

template <typename T>
void put(std::vector<unsigned char>& b, size_t at, const T& v) {
  memcpy(b.data() + at, &v, sizeof(T));
}
template <typename T>
T get(const unsigned char* b, size_t at) {
  T v;
  memcpy(&v, b + at, sizeof(T));
  return v;
}

struct Coord {
  int32_t x, y, z;
};

// Running bbox and min/max of active values.
struct Stats {
  int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX};
  int32_t hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  float vmin = std::numeric_limits<float>::infinity();
  float vmax = -std::numeric_limits<float>::infinity();
  uint64_t active = 0;
  void add(const Stats& o) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], o.lo[a]);
      hi[a] = std::max(hi[a], o.hi[a]);
    }
    vmin = std::min(vmin, o.vmin);
    vmax = std::max(vmax, o.vmax);
    active += o.active;
  }
  // NanoVDB's min/max of a node without active values stay at the
  // background
  float min_or(float bg) const { return active ? vmin : bg; }
  float max_or(float bg) const { return active ? vmax : bg; }
};

// bbox as 6 int32 (min xyz, max xyz); an empty one is (INT_MAX, INT_MIN)
void put_bbox(std::vector<unsigned char>& b, size_t at, const Stats& s) {
  for (int a = 0; a < 3; ++a) put<int32_t>(b, at + 4 * a, s.lo[a]);
  for (int a = 0; a < 3; ++a) put<int32_t>(b, at + 12 + 4 * a, s.hi[a]);
}

uint64_t root_key(int32_t x, int32_t y, int32_t z) {
  return ((uint64_t)((uint32_t)z >> 12)) |
         ((uint64_t)((uint32_t)y >> 12) << 21) |
         ((uint64_t)((uint32_t)x >> 12) << 42);
}

uint32_t child_slot(int32_t x, int32_t y, int32_t z, int log2dim,
                    int child_total) {
  const uint32_t m = (1u << (log2dim + child_total)) - 1u;
  return ((((uint32_t)x & m) >> child_total) << (2 * log2dim)) |
         ((((uint32_t)y & m) >> child_total) << log2dim) |
         (((uint32_t)z & m) >> child_total);
}

// NanoVDB's stringHash of a grid name (the file record's name key).
uint64_t string_hash(const char* s) {
  uint64_t h = 0;
  for (const unsigned char* p = (const unsigned char*)s; *p; ++p) {
    const uint64_t overflow = h >> 56;
    h *= 67;
    h += *p + overflow;
  }
  return h;
}

bool bad_name(const char* name) {
  return !name || strlen(name) > 255;
}

// The grid buffer of n leaves (origins multiples of 8, values z + 8y + 64x,
// masks 16 words with bit v = voxel v active).
int build_grid(const int32_t* origins, const float* values,
               const uint32_t* masks, int64_t n, double voxel_size,
               const char* name, std::vector<unsigned char>& b,
               Stats* total_out, uint32_t counts[3]) {
  if (n < 0 || (n > 0 && (!origins || !values || !masks)))
    return fail(NM_EINVAL, "nm_nvdb: bad leaf arrays");
  if (!(voxel_size > 0) || !(voxel_size < 1e30))
    return fail(NM_EINVAL, "nm_nvdb: voxel size must be positive and finite");
  if (bad_name(name)) return fail(NM_EINVAL, "nm_nvdb: grid name over 255 bytes");
  // leaves in tree order: root key, upper child slot, lower child slot
  std::vector<int64_t> order((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t* o = origins + 3 * i;
    if ((o[0] & 7) || (o[1] & 7) || (o[2] & 7))
      return fail(NM_EINVAL, "nm_nvdb: leaf origin not a multiple of 8");
    order[(size_t)i] = i;
  }
  auto sort_key = [&](int64_t i) {
    const int32_t* o = origins + 3 * i;
    return std::make_tuple(root_key(o[0], o[1], o[2]),
                           child_slot(o[0], o[1], o[2], kUpperLog2, 7),
                           child_slot(o[0], o[1], o[2], kLowerLog2, 3));
  };
  std::sort(order.begin(), order.end(),
            [&](int64_t a, int64_t c) { return sort_key(a) < sort_key(c); });
  for (size_t i = 1; i < order.size(); ++i)
    if (sort_key(order[i]) == sort_key(order[i - 1]))
      return fail(NM_EINVAL, "nm_nvdb: two leaves with the same origin");
  // node tables: upper index of each lower, lower index of each leaf
  std::vector<int64_t> lower_of((size_t)n), upper_of_lower;
  std::vector<Coord> lower_org, upper_org;
  for (size_t r = 0; r < order.size(); ++r) {
    const int32_t* o = origins + 3 * order[r];
    const Coord lo{o[0] & ~127, o[1] & ~127, o[2] & ~127};
    const Coord up{o[0] & ~4095, o[1] & ~4095, o[2] & ~4095};
    if (upper_org.empty() || upper_org.back().x != up.x ||
        upper_org.back().y != up.y || upper_org.back().z != up.z)
      upper_org.push_back(up);
    if (lower_org.empty() || lower_org.back().x != lo.x ||
        lower_org.back().y != lo.y || lower_org.back().z != lo.z) {
      lower_org.push_back(lo);
      upper_of_lower.push_back((int64_t)upper_org.size() - 1);
    }
    lower_of[r] = (int64_t)lower_org.size() - 1;
  }
  const size_t nu = upper_org.size(), nl = lower_org.size(), nf = order.size();
  if (nu > 0xffffffffull || nl > 0xffffffffull || nf > 0xffffffffull)
    return fail(NM_EINVAL, "nm_nvdb: too many nodes");
  const size_t tree_at = kGridSize, root_at = tree_at + kTreeSize;
  const size_t upper_at = root_at + kRootSize + kTileSize * nu;
  const size_t lower_at = upper_at + kUpperSize * nu;
  const size_t leaf_at = lower_at + kLowerSize * nl;
  const size_t end = leaf_at + kLeafSize * nf;
  b.assign(end, 0);
  const float bg = 0.0f;
  std::vector<Stats> lst(nl), ust(nu);
  Stats total;
  // leaves
  for (size_t r = 0; r < nf; ++r) {
    const int64_t i = order[r];
    const int32_t* o = origins + 3 * i;
    const float* v = values + 512 * i;
    const uint32_t* k = masks + 16 * i;
    const size_t at = leaf_at + kLeafSize * r;
    Stats s;
    for (int q = 0; q < 512; ++q) {
      if (!((k[q >> 5] >> (q & 31)) & 1u)) continue;
      const int32_t c[3] = {o[0] + (q >> 6), o[1] + ((q >> 3) & 7), o[2] + (q & 7)};
      for (int a = 0; a < 3; ++a) {
        s.lo[a] = std::min(s.lo[a], c[a]);
        s.hi[a] = std::max(s.hi[a], c[a]);
      }
      s.vmin = std::min(s.vmin, v[q]);
      s.vmax = std::max(s.vmax, v[q]);
      ++s.active;
    }
    // bbox min (inside the leaf; its origin is bbox min & ~7), extent,
    // flags (bit 1: the leaf has a bbox, i.e. active voxels)
    for (int a = 0; a < 3; ++a) {
      put<int32_t>(b, at + 4 * a, s.active ? s.lo[a] : o[a]);
      b[at + 12 + a] = (uint8_t)(s.active ? s.hi[a] - s.lo[a] : 0);
    }
    b[at + 15] = s.active ? 2 : 0;
    for (int w = 0; w < 16; ++w) put<uint32_t>(b, at + 16 + 4 * w, k[w]);
    put<float>(b, at + 80, s.min_or(bg));
    put<float>(b, at + 84, s.max_or(bg));
    memcpy(b.data() + at + kLeafValues, v, 2048);
    lst[(size_t)lower_of[r]].add(s);
    // child entry in the lower node
    const size_t lat = lower_at + kLowerSize * (size_t)lower_of[r];
    const uint32_t slot = child_slot(o[0], o[1], o[2], kLowerLog2, 3);
    b[lat + 544 + (slot >> 3)] |= (unsigned char)(1u << (slot & 7));
    put<int64_t>(b, lat + kLowerTable + 8 * slot, (int64_t)at - (int64_t)lat);
  }
  // lower nodes
  for (size_t l = 0; l < nl; ++l) {
    const size_t at = lower_at + kLowerSize * l;
    const Coord& lo = lower_org[l];
    put_bbox(b, at, lst[l]);
    // tiles that are not children hold the background value
    for (uint32_t q = 0; q < 4096; ++q)
      if (!((b[at + 544 + (q >> 3)] >> (q & 7)) & 1u))
        put<float>(b, at + kLowerTable + 8 * q, bg);
    put<float>(b, at + 1056, lst[l].min_or(bg));
    put<float>(b, at + 1060, lst[l].max_or(bg));
    const size_t ui = (size_t)upper_of_lower[l];
    ust[ui].add(lst[l]);
    const size_t uat = upper_at + kUpperSize * ui;
    const uint32_t slot = child_slot(lo.x, lo.y, lo.z, kUpperLog2, 7);
    b[uat + 4128 + (slot >> 3)] |= (unsigned char)(1u << (slot & 7));
    put<int64_t>(b, uat + kUpperTable + 8 * slot, (int64_t)at - (int64_t)uat);
  }
  // upper nodes and root tiles (in key order: the uppers are sorted)
  for (size_t u = 0; u < nu; ++u) {
    const size_t at = upper_at + kUpperSize * u;
    put_bbox(b, at, ust[u]);
    for (uint32_t q = 0; q < 32768; ++q)
      if (!((b[at + 4128 + (q >> 3)] >> (q & 7)) & 1u))
        put<float>(b, at + kUpperTable + 8 * q, bg);
    put<float>(b, at + 8224, ust[u].min_or(bg));
    put<float>(b, at + 8228, ust[u].max_or(bg));
    total.add(ust[u]);
    const size_t tat = root_at + kRootSize + kTileSize * u;
    const Coord& uo = upper_org[u];
    put<uint64_t>(b, tat, root_key(uo.x, uo.y, uo.z));
    put<int64_t>(b, tat + 8, (int64_t)at - (int64_t)root_at);
    put<uint32_t>(b, tat + 16, 0u);
    put<float>(b, tat + 20, bg);
  }
  // root
  put_bbox(b, root_at, total);
  put<uint32_t>(b, root_at + 24, (uint32_t)nu);
  put<float>(b, root_at + 28, bg);
  put<float>(b, root_at + 32, total.min_or(bg));
  put<float>(b, root_at + 36, total.max_or(bg));
  // tree
  put<uint64_t>(b, tree_at + 0, nf ? leaf_at - tree_at : 0);
  put<uint64_t>(b, tree_at + 8, nl ? lower_at - tree_at : 0);
  put<uint64_t>(b, tree_at + 16, nu ? upper_at - tree_at : 0);
  put<uint64_t>(b, tree_at + 24, root_at - tree_at);
  put<uint32_t>(b, tree_at + 32, (uint32_t)nf);
  put<uint32_t>(b, tree_at + 36, (uint32_t)nl);
  put<uint32_t>(b, tree_at + 40, (uint32_t)nu);
  put<uint64_t>(b, tree_at + 56, total.active);
  // grid
  put<uint64_t>(b, 0, kMagic);
  put<uint64_t>(b, 8, kNoChecksum);
  put<uint32_t>(b, 16, kVersion);
  put<uint32_t>(b, 20, kFlagBBox | kFlagMinMax | kFlagBreadthFirst);
  put<uint32_t>(b, 24, 0u);  // grid index
  put<uint32_t>(b, 28, 1u);  // grid count
  put<uint64_t>(b, 32, (uint64_t)end);
  memcpy(b.data() + 40, name, strlen(name));
  // Map: index -> world is a uniform scale by the voxel size (the map's
  // linear transform, OccupancyMap.h:35-61), single then double precision
  const double dx = voxel_size, inv = 1.0 / voxel_size;
  for (int a = 0; a < 3; ++a) {
    put<float>(b, 296 + 4 * (4 * a), (float)dx);       // mMatF diagonal
    put<float>(b, 332 + 4 * (4 * a), (float)inv);      // mInvMatF
    put<double>(b, 384 + 8 * (4 * a), dx);             // mMatD
    put<double>(b, 456 + 8 * (4 * a), inv);            // mInvMatD
  }
  put<float>(b, 380, 1.0f);   // mTaperF (after mVecF at 368)
  put<double>(b, 552, 1.0);   // mTaperD (after mVecD at 528)
  // world bbox: the index bbox's corners times the voxel size
  for (int a = 0; a < 3; ++a) {
    put<double>(b, 560 + 8 * a, total.active ? total.lo[a] * dx : 0.0);
    put<double>(b, 584 + 8 * a, total.active ? total.hi[a] * dx : 0.0);
    put<double>(b, 608 + 8 * a, dx);
  }
  put<uint32_t>(b, 632, 0u);  // grid class: Unknown
  put<uint32_t>(b, 636, kTypeFloat);
  if (total_out) *total_out = total;
  if (counts) {
    counts[0] = (uint32_t)nf;
    counts[1] = (uint32_t)nl;
    counts[2] = (uint32_t)nu;
  }
  return NM_OK;
}

// Leaves of a FloatGrid buffer of `size` bytes: every leaf node, and every
// active tile of the internal levels expanded into leaves (at most
// max_tile_leaves of them).  Every offset is checked against the buffer.
int parse_grid(const unsigned char* g, uint64_t size, std::vector<int32_t>& o,
               std::vector<float>& v, std::vector<uint32_t>& k,
               double* voxel_size) {
  const int64_t max_tile_leaves = (int64_t)1 << 22;
  auto fits = [&](uint64_t at, uint64_t len) {
    return at <= size && len <= size - at;
  };
  if (!fits(0, kGridSize + kTreeSize) || get<uint64_t>(g, 0) != kMagic)
    return fail(NM_EINVAL, "nm_nvdb: not a NanoVDB grid");
  if ((get<uint32_t>(g, 16) >> 21) != 32u)
    return fail(NM_EINVAL, "nm_nvdb: unsupported NanoVDB major version");
  if (get<uint32_t>(g, 636) != kTypeFloat)
    return fail(NM_EINVAL, "nm_nvdb: the grid is not a FloatGrid");
  if (get<uint64_t>(g, 32) > size)
    return fail(NM_EINVAL, "nm_nvdb: grid larger than its buffer");
  *voxel_size = get<double>(g, 608);
  const uint64_t tree_at = kGridSize;
  const uint64_t root_at = tree_at + get<uint64_t>(g, tree_at + 24);
  if (root_at < tree_at || !fits(root_at, kRootSize))
    return fail(NM_EINVAL, "nm_nvdb: bad root offset");
  const uint32_t ntiles = get<uint32_t>(g, root_at + 24);
  const float bg = get<float>(g, root_at + 28);
  if (!fits(root_at + kRootSize, (uint64_t)kTileSize * ntiles))
    return fail(NM_EINVAL, "nm_nvdb: bad root table");
  int64_t tile_leaves = 0;
  auto add_leaf = [&](int32_t x, int32_t y, int32_t z, const float* vals,
                      const unsigned char* mask8, float fill, bool on) {
    o.push_back(x);
    o.push_back(y);
    o.push_back(z);
    const size_t vi = v.size(), ki = k.size();
    v.resize(vi + 512, fill);
    k.resize(ki + 16, on ? 0xffffffffu : 0u);
    if (vals) memcpy(&v[vi], vals, 2048);
    if (mask8) memcpy(&k[ki], mask8, 64);
  };
  // an active tile of `edge` voxels per axis at (x, y, z): its leaves,
  // all voxels active with the tile's value
  auto add_tile = [&](int32_t x, int32_t y, int32_t z, int edge,
                      float val) -> bool {
    const int64_t per = (int64_t)(edge / 8) * (edge / 8) * (edge / 8);
    if (tile_leaves + per > max_tile_leaves) return false;
    tile_leaves += per;
    for (int a = 0; a < edge; a += 8)
      for (int c = 0; c < edge; c += 8)
        for (int d = 0; d < edge; d += 8)
          add_leaf(x + a, y + c, z + d, nullptr, nullptr, val, true);
    return true;
  };
  for (uint32_t t = 0; t < ntiles; ++t) {
    const uint64_t tat = root_at + kRootSize + (uint64_t)kTileSize * t;
    const uint64_t key = get<uint64_t>(g, tat);
    const int64_t child = get<int64_t>(g, tat + 8);
    // the key's fields are the upper node's origin >> 12 (21 bits each)
    const int32_t ux = (int32_t)((uint32_t)((key >> 42) & 0x1fffff) << 12);
    const int32_t uy = (int32_t)((uint32_t)((key >> 21) & 0x1fffff) << 12);
    const int32_t uz = (int32_t)((uint32_t)(key & 0x1fffff) << 12);
    if (child == 0) {
      if (get<uint32_t>(g, tat + 16) &&
          !add_tile(ux, uy, uz, 4096, get<float>(g, tat + 20)))
        return fail(NM_EINVAL, "nm_nvdb: active root tiles too large to expand");
      continue;
    }
    const uint64_t uat = root_at + (uint64_t)child;
    if (child < 0 || !fits(uat, kUpperSize))
      return fail(NM_EINVAL, "nm_nvdb: bad upper node offset");
    for (uint32_t q = 0; q < 32768; ++q) {
      const bool is_child = (g[uat + 4128 + (q >> 3)] >> (q & 7)) & 1u;
      const bool is_on = (g[uat + 32 + (q >> 3)] >> (q & 7)) & 1u;
      const int32_t lx = ux + (int32_t)((q >> 10) << 7);
      const int32_t ly = uy + (int32_t)(((q >> 5) & 31) << 7);
      const int32_t lz = uz + (int32_t)((q & 31) << 7);
      const uint64_t e = uat + kUpperTable + 8ull * q;
      if (!is_child) {
        if (is_on && !add_tile(lx, ly, lz, 128, get<float>(g, e)))
          return fail(NM_EINVAL, "nm_nvdb: active tiles too large to expand");
        continue;
      }
      const int64_t lc = get<int64_t>(g, e);
      const uint64_t lat = uat + (uint64_t)lc;
      if (lc <= 0 || !fits(lat, kLowerSize))
        return fail(NM_EINVAL, "nm_nvdb: bad lower node offset");
      for (uint32_t r = 0; r < 4096; ++r) {
        const bool c2 = (g[lat + 544 + (r >> 3)] >> (r & 7)) & 1u;
        const bool on2 = (g[lat + 32 + (r >> 3)] >> (r & 7)) & 1u;
        const int32_t fx = lx + (int32_t)((r >> 8) << 3);
        const int32_t fy = ly + (int32_t)(((r >> 4) & 15) << 3);
        const int32_t fz = lz + (int32_t)((r & 15) << 3);
        const uint64_t e2 = lat + kLowerTable + 8ull * r;
        if (!c2) {
          if (on2 && !add_tile(fx, fy, fz, 8, get<float>(g, e2)))
            return fail(NM_EINVAL, "nm_nvdb: active tiles too large to expand");
          continue;
        }
        const int64_t fc = get<int64_t>(g, e2);
        const uint64_t fat = lat + (uint64_t)fc;
        if (fc <= 0 || !fits(fat, kLeafSize))
          return fail(NM_EINVAL, "nm_nvdb: bad leaf offset");
        // the leaf's origin is its bbox min with the low three bits cleared
        const int32_t bx = get<int32_t>(g, fat) & ~7, by = get<int32_t>(g, fat + 4) & ~7,
                      bz = get<int32_t>(g, fat + 8) & ~7;
        if (bx != fx || by != fy || bz != fz)
          return fail(NM_EINVAL, "nm_nvdb: leaf origin does not match its slot");
        add_leaf(fx, fy, fz, reinterpret_cast<const float*>(g + fat + kLeafValues),
                 g + fat + 16, bg, false);
      }
    }
  }
  return NM_OK;
}

int write_file(const char* path, const char* name,
               const std::vector<unsigned char>& b, const Stats& total,
               const uint32_t counts[3], double voxel_size) {
  FILE* f = fopen(path, "wb");
  if (!f) return fail(NM_EINVAL, std::string("nm_nvdb: cannot open ") + path);
  unsigned char head[16] = {0};
  memcpy(head, &kMagic, 8);
  memcpy(head + 8, &kVersion, 4);
  const uint16_t grids = 1, codec = 0;
  memcpy(head + 12, &grids, 2);
  memcpy(head + 14, &codec, 2);
  std::vector<unsigned char> meta(176, 0);
  const uint64_t gsize = b.size();
  put<uint64_t>(meta, 0, gsize);            // grid size
  put<uint64_t>(meta, 8, gsize);            // bytes in the file (no codec)
  put<uint64_t>(meta, 16, string_hash(name));
  put<uint64_t>(meta, 24, total.active);    // active voxels
  put<uint32_t>(meta, 32, kTypeFloat);
  put<uint32_t>(meta, 36, 0u);              // class Unknown
  memcpy(meta.data() + 40, b.data() + 560, 48);  // world bbox
  put_bbox(meta, 88, total);                // index bbox
  put<double>(meta, 112, voxel_size);
  put<double>(meta, 120, voxel_size);
  put<double>(meta, 128, voxel_size);
  const uint32_t name_size = (uint32_t)strlen(name) + 1;
  put<uint32_t>(meta, 136, name_size);
  put<uint32_t>(meta, 140, counts[0]);      // leaves
  put<uint32_t>(meta, 144, counts[1]);      // lower nodes
  put<uint32_t>(meta, 148, counts[2]);      // upper nodes
  put<uint32_t>(meta, 152, 1u);             // root
  // tile counts (156..167) stay 0: the grid holds no tiles
  put<uint16_t>(meta, 168, 0);              // codec NONE
  put<uint32_t>(meta, 172, kVersion);
  bool ok = fwrite(head, 1, 16, f) == 16 && fwrite(meta.data(), 1, 176, f) == 176 &&
            fwrite(name, 1, name_size, f) == name_size &&
            fwrite(b.data(), 1, b.size(), f) == b.size();
  ok = (fclose(f) == 0) && ok;
  if (!ok) return fail(NM_EINVAL, std::string("nm_nvdb: write failed: ") + path);
  return NM_OK;
}

}  // namespace

extern "C" {

int nm_nvdb_build(const int32_t* origins, const float* values,
                  const uint32_t* masks, int64_t n_leaves, double voxel_size,
                  const char* grid_name, void* buffer, uint64_t* bytes) {
  if (!bytes) return fail(NM_EINVAL, "nm_nvdb_build: null size");
  std::vector<unsigned char> b;
  const int st = build_grid(origins, values, masks, n_leaves, voxel_size,
                            grid_name, b, nullptr, nullptr);
  if (st != NM_OK) return st;
  if (buffer) {
    if (*bytes < b.size())
      return fail(NM_EINVAL, "nm_nvdb_build: buffer too small");
    memcpy(buffer, b.data(), b.size());
  }
  *bytes = b.size();
  return NM_OK;
}

int nm_nvdb_write(const char* path, const int32_t* origins,
                  const float* values, const uint32_t* masks, int64_t n_leaves,
                  double voxel_size, const char* grid_name) {
  if (!path) return fail(NM_EINVAL, "nm_nvdb_write: null path");
  std::vector<unsigned char> b;
  Stats total;
  uint32_t counts[3];
  const int st = build_grid(origins, values, masks, n_leaves, voxel_size,
                            grid_name, b, &total, counts);
  if (st != NM_OK) return st;
  return write_file(path, grid_name, b, total, counts, voxel_size);
}

int nm_nvdb_read(const char* path, const char* grid_name, int64_t* n_leaves,
                 int32_t* origins, float* values, uint32_t* masks,
                 double* voxel_size) {
  if (!path || !n_leaves) return fail(NM_EINVAL, "nm_nvdb_read: null argument");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(NM_EINVAL, std::string("nm_nvdb_read: cannot open ") + path);
  std::vector<unsigned char> all;
  unsigned char chunk[1 << 16];
  size_t got;
  while ((got = fread(chunk, 1, sizeof(chunk), f)) > 0)
    all.insert(all.end(), chunk, chunk + got);
  fclose(f);
  const uint64_t size = all.size();
  if (size < 16 || get<uint64_t>(all.data(), 0) != kMagic)
    return fail(NM_EINVAL, "nm_nvdb_read: not a NanoVDB file");
  const uint16_t grids = get<uint16_t>(all.data(), 12);
  if (get<uint16_t>(all.data(), 14) != 0)
    return fail(NM_EINVAL, "nm_nvdb_read: compressed files are not supported");
  uint64_t at = 16;
  for (uint16_t i = 0; i < grids; ++i) {
    if (size - at < 176) break;
    const uint64_t gsize = get<uint64_t>(all.data(), at);
    const uint64_t fsize = get<uint64_t>(all.data(), at + 8);
    const uint32_t type = get<uint32_t>(all.data(), at + 32);
    const uint32_t nsize = get<uint32_t>(all.data(), at + 136);
    const uint16_t codec = get<uint16_t>(all.data(), at + 168);
    at += 176;
    if (codec != 0 || fsize != gsize || nsize == 0 || nsize > size - at)
      return fail(NM_EINVAL, "nm_nvdb_read: unsupported or damaged grid record");
    const std::string name((const char*)all.data() + at, nsize - 1);
    at += nsize;
    if (gsize > size - at) return fail(NM_EINVAL, "nm_nvdb_read: truncated grid");
    const bool want = type == kTypeFloat &&
                      (!grid_name || !grid_name[0] || name == grid_name);
    if (want) {
      std::vector<int32_t> o;
      std::vector<float> v;
      std::vector<uint32_t> k;
      double dx = 0;
      const int st = parse_grid(all.data() + at, gsize, o, v, k, &dx);
      if (st != NM_OK) return st;
      const int64_t n = (int64_t)(o.size() / 3);
      if (origins || values || masks) {
        if (*n_leaves < n) return fail(NM_EINVAL, "nm_nvdb_read: arrays too small");
        if (origins && n) memcpy(origins, o.data(), o.size() * 4);
        if (values && n) memcpy(values, v.data(), v.size() * 4);
        if (masks && n) memcpy(masks, k.data(), k.size() * 4);
      }
      *n_leaves = n;
      if (voxel_size) *voxel_size = dx;
      return NM_OK;
    }
    at += gsize;
  }
  return fail(NM_EINVAL, "nm_nvdb_read: no such FloatGrid in the file");
}

}  // extern "C"
