// nm_oracle.cpp -- CPU restatement of NanoMap's frustum-sensor mapping path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Compiled with -frounding-math -ffp-contract=off so every floating-point
// operation below is one IEEE operation in the rounding mode selected with
// fesetround(), exactly as the reference executes it on the host.
//
// Parity status of each function (what pins it) is listed in the header of
// tests/test_oracle.py.

#include "nm_oracle.h"

#include <algorithm>
#include <array>
#include <climits>
#include <cfenv>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <vector>

namespace {

struct RoundGuard {
  int saved;
  explicit RoundGuard(int rounding) : saved(std::fegetround()) {
    std::fesetround(rounding == 1 ? FE_DOWNWARD : FE_TONEAREST);
  }
  ~RoundGuard() { std::fesetround(saved); }
};

inline uint64_t leaf_key(int i, int j, int k) {
  // Leaf origin coordinates (voxel >> 3) packed in 21 bits each.
  const uint64_t m = (1ull << 21) - 1;
  return ((uint64_t)((i >> 3) & m) << 42) | ((uint64_t)((j >> 3) & m) << 21) |
         (uint64_t)((k >> 3) & m);
}

inline int local_index(int i, int j, int k) {
  // In-leaf offset order of the reference: z + 8y + 64x
  // (frustumCastCloud.cu:215, Handler.cpp:171).
  return (k & 7) + ((j & 7) << 3) + ((i & 7) << 6);
}

struct TempLeaf {
  int origin[3];
  float v[512];
  uint32_t miss[512];
  uint32_t hit[512];
  uint64_t on[8];
};

struct MapLeaf {
  int origin[3];
  float v[512];
  uint64_t active[8];
};

template <class LeafT>
struct SparseGrid {
  std::unordered_map<uint64_t, LeafT*> leaves;
  uint64_t last_key = ~0ull;
  LeafT* last = nullptr;

  ~SparseGrid() { clear(); }
  void clear() {
    for (auto& kv : leaves) delete kv.second;
    leaves.clear();
    last_key = ~0ull;
    last = nullptr;
  }
  // Accessor-style lookup with a one-leaf cache (the cost model of
  // openvdb::tree::ValueAccessor).
  LeafT* probe(int i, int j, int k) const {
    const uint64_t key = leaf_key(i, j, k);
    if (key == last_key) return last;
    auto it = leaves.find(key);
    return it == leaves.end() ? nullptr : it->second;
  }
  LeafT* touch(int i, int j, int k) {
    const uint64_t key = leaf_key(i, j, k);
    if (key == last_key) return last;
    auto it = leaves.find(key);
    LeafT* leaf;
    if (it == leaves.end()) {
      leaf = new LeafT();
      std::memset(leaf, 0, sizeof(LeafT));
      leaf->origin[0] = i & ~7;
      leaf->origin[1] = j & ~7;
      leaf->origin[2] = k & ~7;
      leaves.emplace(key, leaf);
    } else {
      leaf = it->second;
    }
    last_key = key;
    last = leaf;
    return leaf;
  }
};

}  // namespace

struct or_temp {
  SparseGrid<TempLeaf> grid;
  uint64_t visits = 0;
};

struct or_map {
  SparseGrid<MapLeaf> grid;
  float clamp_min, clamp_max, hit_thr, miss_thr;
};

namespace {

// openvdb::math::MinIndex: x only if strictly smallest, else y if y < z,
// else z (ties resolve toward the higher axis).
inline int min_index(const double n[3]) {
  if (n[0] < n[1] && n[0] < n[2]) return 0;
  if (n[1] < n[2]) return 1;
  return 2;
}

// One ray of Handler::populateTempGrid (Handler.cpp:65-84) after the
// direction has been formed: openvdb::math::DDA<Ray<double>,0> from eye
// over [0, ray_length]; a miss on every voxel stepped out of, a hit on the
// end voxel iff its exit time < max_time.
template <class Visit>
inline void dda_cast(const double eye[3], double dx, double dy, double dz,
                     double ray_length, float max_time, Visit&& visit) {
  // Ray::setDir (openvdb math/Ray.h): inv = 1 / dir.
  const double dir[3] = {dx, dy, dz};
  double inv[3];
  for (int a = 0; a < 3; ++a) inv[a] = 1 / dir[a];
  // DDA::init(ray, 0, ray_length): pos = ray(0) = eye + dir * 0.
  double pos[3];
  for (int a = 0; a < 3; ++a) pos[a] = eye[a] + dir[a] * 0.0;
  int voxel[3];
  for (int a = 0; a < 3; ++a) voxel[a] = (int)std::floor(pos[a]);
  int step[3];
  double next[3], delta[3];
  const double t0 = 0.0;
  for (int a = 0; a < 3; ++a) {
    if (dir[a] == 0.0) {  // math::isZero: exact comparison
      step[a] = 0;
      next[a] = std::numeric_limits<double>::max();
      delta[a] = std::numeric_limits<double>::max();
    } else if (inv[a] > 0) {
      step[a] = 1;
      next[a] = t0 + ((double)(voxel[a] + 1) - pos[a]) * inv[a];
      delta[a] = (double)step[a] * inv[a];
    } else {
      step[a] = -1;
      next[a] = t0 + ((double)(voxel[a]) - pos[a]) * inv[a];
      delta[a] = (double)step[a] * inv[a];
    }
  }
  int prev[3] = {voxel[0], voxel[1], voxel[2]};
  double t;
  for (;;) {
    // DDA::step()
    const int a = min_index(next);
    t = next[a];
    next[a] += delta[a];
    voxel[a] += step[a];
    if (!(t <= ray_length)) break;
    visit(prev, false);
    prev[0] = voxel[0];
    prev[1] = voxel[1];
    prev[2] = voxel[2];
  }
  if (t < (double)max_time) visit(prev, true);
}

// Direction set-up of Handler::populateTempGrid (Handler.cpp:61-75).
// Returns false when the pixel is skipped (ray_length >= max_time or NaN).
inline bool ray_setup(const float R[9], float px, float py, float pz,
                      float res, float max_time, double* d, double* L) {
  // Eigen float products, evaluated left to right (Handler.cpp:65-68).
  const float fx = R[0] * px + R[1] * py + R[2] * pz;
  const float fy = R[3] * px + R[4] * py + R[5] * pz;
  const float fz = R[6] * px + R[7] * py + R[8] * pz;
  double x = fx, y = fy, z = fz;
  // Vec3d::length() / gridRes (Handler.cpp:70).
  const double len = std::sqrt(x * x + y * y + z * z);
  const double ray_length = len / (double)res;
  if (!(ray_length < (double)max_time)) return false;  // Handler.cpp:71
  // Vec3d::normalize(eps = 1e-7): skip if |len - 0| <= eps.
  if (!(std::fabs(len - 0.0) <= 1.0e-7)) {
    const double s = 1 / len;
    x *= s;
    y *= s;
    z *= s;
  }
  d[0] = x;
  d[1] = y;
  d[2] = z;
  *L = ray_length;
  return true;
}

inline void temp_visit(or_temp* t, const int v[3], bool is_hit, float hit,
                       float miss) {
  TempLeaf* leaf = t->grid.touch(v[0], v[1], v[2]);
  const int li = local_index(v[0], v[1], v[2]);
  // modifyValueAndActiveState(voxel, miss/hit): value += logodds; on.
  leaf->v[li] += is_hit ? hit : miss;
  if (is_hit)
    leaf->hit[li] += 1;
  else
    leaf->miss[li] += 1;
  leaf->on[li >> 6] |= 1ull << (li & 63);
  t->visits += 1;
}

template <class LeafT>
std::vector<const LeafT*> sorted_leaves(const SparseGrid<LeafT>& g) {
  std::vector<const LeafT*> out;
  out.reserve(g.leaves.size());
  for (auto& kv : g.leaves) out.push_back(kv.second);
  std::sort(out.begin(), out.end(), [](const LeafT* a, const LeafT* b) {
    if (a->origin[0] != b->origin[0]) return a->origin[0] < b->origin[0];
    if (a->origin[1] != b->origin[1]) return a->origin[1] < b->origin[1];
    return a->origin[2] < b->origin[2];
  });
  return out;
}

}  // namespace

extern "C" {

or_temp* or_temp_create(void) { return new or_temp(); }
void or_temp_destroy(or_temp* t) { delete t; }
void or_temp_clear(or_temp* t) {
  t->grid.clear();
  t->visits = 0;
}

// Handler::populateTempGrid (Handler.cpp:28-86), with the hit/miss
// log-odds passed in directly (the reference derives the hit value from
// _probMiss by mistake, Handler.cpp:43; SimHandler.cpp:42 is correct).
int or_cast_cloud(or_temp* t, const unsigned char* cloud, int n_points,
                  int step, const float R[9], const double eye[3], float res,
                  float max_time, float logodds_hit, float logodds_miss,
                  int rounding) {
  RoundGuard guard(rounding);
  int rays = 0;
  for (int i = 0; i < n_points; ++i) {
    const unsigned char* p = cloud + (size_t)i * step;
    float px, py, pz;
    std::memcpy(&px, p + 0, 4);
    std::memcpy(&py, p + 4, 4);
    std::memcpy(&pz, p + 8, 4);
    double d[3], L;
    if (!ray_setup(R, px, py, pz, res, max_time, d, &L)) continue;
    ++rays;
    dda_cast(eye, d[0], d[1], d[2], L, max_time,
             [&](const int v[3], bool is_hit) {
               temp_visit(t, v, is_hit, logodds_hit, logodds_miss);
             });
  }
  return rays;
}

int or_cast_rays(or_temp* t, const float* rays4, int n_rays,
                 const double eye[3], float max_time, float logodds_hit,
                 float logodds_miss, int rounding) {
  RoundGuard guard(rounding);
  int rays = 0;
  for (int i = 0; i < n_rays; ++i) {
    const float* r = rays4 + 4 * (size_t)i;
    const double L = r[3];
    if (!(L > 0.0 && L < (double)max_time)) continue;
    ++rays;
    dda_cast(eye, (double)r[0], (double)r[1], (double)r[2], L, max_time,
             [&](const int v[3], bool is_hit) {
               temp_visit(t, v, is_hit, logodds_hit, logodds_miss);
             });
  }
  return rays;
}

int64_t or_temp_size(const or_temp* t) {
  int64_t n = 0;
  for (auto& kv : t->grid.leaves)
    for (int w = 0; w < 8; ++w) n += __builtin_popcountll(kv.second->on[w]);
  return n;
}

uint64_t or_temp_visits(const or_temp* t) { return t->visits; }

void or_temp_export(const or_temp* t, int32_t* ijk, uint32_t* miss,
                    uint32_t* hit, float* value) {
  int64_t n = 0;
  for (const TempLeaf* leaf : sorted_leaves(t->grid)) {
    // local index li = z + 8y + 64x ascends in (x, y, z) lexicographic order
    for (int li = 0; li < 512; ++li) {
      if (!((leaf->on[li >> 6] >> (li & 63)) & 1)) continue;
      if (ijk) {
        ijk[3 * n + 0] = leaf->origin[0] + (li >> 6);
        ijk[3 * n + 1] = leaf->origin[1] + ((li >> 3) & 7);
        ijk[3 * n + 2] = leaf->origin[2] + (li & 7);
      }
      if (miss) miss[n] = leaf->miss[li];
      if (hit) hit[n] = leaf->hit[li];
      if (value) value[n] = leaf->v[li];
      ++n;
    }
  }
}

or_map* or_map_create(float clamp_min, float clamp_max, float hit_thr,
                      float miss_thr) {
  or_map* m = new or_map();
  m->clamp_min = clamp_min;
  m->clamp_max = clamp_max;
  m->hit_thr = hit_thr;
  m->miss_thr = miss_thr;
  return m;
}

void or_map_destroy(or_map* m) { delete m; }

// Handler::integrateTempGrid (Handler.cpp:88-122): for each active temp
// voxel with a nonzero value: m += tv, clamp to [empty, occ], then active =
// true if m > hitThr, false if m < missThr, unchanged otherwise.
void or_map_integrate(or_map* m, const or_temp* t, int rounding) {
  RoundGuard guard(rounding);
  for (const TempLeaf* tl : sorted_leaves(t->grid)) {
    for (int li = 0; li < 512; ++li) {
      if (!((tl->on[li >> 6] >> (li & 63)) & 1)) continue;
      const float tv = tl->v[li];
      if (!(tv != 0.0f)) continue;
      const int i = tl->origin[0] + (li >> 6);
      const int j = tl->origin[1] + ((li >> 3) & 7);
      const int k = tl->origin[2] + (li & 7);
      MapLeaf* ml = m->grid.touch(i, j, k);
      float value = ml->v[li];
      bool active = (ml->active[li >> 6] >> (li & 63)) & 1;
      value += tv;
      if (value > m->clamp_max) {
        value = m->clamp_max;
      } else if (value < m->clamp_min) {
        value = m->clamp_min;
      }
      if (value > m->hit_thr) {
        active = true;
      } else if (value < m->miss_thr) {
        active = false;
      }
      ml->v[li] = value;
      if (active)
        ml->active[li >> 6] |= 1ull << (li & 63);
      else
        ml->active[li >> 6] &= ~(1ull << (li & 63));
    }
  }
}

int64_t or_map_leaf_count(const or_map* m) {
  return (int64_t)m->grid.leaves.size();
}

void or_map_export(const or_map* m, int32_t* origins, float* values,
                   uint32_t* masks) {
  int64_t n = 0;
  for (const MapLeaf* leaf : sorted_leaves(m->grid)) {
    if (origins) std::memcpy(origins + 3 * n, leaf->origin, 12);
    if (values) std::memcpy(values + 512 * n, leaf->v, 2048);
    if (masks)
      for (int w = 0; w < 16; ++w)
        masks[16 * n + w] = (uint32_t)(leaf->active[w >> 1] >> (32 * (w & 1)));
    ++n;
  }
}

int or_map_get(const or_map* m, int i, int j, int k, float* value,
               int* active) {
  const MapLeaf* leaf = m->grid.probe(i, j, k);
  if (!leaf) {
    *value = 0.0f;
    *active = 0;
    return 0;
  }
  const int li = local_index(i, j, k);
  *value = leaf->v[li];
  *active = (int)((leaf->active[li >> 6] >> (li & 63)) & 1);
  return 1;
}

int or_dda_trace(const double eye[3], const float d[3], float res,
                 float max_time, int rounding, int32_t* out_ijk, int max_out,
                 int* n_out, int* hit) {
  RoundGuard guard(rounding);
  const float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double dir[3], L;
  *n_out = 0;
  *hit = 0;
  if (!ray_setup(R, d[0], d[1], d[2], res, max_time, dir, &L)) return 0;
  int n = 0;
  dda_cast(eye, dir[0], dir[1], dir[2], L, max_time,
           [&](const int v[3], bool is_hit) {
             if (n < max_out) {
               out_ijk[3 * n + 0] = v[0];
               out_ijk[3 * n + 1] = v[1];
               out_ijk[3 * n + 2] = v[2];
             }
             ++n;
             if (is_hit) *hit = 1;
           });
  *n_out = n;
  return 1;
}

float or_fold(float v0, float a, uint32_t k, int rounding) {
  RoundGuard guard(rounding);
  volatile float v = v0;
  for (uint32_t i = 0; i < k; ++i) v = v + a;
  return v;
}

// Deterministic restatement of voxelFilter + rayFromFilter
// (filterCloud.cu:84-133, 279-318), FP32 in round-to-nearest (the
// reference runs it on the GPU).  Points are binned by
// floor(r/res + voxelOriginOffset), r = R p; per bin the fractions
// frac(r/res) are summed in ascending pixel order; each bin emits the ray
// (mean + bin) normalised with length |mean + bin| if 0 < length < max_time.
// Validity (filterCloud.cu:98-112): no NaN coordinate, inside the
// sensor-frame box [worldMin, worldMax] on every axis (when given), and
// 0 < |p| < maxRange.
// Bins are relative to floor(eye) and emitted in ascending
// (x, y, z) order.
//
// precision (PrecisionType, filterCloud.cu:476-553): 0 the FP32 centroid
// above; 1 the same with each fraction rounded to FP16 and every sum,
// count included, rounded to FP16 (voxelFilterHalf2 / rayFromFilterHalf2,
// filterCloud.cu:134-182, 319-351); 2 the voxel centre bin + 0.5 for every
// occupied bin (voxelFilterSimple / rayFromFilterSimple, 183-224,
// 352-384); 3 the same with an 8-bit count, so a bin holding a multiple of
// 256 points wraps to 0 and emits no ray (SimpleQuad, 225-277, 385-434).

// x rounded to the nearest FP16 value, ties to even (subnormals included;
// the values here stay far below the FP16 maximum).
static float half_rn(double x) {
  if (x == 0.0) return (float)x;
  int e = 0;
  std::frexp(std::fabs(x), &e);  // |x| in [2^(e-1), 2^e)
  const int q = (e - 1 < -14 ? -14 : e - 1) - 10;  // exponent of the ulp
  const double u = std::ldexp(1.0, q);
  return (float)(std::nearbyint(x / u) * u);  // exact scaling; RN mode
}

int or_voxel_filter(const unsigned char* cloud, int n_points, int step,
                    const float R[9], const double eye[3], float res,
                    float max_time, float max_range, const float* box,
                    float* rays4, int max_rays, int precision) {
  RoundGuard guard(0);
  struct Bin {
    float sx, sy, sz;
    uint32_t n;
    float hx, hy, hz, hn;  // FP16 sums (precision 1)
  };
  float origin[3];
  for (int a = 0; a < 3; ++a)
    origin[a] = (float)(eye[a] - std::floor(eye[a]));
  struct Key {
    int x, y, z;
    bool operator<(const Key& o) const {
      if (x != o.x) return x < o.x;
      if (y != o.y) return y < o.y;
      return z < o.z;
    }
  };
  std::vector<std::pair<Key, Bin>> bins;
  std::unordered_map<uint64_t, size_t> index;
  for (int i = 0; i < n_points; ++i) {
    const unsigned char* p = cloud + (size_t)i * step;
    float px, py, pz;
    std::memcpy(&px, p + 0, 4);
    std::memcpy(&py, p + 4, 4);
    std::memcpy(&pz, p + 8, 4);
    if (std::isnan(px) || std::isnan(py) || std::isnan(pz)) continue;
    // filterCloud.cu:100-111: outside the sensor-frame box on any axis
    if (box && (px < box[0] || px > box[3] || py < box[1] || py > box[4] ||
                pz < box[2] || pz > box[5]))
      continue;
    const float norm = std::sqrt(px * px + py * py + pz * pz);
    if (!(norm > 0.0f && norm < max_range)) continue;
    const float r[3] = {R[0] * px + R[1] * py + R[2] * pz,
                        R[3] * px + R[4] * py + R[5] * pz,
                        R[6] * px + R[7] * py + R[8] * pz};
    int b[3];
    float f[3];
    for (int a = 0; a < 3; ++a) {
      const float q = r[a] / res;
      b[a] = (int)std::floor(q + origin[a]);
      f[a] = q - (float)(int)std::floor(q);
    }
    const uint64_t key = ((uint64_t)(uint32_t)(b[0] + (1 << 20)) << 42) |
                         ((uint64_t)(uint32_t)(b[1] + (1 << 20)) << 21) |
                         (uint64_t)(uint32_t)(b[2] + (1 << 20));
    auto it = index.find(key);
    if (it == index.end()) {
      index.emplace(key, bins.size());
      bins.push_back({Key{b[0], b[1], b[2]},
                      Bin{0.0f, 0.0f, 0.0f, 0u, 0.0f, 0.0f, 0.0f, 0.0f}});
      it = index.find(key);
    }
    Bin& bin = bins[it->second].second;
    bin.sx += f[0];
    bin.sy += f[1];
    bin.sz += f[2];
    bin.n += 1;
    bin.hx = half_rn((double)bin.hx + (double)half_rn(f[0]));
    bin.hy = half_rn((double)bin.hy + (double)half_rn(f[1]));
    bin.hz = half_rn((double)bin.hz + (double)half_rn(f[2]));
    bin.hn = half_rn((double)bin.hn + 1.0);
  }
  std::sort(bins.begin(), bins.end(),
            [](const std::pair<Key, Bin>& a, const std::pair<Key, Bin>& b) {
              return a.first < b.first;
            });
  int n = 0;
  for (const auto& kb : bins) {
    const Bin& bn = kb.second;
    float pt[3];
    if (precision <= 1) {
      const float cnt = precision == 0 ? (float)bn.n : bn.hn;
      pt[0] = (precision == 0 ? bn.sx : bn.hx) / cnt + (float)kb.first.x;
      pt[1] = (precision == 0 ? bn.sy : bn.hy) / cnt + (float)kb.first.y;
      pt[2] = (precision == 0 ? bn.sz : bn.hz) / cnt + (float)kb.first.z;
    } else {
      if (precision == 3 && bn.n % 256u == 0u) continue;
      pt[0] = (float)kb.first.x + 0.5f;
      pt[1] = (float)kb.first.y + 0.5f;
      pt[2] = (float)kb.first.z + 0.5f;
    }
    const float len = std::sqrt(pt[0] * pt[0] + pt[1] * pt[1] + pt[2] * pt[2]);
    if (!(len > 0.0f && len < max_time)) continue;
    if (n < max_rays) {
      rays4[4 * n + 0] = pt[0] / len;
      rays4[4 * n + 1] = pt[1] / len;
      rays4[4 * n + 2] = pt[2] / len;
      rays4[4 * n + 3] = len;
    }
    ++n;
  }
  return n;
}

// ---------------------------------------------------------------------
// Sensor-simulation cast: restatement of viewSimKernel + getActiveCrossing
// (generateCloud.cu:20-82) over a NanoVDB-style 5-4-3 tree, with the
// rays of gpu::Sensor::updateFrustum/getRay (Sensor.h:62-77, 119-156)
// taking an independent horizontal FOV (the reference derives it from the
// pixel aspect, FrustumData.h:24, Sensor.h:64-67), and the spherical
// laser branch (Sensor.h:134-155).  FP32, round to
// nearest (the reference runs it on the GPU).
// ---------------------------------------------------------------------
}  // extern "C"

struct or_simgrid {
  std::unordered_map<uint64_t, std::array<uint64_t, 8>> leaves;  // mask
  std::unordered_map<uint64_t, int> lowers;  // origin key -> 1
  std::unordered_map<uint64_t, int> uppers;
  int bmin[3], bmax[3];
  int64_t active = 0;
};

namespace {

inline uint64_t node_key(int x, int y, int z, int mask) {
  const uint64_t m = (1ull << 21) - 1;
  return ((uint64_t)(((x & ~mask) >> 3) & m) << 42) |
         ((uint64_t)(((y & ~mask) >> 3) & m) << 21) |
         (uint64_t)(((z & ~mask) >> 3) & m);
}

// ReadAccessor::getDim: 1 inside a leaf, 8 in a lower node without the
// leaf, 128 in an upper node without the lower node, 4096 outside.
inline int sim_get_dim(const or_simgrid* g, int x, int y, int z) {
  if (!g->uppers.count(node_key(x, y, z, 4095))) return 4096;
  if (!g->lowers.count(node_key(x, y, z, 127))) return 128;
  if (!g->leaves.count(node_key(x, y, z, 7))) return 8;
  return 1;
}

inline bool sim_active(const or_simgrid* g, int x, int y, int z) {
  auto it = g->leaves.find(node_key(x, y, z, 7));
  if (it == g->leaves.end()) return false;
  const int li = local_index(x, y, z);
  return (it->second[li >> 6] >> (li & 63)) & 1;
}

inline int min_index_f(const float n[3]) {
  if (n[0] < n[1] && n[0] < n[2]) return 0;
  if (n[1] < n[2]) return 1;
  return 2;
}

struct FRay {
  float eye[3], dir[3], inv[3];
  float t0, t1;
  void at(float t, float p[3]) const {
    for (int a = 0; a < 3; ++a) p[a] = eye[a] + dir[a] * t;
  }
};

// nanovdb::HDDA over FP32 rays.
struct HDda {
  int dim;
  float t0, t1;
  int voxel[3], step[3];
  float next[3], delta[3];
  void init(const FRay& r, float start, float maxt, int d) {
    dim = d;
    t0 = start;
    t1 = maxt;
    float pos[3];
    r.at(t0, pos);
    for (int a = 0; a < 3; ++a) {
      voxel[a] = ((int)std::floor(pos[a])) & ~(d - 1);
      if (r.dir[a] == 0.0f) {
        next[a] = FLT_MAX;
        step[a] = 0;
        delta[a] = 0.0f;
      } else if (r.inv[a] > 0) {
        step[a] = 1;
        next[a] = t0 + ((float)(voxel[a] + d) - pos[a]) * r.inv[a];
        delta[a] = r.inv[a];
      } else {
        step[a] = -1;
        next[a] = t0 + ((float)voxel[a] - pos[a]) * r.inv[a];
        delta[a] = -r.inv[a];
      }
    }
  }
  void update(const FRay& r, int d) {
    if (dim == d) return;
    dim = d;
    float pos[3];
    r.at(t0, pos);
    for (int a = 0; a < 3; ++a) voxel[a] = ((int)std::floor(pos[a])) & ~(d - 1);
    for (int a = 0; a < 3; ++a) {
      if (step[a] == 0) continue;
      next[a] = t0 + ((float)voxel[a] - pos[a]) * r.inv[a];
      if (step[a] > 0) next[a] += (float)d * r.inv[a];
    }
  }
  bool step_() {
    const int a = min_index_f(next);
    t0 = next[a];
    next[a] += (float)dim * delta[a];
    voxel[a] += dim * step[a];
    return t0 <= t1;
  }
};

// Ray::clip(CoordBBox): slabs of [min, max + 1).
bool clip_bbox(FRay& r, const int bmin[3], const int bmax[3]) {
  float t0 = r.t0, t1 = r.t1;
  for (int a = 0; a < 3; ++a) {
    float lo = (float)bmin[a], hi = (float)(bmax[a] + 1);
    if (lo >= hi) return false;
    lo = (lo - r.eye[a]) * r.inv[a];
    hi = (hi - r.eye[a]) * r.inv[a];
    if (lo > hi) std::swap(lo, hi);
    if (lo > t0) t0 = lo;
    if (hi < t1) t1 = hi;
    if (t0 > t1) return false;
  }
  r.t0 = t0;
  r.t1 = t1;
  return true;
}

bool active_crossing(const or_simgrid* g, FRay r, bool hier, int ijk[3],
                     float* t) {
  if (g->active == 0) return false;
  if (!clip_bbox(r, g->bmin, g->bmax) || r.t1 > 1e20f) return false;
  HDda h;
  h.init(r, r.t0, r.t1, 1);
  // Livelock breaker: a ray through an edge where two node boundaries
  // meet can re-anchor back and forth forever with t frozen (refine to a
  // leaf, step, coarsen back to the same anchor).  After kStall steps at
  // one t, coarsening is suppressed until t advances.  Rays that never
  // stall are unaffected.  Steps are also capped (never reached by rays
  // that make progress).
  const int kStall = 8;
  int stall = 0;
  float tprev = -1.0f;
  const double cap = 6.0 * ((double)r.t1 - (double)r.t0 + 16.0) + 4096.0;
  double steps = 0;
  while (h.step_()) {
    if (++steps > cap) return false;
    if (hier) {
      stall = (h.t0 == tprev) ? stall + 1 : 0;
      tprev = h.t0;
      const int d = sim_get_dim(g, h.voxel[0], h.voxel[1], h.voxel[2]);
      if (h.dim != d && !(stall >= kStall && d > h.dim)) h.update(r, d);
      if (h.dim > 1) continue;
    }
    if (sim_active(g, h.voxel[0], h.voxel[1], h.voxel[2])) {
      for (int a = 0; a < 3; ++a) ijk[a] = h.voxel[a];
      *t = h.t0;
      return true;
    }
  }
  return false;
}

}  // namespace

extern "C" {

or_simgrid* or_simgrid_build(const int32_t* origins, const float* values,
                             int64_t n_leaves, float threshold) {
  or_simgrid* g = new or_simgrid();
  for (int a = 0; a < 3; ++a) {
    g->bmin[a] = INT32_MAX;
    g->bmax[a] = INT32_MIN;
  }
  for (int64_t l = 0; l < n_leaves; ++l) {
    std::array<uint64_t, 8> mask{};
    bool any = false;
    const int* o = origins + 3 * l;
    for (int li = 0; li < 512; ++li) {
      if (values[512 * l + li] > threshold) {
        mask[li >> 6] |= 1ull << (li & 63);
        any = true;
        const int v[3] = {o[0] + (li >> 6), o[1] + ((li >> 3) & 7), o[2] + (li & 7)};
        for (int a = 0; a < 3; ++a) {
          g->bmin[a] = std::min(g->bmin[a], v[a]);
          g->bmax[a] = std::max(g->bmax[a], v[a]);
        }
        g->active += 1;
      }
    }
    if (!any) continue;
    g->leaves[node_key(o[0], o[1], o[2], 7)] = mask;
    g->lowers[node_key(o[0], o[1], o[2], 127)] = 1;
    g->uppers[node_key(o[0], o[1], o[2], 4095)] = 1;
  }
  return g;
}

void or_simgrid_destroy(or_simgrid* g) { delete g; }

int64_t or_simgrid_active_count(const or_simgrid* g) { return g->active; }

int or_sim_cast(const or_simgrid* g, const float pos[3], const float R[9],
                float res, int width, int height, float hfov_deg,
                float vfov_deg, float min_range, float max_range, int type,
                const float* rays, int hierarchical, int32_t* ijk,
                float* dist, float* xyz) {
  const bool laser = type == 1;
  RoundGuard guard(0);
  // Laser beams (Sensor.h:134-143, scale from updateLaser 57-58): yaw and
  // pitch in FP32 degrees, -fov/2 + i fov/n, scaled to radians through
  // double; sin and cos of the FP32 angle rounded from double.
  std::vector<float> ls, lc;  // W yaws then H pitches
  if (laser) {
    const float fov[2] = {hfov_deg, vfov_deg};
    const int n[2] = {width, height};
    for (int a = 0; a < 2; ++a) {
      const float step = fov[a] / (float)n[a];
      for (int i = 0; i < n[a]; ++i) {
        volatile float ang = -fov[a] / 2 + (float)i * step;
        ang = (float)((double)ang * (3.14159265358979323846 / 180));
        ls.push_back((float)std::sin((double)ang));
        lc.push_back((float)std::cos((double)ang));
      }
    }
  }
  // Sensor::updateFrustum (Sensor.h:62-77)
  const float hh = (float)std::tan(vfov_deg * 3.14159265358979323846 / 360);
  const float hw = (float)std::tan(hfov_deg * 3.14159265358979323846 / 360);
  const float s0 = (float)(1.0 / width), s1 = (float)(1.0 / height);
  float T[3][3], I[3][3];  // T[c][i]: column c, component i
  for (int i = 0; i < 3; ++i) {
    const float r0 = R[3 * i + 0], r1 = R[3 * i + 1], r2 = R[3 * i + 2];
    T[2][i] = hw * r0 + hh * r1 - r2;
    T[0][i] = r0 * (2 * hw);
    T[1][i] = r1 * (2 * hh);
  }
  const float e0[3] = {1, 0, 0}, e1[3] = {0, 1, 0}, e2[3] = {0, 0, 1};
  for (int i = 0; i < 3; ++i) {
    I[2][i] = hw * e0[i] + hh * e1[i] - e2[i];
    I[0][i] = e0[i] * 2 * hw;
    I[1][i] = e1[i] * 2 * hh;
  }
  int returns = 0;
  for (int h = 0; h < height; ++h) {
    for (int w = 0; w < width; ++w) {
      const size_t pix = (size_t)h * width + w;
      float md[3], dd[3], min_t, max_t;
      auto len = [](const float* x) {
        return std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
      };
      if (type != 0) {
        // Sensor::getRay, laser branch (Sensor.h:134-155): the beam
        // (sin u cos v, sin v, cos u cos v), or a given direction, is
        // normalised; the world ray is R times it, normalised; the point
        // direction is normalised again
        float b[3];
        if (laser) {
          const float su = ls[w], cu = lc[w], sv = ls[width + h], cv = lc[width + h];
          b[0] = su * cv;
          b[1] = sv;
          b[2] = cu * cv;
        } else {
          for (int i = 0; i < 3; ++i) b[i] = rays[3 * (size_t)w + i];
        }
        const float nb = len(b);
        for (int i = 0; i < 3; ++i) b[i] = b[i] / nb;
        for (int i = 0; i < 3; ++i)
          md[i] = R[3 * i] * b[0] + R[3 * i + 1] * b[1] + R[3 * i + 2] * b[2];
        const float nm = len(md), nd = len(b);
        for (int i = 0; i < 3; ++i) md[i] = md[i] / nm;
        for (int i = 0; i < 3; ++i) dd[i] = b[i] / nd;
        min_t = min_range / res;
        max_t = max_range / res;
      } else {
        // Sensor::getRay, frustum branch (Sensor.h:121-133)
        const float u = (float)w * s0, v = (float)h * s1;
        for (int i = 0; i < 3; ++i) {
          md[i] = u * T[0][i] + v * T[1][i] - T[2][i];
          dd[i] = u * I[0][i] + v * I[1][i] - I[2][i];
        }
        const float norm = len(md);
        const float dn = len(dd);
        for (int i = 0; i < 3; ++i) dd[i] = dd[i] / dn;
        for (int i = 0; i < 3; ++i) md[i] = md[i] / norm;
        min_t = (min_range / res) * norm;
        max_t = (max_range / res) * norm;
      }
      FRay r;
      for (int a = 0; a < 3; ++a) {
        r.eye[a] = pos[a] / res;
        r.dir[a] = md[a];
        r.inv[a] = 1.0f / md[a];
      }
      r.t0 = 1.0e-5f;
      r.t1 = FLT_MAX;
      int vox[3];
      float t = 0;
      const bool hit = active_crossing(g, r, hierarchical != 0, vox, &t);
      if (hit && t >= min_t && t <= max_t) {
        const float d = t * res;
        if (ijk) for (int a = 0; a < 3; ++a) ijk[3 * pix + a] = vox[a];
        if (dist) dist[pix] = d;
        if (xyz) for (int a = 0; a < 3; ++a) xyz[3 * pix + a] = dd[a] * d;
        ++returns;
      } else {
        if (ijk) for (int a = 0; a < 3; ++a) ijk[3 * pix + a] = INT32_MIN;
        if (dist) dist[pix] = -1.0f;
        if (xyz) for (int a = 0; a < 3; ++a) xyz[3 * pix + a] = NAN;
      }
    }
  }
  return returns;
}

}  // extern "C"
