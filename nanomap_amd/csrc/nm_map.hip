// nm_map.hip -- operations on a whole device map: growth, batched point
// queries, snapshot and restore.
//
// The reference's map is an openvdb::FloatGrid (include/nanomap/map/
// OccupancyMap.h:35-61) that grows a leaf at a time and is read through a
// value accessor (occAccessor(), OccupancyMap.h:60).  The device map here
// is a leaf pool plus an open-addressing hash (nm_internal.h: MapView):
//   grow     new pool and hash, the pool copied slot for slot (pool
//            indices do not change) and the hash rebuilt from the origins;
//   query    one lane per voxel: hash probe of its leaf, value and mask bit;
//   save     the leaves as nm_map_download gives them, with the map's
//            parameters; load: the same leaves as a fresh pool and hash.
#include <limits.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "nm_internal.h"

namespace nm {
namespace {

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? NM_ENOMEM : NM_EHIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

#define NM_HIP(call)                                 \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

// Inserts every non-hole slot [0, slots) of the pool under its origin.
// The keys are distinct, so each CAS that claims an empty entry is final.
__global__ __launch_bounds__(256) void k_map_rehash(MapView m, int slots,
                                                    int32_t* __restrict__ err) {
  for (int s = blockIdx.x * 256 + threadIdx.x; s < slots;
       s += gridDim.x * 256) {
    const int x = m.origins[3 * s], y = m.origins[3 * s + 1],
              z = m.origins[3 * s + 2];
    if (x == INT_MIN) continue;  // hole
    const unsigned long long key = pack_leaf(x, y, z);
    const unsigned long long mask = (unsigned long long)m.hash_cap - 1;
    unsigned long long h = mix64(key) & mask;
    bool done = false;
    for (int64_t probe = 0; probe < m.hash_cap && !done; ++probe) {
      if (m.hash_keys[h] == ~0ull &&
          atomicCAS(&m.hash_keys[h], ~0ull, key) == ~0ull) {
        m.hash_vals[h] = s;
        done = true;
      }
      h = (h + 1) & mask;
    }
    if (!done) atomicAdd(err, 1);
  }
}

__global__ __launch_bounds__(256) void k_map_query(
    MapView m, const int32_t* __restrict__ ijk, int64_t n,
    float* __restrict__ values, uint8_t* __restrict__ active) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const int x = ijk[3 * i], y = ijk[3 * i + 1], z = ijk[3 * i + 2];
    const int leaf = map_find(m, x, y, z);
    float v = 0.0f;
    uint8_t a = 0;
    if (leaf >= 0) {
      const int li = ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
      v = m.values[(size_t)leaf * 512 + li];
      a = (uint8_t)((m.masks[(size_t)leaf * kMaskWords + (li >> 5)] >>
                     (li & 31)) & 1u);
    }
    if (values) values[i] = v;
    if (active) active[i] = a;
  }
}

// Dense copy of the box [x0, x0 + nx) x [y0, y0 + ny) x [z0, z0 + nz): out
// [(i * ny + j) * nz + k] = value of voxel (x0 + i, y0 + j, z0 + k), or
// `fill` where its leaf is absent.  Consecutive lanes take consecutive k,
// so a wavefront's lookups hit one or two leaves.
__global__ __launch_bounds__(256) void k_map_dense(MapView m, int x0, int y0,
                                                   int z0, int nx, int ny,
                                                   int nz, float fill,
                                                   float* __restrict__ out) {
  const int64_t n = (int64_t)nx * ny * nz;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * 256) {
    const int k = (int)(t % nz);
    const int64_t r = t / nz;
    const int j = (int)(r % ny);
    const int i = (int)(r / ny);
    const int x = x0 + i, y = y0 + j, z = z0 + k;
    const int leaf = map_find(m, x, y, z);
    out[t] = leaf >= 0
                 ? m.values[(size_t)leaf * 512 +
                            (((x & 7) << 6) | ((y & 7) << 3) | (z & 7))]
                 : fill;
  }
}

template <class T>
hipError_t dalloc(T** p, size_t count) {
  return hipMalloc((void**)p, count * sizeof(T) + 64);
}

// Pool slots in use (capped at the capacity) and holes; synchronous.
int map_counts(nm_map* m, int32_t n[2]) {
  NM_HIP(hipStreamSynchronize(m->stream));
  NM_HIP(hipMemcpy(n, m->v.leaf_count, 2 * sizeof(int32_t),
                   hipMemcpyDeviceToHost));
  n[0] = std::min(n[0], m->v.leaf_cap);
  return NM_OK;
}

// Allocates the arrays of a map view of `cap` leaves: pool zeroed, hash
// empty.  Synchronous.
int alloc_view(MapView* v, int64_t cap) {
  int64_t h = 1;
  while (h < 2 * cap) h <<= 1;
  MapView n = *v;
  n.hash_keys = nullptr;
  n.hash_vals = nullptr;
  n.values = nullptr;
  n.masks = nullptr;
  n.origins = nullptr;
  n.leaf_count = nullptr;
  n.leaf_cap = (int32_t)cap;
  n.hash_cap = h;
  hipError_t e = hipSuccess;
  auto A = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
  };
  A(dalloc(&n.hash_keys, (size_t)h));
  A(dalloc(&n.hash_vals, (size_t)h));
  A(dalloc(&n.values, (size_t)cap * 512));
  A(dalloc(&n.masks, (size_t)cap * kMaskWords));
  A(dalloc(&n.origins, (size_t)cap * 3));
  A(dalloc(&n.leaf_count, 2));
  if (e == hipSuccess) e = launch_map_clear(n, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    void* p[] = {n.hash_keys, n.hash_vals, n.values,
                 n.masks,     n.origins,   n.leaf_count};
    for (void* x : p)
      if (x) (void)hipFree(x);
    return hip_fail(e, "map allocation");
  }
  *v = n;
  return NM_OK;
}

void free_view(MapView* v) {
  void* p[] = {v->hash_keys, v->hash_vals, v->values,
               v->masks,     v->origins,   v->leaf_count};
  for (void* x : p)
    if (x) (void)hipFree(x);
}

// Rebuilds v's hash from its pool slots [0, slots); synchronous.
int rehash(MapView& v, int slots) {
  int32_t* err = nullptr;
  NM_HIP(hipMalloc((void**)&err, sizeof(int32_t)));
  hipError_t e = hipMemset(err, 0, sizeof(int32_t));
  if (e == hipSuccess && slots > 0) {
    k_map_rehash<<<std::min((slots + 255) / 256, 4096), 256>>>(v, slots, err);
    e = hipGetLastError();
  }
  int32_t h_err = 0;
  if (e == hipSuccess) e = hipMemcpy(&h_err, err, sizeof(int32_t),
                                     hipMemcpyDeviceToHost);
  (void)hipFree(err);
  if (e != hipSuccess) return hip_fail(e, "map rehash");
  if (h_err) return fail(NM_EOVERFLOW, "map rehash: hash table full");
  return NM_OK;
}

}  // namespace

int map_reserve(nm_map* m, int64_t capacity) {
  if (capacity <= m->v.leaf_cap) return NM_OK;
  if (capacity > (INT_MAX - 64) / 512)
    return fail(NM_EINVAL, "nm_map_reserve: capacity too large");
  NM_HIP(hipSetDevice(m->device));
  int32_t n[2];
  int st = map_counts(m, n);
  if (st != NM_OK) return st;
  MapView nv = m->v;
  st = alloc_view(&nv, capacity);
  if (st != NM_OK) return st;
  const size_t slots = (size_t)n[0];
  hipError_t e = hipSuccess;
  if (slots) {
    e = hipMemcpy(nv.values, m->v.values, slots * 512 * sizeof(float),
                  hipMemcpyDeviceToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(nv.masks, m->v.masks, slots * kMaskWords * sizeof(uint32_t),
                    hipMemcpyDeviceToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(nv.origins, m->v.origins, slots * 3 * sizeof(int32_t),
                    hipMemcpyDeviceToDevice);
  }
  if (e == hipSuccess)
    e = hipMemcpy(nv.leaf_count, n, 2 * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    free_view(&nv);
    return hip_fail(e, "nm_map_reserve");
  }
  st = rehash(nv, (int)slots);
  if (st != NM_OK) {
    free_view(&nv);
    return st;
  }
  free_view(&m->v);
  m->v = nv;
  return NM_OK;
}

}  // namespace nm

using nm::fail;
using nm::hip_fail;

extern "C" {

int nm_map_capacity(nm_map* m, int64_t* cap) {
  if (!m || !cap) return fail(NM_EINVAL, "nm_map_capacity: null argument");
  *cap = m->v.leaf_cap;
  return NM_OK;
}

int nm_map_params(nm_map* m, float* grid_res, float* clamp_min,
                  float* clamp_max, float* hit_thr, float* miss_thr) {
  if (!m) return fail(NM_EINVAL, "nm_map_params: null map");
  if (grid_res) *grid_res = m->grid_res;
  if (clamp_min) *clamp_min = m->v.clamp_min;
  if (clamp_max) *clamp_max = m->v.clamp_max;
  if (hit_thr) *hit_thr = m->v.hit_thr;
  if (miss_thr) *miss_thr = m->v.miss_thr;
  return NM_OK;
}

int nm_map_reserve(nm_map* m, int64_t capacity) {
  if (!m) return fail(NM_EINVAL, "nm_map_reserve: null map");
  return nm::map_reserve(m, capacity);
}

int nm_map_set_auto_grow(nm_map* m, int enable) {
  if (!m) return fail(NM_EINVAL, "nm_map_set_auto_grow: null map");
  m->auto_grow = enable != 0;
  return NM_OK;
}

int nm_map_query(nm_map* m, const int32_t* ijk, int64_t n, float* values,
                 uint8_t* active, int on_device, void* stream) {
  if (!m || n < 0 || (n > 0 && !ijk))
    return fail(NM_EINVAL, "nm_map_query: bad argument");
  if (n == 0) return NM_OK;
  NM_HIP(hipSetDevice(m->device));
  hipStream_t s = (hipStream_t)stream;
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 8192);
  if (on_device) {
    nm::k_map_query<<<grid, 256, 0, s>>>(m->v, ijk, n, values, active);
    NM_HIP(hipGetLastError());
    return NM_OK;
  }
  // host arrays: 12 B of coordinates, 4 B of value and 1 B of state per
  // point, staged in the map's buffer
  const size_t per = 12 + 4 + 1;
  if (n > m->query_cap) {
    NM_HIP(hipStreamSynchronize(s));
    if (m->query_buf) NM_HIP(hipFree(m->query_buf));
    m->query_buf = nullptr;
    m->query_cap = 0;
    NM_HIP(hipMalloc(&m->query_buf, (size_t)n * per + 256));
    m->query_cap = n;
  }
  char* base = (char*)m->query_buf;
  int32_t* d_ijk = (int32_t*)base;
  float* d_val = (float*)(base + (size_t)n * 12);
  uint8_t* d_act = (uint8_t*)(base + (size_t)n * 16);
  NM_HIP(hipMemcpyAsync(d_ijk, ijk, (size_t)n * 12, hipMemcpyHostToDevice, s));
  nm::k_map_query<<<grid, 256, 0, s>>>(m->v, d_ijk, n, d_val, d_act);
  NM_HIP(hipGetLastError());
  if (values)
    NM_HIP(hipMemcpyAsync(values, d_val, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  if (active)
    NM_HIP(hipMemcpyAsync(active, d_act, (size_t)n, hipMemcpyDeviceToHost, s));
  NM_HIP(hipStreamSynchronize(s));
  return NM_OK;
}

int nm_map_dense(nm_map* m, const int32_t lo[3], int nx, int ny, int nz,
                 float fill, float* out, void* stream) {
  if (!m || !lo || !out || nx <= 0 || ny <= 0 || nz <= 0)
    return fail(NM_EINVAL, "nm_map_dense: bad argument");
  // the box's voxels must have leaf keys (21 bits per axis around 0)
  const int64_t lim = (int64_t)1 << 23;
  for (int a = 0; a < 3; ++a) {
    const int64_t d = a == 0 ? nx : (a == 1 ? ny : nz);
    if ((int64_t)lo[a] < -lim || (int64_t)lo[a] + d > lim)
      return fail(NM_EINVAL, "nm_map_dense: box outside the map's key range");
  }
  NM_HIP(hipSetDevice(m->device));
  const int64_t n = (int64_t)nx * ny * nz;
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 65536);
  nm::k_map_dense<<<grid, 256, 0, (hipStream_t)stream>>>(
      m->v, lo[0], lo[1], lo[2], nx, ny, nz, fill, out);
  NM_HIP(hipGetLastError());
  return NM_OK;
}

// ---- snapshot ----------------------------------------------------------
namespace {

struct SnapHeader {
  char magic[8];       // "NMAPLEAF"
  uint32_t version;    // 1
  uint32_t leaf_edge;  // 8
  int64_t leaves;
  float grid_res, clamp_min, clamp_max, hit_thr, miss_thr;
  uint32_t pad[5];
};
static_assert(sizeof(SnapHeader) == 64, "snapshot header is 64 bytes");

// A map on `device` holding n host leaves (nm_map_download layout), with
// room for max(leaf_capacity, n) leaves, or for max(2 n, 2^18) with
// leaf_capacity 0.
int map_from_leaves(int device, float grid_res, float clamp_min,
                    float clamp_max, float hit_thr, float miss_thr, int64_t n,
                    const int32_t* o, const float* v, const uint32_t* k,
                    int64_t leaf_capacity, nm_map** out, const char* who) {
  const int64_t cap = leaf_capacity > 0
                          ? std::max<int64_t>(leaf_capacity, std::max<int64_t>(n, 1))
                          : std::max<int64_t>(2 * n, (int64_t)1 << 18);
  if (cap > (INT_MAX - 64) / 512)
    return fail(NM_EINVAL, std::string(who) + ": capacity too large");
  nm_map* m = nullptr;
  int st = nm_map_create(device, grid_res, clamp_min, clamp_max, hit_thr,
                         miss_thr, (int)cap, &m);
  if (st != NM_OK) return st;
  hipError_t e = hipSuccess;
  if (n) {
    e = hipMemcpy(m->v.origins, o, (size_t)n * 12, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(m->v.values, v, (size_t)n * 2048, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(m->v.masks, k, (size_t)n * nm::kMaskWords * 4,
                    hipMemcpyHostToDevice);
  }
  const int32_t cnt[2] = {(int32_t)n, 0};
  if (e == hipSuccess)
    e = hipMemcpy(m->v.leaf_count, cnt, sizeof(cnt), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    nm_map_destroy(m);
    return hip_fail(e, who);
  }
  st = nm::rehash(m->v, (int)n);
  if (st != NM_OK) {
    nm_map_destroy(m);
    return st;
  }
  *out = m;
  return NM_OK;
}

}  // namespace

int nm_map_save(nm_map* m, const char* path) {
  if (!m || !path) return fail(NM_EINVAL, "nm_map_save: null argument");
  int64_t n = 0;
  int st = nm_map_leaf_count(m, &n);
  if (st != NM_OK) return st;
  std::vector<int32_t> o((size_t)std::max<int64_t>(n, 1) * 3);
  std::vector<float> v((size_t)std::max<int64_t>(n, 1) * 512);
  std::vector<uint32_t> k((size_t)std::max<int64_t>(n, 1) * nm::kMaskWords);
  st = nm_map_download(m, o.data(), v.data(), k.data());
  if (st != NM_OK) return st;
  SnapHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, "NMAPLEAF", 8);
  h.version = 1;
  h.leaf_edge = 8;
  h.leaves = n;
  h.grid_res = m->grid_res;
  h.clamp_min = m->v.clamp_min;
  h.clamp_max = m->v.clamp_max;
  h.hit_thr = m->v.hit_thr;
  h.miss_thr = m->v.miss_thr;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(NM_EINVAL, std::string("nm_map_save: cannot open ") + path);
  bool ok = fwrite(&h, sizeof(h), 1, f) == 1;
  for (int64_t l = 0; ok && l < n; ++l) {
    ok = fwrite(&o[3 * l], 4, 3, f) == 3 &&
         fwrite(&v[512 * l], 4, 512, f) == 512 &&
         fwrite(&k[nm::kMaskWords * l], 4, nm::kMaskWords, f) ==
             (size_t)nm::kMaskWords;
  }
  ok = (fclose(f) == 0) && ok;
  if (!ok) return fail(NM_EINVAL, std::string("nm_map_save: write failed: ") + path);
  return NM_OK;
}

int nm_map_load(int device, const char* path, int64_t leaf_capacity,
                nm_map** out) {
  if (!path || !out) return fail(NM_EINVAL, "nm_map_load: null argument");
  *out = nullptr;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(NM_EINVAL, std::string("nm_map_load: cannot open ") + path);
  SnapHeader h;
  bool ok = fread(&h, sizeof(h), 1, f) == 1 && memcmp(h.magic, "NMAPLEAF", 8) == 0 &&
            h.version == 1 && h.leaf_edge == 8 && h.leaves >= 0 &&
            h.leaves <= (INT_MAX - 64) / 512;
  const int64_t n = ok ? h.leaves : 0;
  std::vector<int32_t> o((size_t)std::max<int64_t>(n, 1) * 3);
  std::vector<float> v((size_t)std::max<int64_t>(n, 1) * 512);
  std::vector<uint32_t> k((size_t)std::max<int64_t>(n, 1) * nm::kMaskWords);
  for (int64_t l = 0; ok && l < n; ++l) {
    ok = fread(&o[3 * l], 4, 3, f) == 3 && fread(&v[512 * l], 4, 512, f) == 512 &&
         fread(&k[nm::kMaskWords * l], 4, nm::kMaskWords, f) ==
             (size_t)nm::kMaskWords &&
         o[3 * l] != INT_MIN && (o[3 * l] & 7) == 0 && (o[3 * l + 1] & 7) == 0 &&
         (o[3 * l + 2] & 7) == 0;
  }
  fclose(f);
  if (!ok) return fail(NM_EINVAL, std::string("nm_map_load: not a valid map snapshot: ") + path);
  return map_from_leaves(device, h.grid_res, h.clamp_min, h.clamp_max,
                         h.hit_thr, h.miss_thr, n, o.data(), v.data(),
                         k.data(), leaf_capacity, out, "nm_map_load");
}

int nm_map_write_nvdb(nm_map* m, const char* path, const char* grid_name) {
  if (!m || !path || !grid_name)
    return fail(NM_EINVAL, "nm_map_write_nvdb: null argument");
  int64_t n = 0;
  int st = nm_map_leaf_count(m, &n);
  if (st != NM_OK) return st;
  std::vector<int32_t> o((size_t)std::max<int64_t>(n, 1) * 3);
  std::vector<float> v((size_t)std::max<int64_t>(n, 1) * 512);
  std::vector<uint32_t> k((size_t)std::max<int64_t>(n, 1) * nm::kMaskWords);
  st = nm_map_download(m, o.data(), v.data(), k.data());
  if (st != NM_OK) return st;
  return nm_nvdb_write(path, o.data(), v.data(), k.data(), n,
                       (double)m->grid_res, grid_name);
}

int nm_map_load_nvdb(int device, const char* path, const char* grid_name,
                     float clamp_min, float clamp_max, float hit_thr,
                     float miss_thr, int64_t leaf_capacity, nm_map** out) {
  if (!path || !out) return fail(NM_EINVAL, "nm_map_load_nvdb: null argument");
  *out = nullptr;
  int64_t n = 0;
  double dx = 0;
  int st = nm_nvdb_read(path, grid_name, &n, nullptr, nullptr, nullptr, &dx);
  if (st != NM_OK) return st;
  if (n > (INT_MAX - 64) / 512)
    return fail(NM_EINVAL, "nm_map_load_nvdb: too many leaves");
  std::vector<int32_t> o((size_t)std::max<int64_t>(n, 1) * 3);
  std::vector<float> v((size_t)std::max<int64_t>(n, 1) * 512);
  std::vector<uint32_t> k((size_t)std::max<int64_t>(n, 1) * nm::kMaskWords);
  st = nm_nvdb_read(path, grid_name, &n, o.data(), v.data(), k.data(), &dx);
  if (st != NM_OK) return st;
  if (!(dx > 0) || !(dx < 1e30))
    return fail(NM_EINVAL, "nm_map_load_nvdb: the grid's voxel size is not positive");
  return map_from_leaves(device, (float)dx, clamp_min, clamp_max, hit_thr,
                         miss_thr, n, o.data(), v.data(), k.data(),
                         leaf_capacity, out, "nm_map_load_nvdb");
}

}  // extern "C"
