// nm_sim.hip -- sensor-simulation cast (generateCloud).
//
// Replaces viewSimKernel + getActiveCrossing (src/nanomap/kernels/
// generateCloud.cu:20-82) and the frustum ray of gpu::Sensor::getRay
// (include/nanomap/gpu/Sensor.h:213-227, set up by updateFrustum 156-171),
// with an independent horizontal FOV (the reference derives it from the
// pixel aspect, FrustumData.h:24).  FP32, IEEE round to nearest.
//
// The occupancy grid mirrors the NanoVDB tree the reference converts its
// map into (SimHandler.cpp:22): 8^3 leaves with 512-bit active masks,
// lower nodes of 16^3 leaves (128^3 voxels), upper nodes of 32^3 lower
// nodes (4096^3 voxels), a root list of upper nodes.  A node exists iff it
// holds an active voxel.  The hierarchical DDA steps at the dimension of
// the node containing the current voxel (1 in a leaf, 8 in a lower node
// without that leaf, 128 in an upper node without that lower node, 4096
// outside every upper node), re-anchoring at ray(t) when it changes, and
// returns the entry time of the first active voxel after the start voxel.
//
// Work layout: one wavefront per 8x8 pixel tile (neighbouring rays share
// empty-space skips and node lookups), tiles of all sensors in one launch.
#include <math.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/nanomap_amd_sim.h"
#include "nm_internal.h"

struct nm_simgrid {
  int device = 0;
  int bmin[3] = {0, 0, 0}, bmax[3] = {-1, -1, -1};
  int64_t active = 0;
  int n_upper = 0, n_lower = 0, n_leaf = 0;
  int32_t* upper_origin = nullptr;  // [n_upper*3]
  int32_t* upper_child = nullptr;   // [n_upper*32768] lower index or -1
  int32_t* lower_child = nullptr;   // [n_lower*4096]  leaf index or -1
  uint32_t* leaf_mask = nullptr;    // [n_leaf*16]
};

namespace nm {
namespace {

struct SimGridView {
  int bmin[3], bmax[3];
  int64_t active;
  int n_upper;
  const int32_t* upper_origin;
  const int32_t* upper_child;
  const int32_t* lower_child;
  const uint32_t* leaf_mask;
};

// Per-sensor constants, computed on the host exactly as the oracle does.
struct SimDesc {
  float T[3][3];  // transformed view vectors, T[column][component]
  float I[3][3];  // identity view vectors
  float s0, s1;   // 1/W, 1/H
  float eye[3];   // position / res
  float res, min_range, max_range;
  int width, height;
  int tile_off;    // first global tile of this sensor
  int tiles_x;     // tiles per row
  int64_t pix_off; // first output pixel of this sensor
};

// Lookups with a one-entry cache per level (nanovdb::ReadAccessor).
struct Acc {
  int up_key[3] = {INT_MIN, 0, 0};
  int up = -1;
  int lo_key[3] = {INT_MIN, 0, 0};
  int lo = -1;
  int lf_key[3] = {INT_MIN, 0, 0};
  int lf = -1;
};

__device__ __forceinline__ bool same3(const int a[3], int x, int y, int z) {
  return a[0] == x && a[1] == y && a[2] == z;
}

// Returns the node dimension at (x,y,z); *leaf = leaf index if dim == 1.
__device__ int get_dim(const SimGridView& g, Acc& c, int x, int y, int z,
                       int* leaf) {
  const int ux = x & ~4095, uy = y & ~4095, uz = z & ~4095;
  if (!same3(c.up_key, ux, uy, uz)) {
    c.up_key[0] = ux;
    c.up_key[1] = uy;
    c.up_key[2] = uz;
    c.up = -1;
    for (int u = 0; u < g.n_upper; ++u) {
      if (g.upper_origin[3 * u] == ux && g.upper_origin[3 * u + 1] == uy &&
          g.upper_origin[3 * u + 2] == uz) {
        c.up = u;
        break;
      }
    }
    c.lo_key[0] = INT_MIN;
    c.lf_key[0] = INT_MIN;
  }
  if (c.up < 0) return 4096;
  const int lx = x & ~127, ly = y & ~127, lz = z & ~127;
  if (!same3(c.lo_key, lx, ly, lz)) {
    c.lo_key[0] = lx;
    c.lo_key[1] = ly;
    c.lo_key[2] = lz;
    const int n = (((x & 4095) >> 7) << 10) | (((y & 4095) >> 7) << 5) |
                  ((z & 4095) >> 7);
    c.lo = g.upper_child[(int64_t)c.up * 32768 + n];
    c.lf_key[0] = INT_MIN;
  }
  if (c.lo < 0) return 128;
  const int fx = x & ~7, fy = y & ~7, fz = z & ~7;
  if (!same3(c.lf_key, fx, fy, fz)) {
    c.lf_key[0] = fx;
    c.lf_key[1] = fy;
    c.lf_key[2] = fz;
    const int n = (((x & 127) >> 3) << 8) | (((y & 127) >> 3) << 4) |
                  ((z & 127) >> 3);
    c.lf = g.lower_child[(int64_t)c.lo * 4096 + n];
  }
  if (c.lf < 0) return 8;
  *leaf = c.lf;
  return 1;
}

__device__ __forceinline__ int fmin_index(float a, float b, float c) {
  if (a < b && a < c) return 0;
  if (b < c) return 1;
  return 2;
}

__global__ __launch_bounds__(256) void k_sim_cast(
    SimGridView g, const SimDesc* __restrict__ descs, int n_desc,
    int total_tiles, int32_t* __restrict__ out_ijk,
    float* __restrict__ out_dist, float* __restrict__ out_xyz,
    int* __restrict__ returns) {
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (tile >= total_tiles) return;
  // sensor of this tile (few sensors: linear scan)
  int s = 0;
  while (s + 1 < n_desc && descs[s + 1].tile_off <= tile) ++s;
  const SimDesc& d = descs[s];
  const int t = tile - d.tile_off;
  const int w = (t % d.tiles_x) * 8 + (lane & 7);
  const int h = (t / d.tiles_x) * 8 + (lane >> 3);
  if (w >= d.width || h >= d.height) return;
  const int64_t pix = d.pix_off + (int64_t)h * d.width + w;

  // gpu::Sensor::getRay, frustum branch
  const float u = (float)w * d.s0, v = (float)h * d.s1;
  float md[3], dd[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    md[i] = u * d.T[0][i] + v * d.T[1][i] - d.T[2][i];
    dd[i] = u * d.I[0][i] + v * d.I[1][i] - d.I[2][i];
  }
  const float norm = __builtin_sqrtf(md[0] * md[0] + md[1] * md[1] + md[2] * md[2]);
  const float dn = __builtin_sqrtf(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
#pragma unroll
  for (int i = 0; i < 3; ++i) dd[i] = dd[i] / dn;
#pragma unroll
  for (int i = 0; i < 3; ++i) md[i] = md[i] / norm;
  const float min_t = (d.min_range / d.res) * norm;
  const float max_t = (d.max_range / d.res) * norm;
  float eye[3], inv[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    eye[a] = d.eye[a];
    inv[a] = 1.0f / md[a];
  }

  bool hit = false;
  int vox[3] = {0, 0, 0};
  float thit = 0.0f;
  if (g.active > 0) {
    // Ray::clip(root bbox) over [min, max + 1)
    float t0 = 1.0e-5f, t1 = FLT_MAX;
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float lo = (float)g.bmin[a], hi = (float)(g.bmax[a] + 1);
      if (lo >= hi) ok = false;
      lo = (lo - eye[a]) * inv[a];
      hi = (hi - eye[a]) * inv[a];
      if (lo > hi) {
        const float tmp = lo;
        lo = hi;
        hi = tmp;
      }
      if (lo > t0) t0 = lo;
      if (hi < t1) t1 = hi;
      if (t0 > t1) ok = false;
    }
    if (ok && !(t1 > 1e20f)) {
      // HDDA::init(ray, t0, t1, 1)
      int dim = 1;
      float pos[3], next[3], delta[3];
      int vx[3], st[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        pos[a] = eye[a] + md[a] * t0;
        vx[a] = (int)__builtin_floorf(pos[a]);
        if (md[a] == 0.0f) {
          next[a] = FLT_MAX;
          st[a] = 0;
          delta[a] = 0.0f;
        } else if (inv[a] > 0) {
          st[a] = 1;
          next[a] = t0 + ((float)(vx[a] + 1) - pos[a]) * inv[a];
          delta[a] = inv[a];
        } else {
          st[a] = -1;
          next[a] = t0 + ((float)vx[a] - pos[a]) * inv[a];
          delta[a] = -inv[a];
        }
      }
      Acc acc;
      float tc = t0;
      // Livelock breaker (same rule as the CPU restatement): after kStall
      // steps at one t, coarsening is suppressed until t advances; steps
      // are capped so every lane exits.
      const int kStall = 8;
      int stall = 0;
      float tprev = -1.0f;
      const double cap = 6.0 * ((double)t1 - (double)t0 + 16.0) + 4096.0;
      double steps = 0;
      for (;;) {
        // HDDA::step()
        const int ax = fmin_index(next[0], next[1], next[2]);
        tc = next[ax];
        next[ax] += (float)dim * delta[ax];
        vx[ax] += dim * st[ax];
        if (!(tc <= t1)) break;
        if (++steps > cap) break;
        stall = (tc == tprev) ? stall + 1 : 0;
        tprev = tc;
        int leaf = -1;
        const int nd = get_dim(g, acc, vx[0], vx[1], vx[2], &leaf);
        if (nd != dim && !(stall >= kStall && nd > dim)) {
          // HDDA::update(ray, nd): re-anchor at ray(t)
          dim = nd;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            pos[a] = eye[a] + md[a] * tc;
            vx[a] = ((int)__builtin_floorf(pos[a])) & ~(nd - 1);
          }
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            if (st[a] == 0) continue;
            next[a] = tc + ((float)vx[a] - pos[a]) * inv[a];
            if (st[a] > 0) next[a] += (float)nd * inv[a];
          }
          leaf = -1;
          if (nd == 1) get_dim(g, acc, vx[0], vx[1], vx[2], &leaf);
        }
        if (dim > 1) continue;
        if (leaf >= 0) {
          const int li = ((vx[0] & 7) << 6) | ((vx[1] & 7) << 3) | (vx[2] & 7);
          if ((g.leaf_mask[(int64_t)leaf * 16 + (li >> 5)] >> (li & 31)) & 1u) {
            hit = true;
            thit = tc;
            vox[0] = vx[0];
            vox[1] = vx[1];
            vox[2] = vx[2];
            break;
          }
        }
      }
    }
  }
  if (hit && thit >= min_t && thit <= max_t) {
    const float dist = thit * d.res;
    if (out_ijk) {
      out_ijk[3 * pix + 0] = vox[0];
      out_ijk[3 * pix + 1] = vox[1];
      out_ijk[3 * pix + 2] = vox[2];
    }
    if (out_dist) out_dist[pix] = dist;
    if (out_xyz) {
      out_xyz[3 * pix + 0] = dd[0] * dist;
      out_xyz[3 * pix + 1] = dd[1] * dist;
      out_xyz[3 * pix + 2] = dd[2] * dist;
    }
    atomicAdd(returns, 1);
  } else {
    if (out_ijk) {
      out_ijk[3 * pix + 0] = INT_MIN;
      out_ijk[3 * pix + 1] = INT_MIN;
      out_ijk[3 * pix + 2] = INT_MIN;
    }
    if (out_dist) out_dist[pix] = -1.0f;
    if (out_xyz) {
      out_xyz[3 * pix + 0] = __builtin_nanf("");
      out_xyz[3 * pix + 1] = __builtin_nanf("");
      out_xyz[3 * pix + 2] = __builtin_nanf("");
    }
  }
}

uint64_t okey(int x, int y, int z) {
  const uint64_t m = (1ull << 21) - 1;
  return ((uint64_t)((x >> 3) & m) << 42) | ((uint64_t)((y >> 3) & m) << 21) |
         (uint64_t)((z >> 3) & m);
}

}  // namespace
}  // namespace nm

namespace {
int sim_fail(int code, const std::string& msg) {
  nm::set_error(msg);
  return code;
}
}  // namespace

extern "C" {

int nm_simgrid_from_leaves(int device, const int32_t* origins,
                           const float* values, int64_t n_leaves,
                           float threshold, nm_simgrid** out) {
  if (!out || (n_leaves > 0 && (!origins || !values)) || n_leaves < 0)
    return sim_fail(NM_EINVAL, "nm_simgrid_from_leaves: bad argument");
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess)
    return sim_fail(NM_ENODEV, "nm_simgrid_from_leaves: bad device");
  nm_simgrid* g = new nm_simgrid();
  g->device = device;
  for (int a = 0; a < 3; ++a) {
    g->bmin[a] = INT_MAX;
    g->bmax[a] = INT_MIN;
  }
  std::vector<int32_t> up_o, up_c, lo_c;
  std::vector<uint32_t> masks;
  std::unordered_map<uint64_t, int> up_idx, lo_idx;
  for (int64_t l = 0; l < n_leaves; ++l) {
    uint32_t m[16] = {0};
    bool any = false;
    const int32_t* o = origins + 3 * l;
    for (int li = 0; li < 512; ++li) {
      if (values[512 * l + li] > threshold) {
        m[li >> 5] |= 1u << (li & 31);
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
    const int ux = o[0] & ~4095, uy = o[1] & ~4095, uz = o[2] & ~4095;
    const uint64_t uk = nm::okey(ux, uy, uz);
    auto ui = up_idx.find(uk);
    int u;
    if (ui == up_idx.end()) {
      u = (int)up_idx.size();
      up_idx.emplace(uk, u);
      up_o.insert(up_o.end(), {ux, uy, uz});
      up_c.resize(up_c.size() + 32768, -1);
    } else {
      u = ui->second;
    }
    const int lx = o[0] & ~127, ly = o[1] & ~127, lz = o[2] & ~127;
    const uint64_t lk = nm::okey(lx, ly, lz);
    auto li_ = lo_idx.find(lk);
    int lo;
    if (li_ == lo_idx.end()) {
      lo = (int)lo_idx.size();
      lo_idx.emplace(lk, lo);
      lo_c.resize(lo_c.size() + 4096, -1);
      const int n = (((o[0] & 4095) >> 7) << 10) | (((o[1] & 4095) >> 7) << 5) |
                    ((o[2] & 4095) >> 7);
      up_c[(size_t)u * 32768 + n] = lo;
    } else {
      lo = li_->second;
    }
    const int leaf = (int)(masks.size() / 16);
    masks.insert(masks.end(), m, m + 16);
    const int n = (((o[0] & 127) >> 3) << 8) | (((o[1] & 127) >> 3) << 4) |
                  ((o[2] & 127) >> 3);
    lo_c[(size_t)lo * 4096 + n] = leaf;
  }
  g->n_upper = (int)up_idx.size();
  g->n_lower = (int)lo_idx.size();
  g->n_leaf = (int)(masks.size() / 16);
  hipError_t e = hipSuccess;
  auto up = [&](void** dst, const void* src, size_t bytes) {
    if (e != hipSuccess) return;
    e = hipMalloc(dst, bytes + 64);
    if (e == hipSuccess && bytes) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
  };
  up((void**)&g->upper_origin, up_o.data(), up_o.size() * 4);
  up((void**)&g->upper_child, up_c.data(), up_c.size() * 4);
  up((void**)&g->lower_child, lo_c.data(), lo_c.size() * 4);
  up((void**)&g->leaf_mask, masks.data(), masks.size() * 4);
  if (e != hipSuccess) {
    nm_simgrid_destroy(g);
    return sim_fail(NM_EHIP, std::string("nm_simgrid_from_leaves: ") + hipGetErrorString(e));
  }
  *out = g;
  return NM_OK;
}

int nm_simgrid_from_map(nm_map* map, float threshold, nm_simgrid** out) {
  if (!map || !out) return sim_fail(NM_EINVAL, "nm_simgrid_from_map: null argument");
  int64_t n = 0;
  int st = nm_map_leaf_count(map, &n);
  if (st != NM_OK) return st;
  std::vector<int32_t> o((size_t)std::max<int64_t>(n, 1) * 3);
  std::vector<float> v((size_t)std::max<int64_t>(n, 1) * 512);
  st = nm_map_download(map, o.data(), v.data(), nullptr);
  if (st != NM_OK) return st;
  int dev = 0;
  (void)hipGetDevice(&dev);
  return nm_simgrid_from_leaves(dev, o.data(), v.data(), n, threshold, out);
}

int nm_simgrid_destroy(nm_simgrid* g) {
  if (!g) return NM_OK;
  (void)hipSetDevice(g->device);
  void* p[] = {g->upper_origin, g->upper_child, g->lower_child, g->leaf_mask};
  for (void* x : p)
    if (x) (void)hipFree(x);
  delete g;
  return NM_OK;
}

int nm_simgrid_info(nm_simgrid* g, int64_t* active, int32_t bbox[6],
                    int32_t counts[3]) {
  if (!g) return sim_fail(NM_EINVAL, "nm_simgrid_info: null grid");
  if (active) *active = g->active;
  if (bbox)
    for (int a = 0; a < 3; ++a) {
      bbox[a] = g->bmin[a];
      bbox[3 + a] = g->bmax[a];
    }
  if (counts) {
    counts[0] = g->n_upper;
    counts[1] = g->n_lower;
    counts[2] = g->n_leaf;
  }
  return NM_OK;
}

int nm_sim_cast(nm_simgrid* g, const nm_sim_sensor* sensors, int n_sensors,
                int32_t* ijk, float* dist, float* xyz, int* returns,
                void* stream) {
  if (!g || !sensors || n_sensors <= 0 || !returns)
    return sim_fail(NM_EINVAL, "nm_sim_cast: bad argument");
  if (hipSetDevice(g->device) != hipSuccess)
    return sim_fail(NM_ENODEV, "nm_sim_cast: bad device");
  std::vector<nm::SimDesc> descs(n_sensors);
  int tiles = 0;
  int64_t pix = 0;
  for (int s = 0; s < n_sensors; ++s) {
    const nm_sim_sensor& in = sensors[s];
    if (in.width <= 0 || in.height <= 0 || !(in.grid_res > 0))
      return sim_fail(NM_EINVAL, "nm_sim_cast: bad sensor");
    nm::SimDesc& d = descs[s];
    const float hh = (float)tan(in.vfov_deg * 3.14159265358979323846 / 360);
    const float hw = (float)tan(in.hfov_deg * 3.14159265358979323846 / 360);
    d.s0 = (float)(1.0 / in.width);
    d.s1 = (float)(1.0 / in.height);
    const float* R = in.rotation;
    for (int i = 0; i < 3; ++i) {
      const float r0 = R[3 * i + 0], r1 = R[3 * i + 1], r2 = R[3 * i + 2];
      d.T[2][i] = hw * r0 + hh * r1 - r2;
      d.T[0][i] = r0 * (2 * hw);
      d.T[1][i] = r1 * (2 * hh);
    }
    const float e0[3] = {1, 0, 0}, e1[3] = {0, 1, 0}, e2[3] = {0, 0, 1};
    for (int i = 0; i < 3; ++i) {
      d.I[2][i] = hw * e0[i] + hh * e1[i] - e2[i];
      d.I[0][i] = e0[i] * 2 * hw;
      d.I[1][i] = e1[i] * 2 * hh;
    }
    for (int a = 0; a < 3; ++a) d.eye[a] = in.position[a] / in.grid_res;
    d.res = in.grid_res;
    d.min_range = in.min_range;
    d.max_range = in.max_range;
    d.width = in.width;
    d.height = in.height;
    d.tiles_x = (in.width + 7) / 8;
    d.tile_off = tiles;
    d.pix_off = pix;
    tiles += d.tiles_x * ((in.height + 7) / 8);
    pix += (int64_t)in.width * in.height;
  }
  hipStream_t s = (hipStream_t)stream;
  nm::SimDesc* ddesc = nullptr;
  int* dret = nullptr;
  hipError_t e = hipMallocAsync((void**)&ddesc, sizeof(nm::SimDesc) * n_sensors, s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&dret, sizeof(int), s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ddesc, descs.data(), sizeof(nm::SimDesc) * n_sensors,
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(dret, 0, sizeof(int), s);
  if (e == hipSuccess) {
    nm::SimGridView v;
    for (int a = 0; a < 3; ++a) {
      v.bmin[a] = g->bmin[a];
      v.bmax[a] = g->bmax[a];
    }
    v.active = g->active;
    v.n_upper = g->n_upper;
    v.upper_origin = g->upper_origin;
    v.upper_child = g->upper_child;
    v.lower_child = g->lower_child;
    v.leaf_mask = g->leaf_mask;
    nm::k_sim_cast<<<(tiles + 3) / 4, 256, 0, s>>>(v, ddesc, n_sensors, tiles,
                                                   ijk, dist, xyz, dret);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(returns, dret, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (ddesc) (void)hipFreeAsync(ddesc, s);
  if (dret) (void)hipFreeAsync(dret, s);
  if (e != hipSuccess)
    return sim_fail(NM_EHIP, std::string("nm_sim_cast: ") + hipGetErrorString(e));
  return NM_OK;
}

}  // extern "C"
