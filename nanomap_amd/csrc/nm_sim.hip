// nm_sim.hip -- sensor-simulation cast (generateCloud).
//
// Replaces viewSimKernel + getActiveCrossing (src/nanomap/kernels/
// generateCloud.cu:20-82) and the rays of gpu::Sensor::getRay
// (include/nanomap/gpu/Sensor.h:119-156, set up by updateFrustum 62-77):
// the frustum branch with an independent horizontal FOV (the reference
// derives it from the pixel aspect, FrustumData.h:24), and the spherical
// laser branch.  FP32, IEEE round to nearest.  A laser beam's sines and
// cosines come from per-sensor tables computed on the host, so every
// arithmetic step on the device is a correctly rounded IEEE operation.
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
#include <cstring>
#include <mutex>
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
  // per-call inputs of nm_sim_cast (return counter, sensor descriptors,
  // laser tables), grown when a call needs more than any before it
  unsigned char* scratch = nullptr;
  size_t scratch_bytes = 0;
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
  float T[3][3];  // frustum: transformed view vectors, T[column][component]
  float I[3][3];  // frustum: identity view vectors
  float R[9];     // laser: rotation, row-major
  float s0, s1;   // frustum: 1/W, 1/H
  float eye[3];   // position / res
  float res, min_range, max_range;
  int width, height;
  int type;        // NM_SENSOR_*
  int tab_off;     // laser: (sin, cos) of the W yaws, then of the H pitches;
                   // rays: W directions (3 floats each)
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

__device__ __forceinline__ float norm3(const float v[3]) {
  return __builtin_sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

__global__ __launch_bounds__(256) void k_zero_i32(int* __restrict__ p) {
  if (threadIdx.x == 0) *p = 0;
}

// One 8x8 pixel tile per wavefront; returns the tile's return count (the
// same value in every lane).
__device__ __forceinline__ int sim_tile(
    const SimGridView& g, const SimDesc* __restrict__ descs,
    const float* __restrict__ tables, int n_desc, int tile,
    int32_t* __restrict__ out_ijk, float* __restrict__ out_dist,
    float* __restrict__ out_xyz) {
  const int lane = threadIdx.x & 63;
  // sensor of this tile (few sensors: linear scan)
  int s = 0;
  while (s + 1 < n_desc && descs[s + 1].tile_off <= tile) ++s;
  const SimDesc& d = descs[s];
  const int t = tile - d.tile_off;
  const int w = (t % d.tiles_x) * 8 + (lane & 7);
  const int h = (t / d.tiles_x) * 8 + (lane >> 3);
  const bool in_image = w < d.width && h < d.height;
  const int64_t pix = d.pix_off + (int64_t)h * d.width + w;

  // gpu::Sensor::getRay.  md: the ray in the world frame, dd: the sensor
  // frame direction the simulated point lies along.
  float md[3], dd[3], min_t, max_t;
  if (d.type != NM_SENSOR_FRUSTUM) {
    // lanes past the image edge read beam 0: they only join the ballot
    const int wc = in_image ? w : 0, hc = in_image ? h : 0;
    if (d.type == NM_SENSOR_LASER) {
      // spherical branch (Sensor.h:134-155): yaw u, pitch v of the beam
      const float su = tables[d.tab_off + 2 * wc];
      const float cu = tables[d.tab_off + 2 * wc + 1];
      const int ph = d.tab_off + 2 * d.width + 2 * hc;
      const float sv = tables[ph], cv = tables[ph + 1];
      dd[0] = su * cv;
      dd[1] = sv;
      dd[2] = cu * cv;
    } else {  // explicit directions
      dd[0] = tables[d.tab_off + 3 * wc];
      dd[1] = tables[d.tab_off + 3 * wc + 1];
      dd[2] = tables[d.tab_off + 3 * wc + 2];
    }
    const float n0 = norm3(dd);
#pragma unroll
    for (int i = 0; i < 3; ++i) dd[i] = dd[i] / n0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
      md[i] = d.R[3 * i] * dd[0] + d.R[3 * i + 1] * dd[1] + d.R[3 * i + 2] * dd[2];
    const float nm_ = norm3(md), nd = norm3(dd);
#pragma unroll
    for (int i = 0; i < 3; ++i) md[i] = md[i] / nm_;
#pragma unroll
    for (int i = 0; i < 3; ++i) dd[i] = dd[i] / nd;
    min_t = d.min_range / d.res;
    max_t = d.max_range / d.res;
  } else {
    // frustum branch (Sensor.h:121-133)
    const float u = (float)w * d.s0, v = (float)h * d.s1;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      md[i] = u * d.T[0][i] + v * d.T[1][i] - d.T[2][i];
      dd[i] = u * d.I[0][i] + v * d.I[1][i] - d.I[2][i];
    }
    const float norm = norm3(md);
    const float dn = norm3(dd);
#pragma unroll
    for (int i = 0; i < 3; ++i) dd[i] = dd[i] / dn;
#pragma unroll
    for (int i = 0; i < 3; ++i) md[i] = md[i] / norm;
    min_t = (d.min_range / d.res) * norm;
    max_t = (d.max_range / d.res) * norm;
  }
  float eye[3], inv[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    eye[a] = d.eye[a];
    inv[a] = 1.0f / md[a];
  }

  bool hit = false;
  int vox[3] = {0, 0, 0};
  float thit = 0.0f;
  if (in_image && g.active > 0) {
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
  const bool ret = in_image && hit && thit >= min_t && thit <= max_t;
  const int count = (int)__popcll(__ballot(ret));
  if (!in_image) return count;
  if (ret) {
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
  return count;
}

// Tiles in a grid-stride loop over kSimBlocks workgroups at most; the
// return count is summed per wavefront, then per workgroup, and added with
// one atomic per workgroup (one per wavefront and tile serialised on the
// counter at the memory side).
constexpr int kSimBlocks = 16384;

__global__ __launch_bounds__(256) void k_sim_cast(
    SimGridView g, const SimDesc* __restrict__ descs,
    const float* __restrict__ tables, int n_desc, int total_tiles,
    int32_t* __restrict__ out_ijk, float* __restrict__ out_dist,
    float* __restrict__ out_xyz, int* __restrict__ returns) {
  __shared__ int s_count[4];
  int count = 0;
  for (int tile = blockIdx.x * 4 + (threadIdx.x >> 6); tile < total_tiles;
       tile += gridDim.x * 4)
    count += sim_tile(g, descs, tables, n_desc, tile, out_ijk, out_dist,
                      out_xyz);
  if ((threadIdx.x & 63) == 0) s_count[threadIdx.x >> 6] = count;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int total = s_count[0] + s_count[1] + s_count[2] + s_count[3];
    if (total) atomicAdd(returns, total);
  }
}

uint64_t okey(int x, int y, int z) {
  const uint64_t m = (1ull << 21) - 1;
  return ((uint64_t)((x >> 3) & m) << 42) | ((uint64_t)((y >> 3) & m) << 21) |
         (uint64_t)((z >> 3) & m);
}

int setup_fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

}  // namespace

// Device layout of one call's inputs: the return counter at 0, the image
// (descriptors, then laser tables) from kSimImageOffset.
constexpr size_t kSimImageOffset = 256;

struct SimLaunch {
  int n_desc = 0;
  int tiles = 0;
  size_t tables_off = 0;  // byte offset of the tables inside the image
};

// Validates the sensors and lays out their descriptors and laser tables.
// With img == null only *bytes (the image size) is computed.
int sim_setup(const nm_sim_sensor* sensors, int n, unsigned char* img,
              size_t cap, size_t* bytes, SimLaunch* L) {
  size_t n_tab = 0;  // floats of laser tables
  for (int k = 0; k < n; ++k) {
    const nm_sim_sensor& in = sensors[k];
    if (in.width <= 0 || in.height <= 0 || !(in.grid_res > 0) ||
        (in.type != NM_SENSOR_FRUSTUM && in.type != NM_SENSOR_LASER &&
         in.type != NM_SENSOR_RAYS) ||
        (in.type == NM_SENSOR_RAYS && (!in.rays || in.height != 1)) ||
        (int64_t)in.width * in.height > (1 << 30))
      return setup_fail(NM_EINVAL, "nm_sim_cast: bad sensor");
    if (in.type == NM_SENSOR_LASER) n_tab += 2 * ((size_t)in.width + in.height);
    if (in.type == NM_SENSOR_RAYS) n_tab += 3 * (size_t)in.width;
  }
  const size_t tab_off = ((size_t)n * sizeof(SimDesc) + 15) & ~(size_t)15;
  *bytes = tab_off + n_tab * sizeof(float);
  if (!img) return NM_OK;
  if (*bytes > cap) return setup_fail(NM_EINVAL, "nm_sim_cast: image too large");
  SimDesc* descs = reinterpret_cast<SimDesc*>(img);
  float* tab = reinterpret_cast<float*>(img + tab_off);
  int tiles = 0, toff = 0;
  int64_t pix = 0;
  for (int k = 0; k < n; ++k) {
    const nm_sim_sensor& in = sensors[k];
    SimDesc d;
    memset(&d, 0, sizeof(d));
    d.type = in.type;
    const float* R = in.rotation;
    for (int i = 0; i < 9; ++i) d.R[i] = R[i];
    if (in.type == NM_SENSOR_FRUSTUM) {
      // Sensor::updateFrustum (Sensor.h:62-77) with an independent hfov
      const float hh = (float)tan(in.vfov_deg * 3.14159265358979323846 / 360);
      const float hw = (float)tan(in.hfov_deg * 3.14159265358979323846 / 360);
      d.s0 = (float)(1.0 / in.width);
      d.s1 = (float)(1.0 / in.height);
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
    } else if (in.type == NM_SENSOR_RAYS) {
      d.tab_off = toff;
      for (int i = 0; i < 3 * in.width; ++i) tab[toff++] = in.rays[i];
    } else {
      // Sensor::getRay, laser branch (Sensor.h:134-143; scale from
      // updateLaser, Sensor.h:57-58): FP32 angles in degrees, scaled to
      // radians through double; sin and cos of the FP32 angle, correctly
      // rounded through double
      d.tab_off = toff;
      const float scale[2] = {in.hfov_deg / (float)in.width,
                              in.vfov_deg / (float)in.height};
      const float half[2] = {-in.hfov_deg / 2, -in.vfov_deg / 2};
      const int count[2] = {in.width, in.height};
      for (int a = 0; a < 2; ++a)
        for (int i = 0; i < count[a]; ++i) {
          volatile float ang = half[a] + (float)i * scale[a];
          ang = (float)((double)ang * (3.14159265358979323846 / 180));
          tab[toff++] = (float)sin((double)ang);
          tab[toff++] = (float)cos((double)ang);
        }
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
    if ((int64_t)tiles + (int64_t)d.tiles_x * ((in.height + 7) / 8) > INT_MAX / 4)
      return setup_fail(NM_EINVAL, "nm_sim_cast: too many pixels");
    tiles += d.tiles_x * ((in.height + 7) / 8);
    pix += (int64_t)in.width * in.height;
    memcpy(&descs[k], &d, sizeof(d));
  }
  L->n_desc = n;
  L->tiles = tiles;
  L->tables_off = tab_off;
  return NM_OK;
}

// Zeroes the return counter and casts every tile; scratch holds the
// counter and, from kSimImageOffset, the image.
hipError_t sim_launch(const nm_simgrid* g, const SimLaunch& L,
                      unsigned char* scratch, int32_t* ijk, float* dist,
                      float* xyz, hipStream_t s) {
  SimGridView v;
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
  int* ret = reinterpret_cast<int*>(scratch);
  const SimDesc* descs = reinterpret_cast<const SimDesc*>(scratch + kSimImageOffset);
  const float* tables =
      reinterpret_cast<const float*>(scratch + kSimImageOffset + L.tables_off);
  k_zero_i32<<<1, 64, 0, s>>>(ret);
  if (L.tiles > 0)
    k_sim_cast<<<std::min((L.tiles + 3) / 4, kSimBlocks), 256, 0, s>>>(
        v, descs, tables, L.n_desc, L.tiles, ijk, dist, xyz, ret);
  return hipGetLastError();
}

}  // namespace nm

namespace {
int sim_fail(int code, const std::string& msg) {
  nm::set_error(msg);
  return code;
}
}  // namespace

// The buffer of nm_sim_cast_static: the return counter and the sensors'
// descriptors and laser tables of one call, laid out as in the grid-owned
// buffer of nm_sim_cast.  Each GPU holds its own copy; host code reaches
// it only through the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[NM_SIM_STATIC_BYTES];

namespace {

constexpr int kStaticDevices = 64;
std::mutex g_static_mu;
unsigned char* g_static_ptr[kStaticDevices] = {};

// Device address of `scratch` on `device` (current), looked up on the
// first call on that device and kept.
int static_scratch(int device, unsigned char** out) {
  if (device < 0 || device >= kStaticDevices)
    return sim_fail(NM_ENODEV, "nm_sim_cast_static: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return sim_fail(NM_EHIP, std::string("nm_sim_cast_static: "
                                           "hipGetSymbolAddress: ") +
                                   hipGetErrorString(e));
    g_static_ptr[device] = (unsigned char*)p;
  }
  *out = g_static_ptr[device];
  return NM_OK;
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
  (void)hipDeviceSynchronize();
  void* p[] = {g->upper_origin, g->upper_child, g->lower_child, g->leaf_mask,
               g->scratch};
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
  size_t bytes = 0;
  int st = nm::sim_setup(sensors, n_sensors, nullptr, 0, &bytes, nullptr);
  if (st != NM_OK) return st;
  std::vector<unsigned char> img(bytes);
  nm::SimLaunch L;
  st = nm::sim_setup(sensors, n_sensors, img.data(), bytes, &bytes, &L);
  if (st != NM_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  const size_t need = nm::kSimImageOffset + bytes;
  if (need > g->scratch_bytes) {
    // a larger sensor set than any before: grow the grid's buffer
    if (hipStreamSynchronize(s) != hipSuccess ||
        (g->scratch && hipFree(g->scratch) != hipSuccess))
      return sim_fail(NM_EHIP, "nm_sim_cast: releasing the old buffer");
    g->scratch = nullptr;
    g->scratch_bytes = 0;
    if (hipMalloc((void**)&g->scratch, need) != hipSuccess)
      return sim_fail(NM_ENOMEM, "nm_sim_cast: descriptor buffer");
    g->scratch_bytes = need;
  }
  hipError_t e = hipMemcpyAsync(g->scratch + nm::kSimImageOffset, img.data(),
                                bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = nm::sim_launch(g, L, g->scratch, ijk, dist, xyz, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(returns, g->scratch, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess)
    return sim_fail(NM_EHIP, std::string("nm_sim_cast: ") + hipGetErrorString(e));
  return NM_OK;
}

int nm_sim_cast_static(nm_simgrid* g, const nm_sim_sensor* sensors,
                       int n_sensors, int32_t* ijk, float* dist, float* xyz,
                       int* returns, void* stream) {
  if (!g || !sensors || n_sensors <= 0 || !returns)
    return sim_fail(NM_EINVAL, "nm_sim_cast_static: bad argument");
  if (hipSetDevice(g->device) != hipSuccess)
    return sim_fail(NM_ENODEV, "nm_sim_cast_static: bad device");
  size_t bytes = 0;
  int st = nm::sim_setup(sensors, n_sensors, nullptr, 0, &bytes, nullptr);
  if (st != NM_OK) return st;
  if (nm::kSimImageOffset + bytes > (size_t)NM_SIM_STATIC_BYTES)
    return sim_fail(NM_EINVAL,
                    "nm_sim_cast_static: the sensors need " +
                        std::to_string(nm::kSimImageOffset + bytes) +
                        " bytes of descriptors, more than the fixed " +
                        std::to_string(NM_SIM_STATIC_BYTES));
  // host image of the inputs: a fixed per-thread array (each call
  // synchronizes before it returns, so the next call may reuse it)
  alignas(16) static thread_local unsigned char img[NM_SIM_STATIC_BYTES];
  nm::SimLaunch L;
  st = nm::sim_setup(sensors, n_sensors, img, sizeof(img), &bytes, &L);
  if (st != NM_OK) return st;
  unsigned char* dev = nullptr;
  st = static_scratch(g->device, &dev);
  if (st != NM_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(dev + nm::kSimImageOffset, img, bytes,
                                hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = nm::sim_launch(g, L, dev, ijk, dist, xyz, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(returns, dev, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess)
    return sim_fail(NM_EHIP, std::string("nm_sim_cast_static: ") +
                                 hipGetErrorString(e));
  return NM_OK;
}

}  // extern "C"
