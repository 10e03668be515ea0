// nm_capi.hip -- the extern "C" entry points of include/nanomap_amd.h.
//
// Host orchestration of one frame, replacing Handler::processPointCloud's
// GPU branch (src/nanomap/handler/Handler.cpp:250-278): the reference
// synchronizes the device three times per frame to learn the ray and leaf
// counts (filterCloud.cu:550-551, frustumCastCloud.cu:347-356) and
// allocates CUB scratch every frame (frustumCastCloud.cu:64-68).  Here
// every count stays on the device, kernels are sized by worst-case
// capacities allocated once per context, and a frame is a fixed sequence
// of launches on one stream with no host synchronization.
#include <limits.h>
#include <math.h>

#include <algorithm>
#include <cmath>
#include <stdio.h>
#include <string.h>

#include <string>
#include <utility>
#include <vector>

#include "nm_fp.h"
#include <climits>
#include <cstring>
#include <utility>
#include "nm_internal.h"

struct nm_ctx {
  int device = 0;
  int max_points = 0;
  float max_time = 0;
  int cube_side_max = 0;
  int64_t cube_cap = 0;
  int tile_cap = 0;
  int cu_count = 256;
  int resident = 1536;  // accumulate workgroups resident at once
  size_t staging_bytes = 0;
  nm::Workspace b{};
  nm::FrameCounters* host_counters = nullptr;
  nm::JobStats* host_stats = nullptr;
  // state of the last nm_cast
  bool pending = false;
  nm_frame frame{};
  nm::Cube cube{};
  nm::Job job{};
  nm::Pipeline pipe{};
  int n = 0;
  int sorted_buf = 0;
  hipStream_t stream = nullptr;
  nm_map* last_map = nullptr;
  bool last_update_valid = false;  // slot_stats hold the last frame's sums
  // branch of the update for the heavy voxels (see launch_update)
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev = nullptr;
  // the workspace's miss-run table: valid for (miss bits, rounding)
  bool runs_valid = false;
  uint64_t runs_key = 0;
  // depth-image factors (ax[0, W), ay[0, H)) of the last depth camera
  float* depth_tab = nullptr;
  int depth_tab_cap = 0;
  bool depth_valid = false;
  nm_depth_camera depth_cam{};
  std::vector<float> depth_host;
  // stage timing
  bool timing = false;
  unsigned timing_mask = ~0u;  // stages timed while timing is on
  struct Mark {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;          // recorded, not yet read
  std::vector<hipEvent_t> free_ev;  // event pool
  double stage_ms[NM_STAGE_COUNT] = {};
  int64_t stage_n[NM_STAGE_COUNT] = {};
};

namespace nm {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

}  // namespace nm

// Sliced accumulate: a frame's work items all start at once when they fit
// on the GPU, so the pass lasts as long as its longest rays.  The blocks
// that hold them are cut into up to kAccumSlices time slices, run as extra
// work items: as many as kSliceOvercommit / 4 times the resident
// accumulate workgroups leave beside the frame's blocks (none when the
// blocks alone fill it), shared out by the slice plan.
constexpr int kAccumSlices = 6;
constexpr int kSliceOvercommit = 8;
static_assert(kAccumSlices <= nm::kMaxSlices, "slice index takes 3 bits");

namespace {

int fail(int code, const std::string& msg) {
  nm::set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(NM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define NM_HIP(call)                                 \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

template <class T>
hipError_t dalloc(T** p, size_t count) {
  return hipMalloc((void**)p, count * sizeof(T) + 64);
}

hipEvent_t take_event(nm_ctx* c) {
  if (!c->free_ev.empty()) {
    hipEvent_t e = c->free_ev.back();
    c->free_ev.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// RAII event pair around one stage (no-op unless timing is enabled).
struct StageTimer {
  nm_ctx* c;
  int stage;
  hipStream_t s;
  hipEvent_t a = nullptr;
  StageTimer(nm_ctx* c_, int stage_, hipStream_t s_) : c(c_), stage(stage_), s(s_) {
    if (!c->timing || !((c->timing_mask >> stage) & 1u)) return;
    a = take_event(c);
    if (a) (void)hipEventRecord(a, s);
  }
  ~StageTimer() {
    if (!a) return;
    hipEvent_t b = take_event(c);
    if (!b) return;
    (void)hipEventRecord(b, s);
    c->marks.push_back({stage, a, b});
  }
};

int drain_marks(nm_ctx* c) {
  for (auto& m : c->marks) {
    float ms = 0;
    hipError_t e = hipEventSynchronize(m.b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, m.a, m.b);
    if (e != hipSuccess) return hip_fail(e, "stage timing");
    c->stage_ms[m.stage] += ms;
    c->stage_n[m.stage] += 1;
    c->free_ev.push_back(m.a);
    c->free_ev.push_back(m.b);
  }
  c->marks.clear();
  return NM_OK;
}

void free_ctx(nm_ctx* c) {
  for (auto& m : c->marks) {
    (void)hipEventDestroy(m.a);
    (void)hipEventDestroy(m.b);
  }
  for (hipEvent_t e : c->free_ev) (void)hipEventDestroy(e);
  if (c->depth_tab) (void)hipFree(c->depth_tab);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  if (c->join_ev) (void)hipEventDestroy(c->join_ev);
  nm::workspace_free(&c->b);
  if (c->host_counters) (void)hipHostFree(c->host_counters);
  if (c->host_stats) (void)hipHostFree(c->host_stats);
  delete c;
}

// The voxel filter's point bounds of a frame: its box unless all six
// bounds are zero, and max_range (max_time * grid_res when 0).
int filter_bounds(const nm_frame& f, nm::FilterBounds* fb) {
  bool any = false;
  for (int a = 0; a < 3; ++a) {
    if (!std::isfinite(f.box_min[a]) || !std::isfinite(f.box_max[a]) ||
        !(f.box_min[a] <= f.box_max[a]))
      return fail(NM_EINVAL, "nm_cast: filter box must be finite with min <= max");
    any = any || f.box_min[a] != 0.0f || f.box_max[a] != 0.0f;
    fb->lo[a] = f.box_min[a];
    fb->hi[a] = f.box_max[a];
  }
  fb->use_box = any ? 1 : 0;
  if (!(f.max_range >= 0.0f) || !std::isfinite(f.max_range))
    return fail(NM_EINVAL, "nm_cast: max_range must be finite and >= 0");
  fb->max_range = f.max_range > 0.0f ? f.max_range : f.max_time * f.grid_res;
  return NM_OK;
}

}  // namespace

extern "C" {

const char* nm_last_error(void) { return nm::g_error.c_str(); }

int nm_version(void) { return 1; }

int nm_device_alloc(int device, int64_t bytes, void** out) {
  if (!out || bytes <= 0) return fail(NM_EINVAL, "nm_device_alloc: bad argument");
  *out = nullptr;
  NM_HIP(hipSetDevice(device));
  const hipError_t e = hipMalloc(out, (size_t)bytes);
  if (e != hipSuccess) {
    *out = nullptr;
    return fail(e == hipErrorOutOfMemory ? NM_ENOMEM : NM_EHIP,
                std::string("nm_device_alloc: ") + hipGetErrorString(e));
  }
  return NM_OK;
}

int nm_device_free(int device, void* ptr) {
  if (!ptr) return NM_OK;
  NM_HIP(hipSetDevice(device));
  NM_HIP(hipDeviceSynchronize());
  NM_HIP(hipFree(ptr));
  return NM_OK;
}

int nm_device_download(int device, void* dst, const void* src, int64_t bytes) {
  if (bytes < 0 || (bytes > 0 && (!dst || !src)))
    return fail(NM_EINVAL, "nm_device_download: bad argument");
  if (bytes == 0) return NM_OK;
  NM_HIP(hipSetDevice(device));
  NM_HIP(hipDeviceSynchronize());
  NM_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return NM_OK;
}

float nm_fold_host(float v, float a, uint32_t k, int rounding) {
  return rounding == NM_ROUND_DOWN ? nm::fold<nm::kDown>(v, a, k)
                                   : nm::fold<nm::kNearest>(v, a, k);
}

int nm_ctx_create(const nm_ctx_config* cfg, nm_ctx** out) {
  if (!cfg || !out) return fail(NM_EINVAL, "nm_ctx_create: null argument");
  *out = nullptr;
  if (cfg->max_points <= 0 || !(cfg->max_time > 0) || cfg->max_time > 4000)
    return fail(NM_EINVAL, "nm_ctx_create: bad max_points or max_time");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(NM_ENODEV, "nm_ctx_create: no HIP device");
  if (cfg->device < 0 || cfg->device >= ndev)
    return fail(NM_EINVAL, "nm_ctx_create: bad device ordinal");
  NM_HIP(hipSetDevice(cfg->device));
  nm_ctx* c = new nm_ctx();
  c->device = cfg->device;
  c->max_points = cfg->max_points;
  c->max_time = cfg->max_time;
  const int r = (int)ceil(((double)cfg->max_time + 2.0) / 8.0) + 1;
  c->cube_side_max = 2 * r + 1;
  c->cube_cap = (int64_t)c->cube_side_max * c->cube_side_max * c->cube_side_max;
  if (c->cube_cap >= (1 << 22)) {
    delete c;
    return fail(NM_EINVAL, "nm_ctx_create: max_time too large for the leaf cube");
  }
  c->tile_cap = cfg->tile_capacity > 0 ? cfg->tile_capacity : (int)c->cube_cap;
  if ((int64_t)c->tile_cap > c->cube_cap) c->tile_cap = (int)c->cube_cap;
  const size_t P = (size_t)cfg->max_points;
  nm::WorkspaceDims d;
  d.cube_cap = c->cube_cap;
  d.tile_cap = c->tile_cap;
  d.ray_cap = (int64_t)P;
  // 16x16 items: ceil(W/16)*ceil(H/16) <= P/256 + (W+H)/16 + 1, W+H <= P+1
  d.item_cap = (int)(P / 16 + P / 256 + 64);
  d.max_jobs = 1;
  d.host_clouds = true;
  c->staging_bytes = P * 16;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, cfg->device) == hipSuccess &&
      prop.multiProcessorCount > 0)
    c->cu_count = prop.multiProcessorCount;
  c->resident = nm::accumulate_resident(c->cu_count);
  hipError_t e = nm::workspace_alloc(d, &c->b);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&c->host_counters, sizeof(nm::FrameCounters));
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&c->host_stats, sizeof(nm::JobStats));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    free_ctx(c);
    return fail(e == hipErrorOutOfMemory ? NM_ENOMEM : NM_EHIP,
                std::string("nm_ctx_create: ") + hipGetErrorString(e));
  }
  *out = c;
  return NM_OK;
}

int nm_ctx_destroy(nm_ctx* ctx) {
  if (!ctx) return NM_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  free_ctx(ctx);
  return NM_OK;
}

int nm_map_create(int device, float grid_res, float clamp_min,
                  float clamp_max, float hit_thr, float miss_thr,
                  int leaf_capacity, nm_map** out) {
  if (!out) return fail(NM_EINVAL, "nm_map_create: null out");
  *out = nullptr;
  if (!(grid_res > 0) || leaf_capacity <= 0 || !(clamp_min <= clamp_max) ||
      !nm::logodds_ok(clamp_min) || !nm::logodds_ok(clamp_max))
    return fail(NM_EINVAL, "nm_map_create: bad argument");
  NM_HIP(hipSetDevice(device));
  nm_map* m = new nm_map();
  m->device = device;
  m->grid_res = grid_res;
  nm::MapView& v = m->v;
  v.leaf_cap = leaf_capacity;
  int64_t h = 1;
  while (h < 2 * (int64_t)leaf_capacity) h <<= 1;
  v.hash_cap = h;
  v.clamp_min = clamp_min;
  v.clamp_max = clamp_max;
  v.hit_thr = hit_thr;
  v.miss_thr = miss_thr;
  hipError_t e = hipSuccess;
  auto A = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
  };
  A(dalloc(&v.hash_keys, (size_t)h));
  A(dalloc(&v.hash_vals, (size_t)h));
  A(dalloc(&v.values, (size_t)leaf_capacity * 512));
  A(dalloc(&v.masks, (size_t)leaf_capacity * nm::kMaskWords));
  A(dalloc(&v.origins, (size_t)leaf_capacity * 3));
  A(dalloc(&v.leaf_count, 2));
  if (e == hipSuccess) e = nm::launch_map_clear(v, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    nm_map_destroy(m);
    return fail(e == hipErrorOutOfMemory ? NM_ENOMEM : NM_EHIP,
                std::string("nm_map_create: ") + hipGetErrorString(e));
  }
  *out = m;
  return NM_OK;
}

int nm_map_destroy(nm_map* m) {
  if (!m) return NM_OK;
  (void)hipSetDevice(m->device);
  (void)hipDeviceSynchronize();
  void* ptrs[] = {m->v.hash_keys, m->v.hash_vals, m->v.values,
                  m->v.masks,     m->v.origins,   m->v.leaf_count,
                  m->query_buf};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete m;
  return NM_OK;
}

int nm_map_clear(nm_map* m, void* stream) {
  if (!m) return fail(NM_EINVAL, "nm_map_clear: null map");
  NM_HIP(hipSetDevice(m->device));
  NM_HIP(nm::launch_map_clear(m->v, (hipStream_t)stream));
  m->stream = (hipStream_t)stream;
  return NM_OK;
}

// Pool slots in use and holes among them (slots reserved for a leaf that
// was then not inserted because none of its voxels changed).
static int map_slots(nm_map* m, int64_t* slots, int64_t* holes) {
  NM_HIP(hipSetDevice(m->device));
  NM_HIP(hipStreamSynchronize(m->stream));
  int32_t n[2] = {0, 0};
  NM_HIP(hipMemcpy(n, m->v.leaf_count, sizeof(n), hipMemcpyDeviceToHost));
  *slots = std::min<int64_t>(n[0], m->v.leaf_cap);
  *holes = n[1];
  return NM_OK;
}

int nm_map_leaf_count(nm_map* m, int64_t* count) {
  if (!m || !count) return fail(NM_EINVAL, "nm_map_leaf_count: null argument");
  int64_t slots = 0, holes = 0;
  int st = map_slots(m, &slots, &holes);
  if (st != NM_OK) return st;
  *count = slots - holes;
  return NM_OK;
}

int nm_map_download(nm_map* m, int32_t* origins, float* values,
                    uint32_t* masks) {
  if (!m) return fail(NM_EINVAL, "nm_map_download: null map");
  int64_t slots = 0, holes = 0;
  int st = map_slots(m, &slots, &holes);
  if (st != NM_OK) return st;
  if (slots == 0) return NM_OK;
  // runs [begin, end) of consecutive non-hole slots
  std::vector<std::pair<int64_t, int64_t>> runs;
  std::vector<int32_t> all((size_t)slots * 3);
  NM_HIP(hipMemcpy(all.data(), m->v.origins, all.size() * sizeof(int32_t),
                   hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < slots;) {
    if (all[3 * i] == INT_MIN) {
      ++i;
      continue;
    }
    int64_t j = i;
    while (j < slots && all[3 * j] != INT_MIN) ++j;
    runs.emplace_back(i, j);
    i = j;
  }
  int64_t out = 0;
  for (const auto& r : runs) {
    const int64_t n = r.second - r.first;
    if (origins)
      memcpy(origins + 3 * out, all.data() + 3 * r.first,
             (size_t)n * 3 * sizeof(int32_t));
    if (values)
      NM_HIP(hipMemcpy(values + out * 512, m->v.values + r.first * 512,
                       (size_t)n * 512 * sizeof(float), hipMemcpyDeviceToHost));
    if (masks)
      NM_HIP(hipMemcpy(masks + out * nm::kMaskWords,
                       m->v.masks + r.first * nm::kMaskWords,
                       (size_t)n * nm::kMaskWords * sizeof(uint32_t),
                       hipMemcpyDeviceToHost));
    out += n;
  }
  return NM_OK;
}

int nm_map_device_arrays(nm_map* m, void** values, void** masks,
                         void** origins, void** hash_keys, void** hash_vals,
                         int64_t* hash_capacity) {
  if (!m) return fail(NM_EINVAL, "nm_map_device_arrays: null map");
  if (values) *values = m->v.values;
  if (masks) *masks = m->v.masks;
  if (origins) *origins = m->v.origins;
  if (hash_keys) *hash_keys = m->v.hash_keys;
  if (hash_vals) *hash_vals = m->v.hash_vals;
  if (hash_capacity) *hash_capacity = m->v.hash_cap;
  return NM_OK;
}

namespace {

// ax[u] = ((u + 0.5) - cx) / fx, ay[v] = -((v + 0.5) - cy) / fy, each
// evaluated in double and rounded once to FP32.
void depth_factors(const nm_depth_camera& c, float* ax, float* ay) {
  for (int u = 0; u < c.width; ++u)
    ax[u] = (float)((((double)u + 0.5) - (double)c.cx) / (double)c.fx);
  for (int v = 0; v < c.height; ++v)
    ay[v] = (float)(-(((double)v + 0.5) - (double)c.cy) / (double)c.fy);
}

bool depth_camera_ok(const nm_depth_camera* c) {
  return c && c->width > 0 && c->height > 0 &&
         (int64_t)c->width * c->height <= (1 << 30) && std::isfinite(c->fx) &&
         std::isfinite(c->fy) && c->fx != 0.0f && c->fy != 0.0f &&
         std::isfinite(c->cx) && std::isfinite(c->cy);
}

int cast_impl(nm_ctx* ctx, const nm_frame* frame, const void* cloud,
              int n_points, int step, int cloud_on_device,
              const nm_depth_camera* cam, void* stream);

}  // namespace

int nm_depth_camera_from_fov(int width, int height, float hfov_deg,
                             float vfov_deg, nm_depth_camera* out) {
  if (!out || width <= 0 || height <= 0 || !(hfov_deg > 0 && hfov_deg < 180) ||
      !(vfov_deg > 0 && vfov_deg < 180))
    return fail(NM_EINVAL, "nm_depth_camera_from_fov: bad argument");
  out->width = width;
  out->height = height;
  out->fx = (float)(0.5 * width / tan(hfov_deg * M_PI / 360.0));
  out->fy = (float)(0.5 * height / tan(vfov_deg * M_PI / 360.0));
  out->cx = 0.5f * (float)width;
  out->cy = 0.5f * (float)height;
  return NM_OK;
}

int nm_depth_to_cloud(const nm_depth_camera* cam, const float* depth,
                      float* xyz) {
  if (!depth_camera_ok(cam) || !depth || !xyz)
    return fail(NM_EINVAL, "nm_depth_to_cloud: bad argument");
  std::vector<float> ax(cam->width), ay(cam->height);
  depth_factors(*cam, ax.data(), ay.data());
  for (int v = 0; v < cam->height; ++v)
    for (int u = 0; u < cam->width; ++u) {
      const size_t i = (size_t)v * cam->width + u;
      const float d = depth[i];
      xyz[3 * i + 0] = d * ax[u];
      xyz[3 * i + 1] = d * ay[v];
      xyz[3 * i + 2] = d;
    }
  return NM_OK;
}

int nm_cast_depth(nm_ctx* ctx, const nm_frame* frame,
                  const nm_depth_camera* cam, const float* depth,
                  int depth_on_device, void* stream) {
  if (!ctx || !frame || !depth_camera_ok(cam))
    return fail(NM_EINVAL, "nm_cast_depth: bad argument");
  return cast_impl(ctx, frame, depth, cam->width * cam->height, 4,
                   depth_on_device, cam, stream);
}

int nm_frame_submit_depth(nm_ctx* ctx, nm_map* map, const nm_frame* frame,
                          const nm_depth_camera* cam, const float* depth,
                          int depth_on_device, void* stream) {
  int st = nm_cast_depth(ctx, frame, cam, depth, depth_on_device, stream);
  if (st != NM_OK) return st;
  return nm_update_map(ctx, map, stream);
}

int nm_cast(nm_ctx* ctx, const nm_frame* frame, const void* cloud,
            int n_points, int step, int cloud_on_device, void* stream) {
  return cast_impl(ctx, frame, cloud, n_points, step, cloud_on_device,
                   nullptr, stream);
}

}  // extern "C"

namespace {

int cast_impl(nm_ctx* ctx, const nm_frame* frame, const void* cloud,
              int n_points, int step, int cloud_on_device,
              const nm_depth_camera* cam, void* stream) {
  if (!ctx || !frame) return fail(NM_EINVAL, "nm_cast: null argument");
  if (n_points < 0 || n_points > ctx->max_points)
    return fail(NM_EINVAL, "nm_cast: n_points exceeds the context capacity");
  if (!cam && n_points > 0 && (!cloud || step < 12 || (step & 3)))
    return fail(NM_EINVAL, "nm_cast: cloud must be xyz f32 with a stride >= 12, multiple of 4");
  if (cam && n_points > 0 && !cloud)
    return fail(NM_EINVAL, "nm_cast_depth: null depth image");
  if (!(frame->grid_res > 0) || !(frame->max_time > 0) ||
      frame->max_time > ctx->max_time)
    return fail(NM_EINVAL, "nm_cast: max_time must be in (0, context max_time]");
  if (!nm::logodds_ok(frame->logodds_hit) || !nm::logodds_ok(frame->logodds_miss))
    return fail(NM_EINVAL, "nm_cast: log-odds must be finite, magnitude <= 1e27");
  if (frame->rounding != NM_ROUND_NEAREST && frame->rounding != NM_ROUND_DOWN)
    return fail(NM_EINVAL, "nm_cast: bad rounding mode");
  if (frame->filter < NM_FILTER_NONE || frame->filter > NM_FILTER_VOXEL_COUNT8)
    return fail(NM_EINVAL, "nm_cast: bad filter type");
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(frame->eye[a]) || fabs(frame->eye[a]) > 1e8)
      return fail(NM_EINVAL, "nm_cast: eye must be finite");
  nm::FilterBounds fb{};
  if (frame->filter != NM_FILTER_NONE) {
    int st = filter_bounds(*frame, &fb);
    if (st != NM_OK) return st;
  }
  if (ctx->pending)
    return fail(NM_ESTATE, "nm_cast: the previous cast was not folded into a map (nm_update_map)");
  NM_HIP(hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  const nm::Cube cube = nm::frame_cube(*frame);
  if (cube.side > ctx->cube_side_max)
    return fail(NM_EINVAL, "nm_cast: frame cube exceeds the context");

  const unsigned char* pts = (const unsigned char*)cloud;
  if (!cloud_on_device && n_points > 0) {
    const size_t bytes = (size_t)n_points * (size_t)step;
    if (bytes > ctx->staging_bytes) {
      NM_HIP(hipStreamSynchronize(s));
      NM_HIP(hipFree(ctx->b.cloud));
      ctx->b.cloud = nullptr;
      NM_HIP(hipMalloc((void**)&ctx->b.cloud, bytes));
      ctx->staging_bytes = bytes;
    }
    NM_HIP(hipMemcpyAsync(ctx->b.cloud, cloud, bytes, hipMemcpyHostToDevice, s));
    pts = ctx->b.cloud;
  }

  if (cam) {
    // the camera's factor tables, uploaded when the camera changes
    const bool same = ctx->depth_valid &&
                      memcmp(&ctx->depth_cam, cam, sizeof(*cam)) == 0;
    if (!same) {
      const int need = cam->width + cam->height;
      NM_HIP(hipStreamSynchronize(s));  // the previous upload has finished
      if (need > ctx->depth_tab_cap) {
        if (ctx->depth_tab) NM_HIP(hipFree(ctx->depth_tab));
        ctx->depth_tab = nullptr;
        ctx->depth_tab_cap = 0;
        NM_HIP(hipMalloc((void**)&ctx->depth_tab, need * sizeof(float)));
        ctx->depth_tab_cap = need;
      }
      ctx->depth_host.resize(need);
      depth_factors(*cam, ctx->depth_host.data(),
                    ctx->depth_host.data() + cam->width);
      NM_HIP(hipMemcpyAsync(ctx->depth_tab, ctx->depth_host.data(),
                            need * sizeof(float), hipMemcpyHostToDevice, s));
      ctx->depth_cam = *cam;
      ctx->depth_valid = true;
    }
  }

  nm::Job p{};
  memcpy(p.R, frame->rotation, sizeof(p.R));
  memcpy(p.eye, frame->eye, sizeof(p.eye));
  p.res = frame->grid_res;
  p.max_time = frame->max_time;
  p.hit = frame->logodds_hit;
  p.miss = frame->logodds_miss;
  p.cube = cube;
  p.cloud = pts;
  p.n = n_points;
  p.step = step;
  p.width = cam ? cam->width : frame->width;
  p.height = cam ? cam->height : frame->height;
  p.depth = cam ? 1 : 0;
  p.depth_ax = cam ? ctx->depth_tab : nullptr;
  p.depth_ay = cam ? ctx->depth_tab + cam->width : nullptr;
  p.tiled = (p.width > 0 && p.height > 0 &&
             (int64_t)p.width * p.height == n_points &&
             frame->filter == NM_FILTER_NONE)
                ? 1
                : 0;
  p.layout = nm::kLayout16x16;
  p.use_rays = frame->filter != NM_FILTER_NONE ? 1 : 0;
  p.rays4 = p.use_rays ? ctx->b.filter_rays : nullptr;
  p.n_dev = p.use_rays ? &ctx->b.counters->n_filter_rays : nullptr;
  p.map = 0;
  p.cube_off = 0;
  p.ray_off = 0;
  p.item_off = 0;
  p.nitems = nm::job_items(p, p.layout);

  nm::Pipeline& pl = ctx->pipe;
  pl = nm::Pipeline{};
  pl.host_jobs = &ctx->job;
  pl.njobs = 1;
  pl.cube_vol = (int64_t)cube.side * cube.side * cube.side;
  pl.n_rays = n_points;
  pl.nitems = p.nitems;
  {
    const int64_t room =
        (int64_t)ctx->resident * kSliceOvercommit / 4 - p.nitems;
    const int64_t extra = std::min<int64_t>(
        std::max<int64_t>(room, 0), (int64_t)p.nitems * (kAccumSlices - 1));
    pl.accum_items = p.nitems + (int)extra;
    p.slices = extra > 0 ? kAccumSlices : 1;
  }
  pl.tile_cap = ctx->tile_cap;
  pl.sort_bits = nm::sort_bits_for((int64_t)ctx->tile_cap * 512);
  pl.heavy_hits = nm::kFrameHeavyHits;
  pl.rounding = frame->rounding;
  pl.cu_count = ctx->cu_count;
  pl.side = ctx->side;
  pl.fork = ctx->fork_ev;
  pl.join = ctx->join_ev;
  ctx->job = p;

  NM_HIP(nm::launch_begin(ctx->b, &p, s));
  if (frame->filter != NM_FILTER_NONE) {
    StageTimer t(ctx, NM_STAGE_FILTER, s);
    float origin[3];
    for (int a = 0; a < 3; ++a)
      origin[a] = (float)(frame->eye[a] - floor(frame->eye[a]));
    NM_HIP(nm::launch_voxel_filter(p, origin, fb,
                                   frame->filter - NM_FILTER_VOXEL, ctx->b,
                                   s));
  }
  {
    StageTimer t(ctx, NM_STAGE_DISCOVER, s);
    NM_HIP(nm::launch_discover(pl, ctx->b, s));
  }
  {
    StageTimer t(ctx, NM_STAGE_COMPACT, s);
    NM_HIP(nm::launch_compact_leaves(pl, ctx->b, s));
  }
  int sorted = 0;
  {
    StageTimer t(ctx, NM_STAGE_HITSORT, s);
    NM_HIP(nm::launch_hit_keys(pl, ctx->b, s));
    NM_HIP(nm::radix_sort_pairs(ctx->b.keys, ctx->b.vals, n_points,
                                pl.sort_bits, ctx->b.radix_hist, &sorted, s,
                                &ctx->b.counters->n_hit_keys));
    NM_HIP(nm::launch_hit_segments(pl, ctx->b, sorted, s));
  }
  {
    StageTimer t(ctx, NM_STAGE_ACCUMULATE, s);
    NM_HIP(nm::launch_accumulate(pl, ctx->b, sorted, s));
  }
  ctx->pending = true;
  ctx->last_update_valid = false;
  ctx->frame = *frame;
  ctx->cube = cube;
  ctx->n = n_points;
  ctx->sorted_buf = sorted;
  ctx->stream = s;
  return NM_OK;
}

}  // namespace

extern "C" {

int nm_update_map(nm_ctx* ctx, nm_map* map, void* stream) {
  if (!ctx || !map) return fail(NM_EINVAL, "nm_update_map: null argument");
  if (!ctx->pending) return fail(NM_ESTATE, "nm_update_map: no pending cast");
  if (map->device != ctx->device)
    return fail(NM_EINVAL, "nm_update_map: map and context on different devices");
  NM_HIP(hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  if (map->auto_grow) {
    // wait for the cast and the map's last work, then make room for every
    // leaf the cast touched (an upper bound on the leaves it adds)
    NM_HIP(hipStreamSynchronize(ctx->stream));
    NM_HIP(hipStreamSynchronize(map->stream));
    nm::FrameCounters c;
    NM_HIP(hipMemcpy(&c, ctx->b.counters, sizeof(c), hipMemcpyDeviceToHost));
    int32_t n[2] = {0, 0};
    NM_HIP(hipMemcpy(n, map->v.leaf_count, sizeof(n), hipMemcpyDeviceToHost));
    const int64_t slots = std::min<int64_t>(n[0], map->v.leaf_cap);
    const int64_t need = slots + std::min<int64_t>(c.n_leaves, ctx->tile_cap);
    if (need > map->v.leaf_cap) {
      const int st = nm::map_reserve(
          map, std::max<int64_t>(need, 2 * (int64_t)map->v.leaf_cap));
      if (st != NM_OK) return st;
    }
  }
  {
    StageTimer t(ctx, NM_STAGE_UPDATE, s);
    // the table of miss runs is rebuilt only when the miss log-odds or the
    // rounding mode change
    const uint64_t key = ((uint64_t)nm::bits_of(ctx->job.miss) << 8) |
                         (uint64_t)(ctx->pipe.rounding & 0xff);
    ctx->pipe.miss_runs_ready = ctx->runs_valid && ctx->runs_key == key;
    ctx->runs_valid = false;
    NM_HIP(nm::launch_update(ctx->pipe, ctx->b, ctx->sorted_buf, nullptr,
                             map->v, s));
    ctx->runs_valid = true;
    ctx->runs_key = key;
  }
  ctx->pending = false;
  ctx->last_update_valid = true;
  ctx->last_map = map;
  ctx->stream = s;
  map->stream = s;
  return NM_OK;
}

int nm_frame_submit(nm_ctx* ctx, nm_map* map, const nm_frame* frame,
                    const void* cloud, int n_points, int step,
                    int cloud_on_device, void* stream) {
  int st = nm_cast(ctx, frame, cloud, n_points, step, cloud_on_device, stream);
  if (st != NM_OK) return st;
  return nm_update_map(ctx, map, stream);
}

int nm_frame_stats_get(nm_ctx* ctx, nm_frame_stats* out) {
  if (!ctx || !out) return fail(NM_EINVAL, "nm_frame_stats_get: null argument");
  NM_HIP(hipSetDevice(ctx->device));
  NM_HIP(hipMemcpyAsync(ctx->host_counters, ctx->b.counters,
                        sizeof(nm::FrameCounters), hipMemcpyDeviceToHost,
                        ctx->stream));
  NM_HIP(hipStreamSynchronize(ctx->stream));
  const nm::FrameCounters& c = *ctx->host_counters;
  memset(out, 0, sizeof(*out));
  out->points = ctx->n;
  out->rays = (int64_t)c.rays;
  out->visits = (int64_t)c.visits;
  out->leaves = c.n_leaves;
  out->errors = c.errors;
  out->overflow = c.overflow;
  if (ctx->last_update_valid) {
    NM_HIP(nm::launch_job_stats(ctx->pipe, ctx->b, ctx->stream));
    NM_HIP(hipMemcpyAsync(ctx->host_stats, ctx->b.job_stats,
                          sizeof(nm::JobStats), hipMemcpyDeviceToHost,
                          ctx->stream));
    NM_HIP(hipStreamSynchronize(ctx->stream));
    out->voxels = (int64_t)ctx->host_stats->voxels;
    out->hit_voxels = (int64_t)ctx->host_stats->hit_voxels;
    out->mixed_voxels = (int64_t)ctx->host_stats->mixed_voxels;
  }
  if (ctx->last_map) {
    int32_t n[2] = {0, 0};
    NM_HIP(hipMemcpy(n, ctx->last_map->v.leaf_count, sizeof(n),
                     hipMemcpyDeviceToHost));
    out->map_leaves = std::min<int64_t>(n[0], ctx->last_map->v.leaf_cap) - n[1];
  }
  return NM_OK;
}

int nm_counts_download(nm_ctx* ctx, int64_t max_leaves, int64_t* n_leaves,
                       int32_t* origins, uint32_t* miss, uint32_t* hit) {
  if (!ctx || !n_leaves) return fail(NM_EINVAL, "nm_counts_download: null argument");
  if (!ctx->pending)
    return fail(NM_ESTATE, "nm_counts_download: call between nm_cast and nm_update_map");
  NM_HIP(hipSetDevice(ctx->device));
  NM_HIP(hipStreamSynchronize(ctx->stream));
  nm::FrameCounters c;
  NM_HIP(hipMemcpy(&c, ctx->b.counters, sizeof(c), hipMemcpyDeviceToHost));
  const int64_t L = std::min<int64_t>(c.n_leaves, ctx->tile_cap);
  *n_leaves = L;
  if (L > max_leaves) return fail(NM_EOVERFLOW, "nm_counts_download: arrays too small");
  if (L == 0) return NM_OK;
  std::string err;
  int32_t* slot_leaf = new int32_t[L];
  uint32_t* tiles = new uint32_t[(size_t)L * 1024];
  uint32_t* starts = new uint32_t[(size_t)L * 512];
  hipError_t e = hipMemcpy(slot_leaf, ctx->b.slot_leaf, L * sizeof(int32_t),
                           hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(tiles, ctx->b.tiles, (size_t)L * 1024 * sizeof(uint32_t),
                  hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(starts, ctx->b.hit_start, (size_t)L * 512 * sizeof(uint32_t),
                  hipMemcpyDeviceToHost);
  if (e == hipSuccess) {
    const int side = ctx->cube.side;
    for (int64_t s = 0; s < L; ++s) {
      const int cidx = slot_leaf[s];
      const int lz = cidx % side, ly = (cidx / side) % side,
                lx = cidx / (side * side);
      if (origins) {
        origins[3 * s + 0] = (lx + ctx->cube.cmin[0]) * 8;
        origins[3 * s + 1] = (ly + ctx->cube.cmin[1]) * 8;
        origins[3 * s + 2] = (lz + ctx->cube.cmin[2]) * 8;
      }
      for (int v = 0; v < 512; ++v) {
        if (miss) miss[s * 512 + v] = tiles[(s * 512 + v) * 2];
        // the hit word holds the end of the voxel's run of sorted hits
        const uint32_t end = tiles[(s * 512 + v) * 2 + 1];
        if (hit) hit[s * 512 + v] = end ? end - starts[s * 512 + v] : 0u;
      }
    }
  }
  delete[] slot_leaf;
  delete[] tiles;
  delete[] starts;
  if (e != hipSuccess) return hip_fail(e, "nm_counts_download");
  return NM_OK;
}

int nm_filter_rays_download(nm_ctx* ctx, int max_rays, int* n_rays,
                            float* rays4) {
  if (!ctx || !n_rays) return fail(NM_EINVAL, "nm_filter_rays_download: null argument");
  NM_HIP(hipSetDevice(ctx->device));
  NM_HIP(hipStreamSynchronize(ctx->stream));
  nm::FrameCounters c;
  NM_HIP(hipMemcpy(&c, ctx->b.counters, sizeof(c), hipMemcpyDeviceToHost));
  *n_rays = c.n_filter_rays;
  if (c.n_filter_rays > max_rays)
    return fail(NM_EOVERFLOW, "nm_filter_rays_download: array too small");
  if (c.n_filter_rays > 0 && rays4)
    NM_HIP(hipMemcpy(rays4, ctx->b.filter_rays,
                     (size_t)c.n_filter_rays * 4 * sizeof(float),
                     hipMemcpyDeviceToHost));
  return NM_OK;
}

int nm_ctx_set_timing(nm_ctx* ctx, int enable) {
  if (!ctx) return fail(NM_EINVAL, "nm_ctx_set_timing: null context");
  ctx->timing = enable != 0;
  ctx->timing_mask = ~0u;
  return NM_OK;
}

int nm_ctx_set_timing_mask(nm_ctx* ctx, unsigned mask) {
  if (!ctx) return fail(NM_EINVAL, "nm_ctx_set_timing_mask: null context");
  ctx->timing = mask != 0;
  ctx->timing_mask = mask;
  return NM_OK;
}

int nm_ctx_stage_times(nm_ctx* ctx, float* ms, int64_t* launches) {
  if (!ctx) return fail(NM_EINVAL, "nm_ctx_stage_times: null context");
  NM_HIP(hipSetDevice(ctx->device));
  int st = drain_marks(ctx);
  if (st != NM_OK) return st;
  for (int i = 0; i < NM_STAGE_COUNT; ++i) {
    if (ms) ms[i] = (float)ctx->stage_ms[i];
    if (launches) launches[i] = ctx->stage_n[i];
  }
  return NM_OK;
}

int nm_ctx_reset_timing(nm_ctx* ctx) {
  if (!ctx) return fail(NM_EINVAL, "nm_ctx_reset_timing: null context");
  NM_HIP(hipSetDevice(ctx->device));
  int st = drain_marks(ctx);
  if (st != NM_OK) return st;
  for (int i = 0; i < NM_STAGE_COUNT; ++i) {
    ctx->stage_ms[i] = 0;
    ctx->stage_n[i] = 0;
  }
  return NM_OK;
}

}  // extern "C"
