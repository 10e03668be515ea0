// nm_batch.hip -- the batched frame path (include/nanomap_amd_batch.h).
//
// A batch is one job table over a shared workspace: job i owns a
// kScanChunk-aligned region of the cube tables, points [ray_off, ray_off +
// n) of the ray arrays and cast work items [item_off, item_off + nitems).
// The launch sequence is the plain frame's (nm_cast.hip) with the table in
// place of the single job, so each frame's map update is exactly the plain
// one.  The sequence is captured once per shape as a HIP graph whose first
// nodes copy the frame table and the map table from pinned host staging;
// a new batch of the same shape rewrites the staging and replays the graph.
// Two staging buffers (and two graphs) let one batch be prepared while the
// previous one runs.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nanomap_amd_batch.h"
#include "nm_fp.h"
#include "nm_internal.h"

namespace {

int bfail(int code, const std::string& msg) {
  nm::set_error(msg);
  return code;
}

int bhip(hipError_t e, const char* what) {
  return bfail(NM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define BH(call)                                 \
  do {                                           \
    hipError_t e_ = (call);                      \
    if (e_ != hipSuccess) return bhip(e_, #call); \
  } while (0)

// What fixes the layout of a frame in the workspace.
struct ShapeJob {
  int n, width, height, side;
  bool operator==(const ShapeJob& o) const {
    return n == o.n && width == o.width && height == o.height &&
           side == o.side;
  }
};

struct Staging {
  nm::Job* jobs = nullptr;      // pinned
  nm::MapView* maps = nullptr;  // pinned
  hipEvent_t done = nullptr;
  bool in_flight = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

}  // namespace

struct nm_batch {
  int device = 0;
  int max_frames = 0;
  int64_t max_points = 0;
  float max_time = 0;
  int cu_count = 256;
  nm::WorkspaceDims dims;
  nm::Workspace ws{};
  nm::Tuning tuning;
  bool use_graph = true;
  // layout of the current shape
  bool layout_valid = false;
  std::vector<ShapeJob> shape;
  int rounding = -1;
  nm::Tuning layout_tuning;
  std::vector<nm::Job> jobs;  // host table: layout fields
  nm::Pipeline pipe{};
  int sorted_buf = 0;
  Staging st[2];
  int next = 0;
  hipStream_t capture = nullptr;  // graphs are captured on this stream
  // branch of the update for the heavy voxels (see launch_update)
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev = nullptr;
  hipEvent_t early_ev = nullptr;
  hipEvent_t prep_ev = nullptr;
  // last batch
  int last_n = 0;
  std::vector<nm_map*> last_maps;
  hipStream_t last_stream = nullptr;
  nm::FrameCounters* host_counters = nullptr;  // pinned
  nm::JobStats* host_stats = nullptr;          // pinned [max_frames]
};

namespace {

void drop_graphs(nm_batch* b) {
  for (Staging& s : b->st) {
    if (s.exec) (void)hipGraphExecDestroy(s.exec);
    if (s.graph) (void)hipGraphDestroy(s.graph);
    s.exec = nullptr;
    s.graph = nullptr;
  }
}

int wait_idle(nm_batch* b) {
  for (Staging& s : b->st) {
    if (s.in_flight) {
      BH(hipEventSynchronize(s.done));
      s.in_flight = false;
    }
  }
  return NM_OK;
}

void free_batch(nm_batch* b) {
  drop_graphs(b);
  for (Staging& s : b->st) {
    if (s.jobs) (void)hipHostFree(s.jobs);
    if (s.maps) (void)hipHostFree(s.maps);
    if (s.done) (void)hipEventDestroy(s.done);
  }
  if (b->host_counters) (void)hipHostFree(b->host_counters);
  if (b->host_stats) (void)hipHostFree(b->host_stats);
  if (b->capture) (void)hipStreamDestroy(b->capture);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  if (b->join_ev) (void)hipEventDestroy(b->join_ev);
  if (b->early_ev) (void)hipEventDestroy(b->early_ev);
  if (b->prep_ev) (void)hipEventDestroy(b->prep_ev);
  nm::workspace_free(&b->ws);
  delete b;
}

int check_frames(nm_batch* b, const nm_agent_frame* f, int n) {
  if (!f || n <= 0) return bfail(NM_EINVAL, "nm_batch: no frames");
  if (n > b->max_frames)
    return bfail(NM_EINVAL, "nm_batch: more frames than max_frames");
  int64_t total = 0;
  for (int i = 0; i < n; ++i) {
    const nm_frame& fr = f[i].frame;
    if (f[i].n_points < 0 || (f[i].n_points > 0 && !f[i].cloud))
      return bfail(NM_EINVAL, "nm_batch: bad cloud");
    if (f[i].n_points > 0 && (f[i].step < 12 || (f[i].step & 3)))
      return bfail(NM_EINVAL, "nm_batch: cloud stride must be >= 12, multiple of 4");
    if (!(fr.grid_res > 0) || !(fr.max_time > 0) || fr.max_time > b->max_time)
      return bfail(NM_EINVAL, "nm_batch: max_time must be in (0, batch max_time]");
    if (fr.rounding != f[0].frame.rounding ||
        (fr.rounding != NM_ROUND_NEAREST && fr.rounding != NM_ROUND_DOWN))
      return bfail(NM_EINVAL, "nm_batch: frames need one valid rounding mode");
    if (fr.filter != NM_FILTER_NONE)
      return bfail(NM_EINVAL, "nm_batch: frames must use NM_FILTER_NONE");
    if (!nm::logodds_ok(fr.logodds_hit) || !nm::logodds_ok(fr.logodds_miss))
      return bfail(NM_EINVAL, "nm_batch: log-odds must be finite, magnitude <= 1e27");
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(fr.eye[a]) || fabs(fr.eye[a]) > 1e8)
        return bfail(NM_EINVAL, "nm_batch: eye must be finite");
    total += f[i].n_points;
  }
  if (total > b->max_points)
    return bfail(NM_EINVAL, "nm_batch: more points than max_points");
  return NM_OK;
}

int check_maps(nm_batch* b, const nm_agent_frame* f, int n) {
  std::vector<const nm_map*> seen;
  for (int i = 0; i < n; ++i) {
    if (!f[i].map) return bfail(NM_EINVAL, "nm_batch_submit: null map");
    if (f[i].map->device != b->device)
      return bfail(NM_EINVAL, "nm_batch_submit: map on another device");
    seen.push_back(f[i].map);
  }
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
    return bfail(NM_EINVAL, "nm_batch_submit: a map appears twice in the batch");
  return NM_OK;
}

std::vector<ShapeJob> shape_of(const nm_agent_frame* f, int n) {
  std::vector<ShapeJob> s(n);
  for (int i = 0; i < n; ++i)
    s[i] = {f[i].n_points, f[i].frame.width, f[i].frame.height,
            nm::frame_cube(f[i].frame).side};
  return s;
}

// Places the frames in the workspace (growing it if needed), uploads the
// job-of-chunk, job-of-item and queue-order tables.  Synchronizes.
int build_layout(nm_batch* b, const nm_agent_frame* f, int n) {
  int st = wait_idle(b);
  if (st != NM_OK) return st;
  drop_graphs(b);
  b->layout_valid = false;
  std::vector<nm::Job> J(n);
  int64_t cube_off = 0, ray_off = 0;
  int64_t item_off = 0;
  for (int i = 0; i < n; ++i) {
    nm::Job& j = J[i];
    j = nm::Job{};
    const nm::Cube cube = nm::frame_cube(f[i].frame);
    j.cube = cube;
    j.n = f[i].n_points;
    j.width = f[i].frame.width;
    j.height = f[i].frame.height;
    j.tiled = (j.width > 0 && j.height > 0 &&
               (int64_t)j.width * j.height == j.n)
                  ? 1
                  : 0;
    j.layout = b->tuning.layout;
    j.map = i;
    j.cube_off = cube_off;
    cube_off += round_up((int64_t)cube.side * cube.side * cube.side,
                         nm::kScanChunk);
    j.ray_off = (int)ray_off;
    ray_off += j.n;
    j.item_off = (int)item_off;
    j.nitems = nm::job_items(j, j.layout);
    item_off += j.nitems;
  }
  if (item_off > (1 << 30))
    return bfail(NM_EINVAL, "nm_batch: too many work items");
  // grow the workspace if this shape needs more cube entries or items
  if (cube_off > b->dims.cube_cap || item_off > b->dims.item_cap) {
    nm::WorkspaceDims d = b->dims;
    d.cube_cap = std::max<int64_t>(d.cube_cap, cube_off);
    d.item_cap = std::max<int>(d.item_cap, (int)(item_off + item_off / 4));
    BH(hipDeviceSynchronize());
    nm::workspace_free(&b->ws);
    hipError_t e = nm::workspace_alloc(d, &b->ws);
    if (e != hipSuccess)
      return bfail(e == hipErrorOutOfMemory ? NM_ENOMEM : NM_EHIP,
                   std::string("nm_batch: workspace: ") + hipGetErrorString(e));
    b->dims = d;
  }
  // job of every scan chunk and of every work item; queue order by
  // decreasing range (longest rays first), frame order within a range
  std::vector<int32_t> scan_job((size_t)(cube_off / nm::kScanChunk) + 1, 0);
  std::vector<int32_t> item_job((size_t)item_off + 1, 0);
  std::vector<int32_t> order((size_t)item_off + 1, 0);
  std::vector<int> by_range(n);
  for (int i = 0; i < n; ++i) {
    by_range[i] = i;
    for (int64_t c = J[i].cube_off / nm::kScanChunk;
         c < (i + 1 < n ? J[i + 1].cube_off : cube_off) / nm::kScanChunk; ++c)
      scan_job[c] = i;
    for (int t = 0; t < J[i].nitems; ++t) item_job[J[i].item_off + t] = i;
  }
  std::stable_sort(by_range.begin(), by_range.end(), [&](int x, int y) {
    return f[x].frame.max_time > f[y].frame.max_time;
  });
  int64_t k = 0;
  for (int i : by_range)
    for (int t = 0; t < J[i].nitems; ++t) order[k++] = J[i].item_off + t;
  BH(hipSetDevice(b->device));
  BH(hipMemcpy(b->ws.scan_job, scan_job.data(), scan_job.size() * 4,
               hipMemcpyHostToDevice));
  BH(hipMemcpy(b->ws.item_job, item_job.data(), item_job.size() * 4,
               hipMemcpyHostToDevice));
  BH(hipMemcpy(b->ws.item_order, order.data(), order.size() * 4,
               hipMemcpyHostToDevice));
  b->jobs = J;
  b->shape = shape_of(f, n);
  b->rounding = f[0].frame.rounding;
  b->layout_tuning = b->tuning;
  nm::Pipeline& p = b->pipe;
  p = nm::Pipeline{};
  p.host_jobs = b->jobs.data();
  p.njobs = n;
  p.cube_vol = cube_off;
  p.n_rays = ray_off;
  p.nitems = (int)item_off;
  p.tile_cap = b->dims.tile_cap;
  p.sort_bits = nm::sort_bits_for((int64_t)b->dims.tile_cap * 512);
  p.heavy_hits = nm::kBatchHeavyHits;
  p.rounding = b->rounding;
  p.cu_count = b->cu_count;
  p.tuning = b->tuning;
  p.side = b->side;
  p.fork = b->fork_ev;
  p.join = b->join_ev;
  p.early_fork = b->early_ev;
  p.prepared = b->prep_ev;
  b->layout_valid = true;
  return NM_OK;
}

bool layout_matches(nm_batch* b, const nm_agent_frame* f, int n) {
  if (!b->layout_valid || (int)b->shape.size() != n ||
      b->rounding != f[0].frame.rounding)
    return false;
  const nm::Tuning& t = b->layout_tuning;
  if (t.layout != b->tuning.layout || t.queue != b->tuning.queue ||
      t.grid_mult != b->tuning.grid_mult || t.lds_slots != b->tuning.lds_slots)
    return false;
  return shape_of(f, n) == b->shape;
}

// Writes the frame table (and the map table) of a batch into staging s.
void fill_staging(nm_batch* b, Staging& s, const nm_agent_frame* f, int n,
                  bool with_maps) {
  for (int i = 0; i < n; ++i) {
    nm::Job j = b->jobs[i];
    const nm_frame& fr = f[i].frame;
    memcpy(j.R, fr.rotation, sizeof(j.R));
    memcpy(j.eye, fr.eye, sizeof(j.eye));
    j.res = fr.grid_res;
    j.max_time = fr.max_time;
    j.hit = fr.logodds_hit;
    j.miss = fr.logodds_miss;
    j.cube = nm::frame_cube(fr);
    j.cloud = (const unsigned char*)f[i].cloud;
    j.step = f[i].step;
    s.jobs[i] = j;
    if (with_maps) s.maps[i] = f[i].map->v;
  }
}

// The launch sequence of one batch from staging s.  dry: cast only, then
// clear the per-frame state without touching any map.
int enqueue(nm_batch* b, Staging& s, hipStream_t q, bool dry) {
  nm::Pipeline& p = b->pipe;
  nm::Workspace& w = b->ws;
  BH(hipMemcpyAsync(w.jobs, s.jobs, sizeof(nm::Job) * p.njobs,
                    hipMemcpyHostToDevice, q));
  if (!dry)
    BH(hipMemcpyAsync(w.maps, s.maps, sizeof(nm::MapView) * p.njobs,
                      hipMemcpyHostToDevice, q));
  BH(nm::launch_begin(w, nullptr, q));
  BH(nm::launch_discover(p, w, q));
  BH(nm::launch_item_order(p, w, q));
  BH(nm::launch_compact_leaves(p, w, q));
  p.map_prepared = false;
  if (!dry) BH(nm::launch_map_prepare(p, w, w.maps, nm::MapView{}, q));
  BH(nm::launch_hit_keys(p, w, q));
  int sorted = 0;
  BH(nm::radix_sort_pairs(w.keys, w.vals, (int)p.n_rays, p.sort_bits,
                          w.radix_hist, &sorted, q, &w.counters->n_hit_keys));
  BH(nm::launch_hit_segments(p, w, sorted, q));
  BH(nm::launch_accumulate(p, w, sorted, q));
  if (dry)
    BH(nm::launch_clear_slots(p, w, q));
  else
    BH(nm::launch_update(p, w, sorted, w.maps, nm::MapView{}, q));
  b->sorted_buf = sorted;
  return NM_OK;
}

int capture(nm_batch* b, Staging& s) {
  BH(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
  const int st = enqueue(b, s, b->capture, false);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->capture, &g);
  if (st != NM_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return bhip(e, "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    return bhip(ei, "hipGraphInstantiate");
  }
  s.graph = g;
  s.exec = x;
  return NM_OK;
}

}  // namespace

extern "C" {

int nm_batch_create(const nm_batch_config* cfg, nm_batch** out) {
  if (!cfg || !out) return bfail(NM_EINVAL, "nm_batch_create: null argument");
  *out = nullptr;
  if (cfg->max_frames <= 0 || cfg->max_points <= 0 ||
      cfg->max_points > ((int64_t)1 << 30) - 1 || !(cfg->max_time > 0) ||
      cfg->max_time > 4000 || cfg->tile_capacity <= 0 ||
      cfg->tile_capacity > (1 << 22))
    return bfail(NM_EINVAL, "nm_batch_create: bad configuration");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return bfail(NM_ENODEV, "nm_batch_create: no HIP device");
  if (cfg->device < 0 || cfg->device >= ndev)
    return bfail(NM_EINVAL, "nm_batch_create: bad device ordinal");
  BH(hipSetDevice(cfg->device));
  nm_batch* b = new nm_batch();
  b->device = cfg->device;
  b->max_frames = cfg->max_frames;
  b->max_points = cfg->max_points;
  b->max_time = cfg->max_time;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, cfg->device) == hipSuccess &&
      prop.multiProcessorCount > 0)
    b->cu_count = prop.multiProcessorCount;
  nm_frame probe{};
  probe.max_time = cfg->max_time;
  const int side = nm::frame_cube(probe).side;
  nm::WorkspaceDims& d = b->dims;
  d.cube_cap = (int64_t)cfg->max_frames *
               round_up((int64_t)side * side * side, nm::kScanChunk);
  d.tile_cap = (int)cfg->tile_capacity;
  d.ray_cap = cfg->max_points;
  d.item_cap = (int)(cfg->max_points / 64 + 64 * (int64_t)cfg->max_frames);
  d.max_jobs = cfg->max_frames;
  d.host_clouds = false;
  hipError_t e = nm::workspace_alloc(d, &b->ws);
  for (Staging& s : b->st) {
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&s.jobs, sizeof(nm::Job) * cfg->max_frames);
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&s.maps, sizeof(nm::MapView) * cfg->max_frames);
    if (e == hipSuccess)
      e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
  }
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&b->host_counters, sizeof(nm::FrameCounters));
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&b->host_stats,
                      sizeof(nm::JobStats) * cfg->max_frames);
  if (e == hipSuccess)
    e = hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking);
  if (e == hipSuccess)
    e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&b->join_ev, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&b->early_ev, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&b->prep_ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    free_batch(b);
    return bfail(e == hipErrorOutOfMemory ? NM_ENOMEM : NM_EHIP,
                 std::string("nm_batch_create: ") + hipGetErrorString(e));
  }
  *out = b;
  return NM_OK;
}

int nm_batch_destroy(nm_batch* b) {
  if (!b) return NM_OK;
  (void)hipSetDevice(b->device);
  (void)hipDeviceSynchronize();
  free_batch(b);
  return NM_OK;
}

int nm_batch_submit(nm_batch* b, const nm_agent_frame* f, int n,
                    void* stream) {
  if (!b) return bfail(NM_EINVAL, "nm_batch_submit: null batch");
  int st = check_frames(b, f, n);
  if (st == NM_OK) st = check_maps(b, f, n);
  if (st != NM_OK) return st;
  BH(hipSetDevice(b->device));
  hipStream_t q = (hipStream_t)stream;
  if (!layout_matches(b, f, n)) {
    st = build_layout(b, f, n);
    if (st != NM_OK) return st;
  }
  Staging& s = b->st[b->next];
  if (s.in_flight) {
    BH(hipEventSynchronize(s.done));
    s.in_flight = false;
  }
  fill_staging(b, s, f, n, true);
  if (b->use_graph) {
    if (!s.exec) {
      st = capture(b, s);
      if (st != NM_OK) return st;
    }
    BH(hipGraphLaunch(s.exec, q));
  } else {
    st = enqueue(b, s, q, false);
    if (st != NM_OK) return st;
  }
  BH(hipEventRecord(s.done, q));
  s.in_flight = true;
  b->next ^= 1;
  b->last_n = n;
  b->last_maps.assign(n, nullptr);
  for (int i = 0; i < n; ++i) {
    b->last_maps[i] = f[i].map;
    f[i].map->stream = q;
  }
  b->last_stream = q;
  return NM_OK;
}

int nm_batch_stats(nm_batch* b, nm_frame_stats* out, int n) {
  if (!b || !out) return bfail(NM_EINVAL, "nm_batch_stats: null argument");
  if (b->last_n == 0 || n != b->last_n)
    return bfail(NM_ESTATE, "nm_batch_stats: n must match the last batch");
  BH(hipSetDevice(b->device));
  hipStream_t q = b->last_stream;
  BH(nm::launch_job_stats(b->pipe, b->ws, q));
  BH(hipMemcpyAsync(b->host_stats, b->ws.job_stats, sizeof(nm::JobStats) * n,
                    hipMemcpyDeviceToHost, q));
  BH(hipMemcpyAsync(b->host_counters, b->ws.counters,
                    sizeof(nm::FrameCounters), hipMemcpyDeviceToHost, q));
  BH(hipStreamSynchronize(q));
  for (int i = 0; i < n; ++i) {
    const nm::JobStats& js = b->host_stats[i];
    nm_frame_stats& o = out[i];
    memset(&o, 0, sizeof(o));
    o.points = b->jobs[i].n;
    o.rays = (int64_t)js.rays;
    o.visits = (int64_t)js.visits;
    o.voxels = (int64_t)js.voxels;
    o.leaves = js.slot_end - js.slot_begin;
    o.hit_voxels = (int64_t)js.hit_voxels;
    o.mixed_voxels = (int64_t)js.mixed_voxels;
    o.errors = (int32_t)js.errors;
    o.overflow = b->host_counters->overflow;
    int32_t c[2] = {0, 0};
    BH(hipMemcpy(c, b->last_maps[i]->v.leaf_count, sizeof(c),
                 hipMemcpyDeviceToHost));
    o.map_leaves = std::min<int64_t>(c[0], b->last_maps[i]->v.leaf_cap) - c[1];
  }
  return NM_OK;
}

int nm_batch_tune(nm_batch* b, const nm_agent_frame* f, int n, int reps,
                  void* stream, nm_batch_tuning* best, float* best_ms) {
  if (!b) return bfail(NM_EINVAL, "nm_batch_tune: null batch");
  int st = check_frames(b, f, n);
  if (st != NM_OK) return st;
  if (reps <= 0) reps = 1;
  BH(hipSetDevice(b->device));
  hipStream_t q = (hipStream_t)stream;
  std::vector<nm::Tuning> cand;
  for (int lds : {4096, 2048})
    for (int layout = 0; layout < nm::kLayoutCount; ++layout) {
      nm::Tuning t;
      t.layout = layout;
      t.lds_slots = lds;
      t.queue = 0;
      cand.push_back(t);
      for (int queue = 1; queue <= 3; ++queue)
        for (int gm : {2, 4, 6, 8}) {
          t.queue = queue;
          t.grid_mult = gm;
          cand.push_back(t);
        }
    }
  const nm::Tuning saved = b->tuning;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  BH(hipEventCreate(&e0));
  BH(hipEventCreate(&e1));
  // ms per batch of tuning t over `runs` batches after one warm-up, or < 0
  auto time_one = [&](const nm::Tuning& t, int runs) -> float {
    b->tuning = t;
    st = build_layout(b, f, n);
    if (st != NM_OK) return -1.0f;
    Staging& s = b->st[0];
    fill_staging(b, s, f, n, false);
    st = enqueue(b, s, q, true);  // warm-up
    if (st != NM_OK) return -1.0f;
    if (hipEventRecord(e0, q) != hipSuccess) {
      st = bfail(NM_EHIP, "nm_batch_tune: timing failed");
      return -1.0f;
    }
    for (int r = 0; r < runs && st == NM_OK; ++r) st = enqueue(b, s, q, true);
    if (st != NM_OK) return -1.0f;
    float ms = 0;
    if (hipEventRecord(e1, q) != hipSuccess ||
        hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      st = bfail(NM_EHIP, "nm_batch_tune: timing failed");
      return -1.0f;
    }
    return ms / runs;
  };
  // sweep every candidate, then re-time the four fastest in two
  // interleaved rounds of longer runs, so that run-to-run noise of the
  // short sweep does not decide between near-equal candidates
  std::vector<std::pair<float, int>> swept;
  for (size_t k = 0; k < cand.size() && st == NM_OK; ++k) {
    const float ms = time_one(cand[k], reps);
    if (st == NM_OK) swept.push_back({ms, (int)k});
  }
  float best_t = -1.0f;
  nm::Tuning best_c = saved;
  if (st == NM_OK && !swept.empty()) {
    std::sort(swept.begin(), swept.end());
    const size_t top = std::min<size_t>(4, swept.size());
    std::vector<float> total(top, 0.0f);
    const int runs = std::max(2 * reps, 4);
    for (int round = 0; round < 2 && st == NM_OK; ++round)
      for (size_t k = 0; k < top && st == NM_OK; ++k)
        total[k] += time_one(cand[swept[k].second], runs);
    if (st == NM_OK)
      for (size_t k = 0; k < top; ++k)
        if (best_t < 0 || total[k] / 2 < best_t) {
          best_t = total[k] / 2;
          best_c = cand[swept[k].second];
        }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->tuning = st == NM_OK ? best_c : saved;
  b->layout_valid = false;
  drop_graphs(b);
  if (st != NM_OK) return st;
  if (best) {
    best->layout = best_c.layout;
    best->queue = best_c.queue;
    best->grid_mult = best_c.grid_mult;
    best->lds_slots = best_c.lds_slots;
  }
  if (best_ms) *best_ms = best_t;
  return NM_OK;
}

int nm_batch_set_tuning(nm_batch* b, const nm_batch_tuning* t) {
  if (!b || !t) return bfail(NM_EINVAL, "nm_batch_set_tuning: null argument");
  if (t->layout < 0 || t->layout >= nm::kLayoutCount || t->queue < 0 ||
      t->queue > 3 || t->grid_mult < 1 || t->grid_mult > 64 ||
      (t->lds_slots != 4096 && t->lds_slots != 2048))
    return bfail(NM_EINVAL, "nm_batch_set_tuning: bad tuning");
  b->tuning.layout = t->layout;
  b->tuning.queue = t->queue;
  b->tuning.grid_mult = t->grid_mult;
  b->tuning.lds_slots = t->lds_slots;
  return NM_OK;
}

int nm_batch_get_tuning(nm_batch* b, nm_batch_tuning* t) {
  if (!b || !t) return bfail(NM_EINVAL, "nm_batch_get_tuning: null argument");
  t->layout = b->tuning.layout;
  t->queue = b->tuning.queue;
  t->grid_mult = b->tuning.grid_mult;
  t->lds_slots = b->tuning.lds_slots;
  return NM_OK;
}

int nm_batch_set_graph(nm_batch* b, int enable) {
  if (!b) return bfail(NM_EINVAL, "nm_batch_set_graph: null batch");
  b->use_graph = enable != 0;
  return NM_OK;
}

}  // extern "C"
