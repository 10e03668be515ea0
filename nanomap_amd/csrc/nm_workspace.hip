// nm_workspace.hip -- allocation of the device workspace shared by the
// plain frame path (one job) and the batched path (many jobs).
#include "nm_internal.h"

namespace nm {

namespace {

template <class T>
hipError_t dalloc(T** p, size_t count) {
  return hipMalloc((void**)p, count * sizeof(T) + 64);
}

}  // namespace

hipError_t workspace_alloc(const WorkspaceDims& d, Workspace* out) {
  Workspace b{};
  const size_t P = (size_t)(d.ray_cap > 0 ? d.ray_cap : 1);
  const size_t T = (size_t)(d.tile_cap > 0 ? d.tile_cap : 1);
  const size_t C = (size_t)(d.cube_cap > 0 ? d.cube_cap : 1);
  const size_t I = (size_t)(d.item_cap > 0 ? d.item_cap : 1);
  const size_t J = (size_t)(d.max_jobs > 0 ? d.max_jobs : 1);
  const size_t scan_blocks = (C > P ? C : P) / kScanChunk + 2;
  hipError_t e = hipSuccess;
  auto A = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
  };
  A(dalloc(&b.leaf_slot, C));
  A(dalloc(&b.slot_leaf, T));
  A(dalloc(&b.slot_job, T));
  A(dalloc(&b.map_idx, T));
  A(dalloc(&b.map_new, T));
  A(dalloc(&b.slot_stats, T * 4));
  A(dalloc(&b.tiles, T * 1024));
  A(dalloc(&b.hit_mask, T * kMaskWords));
  A(dalloc(&b.hit_start, T * 512));
  A(dalloc(&b.ray_hit, P));
  A(dalloc(&b.keys[0], P));
  A(dalloc(&b.keys[1], P));
  A(dalloc(&b.vals[0], P));
  A(dalloc(&b.vals[1], P));
  A(dalloc(&b.gap, P));
  A(dalloc(&b.radix_hist, (size_t)radix_hist_words((int)P)));
  A(dalloc(&b.scan_blocks, scan_blocks));
  A(dalloc(&b.scan_job, scan_blocks));
  A(dalloc(&b.counters, 1));
  A(dalloc(&b.job_stats, J));
  A(dalloc(&b.jobs, J));
  A(dalloc(&b.maps, J));
  A(dalloc(&b.part_discover, I * 4));
  A(dalloc(&b.item_job, I));
  A(dalloc(&b.item_order, I));
  A(dalloc(&b.item_longest, I));
  A(dalloc(&b.slice_items, I * (kMaxSlices - 1)));
  A(dalloc(&b.cost_keys[0], I));
  A(dalloc(&b.cost_keys[1], I));
  A(dalloc(&b.cost_vals[0], I));
  A(dalloc(&b.cost_vals[1], I));
  A(dalloc(&b.miss_runs, J * kMissRuns));
  A(dalloc(&b.runs_key, J));
  if (d.host_clouds) {
    A(dalloc(&b.cloud, P * 16));
    A(dalloc(&b.filter_rays, P * 4));
    A(dalloc(&b.filter_tmp, P * 4));
    A(dalloc(&b.flags, P));
  }
  if (e == hipSuccess) e = hipMemset(b.leaf_slot, 0xff, C * sizeof(int32_t));
  if (e == hipSuccess) e = hipMemset(b.tiles, 0, T * 1024 * sizeof(uint32_t));
  if (e == hipSuccess)
    e = hipMemset(b.hit_mask, 0, T * kMaskWords * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemset(b.counters, 0, sizeof(FrameCounters));
  if (e == hipSuccess) e = hipMemset(b.runs_key, 0xff, J * sizeof(uint64_t));
  if (e == hipSuccess)
    e = hipMemset(b.scan_job, 0, scan_blocks * sizeof(int32_t));
  if (e == hipSuccess) e = hipMemset(b.item_job, 0, I * sizeof(int32_t));
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    workspace_free(&b);
    return e;
  }
  *out = b;
  return hipSuccess;
}

void workspace_free(Workspace* b) {
  void* ptrs[] = {b->leaf_slot,  b->slot_leaf,   b->slot_job,   b->map_idx,
                  b->map_new,    b->slot_stats,  b->tiles,      b->hit_mask,
                  b->hit_start,  b->ray_hit,     b->keys[0],    b->keys[1],
                  b->vals[0],    b->vals[1],     b->gap,
                  b->radix_hist,
                  b->scan_blocks, b->scan_job,   b->counters,   b->job_stats,
                  b->jobs,       b->maps,        b->part_discover,
                  b->item_job,   b->item_order,  b->item_longest,
                  b->slice_items,
                  b->cost_keys[0],
                  b->cost_keys[1], b->cost_vals[0], b->cost_vals[1],
                  b->cloud,      b->filter_rays,
                  b->filter_tmp, b->flags,       b->miss_runs,
                  b->runs_key};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  *b = Workspace{};
}

}  // namespace nm
