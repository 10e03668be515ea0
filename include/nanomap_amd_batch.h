/* nanomap_amd_batch.h -- many sensors' frames in one launch sequence.
 *
 * The reference processes agents one after another: SimManager::
 * updateAgentViews (include/nanomap/manager/SimManager.h:100-113) loops over
 * the agents and their sensors, and each frame goes through
 * Handler::processPointCloud (src/nanomap/handler/Handler.cpp:250-278): one
 * filterCloud + frustumCastCloud launch sequence with host synchronizations
 * and a host fold into the agent's map per frame.  Here a batch of frames --
 * any mix of image sizes, ranges, voxel sizes and poses, each frame with its
 * own map -- is one sequence of launches over a device table of frames,
 * captured once as a HIP graph and replayed for every later batch of the
 * same shape.  Each map is bit-identical to processing its frame alone with
 * nm_frame_submit (nanomap_amd.h).
 */
#ifndef NANOMAP_AMD_BATCH_H
#define NANOMAP_AMD_BATCH_H

#include "nanomap_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nm_batch nm_batch;

typedef struct nm_batch_config {
  int device;
  int max_frames;         /* frames per batch                               */
  int64_t max_points;     /* points of all frames of a batch, < 2^30        */
  float max_time;         /* largest per-frame max_time (voxels)            */
  int64_t tile_capacity;  /* leaf tiles of all frames of a batch, <= 2^22   */
} nm_batch_config;

/* One frame of one agent.  The cloud is device-resident: xyz f32 at byte
 * offsets 0/4/8 with stride `step`, in the sensor frame (as nm_cast). */
typedef struct nm_agent_frame {
  nm_frame frame;
  const void* cloud;
  int n_points;
  int step;
  nm_map* map;
} nm_agent_frame;

/* How the cast kernels walk their work (256-ray pixel tiles). */
typedef struct nm_batch_tuning {
  int layout;    /* pixel tile: 0 = 16x16, 1 = 32x8, 2 = 64x4, 3 = 8x32     */
  int queue;     /* 0: one workgroup per tile; 1: persistent workgroups      */
                 /* draining a device queue in frame order; 2: the same,     */
                 /* longest-range frames first; 3: as 2, and the voxel pass  */
                 /* takes the tiles in order of their longest ray (measured  */
                 /* by the leaf pass), longest first                         */
  int grid_mult; /* persistent workgroups per compute unit (queue 1-3)       */
  int lds_slots; /* LDS aggregation table of the accumulate pass: 4096 or   */
                 /* 2048 entries (fewer entries, more resident workgroups)  */
} nm_batch_tuning;

int nm_batch_create(const nm_batch_config* cfg, nm_batch** out);
int nm_batch_destroy(nm_batch* batch);

/* Casts every frame and folds it into its map, enqueued on `stream` as one
 * graph launch (or plain launches, see nm_batch_set_graph).  All frames use
 * one rounding mode and NM_FILTER_NONE; a map appears at most once per
 * batch.  The first batch of a new shape (the frames' point counts, image
 * sizes and cube sizes) lays out the workspace and captures the graph
 * (synchronizes); later batches of that shape only copy the frame table
 * (poses, clouds, maps) and replay it.  Up to two batches may be in flight;
 * a third waits for the first. */
int nm_batch_submit(nm_batch* batch, const nm_agent_frame* frames, int n,
                    void* stream);

/* Statistics of each frame of the last batch (as nm_frame_stats_get);
 * synchronizes. */
int nm_batch_stats(nm_batch* batch, nm_frame_stats* out, int n);

/* Autotuner: times the cast of `frames` (without touching their maps)
 * under every candidate tuning, `reps` runs each, re-times the four
 * fastest over longer interleaved runs, keeps the fastest and returns it
 * with its time.  Synchronizes. */
int nm_batch_tune(nm_batch* batch, const nm_agent_frame* frames, int n,
                  int reps, void* stream, nm_batch_tuning* best,
                  float* best_ms);
int nm_batch_set_tuning(nm_batch* batch, const nm_batch_tuning* tuning);
int nm_batch_get_tuning(nm_batch* batch, nm_batch_tuning* tuning);
/* 1 (default): replay a captured graph; 0: enqueue plain launches. */
int nm_batch_set_graph(nm_batch* batch, int enable);

#ifdef __cplusplus
}
#endif

#endif /* NANOMAP_AMD_BATCH_H */
