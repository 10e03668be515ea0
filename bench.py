#!/usr/bin/env python3
"""Benchmark of the MI355X NanoMap frustum-sensor path.

Workload (the headline configuration): the 100-frame 640x480 sequence at
0.05 m voxels and 10 m range in a 50x50x10 m room with 300 boxes and 200
cylinders, 5% NaN dropouts, depth noise 0.002 d^2 (synthetic, seeded).
A step is one frame of the plain path -- ray set-up, voxel DDA of every
pixel, hit/miss counts, clamped log-odds update of the device map -- with
the frame's cloud already resident in HBM.  The map carries over between
frames, as in the reference sequence.

Prints one JSON line: value = valid rays per second over the timed steps
(Mrays/s), plus per-frame latency percentiles, stage times, the roofline of
the dominant kernel and the CPU baseline (the C++ restatement of the
reference CPU path, one core, same host, same run).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Mrays/s + p50 frame latency, 640x480@0.05m; batched & "
          "fwd+adjoint vs plain")
HBM_PEAK_GBS = 8000.0
# VALU issue peak of MI355X: one wave64 VALU instruction per 4 cycles on each
# of the 1024 SIMDs at 2.4 GHz, in wave-instructions per second
VALU_PEAK = 1024 * 2.4e9 / 4
# wave64 VALU instructions per voxel visit of the accumulate pass on this
# frame workload (sliced walks included), from its SQ counters
# (profiles/k_accumulate_sq_counters.md)
VALU_PER_VISIT = 2.57


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=100,
                    help="length of the frame sequence")
    ap.add_argument("--filter-frames", type=int, default=20,
                    help="frames of the voxel-filter legs (0 = skip)")
    ap.add_argument("--cpu-frames", type=int, default=30,
                    help="frames of the CPU baseline sample (0 = skip)")
    ap.add_argument("--rounding", choices=["down", "nearest"], default="down")
    ap.add_argument("--batch-steps", type=int, default=10,
                    help="timed passes of the 32-agent batch (0 = skip)")
    ap.add_argument("--soft-steps", type=int, default=10,
                    help="timed passes of the soft march (0 = skip)")
    ap.add_argument("--sim-steps", type=int, default=5,
                    help="timed passes of the 64-pose sensor simulation "
                         "(0 = skip)")
    ap.add_argument("--only", choices=["all", "batch", "soft", "sim"],
                    default="all",
                    help="run one part only (for profiling): the 32-agent "
                         "batch, the soft march or the sensor simulation; "
                         "prints that part's JSON")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="before the timed steps, run one fill and one "
                         "elementwise copy of this many MiB (the known-size "
                         "kernels that calibrate the HBM byte counters)")
    return ap.parse_args()


def measured_traffic(stage):
    """HBM bytes per launch of the stage's kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE runs, scaled by the
    counters' reading for a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(stage)
    return None if k is None else k.get("bytes_per_launch")


def percentile(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs), q)) if xs else None


def bench_batch(nm, W, torch, steps, rounding):
    """32 agents (uneven image sizes, ranges, voxel sizes; own maps): one
    graph-captured batched launch per pass, against the same frames as 32
    plain launches one after another, on the same stream (`speedup`, the
    target's baseline), and as 32 plain launches each waited for before the
    next (`speedup_vs_sync_per_frame`, a per-frame call that returns once
    its map is updated)."""
    _, agents = W.agents_32(nm)
    clouds = [torch.from_numpy(a["cloud"]).cuda() for a in agents]
    frames = [W.agent_frame(nm, a, rounding) for a in agents]
    P = max(c.shape[0] for c in clouds)
    T = max(f.max_time for f in frames) + 1.0
    total = sum(c.shape[0] for c in clouds)

    def maps():
        cap = {0.05: 1 << 19, 0.1: 1 << 17}
        return [nm.Map(a["res"], *W.CLAMP,
                       leaf_capacity=cap.get(a["res"], 1 << 16))
                for a in agents]

    stream = torch.cuda.current_stream()
    sh = ctypes_stream(stream)
    # tile capacity of both: 2^18 leaves, four times the 68.6 k the 32
    # frames touch together (27-bit sort keys for both)
    TILES = 1 << 18
    ctx = nm.Context(P, T, tile_capacity=TILES)
    pmaps = maps()
    rays = 0
    for f, c, m in zip(frames, clouds, pmaps):  # untimed pass: ray counts
        ctx.submit(m, f, c, stream=sh)
        rays += ctx.stats()["rays"]
    batch = nm.Batch(len(agents), total, T, TILES)
    bmaps = maps()
    table = list(zip(frames, clouds, bmaps))
    t0 = time.perf_counter()
    tuning, tune_ms = batch.tune(table, reps=2, stream=sh)
    tune_s = time.perf_counter() - t0
    batch.submit(table, stream=sh)  # layout + graph capture
    torch.cuda.synchronize()

    def timed(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(steps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    def plain_pass():
        for f, c, m in zip(frames, clouds, pmaps):
            ctx.submit(m, f, c, stream=sh)

    plain_ms = timed(plain_pass)

    def plain_sync_pass():
        # each frame waited for before the next is submitted, as a call
        # that returns once the map is updated (Manager::insertPointCloud,
        # and the reference's agent loop, SimManager.h:100-113)
        for f, c, m in zip(frames, clouds, pmaps):
            ctx.submit(m, f, c, stream=sh)
            stream.synchronize()

    plain_sync_ms = timed(plain_sync_pass)
    batch_ms = timed(lambda: batch.submit(table, stream=sh))
    st = batch.stats()
    assert all(s["errors"] == 0 and s["overflow"] == 0 for s in st)
    visits = sum(s["visits"] for s in st)
    return {
        "workload": "32 agents in a 100x100x10 m room (1200 boxes + 800 "
                    "cylinders): image sizes from {160x120, 320x240, "
                    "640x480, 848x480, 1280x720}, ranges 3-20 m, voxels "
                    "{0.05, 0.1, 0.2} m, own maps; clouds resident in HBM",
        "agents": len(agents),
        "points": int(total),
        "rays": int(rays),
        "plain_ms": plain_ms,
        "batched_ms": batch_ms,
        "visits": int(visits),
        "batched_gvisits_per_s": visits / batch_ms / 1e6,
        "plain_mrays_per_s": rays / plain_ms / 1e3,
        "batched_mrays_per_s": rays / batch_ms / 1e3,
        "speedup": plain_ms / batch_ms,
        "plain_sync_ms": plain_sync_ms,
        "speedup_vs_sync_per_frame": plain_sync_ms / batch_ms,
        "tuning": tuning,
        "tuner_s": tune_s,
        "steps": steps,
    }


def bench_sim(nm, W, torch, steps):
    """Sensor-simulation cast (generateCloud): 64 poses of 848x480 at 90x60
    deg, 15 m range, through the 0.1 m occupancy grid of the box-room map
    thresholded at log-odds > 0; each pass casts all 64 poses in one call
    (which returns the hit count, so it synchronises)."""
    o, v = W.box_room_map_leaves(nm)
    grid = nm.SimGrid(o, v, 0.0)
    w, h, res = 848, 480, 0.1
    sensors = []
    for pos, quat in W.sim_poses(nm, 64):
        fr = nm.frame_from_pose(pos, quat, nm.DEFAULT_F, res, 15.0, 0.85, -0.4,
                                0)
        sensors.append(nm.sim_sensor(pos, list(fr.rotation), res, w, h, 90.0,
                                     60.0, 0.1, 15.0))
    n = 0
    for _ in range(2):
        _, _, _, n = grid.cast(sensors)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        _, _, _, n = grid.cast(sensors)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    rays = len(sensors) * w * h
    return {
        "workload": "64 poses x 848x480 (90x60 deg, 15 m) through the 0.1 m "
                    "grid of the box-room map, log-odds > 0",
        "rays": rays,
        "returns": n,
        "ms_per_pass": ms,
        "mrays_per_s": rays / (ms * 1e-3) / 1e6,
        "active_voxels": grid.info()["active"],
        "steps": steps,
        "dense": bench_sim_dense(nm, W, torch, steps),
    }


def bench_sim_dense(nm, W, torch, steps):
    """The same 64 x 848x480 cast through a dense grid: the 0.05 m map of
    the first 20 hall frames (log-odds > 0), from 64 poses of the hall walk,
    so rays traverse mapped free space and end on mapped surfaces."""
    scene, cam, frames = W.hall_walk_frames(nm, n_frames=20)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    m = nm.Map(0.05, *W.CLAMP)
    for pos, quat, cloud in frames:
        ctx.submit(m, W.frame(nm, pos, quat, 0.05, 10.0, nm.ROUND_DOWN), cloud)
    o, v, _ = m.download()
    grid = nm.SimGrid(o, v, 0.0)
    w, h, res = 848, 480, 0.05
    pos, quat = scene.walk(2, 64, 0.1, 2.0, 15.0, 0.3)
    sensors = []
    for k in range(64):
        fr = nm.frame_from_pose(pos[k], quat[k], nm.DEFAULT_F, res, 15.0,
                                0.85, -0.4, 0)
        sensors.append(nm.sim_sensor(pos[k], list(fr.rotation), res, w, h,
                                     90.0, 60.0, 0.1, 15.0))
    n = 0
    for _ in range(2):
        _, _, _, n = grid.cast(sensors)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        _, _, _, n = grid.cast(sensors)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    rays = len(sensors) * w * h
    return {
        "workload": "64 hall-walk poses x 848x480 (90x60 deg, 15 m) through "
                    "the 0.05 m map of 20 hall frames, log-odds > 0",
        "rays": rays,
        "returns": n,
        "ms_per_pass": ms,
        "mrays_per_s": rays / (ms * 1e-3) / 1e6,
        "active_voxels": grid.info()["active"],
        "leaves": int(len(o)),
    }


def bench_soft(nm, W, torch, steps):
    """Forward + adjoint of the differentiable soft march for 16 cameras of
    320x240 (8 m, 160 samples at 0.05 m) over the box-room world mapped at
    0.05 m: one launch each for all cameras, against one launch pair per
    camera; plus 20 Gauss-Newton pose-refinement steps."""
    import numpy as np
    from nanomap_amd import soft
    res = 0.05
    _, cam, pos, quat, clouds = W.soft_walk(nm)
    ctx = nm.Context(cam.width * cam.height, 8.0 / res + 1)
    m = nm.Map(res, *W.CLAMP, leaf_capacity=1 << 19)
    frames = []
    for k in range(len(clouds)):
        f = W.frame(nm, pos[k], quat[k], res, 8.0, nm.ROUND_DOWN)
        ctx.submit(m, f, clouds[k])
        frames.append(f)
    o, v, _ = m.download()
    grid, lo = W.dense_grid(o, v)
    sel = list(range(0, 48, 3))[:16]
    R = torch.tensor(np.stack([np.array(list(frames[k].rotation),
                                        np.float64).reshape(3, 3)
                               for k in sel])).cuda()
    p = torch.tensor(np.stack([np.asarray(pos[k], np.float64)
                               for k in sel])).cuda()
    march = soft.SoftMarch(grid.shape, lo * res, res, 320, 240, 87.0, 58.0,
                           res, 160)
    g32 = m.dense(lo, grid.shape)  # gathered on the device from the map
    cams = soft.cameras(R, p)
    n = cams.shape[0]
    gout = torch.ones((n, 240, 320), dtype=torch.float32, device="cuda")
    gp = torch.zeros((n, 6), dtype=torch.float64, device="cuda")
    import ctypes
    from nanomap_amd._capi import check, lib
    depth = torch.empty((n, 240, 320), dtype=torch.float32, device="cuda")
    logt = torch.empty((n, 240, 320), dtype=torch.float64, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def fwd_bwd(lo_c, hi_c):
        # cameras [lo_c, hi_c) in one forward and one adjoint launch
        c = cams[lo_c:hi_c]
        k = hi_c - lo_c
        off = lo_c * 240 * 320
        check(lib().nm_soft_forward(ctypes.byref(march.grid),
                                    ctypes.c_void_p(g32.data_ptr()),
                                    ctypes.c_void_p(c.data_ptr()), k,
                                    ctypes.byref(march.params),
                                    ctypes.c_void_p(depth.data_ptr() + 4 * off),
                                    ctypes.c_void_p(logt.data_ptr() + 8 * off), st))
        check(lib().nm_soft_backward(ctypes.byref(march.grid),
                                     ctypes.c_void_p(g32.data_ptr()),
                                     ctypes.c_void_p(c.data_ptr()), k,
                                     ctypes.byref(march.params),
                                     ctypes.c_void_p(logt.data_ptr() + 8 * off),
                                     ctypes.c_void_p(gout.data_ptr() + 4 * off),
                                     ctypes.c_void_p(gp.data_ptr() + 8 * 6 * lo_c),
                                     None, None, st))

    def timed(fn):
        fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for _ in range(steps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    batched_ms = timed(lambda: fwd_bwd(0, n))
    per_ms = timed(lambda: [fwd_bwd(i, i + 1) for i in range(n)])
    rays = n * 240 * 320
    # pose refinement from perturbations of up to 0.1 m and 3 deg
    target = march.forward(g32, cams)
    gen = torch.Generator().manual_seed(11)
    dt = (torch.rand((n, 3), generator=gen, dtype=torch.float64) * 2 - 1) * 0.1
    axis = torch.randn((n, 3), generator=gen, dtype=torch.float64)
    axis = axis / axis.norm(dim=1, keepdim=True)
    ang = (torch.rand((n, 1), generator=gen, dtype=torch.float64) * 2 - 1) * np.radians(3)
    R0 = soft.so3_exp((axis * ang).cuda()) @ R
    p0 = p + dt.cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Rf, pf, hist = soft.refine_poses(march, g32, R0, p0, target, iters=20)
    torch.cuda.synchronize()
    refine_s = time.perf_counter() - t0
    t_err = (pf - p).norm(dim=1)
    cosang = ((Rf.transpose(1, 2) @ R).diagonal(dim1=1, dim2=2).sum(1) - 1) / 2
    r_err = torch.rad2deg(torch.arccos(cosang.clamp(-1, 1)))
    return {
        "workload": "16 cameras 320x240 (87x58 deg), 8 m, 160 samples at "
                    "0.05 m, quadratic B-spline over the box-room world "
                    "mapped at 0.05 m (%dx%dx%d grid); forward + adjoint "
                    "pose gradients" % grid.shape,
        "rays": rays,
        "batched_ms": batched_ms,
        "per_frame_ms": per_ms,
        "batched_mrays_per_s": rays / batched_ms / 1e3,
        "per_frame_mrays_per_s": rays / per_ms / 1e3,
        "speedup": per_ms / batched_ms,
        "refine_20_steps_s": refine_s,
        "refine_translation_err_m_median": float(t_err.median()),
        "refine_rotation_err_deg_median": float(r_err.median()),
        "steps": steps,
    }


def bench_filter(nm, W, torch, ctx, frames, dev, P, cam, nf, rounding,
               stream, sh):
    """The first nf frames through the voxel filter, in each precision mode."""
    # the same frames through the voxel filter (filterCloud precision 0):
    # points outside the sensor's frustum box dropped, the rest binned per
    # voxel, one ray per occupied voxel, into a second map
    box = W.hall_box(nm, rounding)
    fparams = [W.frame(nm, p, q, 0.05, 10.0, rounding, cam.width, cam.height,
                       nm.FILTER_VOXEL, box=box) for p, q, _ in frames[:nf]]
    fmap = nm.Map(0.05, *W.CLAMP, leaf_capacity=1 << 19)
    for k in range(2):
        ctx.submit(fmap, fparams[k], (dev[k].data_ptr(), P, 12), stream=sh)
    torch.cuda.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    fa, fb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fa.record(stream)
    for k in range(nf):
        ctx.submit(fmap, fparams[k], (dev[k].data_ptr(), P, 12), stream=sh)
    fb.record(stream)
    torch.cuda.synchronize()
    ctx.set_timing(False)
    fst = ctx.stage_times()
    filtered = {
        "workload": "the first %d frames of the sequence with the voxel "
                    "filter: one ray per occupied 0.05 m voxel" % nf,
        "frames": nf,
        "ms_per_frame": fa.elapsed_time(fb) / nf,
        "rays_per_frame": ctx.stats()["rays"],
        "stage_ms": {k: round(v / n, 5) if n else 0.0
                     for k, (v, n) in fst.items()},
    }
    fmap.close()
    # the lower-memory variants (PrecisionType 1-3): FP16 sums and count,
    # voxel centre with an int count, voxel centre with an 8-bit count
    modes = {}
    for name, filt in (("fp16_centroid", nm.FILTER_VOXEL_HALF),
                       ("centre_int_count", nm.FILTER_VOXEL_COUNT),
                       ("centre_8bit_count", nm.FILTER_VOXEL_COUNT8)):
        mparams = [W.frame(nm, p, q, 0.05, 10.0, rounding, cam.width,
                           cam.height, filt, box=box)
                   for p, q, _ in frames[:nf]]
        mmap = nm.Map(0.05, *W.CLAMP, leaf_capacity=1 << 19)
        for k in range(2):
            ctx.submit(mmap, mparams[k], (dev[k].data_ptr(), P, 12), stream=sh)
        torch.cuda.synchronize()
        fa.record(stream)
        for k in range(nf):
            ctx.submit(mmap, mparams[k], (dev[k].data_ptr(), P, 12), stream=sh)
        fb.record(stream)
        torch.cuda.synchronize()
        modes[name] = {"ms_per_frame": round(fa.elapsed_time(fb) / nf, 4),
                       "rays_last_frame": ctx.stats()["rays"]}
        mmap.close()
    filtered["precision_modes"] = modes
    return filtered


def cpu_baseline(frames_params, clouds, n):
    """Oracle (C++ restatement of Handler::processPointCloudCPU) on one
    pinned core: valid rays per second over the first n frames."""
    import oracle
    from nanomap_amd import workloads as W
    saved = os.sched_getaffinity(0)
    core = min(saved)
    os.sched_setaffinity(0, {core})
    try:
        omap = oracle.Map(*W.CLAMP)
        rays = 0
        t0 = time.perf_counter()
        for k in range(n):
            t = oracle.Temp()
            rays += t.cast_cloud(frames_params[k], clouds[k])
            omap.integrate(t, frames_params[k].rounding)
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, saved)
    cpu_name = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": 1,
            "kind": "port",
            "sample": "first %d frames of the same sequence, %.2f s, %d rays,"
                      " one core pinned (%s)" % (n, dt, rays, cpu_name),
            "ms_per_frame": dt / n * 1e3}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not "
              "supported: every run here uses one GPU (agents and sensors "
              "would shard across GPUs without exchange)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import nanomap_amd as nm
    from nanomap_amd import workloads as W

    rounding = nm.ROUND_DOWN if args.rounding == "down" else nm.ROUND_NEAREST
    if args.only == "batch":
        print(json.dumps({"batched": bench_batch(nm, W, torch,
                                                 max(args.batch_steps, 1),
                                                 rounding)}))
        return
    if args.only == "soft":
        print(json.dumps({"soft_march": bench_soft(nm, W, torch,
                                                   max(args.soft_steps, 1))}))
        return
    if args.only == "sim":
        print(json.dumps({"sim_cast": bench_sim(nm, W, torch,
                                                max(args.sim_steps, 1))}))
        return
    t_setup = time.perf_counter()
    scene, cam, frames = W.hall_walk_frames(nm, n_frames=args.frames)
    P = cam.width * cam.height
    params = [W.frame(nm, p, q, 0.05, 10.0, rounding, cam.width, cam.height)
              for p, q, _ in frames]
    clouds = [c for _, _, c in frames]
    dev = torch.from_numpy(np.stack(clouds)).to("cuda:0")  # [F, P, 3]
    stream = torch.cuda.current_stream()
    sh = ctypes_stream(stream)
    ctx = nm.Context(P, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP, leaf_capacity=1 << 19)
    print("setup %.1f s" % (time.perf_counter() - t_setup), file=sys.stderr,
          flush=True)
    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib << 18
        a = torch.ones(n, device="cuda:0")  # fill: writes 4n bytes
        b = torch.empty_like(a)
        torch.mul(a, 1.0, out=b)  # copy: reads 4n, writes 4n bytes
        torch.cuda.synchronize()
        del a, b
        print("copy probe %d MiB" % args.copy_probe_mib, file=sys.stderr,
              flush=True)

    def submit(k):
        ptr = dev[k].data_ptr()
        ctx.submit(gmap, params[k], (ptr, P, 12), stream=sh)

    # per-frame workload counters (untimed pass over the sequence)
    per = []
    for k in range(args.frames):
        submit(k)
        per.append(ctx.stats())
    assert all(s["errors"] == 0 and s["overflow"] == 0 for s in per), per[0]
    gmap.clear(sh)
    for k in range(args.warmup):
        submit(k % args.frames)
    torch.cuda.synchronize()
    gmap.clear(sh)
    torch.cuda.synchronize()
    print("warm", file=sys.stderr, flush=True)

    # stage split from an untimed pass with an event pair around every stage
    ctx.reset_timing()
    ctx.set_timing(True)
    for k in range(min(args.steps, 20)):
        submit(k % args.frames)
    torch.cuda.synchronize()
    ctx.set_timing(False)
    stages = ctx.stage_times()
    split = {s: (ms / n if n else 0.0) for s, (ms, n) in stages.items()}
    dom = max(split, key=split.get)
    # the timed steps start from an empty map: each frame reserves and
    # inserts its own new leaves, as the sequence does
    gmap.clear(sh)
    torch.cuda.synchronize()

    K = args.steps
    # HIP events on the submit stream around the dominant stage only, over
    # the timed steps, give its launch duration for the roofline
    ctx.reset_timing()
    ctx.set_timing(True, stages=[dom])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        submit(k % args.frames)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.set_timing(False)
    stages = ctx.stage_times()
    rays = sum(per[k % args.frames]["rays"] for k in range(K))
    visits = sum(per[k % args.frames]["visits"] for k in range(K))
    voxels = sum(per[k % args.frames]["voxels"] for k in range(K))
    leaves = sum(per[k % args.frames]["leaves"] for k in range(K))
    dt = t1 - t0
    print("timed %.3f s" % dt, file=sys.stderr, flush=True)

    # Frame latency over the whole sequence (every frame once, from an
    # empty map, after 3 warm-up frames), whatever --steps is:
    #  - clouds resident in HBM: event pair around each frame's submit;
    #  - from the pinned host cloud to the map updated on the device (the
    #    headline latency): the event pair also spans the H2D copy.
    dcam = nm.depth_camera(cam.width, cam.height, cam.hfov_deg, cam.vfov_deg)

    def sequence_latency(host):
        lat = []
        for k in range(3):
            submit(k)
        gmap.clear(sh)
        torch.cuda.synchronize()
        for k in range(args.frames):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            if host == "cloud":
                ctx.submit(gmap, params[k], pinned[k].numpy(), stream=sh)
            elif host == "depth":
                ctx.submit_depth(gmap, params[k], dcam, pinned_depth[k].numpy(),
                                 stream=sh)
            else:
                submit(k)
            b.record(stream)
            torch.cuda.synchronize()
            lat.append(a.elapsed_time(b))
        gmap.clear(sh)
        torch.cuda.synchronize()
        return lat

    pinned = [torch.from_numpy(c).pin_memory() for c in clouds]
    pinned_depth = [torch.from_numpy(np.ascontiguousarray(c[:, 2])).pin_memory()
                    for c in clouds]
    lat = sequence_latency(None)
    pl = sequence_latency("cloud")
    pd = sequence_latency("depth")

    nf = min(args.filter_frames, args.frames)
    filtered = (bench_filter(nm, W, torch, ctx, frames, dev, P, cam, nf,
                             rounding, stream, sh) if nf > 0 else None)

    # roofline of the dominant kernel stage: its duration over the timed
    # steps; the other stages from the untimed split
    avg = dict(split)
    ms_dom, n_dom = stages[dom]
    avg[dom] = ms_dom / n_dom if n_dom else split[dom]
    Pf = float(P)
    U = voxels / K
    L = leaves / K
    R = rays / K
    alg_bytes = {
        # cloud xyz in, hit cell out, leaf marks
        "discover": 12 * Pf + 4 * Pf + 4 * L,
        # cloud xyz in, one 4-byte miss count per touched voxel
        "accumulate": 12 * Pf + 4 * U,
        # counts in (8 B), map value read+write (8 B) per touched voxel,
        # mask words (128 B per leaf)
        "update": 16 * U + 128 * L,
        "hitsort": 4 * 8 * 2 * R,
        "compact": 4.0 * (2 * (-(-202 // 8) + 1) + 1) ** 3,
        "filter": 12 * Pf,
    }
    achieved = alg_bytes[dom] / (avg[dom] * 1e-3) / 1e9
    frame_bytes = 12 * Pf + 16 * U + 16 * L
    frame_ms = sum(avg.values())
    out = {
        "metric": METRIC,
        "value": rays / dt / 1e6,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": dt / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "640x480 depth frames, 87x58 deg, 10 m range, 0.05 m "
                        "voxels; 100-frame random walk (0.1 m, 2 deg/frame) "
                        "in a 50x50x10 m room, 300 boxes + 200 cylinders, "
                        "5% NaN dropouts, noise 0.002 d^2; plain path, clouds "
                        "resident in HBM",
            "frames": args.frames,
            "pixels_per_frame": P,
            "rounding": args.rounding,
            "parallelism": "single GPU",
        },
        "p50_frame_ms": percentile(pl, 50),
        "p99_frame_ms": percentile(pl, 99),
        "latency_frames": len(pl),
        "latency_definition": "pinned host cloud -> map updated on the "
                              "device (H2D included), each of the %d "
                              "frames once from an empty map after 3 "
                              "warm-up frames" % len(pl),
        "p50_frame_ms_resident": percentile(lat, 50),
        "p99_frame_ms_resident": percentile(lat, 99),
        "p50_frame_ms_depth_image": percentile(pd, 50),
        "p99_frame_ms_depth_image": percentile(pd, 99),
        "depth_image_note": "the same frames uploaded as pinned FP32 z-depth "
                            "images (1.2 MB instead of 3.7 MB of xyz), "
                            "back-projected in the cast's ray set-up",
        "voxel_updates_per_s": visits / dt,
        "rays_per_frame": R,
        "visits_per_frame": visits / K,
        "voxels_per_frame": U,
        "leaves_per_frame": L,
        "stage_ms": {s: round(v, 5) for s, v in avg.items()},
        "roofline": {
            "kernel_stage": dom,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom),
            "alg_bytes_per_launch": alg_bytes[dom],
        },
        # the second bound of the dominant kernel: VALU issue, with the
        # instructions per visit from its counters and the visits and
        # duration of the timed launches
        "roofline_issue": {
            "kernel_stage": "accumulate",
            "bound": "valu_issue",
            "achieved": visits / K * VALU_PER_VISIT / (avg["accumulate"] * 1e-3),
            "peak": VALU_PEAK,
            "unit": "wave64 VALU instructions/s",
            "frac": visits / K * VALU_PER_VISIT / (avg["accumulate"] * 1e-3)
                    / VALU_PEAK,
            "valu_per_visit": VALU_PER_VISIT,
        },
        "roofline_frame": {
            "alg_bytes": frame_bytes,
            "gpu_ms": frame_ms,
            "achieved_gbs": frame_bytes / (frame_ms * 1e-3) / 1e9,
            "frac": frame_bytes / (frame_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "cpu_baseline": None,
        "voxel_filter": filtered,
    }
    del dev, ctx, gmap
    if args.batch_steps > 0:
        out["batched"] = bench_batch(nm, W, torch, args.batch_steps, rounding)
        print("batched done", file=sys.stderr, flush=True)
    if args.soft_steps > 0:
        out["soft_march"] = bench_soft(nm, W, torch, args.soft_steps)
        print("soft march done", file=sys.stderr, flush=True)
    if args.sim_steps > 0:
        out["sim_cast"] = bench_sim(nm, W, torch, args.sim_steps)
        print("sim cast done", file=sys.stderr, flush=True)
    if args.cpu_frames > 0:
        out["cpu_baseline"] = cpu_baseline(params, clouds,
                                           min(args.cpu_frames, args.frames))
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))


def ctypes_stream(stream):
    import ctypes
    return ctypes.c_void_p(stream.cuda_stream)


if __name__ == "__main__":
    main()
