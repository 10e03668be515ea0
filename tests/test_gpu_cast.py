"""GPU parity of the plain frame path against the CPU oracle.

Each test runs the HIP path through the C ABI (nanomap_amd -> libnanomap_amd
.so) and the oracle (Handler::populateTempGrid + integrateTempGrid
restated in oracle/) on identical clouds, and requires bit-exact results:
per-voxel (miss, hit) counts of every frame, and the map (leaf set, FP32
log-odds bit patterns, active masks).  Both rounding modes are covered;
round-down is what the reference executes.
"""
import ctypes

import numpy as np
import pytest

import oracle
from nanomap_amd import workloads as W

pytestmark = pytest.mark.gpu

RN, RD = 0, 1


def _check_counts(ctx, temp):
    g_ijk, g_m, g_h = ctx.counts()
    o_ijk, o_m, o_h, _ = temp.export()
    assert g_ijk.shape == o_ijk.shape, (g_ijk.shape, o_ijk.shape)
    assert np.array_equal(g_ijk, o_ijk)
    assert np.array_equal(g_m, o_m)
    assert np.array_equal(g_h, o_h)


def _check_maps(gmap, omap):
    go, gv, gm = gmap.download()
    oo, ov, om = omap.export()
    assert go.shape == oo.shape, (go.shape, oo.shape)
    assert np.array_equal(go, oo)
    bad = np.nonzero(gv.view(np.uint32) != ov.view(np.uint32))
    assert bad[0].size == 0, (bad[0][:5], bad[1][:5], gv[bad][:5], ov[bad][:5])
    assert np.array_equal(gm, om)


def _run_frame(nm, ctx, gmap, omap, frame, cloud, check_counts=True):
    print("frame: points=%d res=%g max_time=%g rounding=%d filter=%d"
          % (len(cloud), frame.grid_res, frame.max_time, frame.rounding,
             frame.filter), flush=True)
    ctx.cast(frame, cloud)
    temp = oracle.Temp()
    if frame.filter != nm.FILTER_NONE:
        rays = oracle.voxel_filter(frame, cloud)
        temp.cast_rays(frame, rays)
    else:
        temp.cast_cloud(frame, cloud)
    if check_counts:
        _check_counts(ctx, temp)
    ctx.update(gmap)
    omap.integrate(temp, frame.rounding)
    return temp


@pytest.mark.parametrize("rounding", [RD, RN])
def test_box_room_counts_and_map_bit_exact(gpu, rounding):
    nm = gpu
    _, cam, frames = W.box_room_frame(nm)
    pos, quat, cloud = frames[0]
    f = W.frame(nm, pos, quat, 0.1, 5.0, rounding, cam.width, cam.height)
    ctx = nm.Context(cam.width * cam.height, f.max_time)
    gmap = nm.Map(0.1, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    temp = _run_frame(nm, ctx, gmap, omap, f, cloud)
    st = ctx.stats()
    _, m, h, _ = temp.export()
    assert st["errors"] == 0 and st["overflow"] == 0
    assert st["visits"] == temp.visits() == int(m.sum() + h.sum())
    assert st["voxels"] == len(m)
    assert st["hit_voxels"] == int((h > 0).sum())
    assert st["mixed_voxels"] == int(((h > 0) & (m > 0)).sum())
    _check_maps(gmap, omap)


def test_hall_ten_frames_bit_exact(gpu):
    nm = gpu
    _, cam, frames = W.hall_walk_frames(nm, n_frames=10)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    for pos, quat, cloud in frames:
        f = W.frame(nm, pos, quat, 0.05, 10.0, RD, cam.width, cam.height)
        _run_frame(nm, ctx, gmap, omap, f, cloud)
        assert ctx.stats()["errors"] == 0
    _check_maps(gmap, omap)


@pytest.mark.timeout(900)
def test_hall_hundred_frames_bit_exact(gpu, tmp_path):
    """The 100-frame 640x480 hall sequence at 0.05 m; every
    frame's counts and the final map bit-exact versus the CPU path.  The
    map is also saved after frame 50 and restored into a new map on a
    fresh context, which takes frames 50-99 and ends bit-exact too."""
    nm = gpu
    _, cam, frames = W.hall_walk_frames(nm, n_frames=100)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    restored, ctx2 = None, None
    for k, (pos, quat, cloud) in enumerate(frames):
        if k == 50:
            gmap.save(tmp_path / "hall50.nmap")
            restored = nm.Map.load(tmp_path / "hall50.nmap")
            ctx2 = nm.Context(cam.width * cam.height, 200.0)
        f = W.frame(nm, pos, quat, 0.05, 10.0, RD, cam.width, cam.height)
        _run_frame(nm, ctx, gmap, omap, f, cloud)
        if restored is not None:
            ctx2.submit(restored, f, cloud)
    st = ctx.stats()
    assert st["errors"] == 0 and st["overflow"] == 0
    _check_maps(gmap, omap)
    st2 = ctx2.stats()
    assert st2["errors"] == 0 and st2["overflow"] == 0
    _check_maps(restored, omap)


def test_sensor_sizes_bit_exact(gpu):
    """Frames of the 32-agent sensor mix, 160x120 to 1280x720 at 0.05 to
    0.1 m, through one context in turn: small frames have room on the GPU
    for the accumulate pass's time slices of their long-ray blocks, the
    1280x720 frame fills it with its own blocks (no slices), 848x480 sits
    between.  Every frame must give the oracle's counts and map."""
    nm = gpu
    _, agents = W.agents_32(nm)
    picks = []
    for want in [(160, 120, 0.05), (1280, 720, 0.1), (320, 240, 0.05),
                 (848, 480, 0.05), (160, 120, 0.1), (640, 480, 0.05)]:
        a = next(a for a in agents
                 if (a["width"], a["height"], a["res"]) == want)
        picks.append(a)
    frames = [W.agent_frame(nm, a, RD) for a in picks]
    ctx = nm.Context(max(len(a["cloud"]) for a in picks),
                     max(f.max_time for f in frames) + 1.0)
    for a, f in zip(picks, frames):
        print("sensor %dx%d res %g range %.1f" % (a["width"], a["height"],
                                                  a["res"], a["range"]),
              flush=True)
        gmap = nm.Map(a["res"], *W.CLAMP)
        omap = oracle.Map(*W.CLAMP)
        _run_frame(nm, ctx, gmap, omap, f, a["cloud"])
        assert ctx.stats()["errors"] == 0
        _check_maps(gmap, omap)
        gmap.close()


def test_edge_cases(gpu):
    nm = gpu
    I = np.eye(3, dtype=np.float32)
    ctx = nm.Context(4096, 100.0)
    gmap = nm.Map(0.1, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    rng = np.random.default_rng(0)
    cases = {
        "empty": np.zeros((0, 3), np.float32),
        "all_nan": np.full((50, 3), np.nan, np.float32),
        "single": np.array([[0.3, 0.2, 1.1]], np.float32),
        "zero_point": np.array([[0, 0, 0], [0.5, 0, 0]], np.float32),
        "beyond_range": np.array([[20.0, 0, 0], [0, 9.99, 0]], np.float32),
        "axis_and_diagonal": np.array([[1, 0, 0], [0, -1, 0], [0, 0, 1],
                                       [0.4, 0.4, 0.4], [-0.4, -0.4, -0.4]],
                                      np.float32),
        "random": (rng.normal(size=(3000, 3)) * 3).astype(np.float32),
    }
    # integer eye (rays start on voxel corners: ties at t = 0) and a
    # position exactly 0.0
    poses = [([0.0, 0.0, 0.0], [1, 0, 0, 0]), ([1.0, 2.0, 3.0], [1, 0, 0, 0]),
             ([-3.05, 7.33, 0.71], [0.9, 0.1, -0.3, 0.2])]
    for rounding in (RD, RN):
        for pos, q in poses:
            for name, cloud in cases.items():
                f = nm.frame_from_pose(pos, q, I, 0.1, 10.0, W.HIT, W.MISS,
                                       rounding)
                print("case", name, pos, rounding, flush=True)
                _run_frame(nm, ctx, gmap, omap, f, cloud)
                assert ctx.stats()["errors"] == 0
    _check_maps(gmap, omap)


def test_stride_16_and_device_pointer(gpu):
    import torch
    nm = gpu
    _, cam, frames = W.box_room_frame(nm)
    pos, quat, cloud = frames[0]
    f = W.frame(nm, pos, quat, 0.1, 5.0, RD)
    padded = np.concatenate([cloud, np.full((len(cloud), 1), 7.0, np.float32)], 1)
    ctx = nm.Context(len(cloud), f.max_time)
    g1 = nm.Map(0.1, *W.CLAMP)
    ctx.submit(g1, f, padded)             # host pointer, stride 16
    dev = torch.from_numpy(padded).cuda()
    g2 = nm.Map(0.1, *W.CLAMP)
    ctx.submit(g2, f, dev)                # device pointer, stride 16
    torch.cuda.synchronize()
    omap = oracle.Map(*W.CLAMP)
    oracle.process_frame(omap, f, cloud)
    _check_maps(g1, omap)
    _check_maps(g2, omap)


def test_run_twice_is_identical(gpu):
    nm = gpu
    _, cam, frames = W.hall_walk_frames(nm, n_frames=1)
    pos, quat, cloud = frames[0]
    f = W.frame(nm, pos, quat, 0.05, 10.0, RD)
    ctx = nm.Context(len(cloud), f.max_time)
    outs = []
    for _ in range(2):
        g = nm.Map(0.05, *W.CLAMP)
        ctx.submit(g, f, cloud)
        o, v, m = g.download()
        outs.append((o, v.view(np.uint32), m))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


def test_unchanged_leaves_are_not_inserted(gpu):
    """A frame whose every temp value is zero (log-odds hit = miss = 0)
    changes nothing (Handler.cpp:116-118 writes only nonzero temp values):
    the leaves it reserved in the pool stay holes, invisible to leaf_count
    and download, and later frames over the same leaves still match."""
    nm = gpu
    _, cam, frames = W.hall_walk_frames(nm, n_frames=2)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    pos, quat, cloud = frames[0]
    z = nm.frame_from_pose(pos, quat, nm.DEFAULT_F, 0.05, 10.0, 0.0, 0.0, RD)
    _run_frame(nm, ctx, gmap, omap, z, cloud)
    assert gmap.leaf_count() == 0
    assert ctx.stats()["map_leaves"] == 0
    for pos, quat, cloud in frames:
        f = W.frame(nm, pos, quat, 0.05, 10.0, RD)
        _run_frame(nm, ctx, gmap, omap, f, cloud, check_counts=False)
    _check_maps(gmap, omap)
    assert ctx.stats()["map_leaves"] == gmap.leaf_count() > 0


@pytest.mark.parametrize("rounding", [RD, RN])
def test_voxel_filter_bit_exact(gpu, rounding):
    nm = gpu
    _, cam, frames = W.hall_walk_frames(nm, n_frames=2)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    for pos, quat, cloud in frames:
        f = W.frame(nm, pos, quat, 0.05, 10.0, rounding, filt=nm.FILTER_VOXEL,
                    box=W.hall_box(nm, rounding))
        ctx.cast(f, cloud)
        g_rays = ctx.filter_rays()
        o_rays = oracle.voxel_filter(f, cloud)
        assert g_rays.shape == o_rays.shape and len(o_rays) > 1000
        assert np.array_equal(g_rays.view(np.uint32), o_rays.view(np.uint32))
        temp = oracle.Temp()
        temp.cast_rays(f, o_rays)
        _check_counts(ctx, temp)
        ctx.update(gmap)
        omap.integrate(temp, rounding)
    _check_maps(gmap, omap)


@pytest.mark.parametrize("mode", ["half", "count", "count8"])
def test_voxel_filter_precision_modes_bit_exact(gpu, mode):
    # PrecisionType 1-3: two sensor frames through the whole update, then a
    # clustered cloud with 256-, 300- and 2100-point voxels (the 8-bit count
    # wraps to 0 at 256; the FP16 count stops at 2048).
    nm = gpu
    filt = {"half": nm.FILTER_VOXEL_HALF, "count": nm.FILTER_VOXEL_COUNT,
            "count8": nm.FILTER_VOXEL_COUNT8}[mode]
    _, cam, frames = W.hall_walk_frames(nm, n_frames=2)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    for pos, quat, cloud in frames:
        f = W.frame(nm, pos, quat, 0.05, 10.0, RD, filt=filt,
                    box=W.hall_box(nm, RD))
        ctx.cast(f, cloud)
        g_rays = ctx.filter_rays()
        o_rays = oracle.voxel_filter(f, cloud)
        assert g_rays.shape == o_rays.shape and len(o_rays) > 1000
        assert np.array_equal(g_rays.view(np.uint32), o_rays.view(np.uint32))
        temp = oracle.Temp()
        temp.cast_rays(f, o_rays)
        _check_counts(ctx, temp)
        ctx.update(gmap)
        omap.integrate(temp, RD)
    _check_maps(gmap, omap)

    rng = np.random.default_rng(23)
    f = W.frame(nm, (1.23, -0.47, 0.22), (1.0, 0.0, 0.0, 0.0), 0.1, 2.0,
                RD, filt=filt)
    origin = np.array([e - np.floor(e) for e in f.eye])
    R = np.asarray(f.rotation, np.float64).reshape(3, 3)  # sensor -> world
    parts = [rng.normal(size=(400, 3)).astype(np.float32) * np.float32(0.6)]
    for n, k in ((256, (14, 8, 5)), (300, (-3, 4, 3)), (2100, (2, -4, -1))):
        c = (np.asarray(k) + 0.5 - origin) * 0.1  # that voxel's centre
        w = c + rng.uniform(-0.03, 0.03, size=(n, 3))
        parts.append((w @ R).astype(np.float32))  # R^T w, row by row
    cloud = np.concatenate(parts)
    print("clustered cloud: points=%d filter=%d" % (len(cloud), filt), flush=True)
    ctx.cast(f, cloud)
    g_rays = ctx.filter_rays()
    o_rays = oracle.voxel_filter(f, cloud)
    assert g_rays.shape == o_rays.shape and len(o_rays) > 50
    assert np.array_equal(g_rays.view(np.uint32), o_rays.view(np.uint32))
    if mode == "count8":
        assert len(o_rays) == len(oracle.voxel_filter(f, cloud, precision=2)) - 1


def _box_probes(cloud, lo, hi, n, rng):
    """Copies of n valid cloud points moved onto each face of the box, one
    float outside it and one float inside it."""
    ok = ~np.isnan(cloud).any(1)
    base = cloud[ok][rng.choice(int(ok.sum()), n, replace=False)]
    out = []
    for a in range(3):
        for face, away in ((lo[a], -np.inf), (hi[a], np.inf)):
            f32 = np.float32(face)
            for v in (f32, np.nextafter(f32, np.float32(away)),
                      np.nextafter(f32, np.float32(-away))):
                p = base.copy()
                p[:, a] = v
                out.append(p)
    return np.concatenate(out).astype(np.float32)


@pytest.mark.parametrize("mode", ["voxel", "half", "count", "count8"])
def test_voxel_filter_box_edges_bit_exact(gpu, mode):
    # The frustum box of the 640x480, 58 deg, 10 m sensor (half width from
    # the aspect: +-7.39 m at 10 m, while the 87 deg camera sees to +-9.5 m):
    # the frame's own far-edge points fall outside it, and probe points on
    # each face, one float outside and one inside, are kept or dropped as
    # the oracle keeps or drops them, in every precision mode.
    nm = gpu
    filt = {"voxel": nm.FILTER_VOXEL, "half": nm.FILTER_VOXEL_HALF,
            "count": nm.FILTER_VOXEL_COUNT,
            "count8": nm.FILTER_VOXEL_COUNT8}[mode]
    rng = np.random.default_rng(31)
    _, cam, frames = W.hall_walk_frames(nm, n_frames=1)
    pos, quat, cloud = frames[0]
    lo, hi = W.hall_box(nm, RD)
    cloud = np.concatenate([cloud, _box_probes(cloud, lo, hi, 400, rng)])
    f = W.frame(nm, pos, quat, 0.05, 10.0, RD, filt=filt, box=(lo, hi))
    unboxed = W.frame(nm, pos, quat, 0.05, 10.0, RD, filt=filt)
    assert len(oracle.voxel_filter(f, cloud)) < \
        len(oracle.voxel_filter(unboxed, cloud))
    ctx = nm.Context(len(cloud), 200.0)
    gmap = nm.Map(0.05, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    print("box edges: points=%d filter=%d" % (len(cloud), filt), flush=True)
    ctx.cast(f, cloud)
    g_rays = ctx.filter_rays()
    o_rays = oracle.voxel_filter(f, cloud)
    assert g_rays.shape == o_rays.shape and len(o_rays) > 1000
    assert np.array_equal(g_rays.view(np.uint32), o_rays.view(np.uint32))
    temp = oracle.Temp()
    temp.cast_rays(f, o_rays)
    _check_counts(ctx, temp)
    ctx.update(gmap)
    omap.integrate(temp, RD)
    _check_maps(gmap, omap)


def test_overflow_is_reported_not_fatal(gpu):
    nm = gpu
    _, cam, frames = W.box_room_frame(nm)
    pos, quat, cloud = frames[0]
    f = W.frame(nm, pos, quat, 0.1, 5.0, RD)
    ctx = nm.Context(len(cloud), f.max_time, tile_capacity=4)
    g = nm.Map(0.1, *W.CLAMP)
    ctx.submit(g, f, cloud)
    st = ctx.stats()
    assert st["overflow"] > 0
    # the context stays usable and a big enough map/context is exact
    small_map = nm.Map(0.1, *W.CLAMP, leaf_capacity=2)
    ctx2 = nm.Context(len(cloud), f.max_time)
    ctx2.submit(small_map, f, cloud)
    assert ctx2.stats()["overflow"] > 0
    assert small_map.leaf_count() == 2
    g2 = nm.Map(0.1, *W.CLAMP)
    ctx.close()
    ctx2.submit(g2, f, cloud)
    omap = oracle.Map(*W.CLAMP)
    oracle.process_frame(omap, f, cloud)
    _check_maps(g2, omap)


def test_bad_arguments_raise(gpu):
    nm = gpu
    ctx = nm.Context(100, 50.0)
    g = nm.Map(0.1)
    f = nm.frame_from_pose([0, 0, 0], [1, 0, 0, 0], np.eye(3, dtype=np.float32),
                           0.1, 5.0, 0.85, -0.4, RD)
    with pytest.raises(nm.NanoMapError):
        ctx.cast(f, np.zeros((101, 3), np.float32))  # too many points
    big = f.copy()
    big.max_time = 60.0
    with pytest.raises(nm.NanoMapError):
        ctx.cast(big, np.zeros((10, 3), np.float32))  # beyond context range
    for lh, lm in ((float("inf"), -0.4), (0.85, float("nan")), (2e27, -0.4)):
        bad = f.copy()
        bad.logodds_hit, bad.logodds_miss = lh, lm
        with pytest.raises(nm.NanoMapError):
            ctx.cast(bad, np.zeros((10, 3), np.float32))  # unbounded log-odds
    with pytest.raises(nm.NanoMapError):
        nm.Map(0.1, -2.0, float("inf"))
    with pytest.raises(nm.NanoMapError) as e:
        ctx.update(g)  # no pending cast
    assert e.value.code == 6
    ctx.cast(f, np.ones((10, 3), np.float32))
    with pytest.raises(nm.NanoMapError):
        ctx.cast(f, np.ones((10, 3), np.float32))  # previous cast not folded
    ctx.update(g)


def _near_wall_cloud(w, h, seed):
    """A w x h depth image of a tilted plane 0.4-0.9 m in front of the
    sensor (optical frame: z forward), with depth noise and dropouts: 0.2 m
    voxels on the plane collect thousands of hits each, and rays that end
    in front of a voxel or pass its corner leave misses between them."""
    rng = np.random.default_rng(seed)
    u = (np.arange(w) + 0.5) / w * 2 - 1
    v = (np.arange(h) + 0.5) / h * 2 - 1
    U, V = np.meshgrid(u * np.tan(np.radians(43.5)), v * np.tan(np.radians(29)))
    z = 0.4 + 0.25 * (U + 1) + 0.1 * (V + 1)
    z = z * (1 + rng.normal(0, 0.02, z.shape))
    pts = np.stack([U * z, V * z, z], -1).reshape(-1, 3).astype(np.float32)
    pts[rng.random(len(pts)) < 0.03] = np.nan
    return pts


@pytest.mark.parametrize("rounding", [RD, RN])
@pytest.mark.parametrize("hit,miss", [(0.85, -0.4), (0.5, -0.25),
                                      (0.001, -2.0), (1.0, -1e-20)])
def test_many_hit_voxels_bit_exact(gpu, rounding, hit, miss):
    """Voxels with hundreds to thousands of hits in one frame, summed by the
    wavefront-parallel path, against the oracle's sequential sums.  The
    log-odds pairs cover a common fixed-point unit with steps of similar
    size, power-of-two steps (round-to-nearest ties in every binade), steps
    2^11 apart, and steps 2^66 apart (no common unit: the floating-point
    path)."""
    nm = gpu
    w, h = 320, 240
    F = np.array([[0, 0, 1], [-1, 0, 0], [0, -1, 0]], np.float32)
    ctx = nm.Context(w * h, 5.0 / 0.2 + 1)
    gmap = nm.Map(0.2, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    for k in range(2):
        f = nm.frame_from_pose([1.03 + 0.07 * k, 2.11, 0.97], [1, 0, 0, 0], F,
                               0.2, 5.0, hit, miss, rounding, w, h)
        temp = _run_frame(nm, ctx, gmap, omap, f, _near_wall_cloud(w, h, k))
        _, m, hh, _ = temp.export()
        print("hits per hit voxel: max %d, voxels > 256 hits: %d, mixed: %d"
              % (hh.max(), (hh > 256).sum(), ((hh > 256) & (m > 0)).sum()),
              flush=True)
        assert (hh > 256).sum() >= 10 and ((hh > 256) & (m > 0)).sum() >= 3
    _check_maps(gmap, omap)


def test_log_odds_and_rounding_change_between_frames(gpu):
    """One context and map over frames whose log-odds and rounding mode
    change from frame to frame: the context's table of miss-only sums must
    follow the frame (it is rebuilt only when they change)."""
    nm = gpu
    _, cam, frames = W.hall_walk_frames(nm, n_frames=5)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    settings = [(0.85, -0.4, RD), (0.7, -0.3, RD), (0.7, -0.3, RN),
                (0.7, -0.3, RN), (0.85, -0.4, RD)]
    for (pos, quat, cloud), (hit, miss, rounding) in zip(frames, settings):
        f = nm.frame_from_pose(pos, quat, nm.DEFAULT_F, 0.05, 10.0, hit, miss,
                               rounding, cam.width, cam.height)
        _run_frame(nm, ctx, gmap, omap, f, cloud)
        assert ctx.stats()["errors"] == 0
        _check_maps(gmap, omap)


def test_stage_timing_all_and_masked(gpu):
    """Stage timing: every stage of a frame timed once per frame, or only
    the stages asked for; results unchanged either way."""
    nm = gpu
    _, cam, frames = W.box_room_frame(nm)
    pos, quat, cloud = frames[0]
    f = W.frame(nm, pos, quat, 0.1, 5.0, RD, cam.width, cam.height)
    ctx = nm.Context(cam.width * cam.height, f.max_time)
    maps = []
    for stages in (None, ["accumulate"], ["discover", "update"]):
        m = nm.Map(0.1, *W.CLAMP)
        ctx.reset_timing()
        ctx.set_timing(True, stages=stages)
        for _ in range(3):
            ctx.submit(m, f, cloud)
        ctx.set_timing(False)
        st = ctx.stage_times()
        timed = {k for k, (ms, n) in st.items() if n}
        want = set(nm.Context.STAGES) - {"filter"} if stages is None else set(stages)
        assert timed == want, (stages, st)
        assert all(n == 3 and ms > 0 for k, (ms, n) in st.items() if k in want)
        maps.append(m.download())
    for o, v, mk in maps[1:]:
        assert np.array_equal(o, maps[0][0])
        assert np.array_equal(v.view(np.uint32), maps[0][1].view(np.uint32))
        assert np.array_equal(mk, maps[0][2])


def test_depth_images_bit_exact(gpu):
    # the hall frames as z-depth images (a third of the bytes of the xyz
    # clouds), back-projected inside the cast's ray set-up: counts and map
    # bit-exact against the oracle fed the host back-projection of the same
    # images; host and device images, with and without the voxel filter
    import torch
    nm = gpu
    _, cam, frames = W.hall_walk_frames(nm, n_frames=4)
    dc = nm.depth_camera(cam.width, cam.height, cam.hfov_deg, cam.vfov_deg)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    for k, (pos, quat, cloud) in enumerate(frames):
        depth = np.ascontiguousarray(cloud[:, 2])
        filt = nm.FILTER_VOXEL if k == 3 else nm.FILTER_NONE
        f = W.frame(nm, pos, quat, 0.05, 10.0, RD, cam.width, cam.height,
                    filt, box=W.hall_box(nm, RD) if filt else None)
        print("depth image %d: filter %d" % (k, filt), flush=True)
        src = torch.from_numpy(depth).cuda() if k % 2 else depth
        ctx.cast_depth(f, dc, src)
        xyz = nm.depth_to_cloud(dc, depth)
        temp = oracle.Temp()
        if filt:
            rays = oracle.voxel_filter(f, xyz)
            g_rays = ctx.filter_rays()
            assert np.array_equal(g_rays.view(np.uint32), rays.view(np.uint32))
            temp.cast_rays(f, rays)
        else:
            temp.cast_cloud(f, xyz)
        _check_counts(ctx, temp)
        ctx.update(gmap)
        omap.integrate(temp, RD)
    _check_maps(gmap, omap)
