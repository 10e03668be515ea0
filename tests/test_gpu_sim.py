"""GPU parity of the sensor-simulation cast (generateCloud) against the
oracle, and the simulated clouds fed back through the mapping path.

64 poses, 848x480 at 90x60 deg, 15 m range, through the 0.1 m occupancy
grid of the box room's final map thresholded at log-odds > 0.
First-hit ijk and the FP32 hit distance must equal the oracle's bit for
bit (the bar is ijk bit-exact and distance within 1e-5 m).
"""
import numpy as np
import pytest

import oracle
from nanomap_amd import workloads as W

pytestmark = pytest.mark.gpu


def _oracle_box_room_leaves(nm):
    _, cam, frames = W.box_room_frame(nm)
    pos, quat, cloud = frames[0]
    f = W.frame(nm, pos, quat, 0.1, 5.0, nm.ROUND_DOWN, cam.width, cam.height)
    om = oracle.Map(*W.CLAMP)
    oracle.process_frame(om, f, cloud)
    o, v, _ = om.export()
    return o, v


def _sensors(nm, poses, res, w, h, hf, vf, rmin, rmax):
    out = []
    for pos, quat in poses:
        fr = nm.frame_from_pose(pos, quat, nm.DEFAULT_F, res, rmax, 0.85, -0.4, 0)
        out.append((pos, np.array(list(fr.rotation), np.float32),
                    nm.sim_sensor(pos, list(fr.rotation), res, w, h, hf, vf,
                                  rmin, rmax)))
    return out


def test_sim_cast_64_poses_bit_exact(gpu):
    nm = gpu
    o, v = _oracle_box_room_leaves(nm)
    # the device map of the box room equals the oracle's (test_gpu_cast); build
    # the device grid from the device map and the oracle grid from its own
    gm_leaves = W.box_room_map_leaves(nm)
    assert np.array_equal(gm_leaves[0], o)
    gsim = nm.SimGrid(gm_leaves[0], gm_leaves[1], 0.0)
    osim = oracle.SimGrid(o, v, 0.0)
    assert gsim.info()["active"] == osim.active_count()
    poses = W.sim_poses(nm, 64)
    sens = _sensors(nm, poses, 0.1, 848, 480, 90.0, 60.0, 0.1, 15.0)
    print("sim cast: 64 x 848x480", flush=True)
    ijk, dist, xyz, n = gsim.cast([s for _, _, s in sens])
    ijk = ijk.cpu().numpy()
    dist = dist.cpu().numpy()
    xyz = xyz.cpu().numpy()
    P = 848 * 480
    total = 0
    for k, (pos, R, _) in enumerate(sens):
        oi, od, ox, on = osim.cast(pos, R, 0.1, 848, 480, 90.0, 60.0, 0.1, 15.0)
        sl = slice(k * P, (k + 1) * P)
        assert np.array_equal(ijk[sl], oi), k
        assert np.array_equal(dist[sl].view(np.uint32), od.view(np.uint32)), k
        assert np.array_equal(xyz[sl].view(np.uint32), ox.view(np.uint32)), k
        total += on
    assert n == total and total > 1000


def test_sim_on_dense_map_and_feedback(gpu):
    """A denser grid (three 640x480 frames at 0.05 m) and the simulated
    clouds fed back through nm_cast/nm_update_map, bit-exact end to end."""
    nm = gpu
    _, cam, frames = W.hall_walk_frames(nm, n_frames=3)
    om = oracle.Map(*W.CLAMP)
    for pos, quat, cloud in frames:
        f = W.frame(nm, pos, quat, 0.05, 10.0, nm.ROUND_DOWN)
        oracle.process_frame(om, f, cloud)
    o, v, _ = om.export()
    gsim = nm.SimGrid(o, v, 0.0)
    osim = oracle.SimGrid(o, v, 0.0)
    poses = [(p, q) for p, q, _ in frames]
    sens = _sensors(nm, poses, 0.05, 320, 240, 87.0, 58.0, 0.1, 10.0)
    print("dense sim cast: 3 x 320x240, active", gsim.info()["active"], flush=True)
    ijk, dist, xyz, n = gsim.cast([s for _, _, s in sens])
    xyz = xyz.cpu().numpy()
    ijk = ijk.cpu().numpy()
    print("dense sim cast done, returns", n, flush=True)
    P = 320 * 240
    ctx = nm.Context(P, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    for k, (pos, R, _) in enumerate(sens):
        oi, od, ox, _ = osim.cast(pos, R, 0.05, 320, 240, 87.0, 58.0, 0.1, 10.0)
        sl = slice(k * P, (k + 1) * P)
        assert np.array_equal(ijk[sl], oi)
        assert np.array_equal(xyz[sl].view(np.uint32), ox.view(np.uint32))
        f = W.frame(nm, poses[k][0], poses[k][1], 0.05, 10.0, nm.ROUND_DOWN)
        ctx.submit(gmap, f, xyz[sl])
        oracle.process_frame(omap, f, ox)
    go, gv, gm = gmap.download()
    oo, ov, omk = omap.export()
    assert np.array_equal(go, oo)
    assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(gm, omk)
    assert n > 10000


def test_sim_empty_grid(gpu):
    nm = gpu
    g = nm.SimGrid(np.zeros((0, 3), np.int32), np.zeros((0, 512), np.float32))
    s = nm.sim_sensor([0, 0, 0], np.eye(3), 0.1, 16, 16, 60.0, 45.0, 0.1, 10.0)
    ijk, dist, xyz, n = g.cast([s])
    assert n == 0 and bool((dist.cpu() == -1).all())


def test_laser_and_frustum_sensors_bit_exact_and_mapped(gpu):
    """Spherical scanners (the reference's laser getRay, Sensor.h:134-155)
    beside a frustum camera in one call over the dense hall grid: ijk,
    distance and point bit-exact against the oracle; the laser clouds then
    map through the unfiltered cast (the reference maps laser clouds on its
    CPU path, Handler.cpp:247-249) bit-exact against the oracle."""
    nm = gpu
    _, cam, frames = W.hall_walk_frames(nm, n_frames=3)
    om = oracle.Map(*W.CLAMP)
    for pos, quat, cloud in frames:
        f = W.frame(nm, pos, quat, 0.05, 10.0, nm.ROUND_DOWN)
        oracle.process_frame(om, f, cloud)
    o, v, _ = om.export()
    gsim = nm.SimGrid(o, v, 0.0)
    osim = oracle.SimGrid(o, v, 0.0)
    kinds = [(901, 16, 360.0, 30.0, True), (320, 240, 87.0, 58.0, False),
             (1024, 64, 360.0, 45.0, True), (361, 32, 180.0, 40.0, True)]
    sens = []
    for k, (w, h, hf, vf, laser) in enumerate(kinds):
        pos, quat, _ = frames[k % 3]
        fr = nm.frame_from_pose(pos, quat, nm.DEFAULT_F, 0.05, 20.0, 0.85,
                                -0.4, 0)
        R = np.array(list(fr.rotation), np.float32)
        sens.append((pos, quat, R, w, h, hf, vf, laser,
                     nm.sim_sensor(pos, R, 0.05, w, h, hf, vf, 0.3, 20.0,
                                   laser=laser)))
    print("laser + frustum sim cast:", [k[:2] for k in kinds], flush=True)
    ijk, dist, xyz, n = gsim.cast([s[-1] for s in sens])
    ijk, dist, xyz = ijk.cpu().numpy(), dist.cpu().numpy(), xyz.cpu().numpy()
    off = 0
    total = 0
    ctx = nm.Context(1024 * 64, 400.0)
    gmap = nm.Map(0.05, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    for pos, quat, R, w, h, hf, vf, laser, _ in sens:
        oi, od, ox, on = osim.cast(pos, R, 0.05, w, h, hf, vf, 0.3, 20.0,
                                   laser=laser)
        sl = slice(off, off + w * h)
        assert np.array_equal(ijk[sl], oi), (w, h)
        assert np.array_equal(dist[sl].view(np.uint32), od.view(np.uint32))
        assert np.array_equal(xyz[sl].view(np.uint32), ox.view(np.uint32))
        assert on > 1000
        total += on
        off += w * h
        if laser:
            f = W.frame(nm, pos, quat, 0.05, 20.0, nm.ROUND_DOWN)
            ctx.submit(gmap, f, xyz[sl])
            oracle.process_frame(omap, f, ox)
    assert n == total
    go, gv, gm = gmap.download()
    oo, ov, omk = omap.export()
    assert np.array_equal(go, oo) and len(go) > 100
    assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(gm, omk)


def test_sim_64_clouds_fed_back(gpu):
    """The 64 simulated 848x480 clouds of the box room (90x60 deg, 15 m)
    mapped at 0.1 m through the plain path, bit-exact against the oracle."""
    nm = gpu
    gm_leaves = W.box_room_map_leaves(nm)
    gsim = nm.SimGrid(gm_leaves[0], gm_leaves[1], 0.0)
    poses = W.sim_poses(nm, 64)
    sens = _sensors(nm, poses, 0.1, 848, 480, 90.0, 60.0, 0.1, 15.0)
    print("sim feedback: 64 x 848x480", flush=True)
    _, _, xyz, n = gsim.cast([s for _, _, s in sens])
    xyz = xyz.cpu().numpy()
    P = 848 * 480
    ctx = nm.Context(P, 150.0)
    gmap = nm.Map(0.1, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    rays = 0
    for k, (pos, quat) in enumerate(poses):
        f = W.frame(nm, pos, quat, 0.1, 15.0, nm.ROUND_DOWN, 848, 480)
        c = xyz[k * P:(k + 1) * P]
        ctx.submit(gmap, f, c)
        rays += ctx.stats()["rays"]
        oracle.process_frame(omap, f, c)
    # (a return may lie past the mapper's range: the frustum's range limit
    # is a z-depth, Sensor.h:129-130)
    assert 1000 < rays <= n
    go, gv, gm = gmap.download()
    oo, ov, omk = omap.export()
    assert np.array_equal(go, oo)
    assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(gm, omk)
