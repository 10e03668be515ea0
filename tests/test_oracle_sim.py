"""CPU pins of the oracle's sensor-simulation cast (generateCloud.cu:20-82).

- The hierarchical NanoVDB-style HDDA (node dims 1/8/128/4096) finds the
  same first active voxel as a plain dim-1 DDA except at re-anchoring
  near-ties (> 99.99% of pixels).
- Its FP32 hit distance equals the analytic entry distance of the hit
  voxel (exact ray/AABB in FP64) within 1e-5 m for > 99.5% of hits and
  within 5e-4 m for all: the residue is FP32 accumulation of the crossing
  times, inherent to the reference's FP32 HDDA.
- Closed form: an axis-aligned ray toward a single active voxel hits it at
  the analytic entry distance of the voxel's face.
- An empty grid returns nothing.
"""
import numpy as np

import oracle


def _box_room_oracle_map():
    import nanomap_amd as nm
    from nanomap_amd import workloads as W
    _, cam, frames = W.box_room_frame(nm)
    pos, quat, cloud = frames[0]
    f = W.frame(nm, pos, quat, 0.1, 5.0, 1, cam.width, cam.height)
    om = oracle.Map(*W.CLAMP)
    oracle.process_frame(om, f, cloud)
    return om, f


def _rays(pos, R, Wd, Hd, hf, vf, res):
    """Per-pixel eye and unit direction of Sensor::getRay, in numpy FP32."""
    R = np.asarray(R, np.float32).reshape(3, 3)
    hh = np.float32(np.tan(vf * np.pi / 360))
    hw = np.float32(np.tan(hf * np.pi / 360))
    s0, s1 = np.float32(1.0 / Wd), np.float32(1.0 / Hd)
    T2 = (hw * R[:, 0] + hh * R[:, 1]) - R[:, 2]
    T0 = R[:, 0] * np.float32(2 * hw)
    T1 = R[:, 1] * np.float32(2 * hh)
    w, h = np.meshgrid(np.arange(Wd), np.arange(Hd))
    u = (w.astype(np.float32) * s0).reshape(-1, 1)
    v = (h.astype(np.float32) * s1).reshape(-1, 1)
    md = (u * T0 + v * T1) - T2
    n = np.sqrt((md[:, 0] * md[:, 0] + md[:, 1] * md[:, 1]) + md[:, 2] * md[:, 2])
    md = md / n[:, None]
    eye = (np.asarray(pos, np.float32) / np.float32(res)).astype(np.float64)
    return eye, md.astype(np.float64)


def test_hierarchical_matches_dim1_and_analytic(oracle_lib):
    import nanomap_amd as nm
    from nanomap_amd import workloads as W
    om, f = _box_room_oracle_map()
    o, v, _ = om.export()
    g = oracle.SimGrid(o, v, 0.0)
    assert g.active_count() > 100
    agree = total = 0
    errs = []
    Wd, Hd = 424, 240
    for pos, quat in W.sim_poses(nm, 6):
        fr = nm.frame_from_pose(pos, quat, nm.DEFAULT_F, 0.1, 15.0, 0.85, -0.4, 0)
        R = np.array(list(fr.rotation), np.float32)
        a_ijk, a_d, _, _ = g.cast(pos, R, 0.1, Wd, Hd, 90.0, 60.0, 0.1, 15.0, True)
        b_ijk, b_d, _, _ = g.cast(pos, R, 0.1, Wd, Hd, 90.0, 60.0, 0.1, 15.0, False)
        same = np.all(a_ijk == b_ijk, axis=1)
        agree += same.sum()
        total += len(same)
        eye, dirs = _rays(pos, R, Wd, Hd, 90.0, 60.0, 0.1)
        hit = a_d >= 0
        lo = a_ijk[hit].astype(np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            ta = (lo - eye[None, :]) / dirs[hit]
            tb = (lo + 1 - eye[None, :]) / dirs[hit]
        t_entry = np.nanmax(np.minimum(ta, tb), axis=1)
        errs.append(np.abs(t_entry * 0.1 - a_d[hit].astype(np.float64)))
    assert agree / total > 0.9999
    errs = np.concatenate(errs)
    assert len(errs) > 1000
    assert np.mean(errs <= 1e-5) > 0.995 and errs.max() < 5e-4


def test_axis_ray_closed_form(oracle_lib):
    # active voxel (30, 10, 10) in index space (res 0.1); a sensor at
    # x = 0.55 m looking down +x: entry at x = 3.0 m -> t = 30 - 5.5 voxels.
    # A second active voxel widens the root bbox: the ray is clipped to the
    # bbox and the voxel it starts in is never tested (generateCloud.cu:
    # 25-28), so a lone active voxel is invisible.
    o = np.array([[24, 8, 8], [8, 16, 16]], np.int32)
    v = np.zeros((2, 512), np.float32)
    v[0, (30 - 24) * 64 + (10 - 8) * 8 + (10 - 8)] = 1.0
    g1 = oracle.SimGrid(o[:1], v[:1], 0.0)
    v[1, (10 - 8) * 64 + (20 - 16) * 8 + (20 - 16)] = 1.0
    g = oracle.SimGrid(o, v, 0.0)
    # columns: optical x -> world +y, optical y -> world +z, optical z -> +x
    R = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], np.float32)
    # the reference ray has no half-pixel offset: pixel (W/2, H/2) of an
    # even-sized image looks exactly along optical z
    W = H = 8
    ijk, d, xyz, n = g.cast([0.55, 1.05, 1.05], R, 0.1, W, H, 10.0, 10.0, 0.0, 10.0)
    pix = 4 * W + 4
    assert tuple(ijk[pix]) == (30, 10, 10)
    assert abs(d[pix] - (3.0 - 0.55)) < 1e-5
    assert n >= 1
    _, _, _, n1 = g1.cast([0.55, 1.05, 1.05], R, 0.1, W, H, 10.0, 10.0, 0.0, 10.0)
    assert n1 == 0


def test_empty_grid_returns_nothing(oracle_lib):
    g = oracle.SimGrid(np.zeros((0, 3), np.int32), np.zeros((0, 512), np.float32))
    ijk, d, xyz, n = g.cast([0, 0, 0], np.eye(3, dtype=np.float32), 0.1, 16, 16,
                            60.0, 45.0, 0.1, 10.0)
    assert n == 0 and np.all(d == -1) and np.all(np.isnan(xyz))


def test_dense_grid_terminates(oracle_lib):
    """Regression: on a dense 0.05 m grid some rays pass exactly through an
    edge where two node boundaries meet; re-anchoring at ray(t) then
    alternates refine/coarsen with t frozen.  The restated HDDA suppresses
    coarsening after 8 steps at one t, so every ray terminates."""
    import nanomap_amd as nm
    from nanomap_amd import workloads as W
    _, cam, frames = W.hall_walk_frames(nm, n_frames=3)
    om = oracle.Map(*W.CLAMP)
    for pos, quat, cloud in frames:
        f = W.frame(nm, pos, quat, 0.05, 10.0, nm.ROUND_DOWN)
        oracle.process_frame(om, f, cloud)
    o, v, _ = om.export()
    g = oracle.SimGrid(o, v, 0.0)
    for pos, quat, _ in frames:
        fr = nm.frame_from_pose(pos, quat, nm.DEFAULT_F, 0.05, 10.0, 0.85, -0.4, 0)
        R = np.array(list(fr.rotation), np.float32)
        _, d, _, n = g.cast(pos, R, 0.05, 320, 240, 87.0, 58.0, 0.1, 10.0)
        assert n > 40000


def _wall_grid(axis, index, extent=64):
    """Active voxels on the plane coordinate[axis] == index over
    [-extent, extent) in the other two axes, plus one far voxel that keeps
    the root bounding box wide."""
    leaves = {}
    other = [a for a in range(3) if a != axis]
    for p in range(-extent, extent):
        for q in range(-extent, extent):
            v = [0, 0, 0]
            v[axis] = index
            v[other[0]], v[other[1]] = p, q
            key = tuple(c & ~7 for c in v)
            li = ((v[0] & 7) << 6) | ((v[1] & 7) << 3) | (v[2] & 7)
            leaves.setdefault(key, np.zeros(512, np.float32))[li] = 1.0
    far = (-200, -200, -200)
    leaves.setdefault(far, np.zeros(512, np.float32))[0] = 1.0
    o = np.array(list(leaves), np.int32)
    v = np.stack([leaves[k] for k in leaves])
    return oracle.SimGrid(o, v, 0.0)


def test_laser_beams_closed_form(oracle_lib):
    # 360 x 30 deg at 1 deg x 1.875 deg: beam w = 180 has yaw -180 + 180 = 0
    # exactly, row h = 8 pitch -15 + 8 * 1.875 = 0: direction (0, 0, 1) in
    # the sensor frame.  With R = I and a wall at z index 40 (res 0.1), a
    # sensor at z = 0.25 m sees it at 4.0 - 0.25 m; the point is (0, 0, d).
    # Beam w = 270 (yaw 90 deg) looks along +x: a wall at x index 30 from
    # x = 0.55 m is at 3.0 - 0.55 m.  Pitch +-15 deg beams over a floor:
    # d = h / sin(15 deg).
    W, H = 360, 16
    I = np.eye(3, dtype=np.float32)
    g = _wall_grid(2, 40)
    ijk, d, xyz, n = g.cast([0.05, 0.05, 0.25], I, 0.1, W, H, 360.0, 30.0,
                            0.0, 20.0, laser=True)
    pix = 8 * W + 180
    assert tuple(ijk[pix]) == (0, 0, 40)
    assert abs(d[pix] - 3.75) < 1e-5
    assert np.allclose(xyz[pix], [0, 0, d[pix]], atol=1e-6)
    # every return's point lies at its distance along a unit direction
    ok = d >= 0
    assert ok.sum() > 100
    assert np.allclose(np.linalg.norm(xyz[ok], axis=1), d[ok], rtol=1e-6)
    gx = _wall_grid(0, 30)
    ijk, d, _, _ = gx.cast([0.55, 0.05, 0.05], I, 0.1, W, H, 360.0, 30.0,
                           0.0, 20.0, laser=True)
    pix = 8 * W + 270
    assert tuple(ijk[pix]) == (30, 0, 0)
    assert abs(d[pix] - 2.45) < 1e-5
    # the floor below (y axis is "up" for the beam pitch): row 0 is pitch
    # -15 deg; a wall at y index -20 from y = 0.05 m, beam yaw 0
    # (voxel y = -20 spans [-2.0, -1.9) m: the beam enters its top face)
    gy = _wall_grid(1, -20, extent=128)
    ijk, d, _, _ = gy.cast([0.05, 0.05, 0.05], I, 0.1, W, H, 360.0, 30.0,
                           0.0, 20.0, laser=True)
    pix = 0 * W + 180
    assert ijk[pix][1] == -20
    expect = (0.05 + 1.9) / np.sin(np.radians(15.0))
    assert abs(d[pix] - expect) < 1e-4
    # ranges are distances: a max range below the wall gives no return
    _, d2, _, _ = g.cast([0.05, 0.05, 0.25], I, 0.1, W, H, 360.0, 30.0, 0.0,
                         3.5, laser=True)
    assert d2[8 * W + 180] == -1


def _hall_oracle_grid(n=3):
    import nanomap_amd as nm
    from nanomap_amd import workloads as W
    _, cam, frames = W.hall_walk_frames(nm, n_frames=n)
    om = oracle.Map(*W.CLAMP)
    for pos, quat, cloud in frames:
        f = W.frame(nm, pos, quat, 0.05, 10.0, nm.ROUND_DOWN)
        oracle.process_frame(om, f, cloud)
    o, v, _ = om.export()
    return oracle.SimGrid(o, v, 0.0), frames


def test_laser_hierarchical_matches_dim1(oracle_lib):
    # a 901 x 16 beam scanner (0.4 x 2 deg) with the sample frame transform
    # (beam y up = world z up) at the hall poses, over the 0.05 m hall map
    import nanomap_amd as nm
    g, frames = _hall_oracle_grid()
    agree = total = 0
    for pos, quat, _ in frames:
        fr = nm.frame_from_pose(pos, quat, nm.DEFAULT_F, 0.05, 30.0, 0.85,
                                -0.4, 0)
        R = np.array(list(fr.rotation), np.float32)
        a_ijk, a_d, _, na = g.cast(pos, R, 0.05, 901, 16, 360.0, 30.0, 0.5,
                                   30.0, True, laser=True)
        b_ijk, b_d, _, nb = g.cast(pos, R, 0.05, 901, 16, 360.0, 30.0, 0.5,
                                   30.0, False, laser=True)
        same = np.all(a_ijk == b_ijk, axis=1)
        agree += same.sum()
        total += len(same)
        assert na > 1000
    assert agree / total > 0.999


def test_explicit_rays_closed_form(oracle_lib):
    # explicit directions (type 2, the observation rays of SimManager's
    # collision query): the world ray is R times the normalised direction,
    # ranges are distances.  A wall at x index 30 seen from x = 0.55 m is at
    # 3.0 - 0.55 m along +x; -x sees nothing; an unnormalised (2, 0, 0) is
    # the same ray; with R taking sensor +z to world +x, (0, 0, 1) hits it too
    g = _wall_grid(0, 30)
    I = np.eye(3, dtype=np.float32)
    rays = np.array([[1, 0, 0], [-1, 0, 0], [2, 0, 0]], np.float32)
    ijk, d, xyz, n = g.cast([0.55, 0.05, 0.05], I, 0.1, 3, 1, 0.0, 0.0, 0.0,
                            20.0, rays=rays)
    assert tuple(ijk[0]) == (30, 0, 0) and abs(d[0] - 2.45) < 1e-5
    assert d[1] == -1 and np.isnan(xyz[1]).all()
    assert tuple(ijk[2]) == (30, 0, 0) and d[2] == d[0]
    assert n == 2
    assert np.allclose(xyz[0], [d[0], 0, 0], atol=1e-6)
    R = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], np.float32)
    ijk, d, _, _ = g.cast([0.55, 0.05, 0.05], R, 0.1, 1, 1, 0.0, 0.0, 0.0,
                          20.0, rays=np.array([[0, 0, 1]], np.float32))
    assert tuple(ijk[0]) == (30, 0, 0) and abs(d[0] - 2.45) < 1e-5
    # a range below the wall's distance: no return
    _, d, _, n = g.cast([0.55, 0.05, 0.05], I, 0.1, 1, 1, 0.0, 0.0, 0.0, 2.0,
                        rays=rays[:1])
    assert d[0] == -1 and n == 0


def test_explicit_rays_reproduce_the_laser(oracle_lib):
    # the laser's beams (sin u cos v, sin v, cos u cos v), u = -180 + w deg,
    # v = -15 + 1.875 h deg, given as explicit rays: the same first voxels
    # over the hall map (directions from numpy, so equal up to the last bit
    # of the beam tables)
    import nanomap_amd as nm
    g, frames = _hall_oracle_grid(2)
    W, H = 360, 16
    u = np.radians(-180.0 + np.arange(W) * 1.0)
    v = np.radians(-15.0 + np.arange(H) * 1.875)
    U, V = np.meshgrid(u, v)
    beams = np.stack([np.sin(U) * np.cos(V), np.sin(V), np.cos(U) * np.cos(V)],
                     -1).reshape(-1, 3).astype(np.float32)
    agree = total = 0
    for pos, quat, _ in frames:
        fr = nm.frame_from_pose(pos, quat, nm.DEFAULT_F, 0.05, 20.0, 0.85,
                                -0.4, 0)
        R = np.array(list(fr.rotation), np.float32)
        a_ijk, a_d, _, na = g.cast(pos, R, 0.05, W, H, 360.0, 30.0, 0.0,
                                   20.0, laser=True)
        b_ijk, b_d, _, nb = g.cast(pos, R, 0.05, W * H, 1, 0.0, 0.0, 0.0,
                                   20.0, rays=beams)
        same = np.all(a_ijk == b_ijk, axis=1)
        agree += same.sum()
        total += len(same)
        hit = same & (a_d >= 0)
        assert np.allclose(a_d[hit], b_d[hit], atol=1e-4)
        assert na > 500
    assert agree / total > 0.999
