"""The C++ host API (nanomap_amd/host/nanomap/...: Manager, map::Map,
config::Config, sensor::SensorData, Pose), driven by tests/cpp/test_host.

CPU: config and sensor files in the reference's text formats
(src/nanomap/config/Config.cpp) parse to the expected values.
GPU: Manager::insertPointCloud over several frames builds a map equal bit
for bit to the oracle's Handler::processPointCloudCPU on the same inputs,
with and without the voxel filter.
"""
import math
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle
from nanomap_amd import build as nb

DRIVER = nb.HOST_TEST

FRUSTUM = """#frustumconfig
Name:
{name}
Id:
{id}
HRes:
{w}
VRes:
{h}
VFOV:
{vfov}
Rate:
30
MaxRange:
{rmax}
MinRange:
0.1
ProbHit:
{ph}
ProbMiss:
{pm}
FrameTransform:
0.0 -1.0 0.0
0.0 0.0 1.0
1.0 0.0 0.0
#endconfig
"""

LASER = """#laserconfig
Name:
lidar
Id:
2
AngularHRes:
0.4
AngularVRes:
2.0
HFOV:
360
VFOV:
30
Rate:
10
MaxRange:
30
MinRange:
0.5
ProbHit:
0.8
ProbMiss:
0.35
FrameTransform:
1.0 0.0 0.0
0.0 1.0 0.0
0.0 0.0 1.0
#endconfig
"""


def _write_config(tmp, res, filter_type, sensors, precision=0, update_type=0):
    paths = []
    for i, text in enumerate(sensors):
        p = os.path.join(tmp, "sensor%d.txt" % i)
        with open(p, "w") as f:
            f.write(text)
        paths.append(p)
    cfg = os.path.join(tmp, "config.txt")
    with open(cfg, "w") as f:
        f.write("#config\nMappingRes:\n%s\nProbHitThres:\n0.5\nProbMissThres:\n"
                "0.5\nGridRes:\n%s\nNodeEdge:\n8\nFrustumAllocationFactor:\n0.4\n"
                "FilterType:\n%d\nUpdateType:\n%d\nPrecisionType:\n%d\n"
                "PublishSensor:\n1\n" % (res, res, filter_type, update_type,
                                         precision))
        for p in paths:
            f.write("Sensor:\n%s\n" % p)
        f.write("#endconfig\n")
    return cfg


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        nb.build_host_test()
    return DRIVER


def _logit(p):
    return np.float32(math.log(float(np.float32(p))) - math.log(1.0 - float(np.float32(p))))


def test_config_and_sensor_files_parse(driver, tmp_path):
    cfg = _write_config(str(tmp_path), 0.05, 1, [
        FRUSTUM.format(name="cam", id=1000, w=640, h=480, vfov=43, rmax=10,
                       ph=0.95, pm=0.3), LASER])
    out = subprocess.run([driver, "parse", cfg], capture_output=True,
                         text=True, check=True).stdout.split("\n")
    kv = dict(l.split(" ", 1) for l in out if l and not l.startswith("sensor "))
    assert float(kv["MappingRes"]) == np.float32(0.05)
    assert kv["FilterType"] == "1" and kv["NodeEdge"] == "8"
    assert kv["sensors"] == "2" and kv["lookup"] == "1"
    sens = [l.split() for l in out if l.startswith("sensor ")]
    cam = sens[0]
    assert cam[1] == "cam" and cam[3] == "1000" and cam[5] == "0"
    assert cam[7] == "640x480"
    # logit(p) in FP64, rounded to FP32
    assert np.float32(cam[19]) == _logit(0.95) and np.float32(cam[20]) == _logit(0.3)
    # FrameTransform: column i = file line i (Config.cpp:249-256)
    F = np.array([float(x) for x in cam[22:31]]).reshape(3, 3)
    assert np.array_equal(F, [[0, 0, 1], [-1, 0, 0], [0, 1, 0]])
    lidar = sens[1]
    # beams = (int)(FOV / angular resolution + 1) (LaserData.h:23-24)
    assert lidar[5] == "1" and lidar[7] == "%dx%d" % (int(360 / 0.4 + 1), int(30 / 2.0 + 1))
    assert int(kv["maxPCLSize"]) == max(640 * 480, int(360 / 0.4 + 1) * int(30 / 2.0 + 1))


def test_bad_config_is_reported(driver, tmp_path):
    p = os.path.join(str(tmp_path), "bad.txt")
    with open(p, "w") as f:
        f.write("#notaconfig\n")
    r = subprocess.run([driver, "parse", p], capture_output=True, text=True)
    assert r.returncode == 1 and "#config" in r.stderr


def _read_map(path):
    with open(path, "rb") as f:
        n = struct.unpack("<q", f.read(8))[0]
        o = np.frombuffer(f.read(n * 12), np.int32).reshape(n, 3)
        v = np.frombuffer(f.read(n * 2048), np.float32).reshape(n, 512)
        m = np.frombuffer(f.read(n * 64), np.uint32).reshape(n, 16)
    order = np.lexsort((o[:, 2], o[:, 1], o[:, 0]))
    return o[order], v[order], m[order]


@pytest.mark.gpu
@pytest.mark.parametrize("filter_type,precision,update_type,capacity",
                         [(0, 0, 0, 0), (1, 0, 0, 0), (1, 1, 0, 0),
                          (1, 2, 0, 0), (1, 3, 0, 0), (0, 0, 1, 0),
                          (0, 0, 0, 16), (1, 0, 1, 8)])
def test_manager_maps_match_the_oracle(gpu, driver, tmp_path, filter_type,
                                       precision, update_type, capacity):
    # The driver also checks map::Map's occupiedGrid() against leaves() and
    # occAccessor() reads against the leaf values and masks (exit code 5).
    # FilterType 1 with PrecisionType 0-3 selects the four voxel filters,
    # which drop points outside the frustum's sensor-frame box.  UpdateType
    # 1 (the reference's per-leaf BlockWorker fold, BlockWorker.cpp:95-176)
    # adds, clamps and thresholds exactly as UpdateType 0, so its map is the
    # same.  capacity > 0 starts the map with that many leaves: it grows
    # as the frames need, and no update is lost.
    nm = gpu
    res, rmax, w, h = 0.1, 5.0, 160, 120
    cfg = _write_config(str(tmp_path), res, filter_type, [
        FRUSTUM.format(name="cam", id=1, w=w, h=h, vfov=45, rmax=rmax,
                       ph=0.7, pm=0.4)], precision, update_type)
    box = nm.frustum_box(w, h, 45.0, 0.1, rmax, nm.ROUND_DOWN)
    scene = nm.Scene(1, (20.0, 20.0, 5.0), 40, 0.3, 2.0)
    pos, quat = scene.walk(8, 3, 0.1, 2.0, 15.0, 0.3)
    cam = nm.camera(w, h, 60.0, 45.0, rmax, noise_k=0.002, dropout=0.05)
    args = []
    # map::Map's clamps, logit(0.12) and logit(0.97) (OccupancyMap.h:41-42)
    omap = oracle.Map(float(np.float32(math.log(0.12) - math.log(1 - 0.12))),
                      float(np.float32(math.log(0.97) - math.log(1 - 0.97))),
                      0.0, 0.0)
    for k in range(3):
        cloud = scene.render(cam, pos[k], quat[k], 100 + k)
        path = os.path.join(str(tmp_path), "cloud%d.bin" % k)
        cloud.astype(np.float32).tofile(path)
        p, q = pos[k], quat[k]
        args.append("%s,%r,%r,%r,%r,%r,%r,%r" % (path, *[float(x) for x in p],
                                                 *[float(x) for x in q]))
        f = nm.frame_from_pose(p, q, nm.DEFAULT_F, res, rmax, float(_logit(0.7)),
                               float(_logit(0.4)), nm.ROUND_DOWN, w, h,
                               nm.FILTER_VOXEL + precision if filter_type == 1
                               else nm.FILTER_NONE, box=box)
        t = oracle.Temp()
        if filter_type == 1:
            t.cast_rays(f, oracle.voxel_filter(f, cloud, precision=precision))
        else:
            t.cast_cloud(f, cloud)
        omap.integrate(t, nm.ROUND_DOWN)
    out = os.path.join(str(tmp_path), "map.bin")
    print("host driver: 3 frames, filter", filter_type, "precision", precision,
          "update", update_type, "capacity", capacity, flush=True)
    extra = ["--capacity", str(capacity)] if capacity else []
    r = subprocess.run([driver, "map", cfg, "cam", str(w), str(h), out] +
                       extra + args, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    go, gv, gm = _read_map(out)
    oo, ov, omk = omap.export()
    assert go.shape == oo.shape and len(go) > 0
    assert np.array_equal(go, oo)
    assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(gm, omk)
    # the driver's map::Map::writeNanoVDB file (read back by the driver
    # through the .nvdb constructor) holds the oracle's map
    import nvdb_layout as L
    with open(out + ".nvdb", "rb") as f:
        _, grids = L.read_file(f.read())
    leaves = grids[0][2].leaves()
    no = np.array([x[0] for x in leaves], np.int32)
    order = np.lexsort((no[:, 2], no[:, 1], no[:, 0]))
    assert np.array_equal(no[order], oo)
    nv = np.stack([x[1] for x in leaves])[order]
    assert np.array_equal(nv.view(np.uint32), ov.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("filter_type", [0, 1])
def test_manager_laser_maps_match_the_oracle(gpu, driver, tmp_path,
                                             filter_type):
    # A #laserconfig sensor (901 x 16 beams, 30 m) through
    # Manager::insertPointCloud: the reference sends laser clouds down the
    # CPU path, which casts every point whatever the FilterType
    # (Handler.cpp:247-249, processPointCloudCPU), so the map is the
    # oracle's cast of every point of the cloud.  Clouds are beam
    # directions at ranges 0.3-34 m (some beyond the 30 m range, some NaN).
    nm = gpu
    res, rmax, w, h = 0.1, 30.0, 901, 16
    cfg = _write_config(str(tmp_path), res, filter_type, [LASER])
    rng = np.random.default_rng(11 + filter_type)
    az = np.radians(np.arange(w) * 0.4 - 180.0)
    el = np.radians(np.arange(h) * 2.0 - 15.0)
    A, E = np.meshgrid(az, el)
    dirs = np.stack([np.cos(E) * np.cos(A), np.cos(E) * np.sin(A), np.sin(E)],
                    -1).reshape(-1, 3)
    omap = oracle.Map(float(np.float32(math.log(0.12) - math.log(1 - 0.12))),
                      float(np.float32(math.log(0.97) - math.log(1 - 0.97))),
                      0.0, 0.0)
    args = []
    for k in range(3):
        r = rng.uniform(0.3, 34.0, size=(w * h, 1))
        cloud = (dirs * r).astype(np.float32)
        cloud[rng.random(w * h) < 0.05] = np.nan
        path = os.path.join(str(tmp_path), "scan%d.bin" % k)
        cloud.tofile(path)
        p = rng.uniform(-5.0, 5.0, size=3).astype(np.float32)
        q = rng.normal(size=4)
        q = (q / np.linalg.norm(q)).astype(np.float32)
        args.append("%s,%r,%r,%r,%r,%r,%r,%r" % (path, *[float(x) for x in p],
                                                 *[float(x) for x in q]))
        f = nm.frame_from_pose(p, q, np.eye(3, dtype=np.float32), res, rmax,
                               float(_logit(0.8)), float(_logit(0.35)),
                               nm.ROUND_DOWN, w, h, nm.FILTER_NONE)
        t = oracle.Temp()
        t.cast_cloud(f, cloud)
        omap.integrate(t, nm.ROUND_DOWN)
    out = os.path.join(str(tmp_path), "map.bin")
    print("host driver: 3 laser scans, filter", filter_type, flush=True)
    r = subprocess.run([driver, "map", cfg, "lidar", str(w), str(h), out] +
                       args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    go, gv, gm = _read_map(out)
    oo, ov, omk = omap.export()
    assert go.shape == oo.shape and len(go) > 1000
    assert np.array_equal(go, oo)
    assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(gm, omk)


AGENT = """#agentconfig
Spawn:
  Type:
{spawn}
Observations:
  NumObs:
80
  LenObs:
5
  CollisionDist:
0.3
Sensor:
  Name:
cam
  Position:
0 0 0
  Orientation:
1.0 0.0 0.0 0.0
#endconfig
"""


def _agent_config(tmp, cfg_path, snapshot, spawns):
    """Append SimGrid and Agent lines (one agent file per spawn) to a
    config written by _write_config."""
    lines = ["SimGrid:\n%s\n" % snapshot]
    for i, sp in enumerate(spawns):
        p = os.path.join(tmp, "agent%d.txt" % i)
        with open(p, "w") as f:
            f.write(AGENT.format(spawn="Random" if sp is None else
                                 "fixed\n  Position:\n%r %r %r\n  Orientation:\n"
                                 "0 0 0 0" % tuple(float(x) for x in sp)))
        lines.append("Agent:\n%s\n" % p)
    text = open(cfg_path).read().replace("#endconfig\n", "".join(lines) +
                                         "#endconfig\n")
    with open(cfg_path, "w") as f:
        f.write(text)


@pytest.mark.gpu
def test_sim_manager_32_agents_match_a_manager_loop(gpu, driver, tmp_path):
    # SimManager over 32 agents (16 fixed spawns along the hall walk, 16
    # random), two rounds of views: every agent's map equals, bit for bit,
    # the map of a loop of Manager::insertPointCloud calls on the same
    # simulated clouds.  The collision scan of agent 0 agrees with the
    # oracle's cast of its observation rays, and the seeded spawn is an
    # observed-free voxel of the grid.
    from nanomap_amd import workloads as W
    nm = gpu
    tmp = str(tmp_path)
    _, cam, frames = W.hall_walk_frames(nm, n_frames=3)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    sim_map = nm.Map(0.05, *W.CLAMP, hit_thr=0.0, miss_thr=0.0)
    for pos, quat, cloud in frames:
        ctx.submit(sim_map, W.frame(nm, pos, quat, 0.05, 10.0, nm.ROUND_DOWN),
                   cloud)
    snap = os.path.join(tmp, "hall.nmap")
    sim_map.save(snap)
    nvdb = os.path.join(tmp, "hall.nvdb")
    sim_map.write_nvdb(nvdb)
    o, v, _ = sim_map.download()
    sim_map.close()
    cfg = _write_config(tmp, 0.1, 0, [
        FRUSTUM.format(name="cam", id=1, w=160, h=120, vfov=45, rmax=5.0,
                       ph=0.7, pm=0.4)])
    with open(cfg) as f:
        text = f.read()
    with open(cfg, "w") as f:  # GridRes = the snapshot's voxel size
        f.write(text.replace("GridRes:\n0.1\n", "GridRes:\n0.05\n"))
    scene, _ = W.hall_scene(nm)
    walk, _ = scene.walk(2, 18, 0.1, 2.0, 15.0, 0.3)  # the hall walk's poses
    spawns = [None] + [walk[k] for k in range(3, 18)] + [None] * 16
    _agent_config(tmp, cfg, snap, spawns)
    print("SimManager: 32 agents, 2 rounds", flush=True)
    r = subprocess.run([driver, "sim", cfg, "2"], capture_output=True,
                       text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    # the same grid given as a NanoVDB file ("SimGrid:" naming a .nvdb):
    # the same agents, maps, collisions and spawn
    with open(cfg) as f:
        text = f.read()
    cfg_nvdb = os.path.join(tmp, "config_nvdb.txt")
    with open(cfg_nvdb, "w") as f:
        f.write(text.replace(snap, nvdb))
    print("SimManager: the grid from a .nvdb file", flush=True)
    r2 = subprocess.run([driver, "sim", cfg_nvdb, "2"], capture_output=True,
                        text=True, timeout=180)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    assert r2.stdout == r.stdout
    out = r.stdout.split("\n")
    summary = [l for l in out if l.startswith("agents ")][0].split()
    assert summary[1] == "32" and summary[-1] == "1"
    assert int(summary[5]) > 0 and int(summary[7]) > 0  # returns, leaves

    osim = oracle.SimGrid(o, v, 0.0)
    head = [l for l in out if l.startswith("rays ")][0].split()
    n_obs, coll = int(head[1]), np.float32(float.fromhex(head[2]))
    rays = np.array([[float.fromhex(x) for x in l.split()[1:]]
                     for l in out if l.startswith("ray ")], np.float32)
    assert rays.shape == (n_obs, 3) == (80, 3)
    assert np.allclose(np.linalg.norm(rays, axis=1), 1.0, atol=1e-5)
    pose0 = np.array([float.fromhex(x) for x in
                      [l for l in out if l.startswith("pose0 ")][0].split()[1:]],
                     np.float32)
    first = int([l for l in out if l.startswith("collision_step ")][0].split()[1])
    max_range = np.float32(coll * np.float32(0.05))
    eye_R = np.eye(3, dtype=np.float32)

    def hits(k):
        eye = pose0 + np.array([np.float32(0.05) * np.float32(k), 0, 0],
                               np.float32)
        return osim.cast(eye, eye_R, 0.05, n_obs, 1, 0.0, 0.0, 0.0,
                         float(max_range), rays=rays)[3]

    last = first if first >= 0 else 200
    # the moves the library found free are free for the oracle, and the
    # first it found colliding collides
    assert all(hits(k) == 0 for k in range(last))
    if first >= 0:
        assert hits(first) > 0

    sp = np.array([float.fromhex(x) for x in
                   [l for l in out if l.startswith("spawn ")][0].split()[1:]])
    ijk = np.floor(sp / 0.05).astype(np.int32)
    loaded = nm.Map.load(snap)
    val, act = loaded.query(ijk.reshape(1, 3))
    assert not act[0] and val[0] < 0.0
