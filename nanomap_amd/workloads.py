"""Seeded synthetic workloads (nm_scene_*) for tests, benchmarks, examples.

box_room_frame: 20x20x5 m room, 40 boxes (sides 0.3-2 m), one 160x120
      frame at 60x45 deg, 5 m range, 0.1 m voxels, noise sigma = 0.002 d^2.
hall_walk_frames: 50x50x10 m hall, 300 boxes + 200 cylinders (r 0.1-0.5 m), a
      100-frame random walk (0.1 m, 2 deg per frame), 640x480 at 87x58 deg,
      10 m range, 0.05 m voxels, 5% NaN dropouts.
agents_32: 32 agents in a 100x100x10 m room (1200 boxes + 800 cylinders), each
      with its own pose, sensor resolution from {160x120, 320x240, 640x480,
      848x480, 1280x720} (87x58 deg for 4:3, 90x60 deg for 16:9), range
      uniform 3-20 m, voxel size from {0.05, 0.1, 0.2} m, all from the seed;
      depth noise and dropouts as hall_walk_frames.
soft_walk: the box-room world mapped at 0.05 m from a 48-frame random
      walk of 320x240 frames (87x58 deg, 8 m range); 16 soft-march cameras at every
      third pose of the walk (320x240, 8 m: 160 samples 0.05 m apart).
Log-odds hit +0.85, miss -0.4, clamp [-2.0, 3.5], thresholds 0/0.
"""
import numpy as np

HIT = 0.85
MISS = -0.4
CLAMP = (-2.0, 3.5)


def box_room_frame(nm, seed=1):
    """One 160x120 frame at 0.1 m in a 20x20x5 m box room."""
    scene = nm.Scene(seed, (20.0, 20.0, 5.0), 40, 0.3, 2.0)
    pos, quat = scene.sample_pose(seed, 15.0, 0.3)
    cam = nm.camera(160, 120, 60.0, 45.0, 5.0, noise_k=0.002, dropout=0.0)
    cloud = scene.render(cam, pos, quat, seed)
    return scene, cam, [(pos, quat, cloud)]


def hall_scene(nm, seed=2):
    """50x50x10 m room, 640x480 camera at 87x58 deg, 10 m range."""
    scene = nm.Scene(seed, (50.0, 50.0, 10.0), 300, 0.3, 2.0, 200, 0.1, 0.5)
    cam = nm.camera(640, 480, 87.0, 58.0, 10.0, noise_k=0.002, dropout=0.05)
    return scene, cam


def hall_walk_frames(nm, n_frames=100, seed=2, first=0):
    scene, cam = hall_scene(nm, seed)
    pos, quat = scene.walk(seed, first + n_frames, 0.1, 2.0, 15.0, 0.3)
    out = []
    for k in range(first, first + n_frames):
        out.append((pos[k], quat[k], scene.render(cam, pos[k], quat[k], seed * 1000 + k)))
    return scene, cam, out


def frame(nm, pos, quat, res, rng, rounding, w=0, h=0, filt=0, box=None):
    return nm.frame_from_pose(pos, quat, nm.DEFAULT_F, res, rng, HIT, MISS,
                              rounding, w, h, filt, box=box)


def hall_box(nm, rounding):
    """The voxel filter's sensor-frame box of the hall camera (640x480, 58
    deg vertical FOV, 10 m; the half width follows the pixel aspect)."""
    return nm.frustum_box(640, 480, 58.0, 0.0, 10.0, rounding)


def sim_poses(nm, n_poses=64, seed=3):
    """Sensor-simulation poses in the box room (same world seed)."""
    scene = nm.Scene(1, (20.0, 20.0, 5.0), 40, 0.3, 2.0)
    out = []
    for k in range(n_poses):
        out.append(scene.sample_pose(seed * 1000 + k, 15.0, 0.3))
    return out


def box_room_map_leaves(nm, oracle_map=None):
    """(origins, values) of the box-room final map computed on the device."""
    _, cam, frames = box_room_frame(nm)
    pos, quat, cloud = frames[0]
    f = frame(nm, pos, quat, 0.1, 5.0, nm.ROUND_DOWN, cam.width, cam.height)
    ctx = nm.Context(cam.width * cam.height, f.max_time)
    m = nm.Map(0.1, *CLAMP)
    ctx.submit(m, f, cloud)
    o, v, _ = m.download()
    return o, v


AGENT_RESOLUTIONS = ((160, 120), (320, 240), (640, 480), (848, 480),
                    (1280, 720))
AGENT_VOXELS = (0.05, 0.1, 0.2)


def agents_scene(nm, seed=4):
    """100x100x10 m room with 1200 boxes and 800 cylinders."""
    return nm.Scene(seed, (100.0, 100.0, 10.0), 1200, 0.3, 2.0, 800, 0.1, 0.5)


def agents_32(nm, n_agents=32, seed=4, scene=None, frame_index=0):
    """Per agent: dict(width, height, hfov, vfov, range, res, pos, quat,
    cloud).  frame_index > 0 moves every agent one random-walk step per
    index (0.1 m, 2 deg), so consecutive batches are new frames."""
    scene = scene or agents_scene(nm, seed)
    rng = np.random.RandomState(seed)
    out = []
    for k in range(n_agents):
        w, h = AGENT_RESOLUTIONS[rng.randint(len(AGENT_RESOLUTIONS))]
        rmax = float(rng.uniform(3.0, 20.0))
        res = float(AGENT_VOXELS[rng.randint(len(AGENT_VOXELS))])
        hf, vf = (87.0, 58.0) if w * 3 == h * 4 else (90.0, 60.0)
        pos, quat = scene.walk(seed * 1000 + k, frame_index + 1, 0.1, 2.0,
                               15.0, 0.3)
        pos, quat = pos[frame_index], quat[frame_index]
        cam = nm.camera(w, h, hf, vf, rmax, noise_k=0.002, dropout=0.05)
        cloud = scene.render(cam, pos, quat, seed * 100000 + k * 100 +
                             frame_index)
        out.append(dict(width=w, height=h, hfov=hf, vfov=vf, range=rmax,
                        res=res, pos=pos, quat=quat, cloud=cloud))
    return scene, out


def agent_frame(nm, a, rounding):
    return nm.frame_from_pose(a["pos"], a["quat"], nm.DEFAULT_F, a["res"],
                              a["range"], HIT, MISS, rounding, a["width"],
                              a["height"])


def soft_walk(nm, n_walk=48, seed=5):
    """(scene, camera, positions, quaternions, clouds) of the soft_walk walk."""
    scene = nm.Scene(1, (20.0, 20.0, 5.0), 40, 0.3, 2.0)
    pos, quat = scene.walk(seed, n_walk, 0.1, 2.0, 15.0, 0.3)
    cam = nm.camera(320, 240, 87.0, 58.0, 8.0, noise_k=0.002, dropout=0.0)
    clouds = [scene.render(cam, pos[k], quat[k], seed * 1000 + k)
              for k in range(n_walk)]
    return scene, cam, pos, quat, clouds


def dense_grid(origins, values, pad=2):
    """Dense [nx, ny, nz] float32 log-odds over the leaves' bounding box
    (plus pad voxels), and the index of its voxel (0, 0, 0)."""
    o = np.asarray(origins, np.int64)
    lo = o.min(0) - pad
    hi = o.max(0) + 8 + pad
    shape = tuple(int(x) for x in (hi - lo))
    g = np.zeros(shape, np.float32)
    li = np.arange(512)
    off = np.stack([li >> 6, (li >> 3) & 7, li & 7], 1)
    ijk = (o[:, None, :] + off[None, :, :] - lo).reshape(-1, 3)
    g[ijk[:, 0], ijk[:, 1], ijk[:, 2]] = np.asarray(values, np.float32).reshape(-1)
    return g, lo
