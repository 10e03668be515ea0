"""nanomap_amd -- MI355X-native rebuild of NanoMap's frustum-sensor path.

Python front end over the C ABI (include/nanomap_amd.h): a per-sensor
device Context, a device-resident Map, and seeded synthetic scenes.  The
C++ mirror of the reference API (Manager, Map, SensorData, Config, Pose)
lives in nanomap_amd/host; this module exposes the same pipeline to
Python for tests, benchmarks and the differentiable soft march.
"""
import ctypes

import numpy as np

from ._capi import (SENSOR_FRUSTUM, SENSOR_LASER, SENSOR_RAYS,
                    FILTER_NONE, FILTER_VOXEL, FILTER_VOXEL_COUNT,
                    FILTER_VOXEL_COUNT8, FILTER_VOXEL_HALF, ROUND_DOWN,
                    ROUND_NEAREST,
                    AgentFrame, BatchConfig, BatchTuning, Camera, CtxConfig,
                    DepthCamera,
                    Frame, FrameStats, NanoMapError, SimSensor, _ptr, check,
                    lib)

__all__ = ["Context", "Map", "Batch", "Scene", "Camera", "Frame", "FrameStats",
           "SimGrid", "SimSensor", "sim_sensor",
           "frame_from_pose", "frustum_box", "depth_camera",
           "depth_to_cloud", "fold", "NanoMapError", "ROUND_DOWN",
           "ROUND_NEAREST", "FILTER_NONE", "FILTER_VOXEL", "FILTER_VOXEL_HALF",
           "FILTER_VOXEL_COUNT", "FILTER_VOXEL_COUNT8", "DEFAULT_F"]

# FrameTransform of the reference's sample sensor (config/sensors/
# 640x480.txt:22-25, column i = file line i): optical z -> body x.
DEFAULT_F = np.array([[0.0, 0.0, 1.0],
                      [-1.0, 0.0, 0.0],
                      [0.0, 1.0, 0.0]], dtype=np.float32)


def fold(v, a, k, rounding):
    """round(v + a) applied k times in FP32 (host build of nm::fold)."""
    return lib().nm_fold_host(float(v), float(a), int(k), int(rounding))


def frustum_box(width, height, vfov_deg, min_range, max_range,
                rounding=ROUND_DOWN):
    """(box_min, box_max): the voxel filter's sensor-frame point bounds of a
    frustum sensor (FrustumData::calculateSensorBounds)."""
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    check(lib().nm_frustum_box(int(width), int(height), float(vfov_deg),
                               float(min_range), float(max_range),
                               int(rounding), _ptr(lo, ctypes.c_float),
                               _ptr(hi, ctypes.c_float)))
    return lo, hi


def frame_from_pose(pos, quat, F=DEFAULT_F, grid_res=0.1, max_range=5.0,
                    logodds_hit=0.85, logodds_miss=-0.4, rounding=ROUND_DOWN,
                    width=0, height=0, filter=FILTER_NONE, box=None):
    """Per-frame host set-up (R, eye, max_time) as the reference does it.
    box: (box_min, box_max) of the voxel filter, e.g. from frustum_box()."""
    f = Frame()
    p = np.ascontiguousarray(pos, dtype=np.float32)
    q = np.ascontiguousarray(quat, dtype=np.float32)
    Fm = np.ascontiguousarray(F, dtype=np.float32).reshape(9)
    check(lib().nm_frame_from_pose(_ptr(p, ctypes.c_float),
                                   _ptr(q, ctypes.c_float),
                                   _ptr(Fm, ctypes.c_float), grid_res,
                                   max_range, logodds_hit, logodds_miss,
                                   rounding, ctypes.byref(f)))
    f.width, f.height, f.filter = width, height, filter
    if box is not None:
        for a in range(3):
            f.box_min[a] = float(box[0][a])
            f.box_max[a] = float(box[1][a])
    return f


def depth_camera(width, height, hfov_deg, vfov_deg):
    """Pinhole depth camera with fx, fy from the FOVs and the principal
    point at the image centre (nm_depth_camera_from_fov)."""
    c = DepthCamera()
    check(lib().nm_depth_camera_from_fov(int(width), int(height),
                                         float(hfov_deg), float(vfov_deg),
                                         ctypes.byref(c)))
    return c


def depth_to_cloud(cam, depth):
    """Host back-projection of a depth image [H*W] into xyz [H*W, 3]: the
    points the device computes from the same image, bit for bit."""
    d = np.ascontiguousarray(depth, np.float32).reshape(-1)
    out = np.zeros((d.size, 3), np.float32)
    check(lib().nm_depth_to_cloud(ctypes.byref(cam),
                                  d.ctypes.data_as(ctypes.c_void_p),
                                  out.ctypes.data_as(ctypes.c_void_p)))
    return out


class Context:
    """Per-sensor device workspace (SensorAllocator + SensorBucket)."""

    def __init__(self, max_points, max_time, device=0, tile_capacity=0):
        cfg = CtxConfig(device, int(max_points), float(max_time),
                        int(tile_capacity))
        h = ctypes.c_void_p()
        check(lib().nm_ctx_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            lib().nm_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _cloud_args(cloud, step):
        """(pointer, n_points, step, on_device) for a numpy array, a torch
        tensor or a (device_ptr, n_points, step) tuple."""
        if isinstance(cloud, tuple):
            ptr, n, st = cloud
            return ctypes.c_void_p(int(ptr)), int(n), int(st), 1
        if hasattr(cloud, "data_ptr"):  # torch tensor
            n = cloud.shape[0]
            st = step or cloud.stride(0) * cloud.element_size()
            return ctypes.c_void_p(cloud.data_ptr()), n, st, int(cloud.is_cuda)
        a = np.ascontiguousarray(cloud, dtype=np.float32)
        a = a.reshape(a.shape[0], -1) if a.size else a.reshape(0, 3)
        return a.ctypes.data_as(ctypes.c_void_p), a.shape[0], a.shape[1] * 4, 0

    def cast(self, frame, cloud, step=0, stream=None):
        ptr, n, st, dev = self._cloud_args(cloud, step)
        self._keep = cloud
        check(lib().nm_cast(self.handle, ctypes.byref(frame), ptr, n, st, dev,
                            stream))

    def update(self, m, stream=None):
        check(lib().nm_update_map(self.handle, m.handle, stream))

    def submit(self, m, frame, cloud, step=0, stream=None):
        ptr, n, st, dev = self._cloud_args(cloud, step)
        self._keep = cloud
        check(lib().nm_frame_submit(self.handle, m.handle, ctypes.byref(frame),
                                    ptr, n, st, dev, stream))

    @staticmethod
    def _depth_args(depth):
        if hasattr(depth, "data_ptr"):  # torch tensor
            return ctypes.c_void_p(depth.data_ptr()), int(depth.is_cuda)
        a = np.ascontiguousarray(depth, dtype=np.float32)
        return a.ctypes.data_as(ctypes.c_void_p), 0

    def cast_depth(self, frame, cam, depth, stream=None):
        """nm_cast of a depth image (cam.width x cam.height floats)."""
        ptr, dev = self._depth_args(depth)
        self._keep = depth
        check(lib().nm_cast_depth(self.handle, ctypes.byref(frame),
                                  ctypes.byref(cam), ptr, dev, stream))

    def submit_depth(self, m, frame, cam, depth, stream=None):
        ptr, dev = self._depth_args(depth)
        self._keep = depth
        check(lib().nm_frame_submit_depth(self.handle, m.handle,
                                          ctypes.byref(frame),
                                          ctypes.byref(cam), ptr, dev,
                                          stream))

    def stats(self):
        s = FrameStats()
        check(lib().nm_frame_stats_get(self.handle, ctypes.byref(s)))
        return s.as_dict()

    def counts(self, max_leaves=1 << 20):
        """Per-voxel (miss, hit) counts of the last cast, as sorted arrays
        (ijk [n,3], miss [n], hit [n]) over touched voxels."""
        n = ctypes.c_int64()
        st = lib().nm_counts_download(self.handle, 0, ctypes.byref(n), None,
                                      None, None)
        if st not in (0, 3):
            check(st)
        L = n.value
        origins = np.zeros((max(L, 1), 3), np.int32)
        miss = np.zeros((max(L, 1), 512), np.uint32)
        hit = np.zeros((max(L, 1), 512), np.uint32)
        check(lib().nm_counts_download(self.handle, L, ctypes.byref(n),
                                       _ptr(origins, ctypes.c_int32),
                                       _ptr(miss, ctypes.c_uint32),
                                       _ptr(hit, ctypes.c_uint32)))
        return _leaves_to_voxels(origins[:L], miss[:L], hit[:L])

    STAGES = ("filter", "discover", "compact", "hitsort", "accumulate",
              "update")

    def set_timing(self, enable, stages=None):
        """Per-stage HIP event timing on or off; `stages` (names from
        STAGES) limits it to those stages."""
        if stages is None or not enable:
            check(lib().nm_ctx_set_timing(self.handle, int(bool(enable))))
        else:
            mask = 0
            for st in stages:
                mask |= 1 << self.STAGES.index(st)
            check(lib().nm_ctx_set_timing_mask(self.handle, mask))

    def reset_timing(self):
        check(lib().nm_ctx_reset_timing(self.handle))

    def stage_times(self):
        """{stage: (total_ms, launches)} since the last reset."""
        ms = np.zeros(6, np.float32)
        n = np.zeros(6, np.int64)
        check(lib().nm_ctx_stage_times(self.handle, _ptr(ms, ctypes.c_float),
                                       _ptr(n, ctypes.c_int64)))
        return {s: (float(ms[i]), int(n[i])) for i, s in enumerate(self.STAGES)}

    def filter_rays(self):
        n = ctypes.c_int()
        st = lib().nm_filter_rays_download(self.handle, 0, ctypes.byref(n),
                                           None)
        if st not in (0, 3):
            check(st)
        out = np.zeros((max(n.value, 1), 4), np.float32)
        check(lib().nm_filter_rays_download(self.handle, n.value,
                                            ctypes.byref(n),
                                            _ptr(out, ctypes.c_float)))
        return out[:n.value]


def _local_offsets():
    li = np.arange(512)
    return np.stack([li >> 6, (li >> 3) & 7, li & 7], axis=1).astype(np.int32)


def _leaves_to_voxels(origins, miss, hit):
    """Flatten leaf tiles to touched voxels sorted by (i, j, k)."""
    if origins.shape[0] == 0:
        return (np.zeros((0, 3), np.int32), np.zeros(0, np.uint32),
                np.zeros(0, np.uint32))
    ijk = (origins[:, None, :] + _local_offsets()[None, :, :]).reshape(-1, 3)
    m = miss.reshape(-1)
    h = hit.reshape(-1)
    sel = (m | h) != 0
    ijk, m, h = ijk[sel], m[sel], h[sel]
    order = np.lexsort((ijk[:, 2], ijk[:, 1], ijk[:, 0]))
    return ijk[order], m[order], h[order]


class Map:
    """Device-resident sparse log-odds map (map::Map)."""

    def __init__(self, grid_res, clamp_min=-2.0, clamp_max=3.5, hit_thr=0.0,
                 miss_thr=0.0, leaf_capacity=1 << 18, device=0):
        h = ctypes.c_void_p()
        check(lib().nm_map_create(device, grid_res, clamp_min, clamp_max,
                                  hit_thr, miss_thr, int(leaf_capacity),
                                  ctypes.byref(h)))
        self.handle = h
        self.grid_res = grid_res

    def close(self):
        if self.handle:
            lib().nm_map_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self, stream=None):
        check(lib().nm_map_clear(self.handle, stream))

    def leaf_count(self):
        n = ctypes.c_int64()
        check(lib().nm_map_leaf_count(self.handle, ctypes.byref(n)))
        return n.value

    def download(self):
        """(origins [L,3], values [L,512], masks [L,16]) sorted by origin."""
        L = self.leaf_count()
        o = np.zeros((max(L, 1), 3), np.int32)
        v = np.zeros((max(L, 1), 512), np.float32)
        m = np.zeros((max(L, 1), 16), np.uint32)
        check(lib().nm_map_download(self.handle, _ptr(o, ctypes.c_int32),
                                    _ptr(v, ctypes.c_float),
                                    _ptr(m, ctypes.c_uint32)))
        o, v, m = o[:L], v[:L], m[:L]
        order = np.lexsort((o[:, 2], o[:, 1], o[:, 0]))
        return o[order], v[order], m[order]

    def capacity(self):
        n = ctypes.c_int64()
        check(lib().nm_map_capacity(self.handle, ctypes.byref(n)))
        return n.value

    def reserve(self, leaf_capacity):
        check(lib().nm_map_reserve(self.handle, int(leaf_capacity)))

    def set_auto_grow(self, enable):
        check(lib().nm_map_set_auto_grow(self.handle, int(bool(enable))))

    def query(self, ijk, stream=None):
        """(values [n], active [n] bool) of voxels ijk [n, 3]: host arrays,
        or device tensors (int32 [n,3] -> returns device tensors)."""
        if hasattr(ijk, "data_ptr") and ijk.is_cuda:
            import torch
            n = ijk.shape[0]
            vals = torch.empty(n, dtype=torch.float32, device=ijk.device)
            act = torch.empty(n, dtype=torch.uint8, device=ijk.device)
            check(lib().nm_map_query(self.handle,
                                     ctypes.c_void_p(ijk.data_ptr()), n,
                                     ctypes.c_void_p(vals.data_ptr()),
                                     ctypes.c_void_p(act.data_ptr()), 1,
                                     stream))
            return vals, act.bool()
        a = np.ascontiguousarray(ijk, np.int32).reshape(-1, 3)
        vals = np.zeros(max(len(a), 1), np.float32)
        act = np.zeros(max(len(a), 1), np.uint8)
        check(lib().nm_map_query(self.handle, a.ctypes.data_as(ctypes.c_void_p),
                                 len(a), vals.ctypes.data_as(ctypes.c_void_p),
                                 act.ctypes.data_as(ctypes.c_void_p), 0,
                                 stream))
        return vals[:len(a)], act[:len(a)].astype(bool)

    def dense(self, lo, shape, fill=0.0, stream=None):
        """Device tensor [nx, ny, nz] float32 of the voxels lo + (i, j, k),
        `fill` where the leaf is absent (nm_map_dense)."""
        import torch
        nx, ny, nz = (int(x) for x in shape)
        out = torch.empty((nx, ny, nz), dtype=torch.float32, device="cuda")
        l = (ctypes.c_int32 * 3)(*(int(x) for x in lo))
        check(lib().nm_map_dense(self.handle, l, nx, ny, nz, float(fill),
                                 ctypes.c_void_p(out.data_ptr()), stream))
        return out

    def save(self, path):
        check(lib().nm_map_save(self.handle, str(path).encode()))

    @classmethod
    def load(cls, path, leaf_capacity=0, device=0):
        h = ctypes.c_void_p()
        check(lib().nm_map_load(device, str(path).encode(), int(leaf_capacity),
                                ctypes.byref(h)))
        return cls._wrap(h)

    @classmethod
    def _wrap(cls, h):
        m = cls.__new__(cls)
        m.handle = h
        res = ctypes.c_float()
        check(lib().nm_map_params(h, ctypes.byref(res), None, None, None,
                                  None))
        m.grid_res = res.value
        return m

    def write_nvdb(self, path, grid_name="occupancy"):
        """The map as a NanoVDB FloatGrid file (nm_map_write_nvdb)."""
        check(lib().nm_map_write_nvdb(self.handle, str(path).encode(),
                                      grid_name.encode()))

    @classmethod
    def load_nvdb(cls, path, grid_name="", clamp_min=-2.0, clamp_max=3.5,
                  hit_thr=0.0, miss_thr=0.0, leaf_capacity=0, device=0):
        """A map from a FloatGrid of a .nvdb file (nm_map_load_nvdb)."""
        h = ctypes.c_void_p()
        check(lib().nm_map_load_nvdb(device, str(path).encode(),
                                     grid_name.encode(), clamp_min, clamp_max,
                                     hit_thr, miss_thr, int(leaf_capacity),
                                     ctypes.byref(h)))
        return cls._wrap(h)


def _leaf_arrays(origins, values, masks):
    o = np.ascontiguousarray(origins, np.int32).reshape(-1, 3)
    v = np.ascontiguousarray(values, np.float32).reshape(-1, 512)
    k = np.ascontiguousarray(masks, np.uint32).reshape(-1, 16)
    if not (len(o) == len(v) == len(k)):
        raise ValueError("origins, values and masks need one row per leaf")
    return o, v, k


def nvdb_build(origins, values, masks, voxel_size, grid_name="occupancy"):
    """NanoVDB FloatGrid buffer (bytes) of host leaves (nm_nvdb_build)."""
    o, v, k = _leaf_arrays(origins, values, masks)
    n = ctypes.c_uint64(0)
    args = (o.ctypes.data, v.ctypes.data, k.ctypes.data, len(o),
            float(voxel_size), grid_name.encode())
    check(lib().nm_nvdb_build(*args, None, ctypes.byref(n)))
    buf = (ctypes.c_ubyte * n.value)()
    check(lib().nm_nvdb_build(*args, buf, ctypes.byref(n)))
    return bytes(buf)


def nvdb_write(path, origins, values, masks, voxel_size,
               grid_name="occupancy"):
    """Host leaves as a .nvdb file (nm_nvdb_write)."""
    o, v, k = _leaf_arrays(origins, values, masks)
    check(lib().nm_nvdb_write(str(path).encode(), o.ctypes.data,
                              v.ctypes.data, k.ctypes.data, len(o),
                              float(voxel_size), grid_name.encode()))


def nvdb_read(path, grid_name=""):
    """(origins, values, masks, voxel_size) of a .nvdb FloatGrid
    (nm_nvdb_read)."""
    n = ctypes.c_int64(0)
    dx = ctypes.c_double(0)
    check(lib().nm_nvdb_read(str(path).encode(), grid_name.encode(),
                             ctypes.byref(n), None, None, None,
                             ctypes.byref(dx)))
    o = np.zeros((n.value, 3), np.int32)
    v = np.zeros((n.value, 512), np.float32)
    k = np.zeros((n.value, 16), np.uint32)
    if n.value:
        check(lib().nm_nvdb_read(str(path).encode(), grid_name.encode(),
                                 ctypes.byref(n), o.ctypes.data,
                                 v.ctypes.data, k.ctypes.data,
                                 ctypes.byref(dx)))
    return o, v, k, dx.value


class Batch:
    """Many agents' frames in one graph-captured launch sequence
    (include/nanomap_amd_batch.h).  Frames are (frame, cloud, map) with the
    cloud device-resident: a torch CUDA tensor [n, 3|4] float32 or a
    (device_ptr, n_points, step) tuple."""

    def __init__(self, max_frames, max_points, max_time, tile_capacity,
                 device=0):
        cfg = BatchConfig(device, int(max_frames), int(max_points),
                          float(max_time), int(tile_capacity))
        h = ctypes.c_void_p()
        check(lib().nm_batch_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.n = 0

    def close(self):
        if self.handle:
            lib().nm_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _table(frames):
        arr = (AgentFrame * len(frames))()
        keep = []
        for i, (fr, cloud, m) in enumerate(frames):
            ptr, n, st, dev = Context._cloud_args(cloud, 0)
            if not dev:
                raise NanoMapError(1, "Batch: clouds must be on the device")
            arr[i].frame = fr
            arr[i].cloud = ptr
            arr[i].n_points = n
            arr[i].step = st
            arr[i].map = m.handle
            keep.append(cloud)
        return arr, keep

    def submit(self, frames, stream=None):
        arr, keep = self._table(frames)
        self._keep = keep
        check(lib().nm_batch_submit(self.handle, arr, len(frames), stream))
        self.n = len(frames)

    def stats(self):
        out = (FrameStats * self.n)()
        check(lib().nm_batch_stats(self.handle, out, self.n))
        return [o.as_dict() for o in out]

    def tune(self, frames, reps=2, stream=None):
        """Autotune the cast kernels on these frames; returns (tuning dict,
        ms per batch of the fastest candidate)."""
        arr, keep = self._table(frames)
        best = BatchTuning()
        ms = ctypes.c_float()
        check(lib().nm_batch_tune(self.handle, arr, len(frames), int(reps),
                                  stream, ctypes.byref(best),
                                  ctypes.byref(ms)))
        return best.as_dict(), ms.value

    def set_tuning(self, layout=0, queue=0, grid_mult=4, lds_slots=4096):
        t = BatchTuning(int(layout), int(queue), int(grid_mult),
                        int(lds_slots))
        check(lib().nm_batch_set_tuning(self.handle, ctypes.byref(t)))

    def get_tuning(self):
        t = BatchTuning()
        check(lib().nm_batch_get_tuning(self.handle, ctypes.byref(t)))
        return t.as_dict()

    def set_graph(self, enable):
        check(lib().nm_batch_set_graph(self.handle, int(bool(enable))))


class Scene:
    """Seeded box + cylinder room (nm_scene_*)."""

    def __init__(self, seed, world, n_boxes, box_min=0.3, box_max=2.0,
                 n_cylinders=0, cyl_rmin=0.1, cyl_rmax=0.5):
        h = ctypes.c_void_p()
        w = (ctypes.c_double * 3)(*world)
        check(lib().nm_scene_create(seed, w, n_boxes, box_min, box_max,
                                    n_cylinders, cyl_rmin, cyl_rmax,
                                    ctypes.byref(h)))
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                lib().nm_scene_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def render(self, cam, pos, quat, seed):
        out = np.zeros((cam.height * cam.width, 3), np.float32)
        p = np.ascontiguousarray(pos, dtype=np.float32)
        q = np.ascontiguousarray(quat, dtype=np.float32)
        check(lib().nm_scene_render(self.handle, ctypes.byref(cam),
                                    _ptr(p, ctypes.c_float),
                                    _ptr(q, ctypes.c_float), seed,
                                    _ptr(out, ctypes.c_float)))
        return out

    def sample_pose(self, seed, max_pitch_deg=15.0, clearance=0.3):
        p = np.zeros(3, np.float32)
        q = np.zeros(4, np.float32)
        check(lib().nm_scene_sample_pose(self.handle, seed, max_pitch_deg,
                                         clearance, _ptr(p, ctypes.c_float),
                                         _ptr(q, ctypes.c_float)))
        return p, q

    def walk(self, seed, n, step_m=0.1, step_deg=2.0, max_pitch_deg=15.0,
             clearance=0.3):
        p = np.zeros((n, 3), np.float32)
        q = np.zeros((n, 4), np.float32)
        check(lib().nm_scene_walk(self.handle, seed, n, step_m, step_deg,
                                  max_pitch_deg, clearance,
                                  _ptr(p, ctypes.c_float),
                                  _ptr(q, ctypes.c_float)))
        return p, q


def camera(width, height, hfov_deg, vfov_deg, max_range, noise_k=0.002,
           dropout=0.0, F=DEFAULT_F):
    c = Camera()
    c.width, c.height = width, height
    c.hfov_deg, c.vfov_deg = hfov_deg, vfov_deg
    c.max_range, c.noise_k, c.dropout = max_range, noise_k, dropout
    for i, x in enumerate(np.asarray(F, np.float32).reshape(9)):
        c.frame[i] = float(x)
    return c


def sim_sensor(pos, rotation, grid_res, width, height, hfov_deg, vfov_deg,
               min_range, max_range, laser=False, rays=None):
    """A simulated sensor: a frustum camera of width x height pixels, with
    laser=True a spherical scanner of width x height beams over hfov x vfov
    degrees, or with rays=[n, 3] n explicit directions (width n, height 1)
    (include/nanomap_amd_sim.h).  The sensor keeps a reference to rays."""
    s = SimSensor()
    s.type = SENSOR_LASER if laser else SENSOR_FRUSTUM
    if rays is not None:
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 3)
        s.type = SENSOR_RAYS
        s.rays = r.ctypes.data
        s._rays = r
        width, height = r.shape[0], 1
    for k in range(3):
        s.position[k] = float(pos[k])
    for k, x in enumerate(np.asarray(rotation, np.float32).reshape(9)):
        s.rotation[k] = float(x)
    s.grid_res = grid_res
    s.width, s.height = width, height
    s.hfov_deg, s.vfov_deg = hfov_deg, vfov_deg
    s.min_range, s.max_range = min_range, max_range
    return s


class SimGrid:
    """Device occupancy grid for the sensor-simulation cast (generateCloud)."""

    def __init__(self, origins=None, values=None, threshold=0.0, device=0,
                 from_map=None):
        h = ctypes.c_void_p()
        if from_map is not None:
            check(lib().nm_simgrid_from_map(from_map.handle, threshold,
                                            ctypes.byref(h)))
        else:
            o = np.ascontiguousarray(origins, np.int32).reshape(-1, 3)
            v = np.ascontiguousarray(values, np.float32).reshape(-1, 512)
            check(lib().nm_simgrid_from_leaves(device, _ptr(o, ctypes.c_int32),
                                               _ptr(v, ctypes.c_float),
                                               o.shape[0], threshold,
                                               ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().nm_simgrid_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        a = ctypes.c_int64()
        bb = np.zeros(6, np.int32)
        cn = np.zeros(3, np.int32)
        check(lib().nm_simgrid_info(self.handle, ctypes.byref(a),
                                    _ptr(bb, ctypes.c_int32),
                                    _ptr(cn, ctypes.c_int32)))
        return {"active": a.value, "bbox": bb.tolist(), "nodes": cn.tolist()}

    def cast(self, sensors, stream=None, static=False):
        """Simulate the sensors; returns device torch tensors
        (ijk [N,3] int32, dist [N] f32, xyz [N,3] f32) and the return
        count, N = total pixels.  static=True takes the allocation-free
        entry point (nm_sim_cast_static: fixed per-GPU descriptor buffer)."""
        import torch
        if isinstance(sensors, SimSensor):
            sensors = [sensors]
        arr = (SimSensor * len(sensors))(*sensors)
        N = sum(s.width * s.height for s in sensors)
        ijk = torch.empty((N, 3), dtype=torch.int32, device="cuda")
        dist = torch.empty(N, dtype=torch.float32, device="cuda")
        xyz = torch.empty((N, 3), dtype=torch.float32, device="cuda")
        n = ctypes.c_int()
        fn = lib().nm_sim_cast_static if static else lib().nm_sim_cast
        check(fn(self.handle, arr, len(sensors),
                 ctypes.c_void_p(ijk.data_ptr()),
                 ctypes.c_void_p(dist.data_ptr()),
                 ctypes.c_void_p(xyz.data_ptr()), ctypes.byref(n), stream))
        return ijk, dist, xyz, n.value
