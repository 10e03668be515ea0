"""Python binding of the CPU oracle (oracle/libnm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, always as the checker.  The product
(nanomap_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnm_oracle.so")

_P = ctypes.c_void_p
_F = ctypes.c_float
_I = ctypes.c_int
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PU32 = ctypes.POINTER(ctypes.c_uint32)
_PF = ctypes.POINTER(ctypes.c_float)
_PD = ctypes.POINTER(ctypes.c_double)
_D = ctypes.c_double

SIGNATURES = {
    "or_temp_create": (_P, []),
    "or_temp_destroy": (None, [_P]),
    "or_temp_clear": (None, [_P]),
    "or_cast_cloud": (_I, [_P, _P, _I, _I, _PF, _PD, _F, _F, _F, _F, _I]),
    "or_cast_rays": (_I, [_P, _PF, _I, _PD, _F, _F, _F, _I]),
    "or_temp_size": (ctypes.c_int64, [_P]),
    "or_temp_export": (None, [_P, _PI32, _PU32, _PU32, _PF]),
    "or_temp_visits": (ctypes.c_uint64, [_P]),
    "or_map_create": (_P, [_F, _F, _F, _F]),
    "or_map_destroy": (None, [_P]),
    "or_map_integrate": (None, [_P, _P, _I]),
    "or_map_leaf_count": (ctypes.c_int64, [_P]),
    "or_map_export": (None, [_P, _PI32, _PF, _PU32]),
    "or_map_get": (_I, [_P, _I, _I, _I, _PF, ctypes.POINTER(_I)]),
    "or_dda_trace": (_I, [_PD, _PF, _F, _F, _I, _PI32, _I,
                          ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "or_fold": (_F, [_F, _F, ctypes.c_uint32, _I]),
    "or_voxel_filter": (_I, [_P, _I, _I, _PF, _PD, _F, _F, _F, _PF, _PF, _I,
                             _I]),
    "or_simgrid_build": (_P, [_PI32, _PF, ctypes.c_int64, _F]),
    "or_simgrid_destroy": (None, [_P]),
    "or_simgrid_active_count": (ctypes.c_int64, [_P]),
    "or_sim_cast": (_I, [_P, _PF, _PF, _F, _I, _I, _F, _F, _F, _F, _I, _PF,
                         _I, _PI32, _PF, _PF]),
    "or_soft_forward": (None, [_PD, _I, _I, _I, _PD, _D, _D, _PD, _I, _I, _I,
                               _D, _D, _D, _I, _PD]),
    "or_soft_backward": (None, [_PD, _I, _I, _I, _PD, _D, _D, _PD, _I, _I,
                                _I, _D, _D, _D, _I, _PD, _PD, _PD]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _frame_args(frame):
    R = np.array(list(frame.rotation), np.float32)
    eye = np.array(list(frame.eye), np.float64)
    return R, eye


class Temp:
    """Temporary grid of one or more casts (Handler::populateTempGrid)."""

    def __init__(self):
        self.h = lib().or_temp_create()

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_temp_destroy(self.h)
            self.h = None

    def cast_cloud(self, frame, cloud):
        a = np.ascontiguousarray(cloud, dtype=np.float32)
        a = a.reshape(a.shape[0], -1) if a.size else a.reshape(0, 3)
        R, eye = _frame_args(frame)
        return lib().or_cast_cloud(self.h, a.ctypes.data_as(_P), a.shape[0],
                                   a.shape[1] * 4, _p(R, ctypes.c_float),
                                   _p(eye, ctypes.c_double), frame.grid_res,
                                   frame.max_time, frame.logodds_hit,
                                   frame.logodds_miss, frame.rounding)

    def cast_rays(self, frame, rays4):
        r = np.ascontiguousarray(rays4, dtype=np.float32).reshape(-1, 4)
        _, eye = _frame_args(frame)
        return lib().or_cast_rays(self.h, _p(r, ctypes.c_float), r.shape[0],
                                  _p(eye, ctypes.c_double), frame.max_time,
                                  frame.logodds_hit, frame.logodds_miss,
                                  frame.rounding)

    def visits(self):
        return int(lib().or_temp_visits(self.h))

    def export(self):
        """(ijk [n,3], miss [n], hit [n], value [n]) sorted by (i, j, k)."""
        n = int(lib().or_temp_size(self.h))
        ijk = np.zeros((max(n, 1), 3), np.int32)
        miss = np.zeros(max(n, 1), np.uint32)
        hit = np.zeros(max(n, 1), np.uint32)
        val = np.zeros(max(n, 1), np.float32)
        lib().or_temp_export(self.h, _p(ijk, ctypes.c_int32),
                             _p(miss, ctypes.c_uint32), _p(hit, ctypes.c_uint32),
                             _p(val, ctypes.c_float))
        ijk, miss, hit, val = ijk[:n], miss[:n], hit[:n], val[:n]
        order = np.lexsort((ijk[:, 2], ijk[:, 1], ijk[:, 0]))
        return ijk[order], miss[order], hit[order], val[order]


class Map:
    """Log-odds map (Handler::integrateTempGrid into map::Map)."""

    def __init__(self, clamp_min=-2.0, clamp_max=3.5, hit_thr=0.0,
                 miss_thr=0.0):
        self.h = lib().or_map_create(clamp_min, clamp_max, hit_thr, miss_thr)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_map_destroy(self.h)
            self.h = None

    def integrate(self, temp, rounding):
        lib().or_map_integrate(self.h, temp.h, rounding)

    def leaf_count(self):
        return int(lib().or_map_leaf_count(self.h))

    def export(self):
        L = self.leaf_count()
        o = np.zeros((max(L, 1), 3), np.int32)
        v = np.zeros((max(L, 1), 512), np.float32)
        m = np.zeros((max(L, 1), 16), np.uint32)
        lib().or_map_export(self.h, _p(o, ctypes.c_int32),
                            _p(v, ctypes.c_float), _p(m, ctypes.c_uint32))
        return o[:L], v[:L], m[:L]

    def get(self, i, j, k):
        v = ctypes.c_float()
        a = ctypes.c_int()
        ok = lib().or_map_get(self.h, i, j, k, ctypes.byref(v), ctypes.byref(a))
        return bool(ok), v.value, bool(a.value)


def process_frame(omap, frame, cloud):
    """Handler::processPointCloudCPU: populate a temp grid, integrate it."""
    t = Temp()
    t.cast_cloud(frame, cloud)
    omap.integrate(t, frame.rounding)
    return t


def dda_trace(eye, d, res, max_time, rounding, max_out=100000):
    """Voxel sequence of one ray: ([n,3] voxels, hit flag) or None."""
    e = np.ascontiguousarray(eye, np.float64)
    dv = np.ascontiguousarray(d, np.float32)
    out = np.zeros((max_out, 3), np.int32)
    n = ctypes.c_int()
    h = ctypes.c_int()
    ok = lib().or_dda_trace(_p(e, ctypes.c_double), _p(dv, ctypes.c_float),
                            res, max_time, rounding, _p(out, ctypes.c_int32),
                            max_out, ctypes.byref(n), ctypes.byref(h))
    if not ok:
        return None
    return out[:min(n.value, max_out)], bool(h.value)


def fold(v, a, k, rounding):
    return lib().or_fold(float(v), float(a), int(k), int(rounding))


def filter_bounds(frame):
    """(max_range, box [6] or None) of a frame's voxel filter: the frame's
    box_min/box_max unless all six are zero, and its max_range (metres;
    max_time * grid_res when 0 or absent)."""
    mr = float(getattr(frame, "max_range", 0.0) or 0.0)
    if mr == 0.0:
        mr = float(np.float32(np.float32(frame.max_time) *
                              np.float32(frame.grid_res)))
    lo = list(getattr(frame, "box_min", (0.0, 0.0, 0.0)))
    hi = list(getattr(frame, "box_max", (0.0, 0.0, 0.0)))
    box = np.array(lo + hi, np.float32)
    return mr, (None if not np.any(box) else box)


def voxel_filter(frame, cloud, max_rays=None, precision=None):
    """Rays of the voxel filter; precision 0-3 as PrecisionType (default:
    from frame.filter, NM_FILTER_VOXEL.. -> 0..3)."""
    if precision is None:
        precision = max(int(getattr(frame, "filter", 1)) - 1, 0)
    a = np.ascontiguousarray(cloud, dtype=np.float32)
    a = a.reshape(a.shape[0], -1) if a.size else a.reshape(0, 3)
    R, eye = _frame_args(frame)
    max_range, box = filter_bounds(frame)
    max_rays = max_rays or a.shape[0]
    out = np.zeros((max(max_rays, 1), 4), np.float32)
    n = lib().or_voxel_filter(a.ctypes.data_as(_P), a.shape[0], a.shape[1] * 4,
                              _p(R, ctypes.c_float), _p(eye, ctypes.c_double),
                              frame.grid_res, frame.max_time, max_range,
                              None if box is None else _p(box, ctypes.c_float),
                              _p(out, ctypes.c_float), max_rays, precision)
    return out[:min(n, max_rays)]


class SimGrid:
    """Occupancy grid of the sensor-simulation cast (value > threshold)."""

    def __init__(self, origins, values, threshold=0.0):
        o = np.ascontiguousarray(origins, np.int32).reshape(-1, 3)
        v = np.ascontiguousarray(values, np.float32).reshape(-1, 512)
        self.h = lib().or_simgrid_build(_p(o, ctypes.c_int32),
                                        _p(v, ctypes.c_float), o.shape[0],
                                        threshold)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_simgrid_destroy(self.h)
            self.h = None

    def active_count(self):
        return int(lib().or_simgrid_active_count(self.h))

    def cast(self, pos, R, res, width, height, hfov, vfov, min_range,
             max_range, hierarchical=True, laser=False, rays=None):
        """(ijk [P,3], dist [P], xyz [P,3], returns) of one frustum, of one
        spherical laser scanner (width x height beams), or of explicit ray
        directions (rays [width, 3], height 1)."""
        rr = None if rays is None else np.ascontiguousarray(rays, np.float32)
        P = width * height
        ijk = np.zeros((P, 3), np.int32)
        dist = np.zeros(P, np.float32)
        xyz = np.zeros((P, 3), np.float32)
        p = np.ascontiguousarray(pos, np.float32)
        Rm = np.ascontiguousarray(R, np.float32).reshape(9)
        n = lib().or_sim_cast(self.h, _p(p, ctypes.c_float),
                              _p(Rm, ctypes.c_float), res, width, height,
                              hfov, vfov, min_range, max_range,
                              2 if rays is not None else int(bool(laser)),
                              None if rays is None else _p(rr, ctypes.c_float),
                              int(hierarchical), _p(ijk, ctypes.c_int32),
                              _p(dist, ctypes.c_float), _p(xyz, ctypes.c_float))
        return ijk, dist, xyz, n


class SoftScene:
    """Dense FP64 log-odds grid and camera parameters of the soft march
    (the model of include/nanomap_amd_soft.h, restated in FP64)."""

    def __init__(self, grid, origin, res, background, width, height, hfov,
                 vfov, step, n_samples):
        self.grid = np.ascontiguousarray(grid, np.float64)
        self.origin = np.ascontiguousarray(origin, np.float64)
        self.res, self.bg = float(res), float(background)
        self.width, self.height = int(width), int(height)
        self.hfov, self.vfov = float(hfov), float(vfov)
        self.step, self.ns = float(step), int(n_samples)

    def _args(self, cams):
        c = np.ascontiguousarray(cams, np.float64).reshape(-1, 12)
        nx, ny, nz = self.grid.shape
        return c, [_p(self.grid, ctypes.c_double), nx, ny, nz,
                   _p(self.origin, ctypes.c_double), self.res, self.bg,
                   _p(c, ctypes.c_double), c.shape[0], self.width,
                   self.height, self.hfov, self.vfov, self.step, self.ns]

    def forward(self, cams):
        """Expected depth [n_cams, height, width]."""
        c, args = self._args(cams)
        out = np.zeros((c.shape[0], self.height, self.width), np.float64)
        lib().or_soft_forward(*args, _p(out, ctypes.c_double))
        return out

    def backward(self, cams, grad_depth, want_grid=True):
        """(dL/dxi [n_cams, 6] as (rho, omega), dL/dgrid or None)."""
        c, args = self._args(cams)
        g = np.ascontiguousarray(grad_depth, np.float64)
        gp = np.zeros((c.shape[0], 6), np.float64)
        gg = np.zeros_like(self.grid) if want_grid else None
        lib().or_soft_backward(*args, _p(g, ctypes.c_double),
                               _p(gp, ctypes.c_double),
                               _p(gg, ctypes.c_double) if want_grid else None)
        return gp, gg


def perturb_camera(cam, xi):
    """Camera (12 doubles) moved by the world-frame perturbation
    xi = (rho, omega): R' = exp(omega^) R, p' = p + rho (rotation about the
    camera centre), in FP64 (Rodrigues)."""
    cam = np.asarray(cam, np.float64).reshape(12)
    R = cam[:9].reshape(3, 3)
    w = np.asarray(xi[3:], np.float64)
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-300:
        E = np.eye(3)
    else:
        E = (np.eye(3) + np.sin(th) / th * K
             + (1 - np.cos(th)) / th ** 2 * K @ K)
    out = np.empty(12)
    out[:9] = (E @ R).reshape(9)
    out[9:] = cam[9:] + np.asarray(xi[:3], np.float64)
    return out
