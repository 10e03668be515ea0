"""ctypes binding of the C ABI in include/nanomap_amd.h,
include/nanomap_amd_batch.h, include/nanomap_amd_sim.h,
include/nanomap_amd_soft.h and include/nanomap_amd_scene.h.

The shared library is built in-tree (nanomap_amd/libnanomap_amd.so).  If it
is missing, every entry point raises: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

NM_OK = 0
NM_EINVAL = 1
NM_ENOMEM = 2
NM_EOVERFLOW = 3
NM_EHIP = 4
NM_ENODEV = 5
NM_ESTATE = 6

ROUND_NEAREST = 0
ROUND_DOWN = 1
FILTER_NONE = 0
FILTER_VOXEL = 1         # FP32 centroid (PrecisionType 0)
FILTER_VOXEL_HALF = 2    # FP16 sums and count (PrecisionType 1)
FILTER_VOXEL_COUNT = 3   # voxel centre, int count (PrecisionType 2)
FILTER_VOXEL_COUNT8 = 4  # voxel centre, 8-bit count (PrecisionType 3)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "libnanomap_amd.so")


class NanoMapError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("nanomap_amd error %d: %s" % (code, msg))
        self.code = code


class CtxConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("max_points", ctypes.c_int),
                ("max_time", ctypes.c_float), ("tile_capacity", ctypes.c_int)]


class Frame(ctypes.Structure):
    _fields_ = [("rotation", ctypes.c_float * 9), ("eye", ctypes.c_double * 3),
                ("grid_res", ctypes.c_float), ("max_time", ctypes.c_float),
                ("logodds_hit", ctypes.c_float),
                ("logodds_miss", ctypes.c_float), ("rounding", ctypes.c_int),
                ("filter", ctypes.c_int), ("width", ctypes.c_int),
                ("height", ctypes.c_int), ("box_min", ctypes.c_float * 3),
                ("box_max", ctypes.c_float * 3), ("max_range", ctypes.c_float)]

    def copy(self):
        f = Frame()
        ctypes.memmove(ctypes.byref(f), ctypes.byref(self), ctypes.sizeof(Frame))
        return f


class DepthCamera(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int),
                ("fx", ctypes.c_float), ("fy", ctypes.c_float),
                ("cx", ctypes.c_float), ("cy", ctypes.c_float)]


class FrameStats(ctypes.Structure):
    _fields_ = [("points", ctypes.c_int64), ("rays", ctypes.c_int64),
                ("visits", ctypes.c_int64), ("voxels", ctypes.c_int64),
                ("leaves", ctypes.c_int64), ("hit_voxels", ctypes.c_int64),
                ("mixed_voxels", ctypes.c_int64),
                ("map_leaves", ctypes.c_int64), ("errors", ctypes.c_int32),
                ("overflow", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("max_frames", ctypes.c_int),
                ("max_points", ctypes.c_int64), ("max_time", ctypes.c_float),
                ("tile_capacity", ctypes.c_int64)]


class AgentFrame(ctypes.Structure):
    _fields_ = [("frame", Frame), ("cloud", ctypes.c_void_p),
                ("n_points", ctypes.c_int), ("step", ctypes.c_int),
                ("map", ctypes.c_void_p)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("layout", ctypes.c_int), ("queue", ctypes.c_int),
                ("grid_mult", ctypes.c_int), ("lds_slots", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SoftGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("nz", ctypes.c_int), ("origin", ctypes.c_double * 3),
                ("res", ctypes.c_float), ("background", ctypes.c_float)]


class SoftParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int),
                ("hfov_deg", ctypes.c_float), ("vfov_deg", ctypes.c_float),
                ("step", ctypes.c_float), ("n_samples", ctypes.c_int)]


class SimSensor(ctypes.Structure):
    _fields_ = [("position", ctypes.c_float * 3),
                ("rotation", ctypes.c_float * 9), ("grid_res", ctypes.c_float),
                ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("hfov_deg", ctypes.c_float), ("vfov_deg", ctypes.c_float),
                ("min_range", ctypes.c_float), ("max_range", ctypes.c_float),
                ("type", ctypes.c_int), ("rays", ctypes.c_void_p)]


SENSOR_FRUSTUM = 0
SENSOR_LASER = 1
SENSOR_RAYS = 2
SIM_STATIC_BYTES = 65536  # NM_SIM_STATIC_BYTES


class Camera(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int),
                ("hfov_deg", ctypes.c_float), ("vfov_deg", ctypes.c_float),
                ("max_range", ctypes.c_float), ("noise_k", ctypes.c_float),
                ("dropout", ctypes.c_float), ("frame", ctypes.c_float * 9)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PU32 = ctypes.POINTER(ctypes.c_uint32)
_PF = ctypes.POINTER(ctypes.c_float)
_PD = ctypes.POINTER(ctypes.c_double)
_PI64 = ctypes.POINTER(ctypes.c_int64)

# name -> (restype, argtypes); every symbol of the public headers.
SIGNATURES = {
    "nm_last_error": (ctypes.c_char_p, []),
    "nm_version": (_I, []),
    "nm_ctx_create": (_I, [ctypes.POINTER(CtxConfig), ctypes.POINTER(_P)]),
    "nm_ctx_destroy": (_I, [_P]),
    "nm_map_create": (_I, [_I, _F, _F, _F, _F, _F, _I, ctypes.POINTER(_P)]),
    "nm_map_destroy": (_I, [_P]),
    "nm_map_clear": (_I, [_P, _P]),
    "nm_map_leaf_count": (_I, [_P, _PI64]),
    "nm_map_download": (_I, [_P, _PI32, _PF, _PU32]),
    "nm_map_device_arrays": (_I, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P),
                                  ctypes.POINTER(_P), ctypes.POINTER(_P),
                                  ctypes.POINTER(_P), _PI64]),
    "nm_map_capacity": (_I, [_P, _PI64]),
    "nm_map_params": (_I, [_P, _PF, _PF, _PF, _PF, _PF]),
    "nm_map_reserve": (_I, [_P, ctypes.c_int64]),
    "nm_map_set_auto_grow": (_I, [_P, _I]),
    "nm_map_query": (_I, [_P, _P, ctypes.c_int64, _P, _P, _I, _P]),
    "nm_map_dense": (_I, [_P, _P, _I, _I, _I, ctypes.c_float, _P, _P]),
    "nm_map_save": (_I, [_P, ctypes.c_char_p]),
    "nm_nvdb_build": (_I, [_P, _P, _P, ctypes.c_int64, ctypes.c_double,
                           ctypes.c_char_p, _P,
                           ctypes.POINTER(ctypes.c_uint64)]),
    "nm_nvdb_write": (_I, [ctypes.c_char_p, _P, _P, _P, ctypes.c_int64,
                           ctypes.c_double, ctypes.c_char_p]),
    "nm_nvdb_read": (_I, [ctypes.c_char_p, ctypes.c_char_p,
                          ctypes.POINTER(ctypes.c_int64), _P, _P, _P,
                          ctypes.POINTER(ctypes.c_double)]),
    "nm_map_write_nvdb": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p]),
    "nm_map_load_nvdb": (_I, [_I, ctypes.c_char_p, ctypes.c_char_p,
                              ctypes.c_float, ctypes.c_float, ctypes.c_float,
                              ctypes.c_float, ctypes.c_int64, _P]),
    "nm_map_load": (_I, [_I, ctypes.c_char_p, ctypes.c_int64,
                         ctypes.POINTER(_P)]),
    "nm_cast": (_I, [_P, ctypes.POINTER(Frame), _P, _I, _I, _I, _P]),
    "nm_update_map": (_I, [_P, _P, _P]),
    "nm_frame_submit": (_I, [_P, _P, ctypes.POINTER(Frame), _P, _I, _I, _I,
                             _P]),
    "nm_frame_stats_get": (_I, [_P, ctypes.POINTER(FrameStats)]),
    "nm_depth_camera_from_fov": (_I, [_I, _I, _F, _F,
                                      ctypes.POINTER(DepthCamera)]),
    "nm_depth_to_cloud": (_I, [ctypes.POINTER(DepthCamera), _P, _P]),
    "nm_cast_depth": (_I, [_P, ctypes.POINTER(Frame),
                           ctypes.POINTER(DepthCamera), _P, _I, _P]),
    "nm_frame_submit_depth": (_I, [_P, _P, ctypes.POINTER(Frame),
                                   ctypes.POINTER(DepthCamera), _P, _I, _P]),
    "nm_counts_download": (_I, [_P, ctypes.c_int64, _PI64, _PI32, _PU32,
                                _PU32]),
    "nm_filter_rays_download": (_I, [_P, _I, ctypes.POINTER(_I), _PF]),
    "nm_fold_host": (_F, [_F, _F, ctypes.c_uint32, _I]),
    "nm_device_alloc": (_I, [_I, ctypes.c_int64, ctypes.POINTER(_P)]),
    "nm_device_free": (_I, [_I, _P]),
    "nm_device_download": (_I, [_I, _P, _P, ctypes.c_int64]),
    "nm_ctx_set_timing": (_I, [_P, _I]),
    "nm_ctx_set_timing_mask": (_I, [_P, ctypes.c_uint]),
    "nm_ctx_stage_times": (_I, [_P, _PF, _PI64]),
    "nm_ctx_reset_timing": (_I, [_P]),
    "nm_batch_create": (_I, [ctypes.POINTER(BatchConfig),
                             ctypes.POINTER(_P)]),
    "nm_batch_destroy": (_I, [_P]),
    "nm_batch_submit": (_I, [_P, ctypes.POINTER(AgentFrame), _I, _P]),
    "nm_batch_stats": (_I, [_P, ctypes.POINTER(FrameStats), _I]),
    "nm_batch_tune": (_I, [_P, ctypes.POINTER(AgentFrame), _I, _I, _P,
                           ctypes.POINTER(BatchTuning), _PF]),
    "nm_batch_set_tuning": (_I, [_P, ctypes.POINTER(BatchTuning)]),
    "nm_batch_get_tuning": (_I, [_P, ctypes.POINTER(BatchTuning)]),
    "nm_batch_set_graph": (_I, [_P, _I]),
    "nm_soft_forward": (_I, [ctypes.POINTER(SoftGrid), _P, _P, _I,
                             ctypes.POINTER(SoftParams), _P, _P, _P]),
    "nm_soft_backward": (_I, [ctypes.POINTER(SoftGrid), _P, _P, _I,
                              ctypes.POINTER(SoftParams), _P, _P, _P, _P, _P,
                              _P]),
    "nm_soft_forward_f64": (_I, [ctypes.POINTER(SoftGrid), _P, _P, _I,
                                 ctypes.POINTER(SoftParams), _P, _P, _P]),
    "nm_soft_backward_f64": (_I, [ctypes.POINTER(SoftGrid), _P, _P, _I,
                                  ctypes.POINTER(SoftParams), _P, _P, _P, _P,
                                  _P]),
    "nm_scene_create": (_I, [ctypes.c_uint64, _PD, _I, _D, _D, _I, _D, _D,
                             ctypes.POINTER(_P)]),
    "nm_scene_destroy": (_I, [_P]),
    "nm_scene_render": (_I, [_P, ctypes.POINTER(Camera), _PF, _PF,
                             ctypes.c_uint64, _PF]),
    "nm_scene_sample_pose": (_I, [_P, ctypes.c_uint64, _F, _F, _PF, _PF]),
    "nm_scene_walk": (_I, [_P, ctypes.c_uint64, _I, _F, _F, _F, _F, _PF,
                           _PF]),
    "nm_frame_from_pose": (_I, [_PF, _PF, _PF, _F, _F, _F, _F, _I,
                                ctypes.POINTER(Frame)]),
    "nm_frustum_box": (_I, [_I, _I, _F, _F, _F, _I, _PF, _PF]),
    "nm_simgrid_from_leaves": (_I, [_I, _PI32, _PF, ctypes.c_int64, _F,
                                    ctypes.POINTER(_P)]),
    "nm_simgrid_from_map": (_I, [_P, _F, ctypes.POINTER(_P)]),
    "nm_simgrid_destroy": (_I, [_P]),
    "nm_simgrid_info": (_I, [_P, _PI64, _PI32, _PI32]),
    "nm_sim_cast": (_I, [_P, ctypes.POINTER(SimSensor), _I, _P, _P, _P,
                         ctypes.POINTER(_I), _P]),
    "nm_sim_cast_static": (_I, [_P, ctypes.POINTER(SimSensor), _I, _P, _P,
                                _P, ctypes.POINTER(_I), _P]),
}

_lib = None


def lib():
    """Load libnanomap_amd.so (raises if it has not been built).

    torch, when installed, is imported first: its bundled HIP runtime then
    satisfies this library's libamdhip64.so.7 dependency, so one process
    never holds two HIP runtimes (torch's libraries ask for the unversioned
    libamdhip64.so and would otherwise load a second copy)."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise NanoMapError(NM_ENODEV,
                               "%s is not built: run __graft_entry__.build()"
                               % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != NM_OK:
        msg = lib().nm_last_error()
        raise NanoMapError(status, msg.decode() if msg else "")
    return status


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))
