"""ctypes declarations of the C ABI in include/cqx.h (libcqx.so)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CQX_LIBRARY names another build of the same sources (the sanitizer build of the host code,
# cubiquity_amd/build.py --sanitize); by default the in-tree libcqx.so.
LIB_PATH = os.environ.get("CQX_LIBRARY") or os.path.join(HERE, "libcqx.so")

CQX_SURFACE_PROPERTIES = 0x1
MAX_FOOTPRINT_DISABLED = -1.0
HIT_FLAG = 0x1
HIT_ANOMALY = 0x8000

SUBDAG_DTYPE = np.dtype([("lower", "<i4", (3,)), ("height", "<i4"), ("pad0", "<u4"), ("node", "<u4"),
                         ("pad1", "<u4"), ("pad2", "<u4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("voxel", "<i4", (3,)), ("node", "<u4"), ("info", "<u4"),
                      ("entries", "<u4"), ("iters", "<u4")])
INTERSECTION_DTYPE = np.dtype([("hit", "<i4"), ("material", "<u4"), ("distance", "<f8"),
                               ("position", "<f4", (3,)), ("normal", "<f4", (3,))])
assert SUBDAG_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 32 and INTERSECTION_DTYPE.itemsize == 40


class CqxError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"cqx error {status}: {message}")
        self.status = status


class GenParams(ctypes.Structure):
    _fields_ = [("log2_size", ctypes.c_int32), ("octaves", ctypes.c_int32), ("n_spheres", ctypes.c_int32),
                ("n_materials", ctypes.c_int32), ("seed", ctypes.c_uint64), ("gain", ctypes.c_double),
                ("lacunarity", ctypes.c_double), ("period", ctypes.c_double), ("base_height", ctypes.c_double),
                ("amplitude", ctypes.c_double), ("rock_height", ctypes.c_double), ("r_min", ctypes.c_double),
                ("r_max", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LaunchCfg(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_uint32), ("block", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32),
                ("regroup_interval", ctypes.c_uint32), ("refill_lanes", ctypes.c_uint32),
                ("max_iters", ctypes.c_uint32), ("lds_nodes", ctypes.c_uint32),
                ("bin_bits", ctypes.c_uint32), ("stack_rows", ctypes.c_uint32)]


class PtParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("spp", ctypes.c_uint32),
                ("bounces", ctypes.c_uint32), ("eye", ctypes.c_double * 3), ("target", ctypes.c_double * 3),
                ("fov_y_degrees", ctypes.c_double), ("n_materials", ctypes.c_uint32),
                ("albedo", ctypes.POINTER(ctypes.c_float)), ("emission", ctypes.POINTER(ctypes.c_float)),
                ("sky", ctypes.c_float * 3), ("bounce_offset", ctypes.c_float), ("max_footprint", ctypes.c_float),
                ("seed", ctypes.c_uint64)]


assert ctypes.sizeof(PtParams) == 128


class DemoParams(ctypes.Structure):
    _fields_ = [("colours", ctypes.c_void_p), ("max_footprint", ctypes.c_float), ("gamma", ctypes.c_float),
                ("add_noise", ctypes.c_uint32), ("include_sun", ctypes.c_uint32), ("include_sky", ctypes.c_uint32),
                ("accumulate", ctypes.c_uint32), ("seed", ctypes.c_uint64)]


assert ctypes.sizeof(DemoParams) == 40


class SphereEdit(ctypes.Structure):
    _fields_ = [("centre", ctypes.c_float * 3), ("radius", ctypes.c_float), ("material", ctypes.c_uint32)]


MATERIAL_DTYPE = np.dtype([("name", "S64"), ("base_color", "<f4", (3,)), ("emission", "<f4", (3,)),
                           ("flags", "<u4")])
MATERIAL_HAS_NAME, MATERIAL_HAS_BASE_COLOR, MATERIAL_HAS_EMISSION = 1, 2, 4

SPHERE_EDIT_DTYPE = np.dtype([("centre", "<f4", (3,)), ("radius", "<f4"), ("material", "<u4")])
assert SPHERE_EDIT_DTYPE.itemsize == ctypes.sizeof(SphereEdit) == 20


class Batch(ctypes.Structure):
    _fields_ = [("volume", ctypes.c_void_p), ("rays", ctypes.c_void_p), ("hits", ctypes.c_void_p),
                ("n", ctypes.c_uint64), ("flags", ctypes.c_uint32), ("max_footprint", ctypes.c_float)]


assert ctypes.sizeof(Batch) == 40

_LIB = None

_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32
_I32 = ctypes.c_int32
_ST = ctypes.c_int32

MESSAGE_FN = ctypes.CFUNCTYPE(None, ctypes.c_char_p)
PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int)

# name -> (restype, argtypes); the authoritative list of exported entry points.
SIGNATURES = {
    "cqx_last_error": (ctypes.c_char_p, []),
    "cqx_version": (ctypes.c_char_p, []),
    "cqx_set_debug_handler": (None, [_P]),
    "cqx_set_warning_handler": (None, [_P]),
    "cqx_set_progress_handler": (None, [_P]),
    "cqx_hv_load_dag": (_ST, [ctypes.c_char_p, _P]),
    "cqx_hv_save_dag": (_ST, [_P, ctypes.c_char_p]),
    "cqx_hv_from_nodes": (_ST, [ctypes.POINTER(ctypes.c_uint32), _U64, _U32, _P]),
    "cqx_hv_free": (None, [_P]),
    "cqx_hv_nodes": (_P, [_P]),
    "cqx_hv_node_count": (_U64, [_P]),
    "cqx_hv_root": (_U32, [_P]),
    "cqx_hv_voxel": (_ST, [_P, _I32, _I32, _I32, _P]),
    "cqx_find_subdags": (_ST, [ctypes.POINTER(ctypes.c_uint32), _U64, _U32, _P]),
    "cqx_validate_dag": (_ST, [ctypes.POINTER(ctypes.c_uint32), _U64, _U32]),
    "cqx_hits_to_intersections": (_ST, [_P, _P, _U64, _U32, _P]),
    "cqx_gen_default_params": (_ST, [_I32, _U64, _P]),
    "cqx_hv_generate": (_ST, [_P, ctypes.c_int, _P]),
    "cqx_hv_from_dense": (_ST, [_P, _P, _P, _P]),
    "cqx_hv_edit_spheres": (_ST, [_P, _P, _U64, _P]),
    "cqx_hv_edit_spheres_mt": (_ST, [_P, _P, _U64, ctypes.c_int32, _P]),
    "cqx_hv_compact": (_ST, [_P]),
    "cqx_hv_checkpoint": (_ST, [_P]),
    "cqx_trace_ref": (_ST, [_P, _P, _P, _U64, _P, _P]),
    "cqx_trace_ref_host": (_ST, [_P, _P, _P, _U64, _P]),
    "cqx_materials_load": (_ST, [ctypes.c_char_p, _P, _P, _P, _P]),
    "cqx_materials_save": (_ST, [ctypes.c_char_p, _P, _U32]),
    "cqx_materials_colour_table": (_ST, [_P, _U32, _P]),
    "cqx_hv_undo": (_ST, [_P]),
    "cqx_hv_redo": (_ST, [_P]),
    "cqx_volume_update": (_ST, [_P, _P, ctypes.POINTER(ctypes.c_uint32), _U64, _U64, _U32, _P]),
    "cqx_camera_rays": (_ST, [_P, _P, ctypes.c_double, _U32, _U32, _P]),
    "cqx_incoherent_rays": (_ST, [_P, _I32, _U64, _U64, ctypes.c_int, _P]),
    "cqx_create": (_ST, [ctypes.c_int, _P]),
    "cqx_destroy": (_ST, [_P]),
    "cqx_volume_upload": (_ST, [_P, ctypes.POINTER(ctypes.c_uint32), _U64, _U32, _P, _P]),
    "cqx_volume_free": (_ST, [_P, _P]),
    "cqx_volume_info": (_ST, [_P, _P, _P, _P, _P]),
    "cqx_volume_edit_spheres": (_ST, [_P, _P, _P, _U64, _P]),
    "cqx_volume_checkpoint": (_ST, [_P, _P]),
    "cqx_volume_undo": (_ST, [_P, _P]),
    "cqx_volume_redo": (_ST, [_P, _P]),
    "cqx_volume_download": (_ST, [_P, _P, _U64, _U64, _P]),
    "cqx_intersect": (_ST, [_P, _P, _P, _U64, _U32, ctypes.c_float, _P, _P, _P]),
    "cqx_intersect_noalloc": (_ST, [_P, _P, _P, _U64, _U32, ctypes.c_float, _P, _P, _P]),
    "cqx_intersect_host": (_ST, [_P, _P, _P, _U64, _U32, ctypes.c_float, _P]),
    "cqx_intersect_batched": (_ST, [_P, _P, _U32, _P, _P]),
    "cqx_autotune": (_ST, [_P, _P, _P, _U64, _U32, ctypes.c_float, _P, _P, _P, _P]),
    "cqx_autotune_clear": (_ST, [_P]),
    "cqx_pathtrace": (_ST, [_P, _P, _P, _U32, _P, _P, _P]),
    "cqx_demo_trace": (_ST, [_P, _P, _P, _U64, _P, _P, _P, _P]),
    "cqx_pathtrace_adjoint": (_ST, [_P, _P, _P, _P, _P, _P, _P, _P]),
}


def loaded():
    return _LIB is not None


def lib():
    """The loaded libcqx.so. Raises if it has not been built: there is no fallback."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m cubiquity_amd.build`")
        # One HIP runtime per process: torch bundles its own libamdhip64 (soname libamdhip64.so.7)
        # and opens it by file name. Loading torch first makes libcqx.so's dependency on
        # libamdhip64.so.7 resolve to that same copy instead of a second runtime from /opt/rocm.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(status):
    if status != 0:
        msg = lib().cqx_last_error()
        raise CqxError(status, msg.decode() if msg else "")
    return status
