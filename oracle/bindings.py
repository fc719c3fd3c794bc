"""Test infrastructure only: ctypes bindings of the oracle (liboracle.so, the C restatement) and of
the compiled reference (_ref/libcubiquity_ref.so). Only tests/, the smoke check and bench.py's
CPU-baseline leg import this module; the product never does.

Both libraries are built by ``make -C oracle`` (the reference one only where the reference
checkout exists; the built file travels with the tree).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CQX_ORACLE_LIBRARY names another build of the same sources (make -C oracle asan)
ORACLE_LIB = os.environ.get("CQX_ORACLE_LIBRARY") or os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libcubiquity_ref.so")
REF_SRC = "/root/reference/src/library"

HIT_DTYPE = np.dtype([("t", "<f4"), ("voxel", "<i4", (3,)), ("node", "<u4"), ("info", "<u4"),
                      ("entries", "<u4"), ("iters", "<u4")])
RESULT_DTYPE = np.dtype([("hit", "<i4"), ("material", "<u4"), ("distance", "<f8"),
                         ("position", "<f4", (3,)), ("normal", "<f4", (3,))])
SUBDAG_DTYPE = np.dtype([("lower", "<i4", (3,)), ("height", "<i4"), ("pad0", "<u4"), ("node", "<u4"),
                         ("pad1", "<u4"), ("pad2", "<u4")])

_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64


def build(ref=True):
    targets = ["oracle"] + (["ref", "dropin"] if ref and os.path.isdir(REF_SRC) else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class _DemoParams(ctypes.Structure):
    _fields_ = [("colours", ctypes.c_void_p), ("max_footprint", ctypes.c_float), ("gamma", ctypes.c_float),
                ("add_noise", ctypes.c_uint32), ("include_sun", ctypes.c_uint32), ("include_sky", ctypes.c_uint32),
                ("accumulate", ctypes.c_uint32), ("seed", ctypes.c_uint64)]


class Oracle:
    """The C restatement (cpu_ref.c / cpu_gen.c)."""

    def __init__(self, lib_path=None):
        """lib_path: another build of the same sources (the mutation tests load planted-slip
        builds this way); default liboracle.so."""
        if lib_path is None:
            lib_path = ORACLE_LIB
            if not os.path.exists(lib_path):
                build(ref=False)
        L = ctypes.CDLL(lib_path)
        sig = {
            "orc_find_subdags": (ctypes.c_int, [_P, _U64, _U32, _P]),
            "orc_intersect": (None, [_P, _P, _P, _U64, ctypes.c_int, ctypes.c_float, _U32, _P]),
            "orc_intersect_mt": (None, [_P, _P, _P, _U64, ctypes.c_int, ctypes.c_float, _U32, _P, ctypes.c_int]),
            "orc_hits_to_results": (None, [_P, _P, _U64, ctypes.c_int, _P]),
            "orc_trace_ref": (None, [_P, _U32, _P, _U64, _P]),
            "orc_voxel": (_U32, [_P, _U32, _I32, _I32, _I32]),
            "orc_descend": (None, [_P, _U32, _I32, _I32, _I32, _P]),
            "orc_dense_readout": (None, [_P, _U32, _P, _P, _P]),
            "orc_time_intersect": (ctypes.c_double, [_P, _P, _P, _U64, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                                     _U64, ctypes.c_int, _P]),
            "orc_dense_brute": (None, [_P, ctypes.c_int, _P, _U64, _P, _P, _P, _P, _P]),
            "orc_gen_height": (ctypes.c_double, [_P, _I64, _I64]),
            "orc_gen_spheres": (None, [_P, _P]),
            "orc_gen_material": (_U32, [_P, _P, _I64, _I64, _I64]),
            "orc_gen_dense": (None, [_P, _P, ctypes.c_int]),
            "orc_compare_dense": (_U64, [_P, ctypes.c_int, _P, _U32, ctypes.c_int]),
            "orc_pathtrace": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int]),
            "orc_pathtrace_f64": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, ctypes.c_int]),
            "orc_pcg32": (None, [_U64, _U64, _U32, _P]),
            "orc_pcg_uniform": (None, [_U64, _U64, _U32, _P]),
            "orc_sincos2pi": (None, [_P, _U64, _P, _P]),
            "orc_pt_bounce_rays": (_U32, [_P, _P, ctypes.c_float, _U64, _U64, _U32, _P]),
            "orc_pt_camera_ray": (ctypes.c_int, [_P, _U64, _P, _P]),
            "orc_bit_mix": (_U64, [_U64]),
            "orc_fnv1a": (_U64, [_P, _I64, _U64]),
            "orc_position_noise": (ctypes.c_float, [_P]),
            "orc_random_point": (ctypes.c_int, [_P, _P]),
            "orc_demo_pixel_state": (_U32, [_U64, _U64]),
            "orc_demo_trace": (None, [_P, _P, _P, _U64, _P, ctypes.c_int, _P, _P, ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        self.L = L

    @staticmethod
    def _nodes(nodes):
        return np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 8)

    def find_subdags(self, nodes, root):
        nodes = self._nodes(nodes)
        out = np.zeros(8, dtype=SUBDAG_DTYPE)
        if self.L.orc_find_subdags(_ptr(nodes), nodes.shape[0], root, _ptr(out)) != 0:
            raise ValueError("malformed store")
        return out

    def intersect(self, nodes, subdags, rays, surf=True, max_footprint=-1.0, max_iters=0, threads=0):
        nodes = self._nodes(nodes)
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        subdags = np.ascontiguousarray(subdags, dtype=SUBDAG_DTYPE)
        out = np.zeros(rays.shape[0], dtype=HIT_DTYPE)
        if threads and threads > 1:
            self.L.orc_intersect_mt(_ptr(nodes), _ptr(subdags), _ptr(rays), rays.shape[0], int(surf),
                                    ctypes.c_float(max_footprint), max_iters, _ptr(out), int(threads))
        else:
            self.L.orc_intersect(_ptr(nodes), _ptr(subdags), _ptr(rays), rays.shape[0], int(surf),
                                 ctypes.c_float(max_footprint), max_iters, _ptr(out))
        return out

    def to_results(self, rays, hits, surf=True):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        hits = np.ascontiguousarray(hits).view(HIT_DTYPE).reshape(-1)
        out = np.zeros(hits.shape[0], dtype=RESULT_DTYPE)
        self.L.orc_hits_to_results(_ptr(rays), _ptr(hits), hits.shape[0], int(surf), _ptr(out))
        return out

    def trace_ref(self, nodes, root, rays):
        nodes = self._nodes(nodes)
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        out = np.zeros(rays.shape[0], dtype=RESULT_DTYPE)
        self.L.orc_trace_ref(_ptr(nodes), root, _ptr(rays), rays.shape[0], _ptr(out))
        return out

    def voxel(self, nodes, root, x, y, z):
        nodes = self._nodes(nodes)
        return int(self.L.orc_voxel(_ptr(nodes), root, int(x), int(y), int(z)))

    # ---- pieces of the path tracer (cpu_pt.c), for the pinning tests ----
    def pcg32(self, initstate, initseq, n):
        out = np.zeros(n, np.uint32)
        self.L.orc_pcg32(initstate, initseq, n, _ptr(out))
        return out

    def pcg_uniform(self, initstate, initseq, n):
        out = np.zeros(n, np.float32)
        self.L.orc_pcg_uniform(initstate, initseq, n, _ptr(out))
        return out

    def sincos2pi(self, u):
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1)
        c = np.zeros_like(u)
        s = np.zeros_like(u)
        self.L.orc_sincos2pi(_ptr(u), u.size, _ptr(c), _ptr(s))
        return c, s

    def bounce_rays(self, normal, pos, offset, initstate, initseq, count):
        """(count, 6) rays: cosine-weighted bounces about `normal` from `pos`, one PCG stream."""
        n = np.ascontiguousarray(normal, dtype=np.float32).reshape(3)
        p = np.ascontiguousarray(pos, dtype=np.float32).reshape(3)
        out = np.zeros((count, 6), np.float32)
        k = self.L.orc_pt_bounce_rays(_ptr(n), _ptr(p), ctypes.c_float(offset), initstate, initseq, count,
                                      _ptr(out))
        return out[:k]

    def pt_camera_ray(self, params, path):
        """(ray, jitter) of path number `path`."""
        ray = np.zeros(6, np.float32)
        jit = np.zeros(2, np.float32)
        if self.L.orc_pt_camera_ray(ctypes.byref(params), path, _ptr(ray), _ptr(jit)) != 0:
            raise ValueError("bad path-tracer parameters")
        return ray, jit

    # ---- the reference CPU tracer's shading (cpu_demo.c) ----
    def bit_mix(self, x):
        return int(self.L.orc_bit_mix(int(x) & (2**64 - 1)))

    def fnv1a(self, data, seed=0xCBF29CE484222325):
        b = np.frombuffer(bytes(data), np.uint8)
        return int(self.L.orc_fnv1a(_ptr(b), b.size, seed))

    def position_noise(self, pos):
        p = np.ascontiguousarray(pos, dtype=np.float32).reshape(3)
        return float(self.L.orc_position_noise(_ptr(p)))

    def random_point(self, state):
        """One randomPointInUnitSphere draw: (point, new state, generator steps)."""
        st = np.array([state], np.uint32)
        out = np.zeros(3, np.float32)
        k = self.L.orc_random_point(_ptr(st), _ptr(out))
        return out, int(st[0]), int(k)

    def demo_pixel_state(self, seed, pixel):
        return int(self.L.orc_demo_pixel_state(seed, pixel))

    def demo_trace(self, nodes, subdags, rays, colours, max_footprint=0.0035, gamma=0.0, add_noise=True,
                   include_sun=True, include_sky=True, seed=0, serial_state=None, threads=8, out=None):
        """traceSingleRay per primary ray -> (n, 3) float32. serial_state: None for per-pixel
        generators (the GPU scheme), else the reference's single generator from this state
        (returns (image, final state))."""
        nodes = self._nodes(nodes)
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        subdags = np.ascontiguousarray(subdags, dtype=SUBDAG_DTYPE)
        col = np.ascontiguousarray(colours, dtype=np.float32).reshape(256, 3)
        prm = _DemoParams(col.ctypes.data, max_footprint, gamma, int(add_noise), int(include_sun),
                          int(include_sky), 1 if out is not None else 0, seed)
        img = out if out is not None else np.zeros((rays.shape[0], 3), np.float32)
        st = np.array([serial_state if serial_state is not None else 0], np.uint32)
        self.L.orc_demo_trace(_ptr(nodes), _ptr(subdags), _ptr(rays), rays.shape[0], ctypes.byref(prm),
                              1 if serial_state is not None else 0, _ptr(st), _ptr(img), int(threads))
        return img if serial_state is None else (img, int(st[0]))

    def dense(self, nodes, root, lower, size):
        """Voxels of the box [lower, lower + size) as a [z][y][x] uint8 array."""
        nodes = self._nodes(nodes)
        lo = np.asarray(lower, np.int32)
        sz = np.asarray(size, np.int32)
        out = np.zeros((size[2], size[1], size[0]), np.uint8)
        self.L.orc_dense_readout(_ptr(nodes), root, _ptr(lo), _ptr(sz), _ptr(out))
        return out

    def descend(self, nodes, root, x, y, z):
        nodes = self._nodes(nodes)
        path = np.zeros(33, dtype=np.uint32)
        self.L.orc_descend(_ptr(nodes), root, int(x), int(y), int(z), _ptr(path))
        return path

    def time_intersect(self, nodes, subdags, rays, surf=True, max_footprint=-1.0, threads=1, chunk=4096, reps=3):
        nodes = self._nodes(nodes)
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        subdags = np.ascontiguousarray(subdags, dtype=SUBDAG_DTYPE)
        hits = ctypes.c_uint64()
        t = self.L.orc_time_intersect(_ptr(nodes), _ptr(subdags), _ptr(rays), rays.shape[0], int(surf),
                                      ctypes.c_float(max_footprint), int(threads), chunk, reps, ctypes.byref(hits))
        return t, hits.value

    def dense_brute(self, grid, rays):
        grid = np.ascontiguousarray(grid, dtype=np.uint8)
        n = grid.shape[0]
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        m = rays.shape[0]
        hit = np.zeros(m, np.int32)
        tn = np.zeros(m, np.float64)
        ts = np.zeros(m, np.float64)
        mat = np.zeros(m, np.uint32)
        vox = np.zeros((m, 3), np.int32)
        self.L.orc_dense_brute(_ptr(grid), n, _ptr(rays), m, _ptr(hit), _ptr(tn), _ptr(ts), _ptr(mat), _ptr(vox))
        return hit.astype(bool), tn, ts, mat, vox

    def gen_dense(self, params, threads=8):
        n = 1 << params.log2_size
        grid = np.zeros((n, n, n), dtype=np.uint8)  # [z][y][x]
        self.L.orc_gen_dense(ctypes.byref(params), _ptr(grid), int(threads))
        return grid

    def gen_height(self, params, x, z):
        return self.L.orc_gen_height(ctypes.byref(params), int(x), int(z))

    def pathtrace(self, nodes, root, params, threads=8):
        """FP32 image (height, width, 3); `params` is a ctypes struct laid out as orc_pt_params."""
        nodes = self._nodes(nodes)
        sub = self.find_subdags(nodes, root)
        img = np.zeros((params.height, params.width, 3), np.float32)
        if self.L.orc_pathtrace(_ptr(nodes), _ptr(sub), ctypes.byref(params), _ptr(img), int(threads)) != 0:
            raise ValueError("bad path-tracer parameters")
        return img

    def pathtrace_f64(self, nodes, root, params, d_image=None, threads=8):
        """FP64 image and, if d_image is given, the gradient of sum(d_image * image) with respect to
        the albedo and emission tables (n_materials x 3 each)."""
        nodes = self._nodes(nodes)
        sub = self.find_subdags(nodes, root)
        img = np.zeros((params.height, params.width, 3), np.float64)
        m = params.n_materials
        da = np.zeros((m, 3), np.float64)
        de = np.zeros((m, 3), np.float64)
        w = None if d_image is None else np.ascontiguousarray(d_image, dtype=np.float64)
        rc = self.L.orc_pathtrace_f64(_ptr(nodes), _ptr(sub), ctypes.byref(params),
                                      None if w is None else _ptr(w), _ptr(img), _ptr(da), _ptr(de), int(threads))
        if rc != 0:
            raise ValueError("bad path-tracer parameters")
        return img, da, de

    def compare_dense(self, grid, log2_size, nodes, root, threads=8):
        nodes = self._nodes(nodes)
        grid = np.ascontiguousarray(grid, dtype=np.uint8)
        return int(self.L.orc_compare_dense(_ptr(grid), int(log2_size), _ptr(nodes), root, int(threads)))


class Reference:
    """The unmodified reference library compiled in place (oracle/_ref)."""

    @staticmethod
    def available():
        return os.path.exists(REF_LIB)

    def __init__(self):
        L = ctypes.CDLL(REF_LIB)
        sig = {
            "ref_volume_load": (_P, [ctypes.c_char_p]),
            "ref_volume_new": (_P, []),
            "ref_volume_free": (None, [_P]),
            "ref_volume_root": (_U32, [_P]),
            "ref_volume_set_voxel": (None, [_P, _I32, _I32, _I32, _U32]),
            "ref_volume_save": (None, [_P, ctypes.c_char_p]),
            "ref_fill_sphere": (None, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _U32]),
            "ref_volume_dense": (None, [_P, _I32, _I32, _I32, _I32, _I32, _I32, _P]),
            "ref_volume_bake": (None, [_P]),
            "ref_volume_checkpoint": (None, [_P]),
            "ref_volume_undo": (None, [_P]),
            "ref_volume_redo": (None, [_P]),
            "ref_volume_count_nodes": (_U32, [_P]),
            "ref_volume_voxel": (_U32, [_P, _I32, _I32, _I32]),
            "ref_find_subdags": (None, [_P, _P]),
            "ref_intersect": (None, [_P, _P, _U64, ctypes.c_int, ctypes.c_float, _P]),
            "ref_trace_ref": (None, [_P, _P, _U64, _P]),
            "ref_time_intersect": (ctypes.c_double, [_P, _P, _U64, ctypes.c_int, ctypes.c_float, ctypes.c_int, _U64,
                                                     ctypes.c_int, _P]),
            "ref_bit_mix": (_U64, [_U64]),
            "ref_fnv1a": (_U64, [_P, _I64, _U64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        self.L = L

    def load(self, path):
        h = self.L.ref_volume_load(os.fsencode(path))
        if not h:
            raise IOError(path)
        return RefVolume(self, h)

    def new(self):
        return RefVolume(self, self.L.ref_volume_new())

    def bit_mix(self, x):
        return int(self.L.ref_bit_mix(int(x) & (2**64 - 1)))

    def fnv1a(self, data, seed=0xCBF29CE484222325):
        b = np.frombuffer(bytes(data), np.uint8)
        return int(self.L.ref_fnv1a(_ptr(b), b.size, seed))


class RefVolume:
    def __init__(self, ref, h):
        self.ref = ref
        self.h = h

    def __del__(self):
        if self.h:
            self.ref.L.ref_volume_free(self.h)
            self.h = None

    @property
    def root(self):
        return int(self.ref.L.ref_volume_root(self.h))

    def set_voxel(self, x, y, z, m):
        self.ref.L.ref_volume_set_voxel(self.h, int(x), int(y), int(z), int(m))

    def save(self, path):
        self.ref.L.ref_volume_save(self.h, os.fsencode(path))

    def fill_sphere(self, centre, radius, material):
        self.ref.L.ref_fill_sphere(self.h, float(centre[0]), float(centre[1]), float(centre[2]), float(radius),
                                   int(material))

    def checkpoint(self):
        self.ref.L.ref_volume_checkpoint(self.h)

    def undo(self):
        self.ref.L.ref_volume_undo(self.h)

    def redo(self):
        self.ref.L.ref_volume_redo(self.h)

    def dense(self, lower, size):
        """Voxels of the box [lower, lower + size) as a [z][y][x] uint8 array."""
        out = np.zeros((size[2], size[1], size[0]), np.uint8)
        self.ref.L.ref_volume_dense(self.h, int(lower[0]), int(lower[1]), int(lower[2]), int(size[0]), int(size[1]),
                                    int(size[2]), _ptr(out))
        return out

    def bake(self):
        self.ref.L.ref_volume_bake(self.h)

    def count_nodes(self):
        return int(self.ref.L.ref_volume_count_nodes(self.h))

    def voxel(self, x, y, z):
        return int(self.ref.L.ref_volume_voxel(self.h, int(x), int(y), int(z)))

    def find_subdags(self):
        out = np.zeros(8, dtype=SUBDAG_DTYPE)
        self.ref.L.ref_find_subdags(self.h, _ptr(out))
        return out

    def intersect(self, rays, surf=True, max_footprint=-1.0):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        out = np.zeros(rays.shape[0], dtype=RESULT_DTYPE)
        self.ref.L.ref_intersect(self.h, _ptr(rays), rays.shape[0], int(surf), ctypes.c_float(max_footprint), _ptr(out))
        return out

    def trace_ref(self, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        out = np.zeros(rays.shape[0], dtype=RESULT_DTYPE)
        self.ref.L.ref_trace_ref(self.h, _ptr(rays), rays.shape[0], _ptr(out))
        return out

    def time_intersect(self, rays, surf=True, max_footprint=-1.0, threads=1, chunk=4096, reps=3):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        hits = ctypes.c_uint64()
        t = self.ref.L.ref_time_intersect(self.h, _ptr(rays), rays.shape[0], int(surf), ctypes.c_float(max_footprint),
                                          int(threads), chunk, reps, ctypes.byref(hits))
        return t, hits.value
