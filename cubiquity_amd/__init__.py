"""cubiquity_amd: MI355X-native sparse-voxel-DAG ray traversal (Cubiquity's raycast path).

Python host layer over the C ABI of ``libcqx.so`` (``include/cqx.h``). It mirrors the
reference's raycast API (``findSubDAGs``, ``intersectVolume``, ``Volume``,
src/library/raytracing.h:57-81, storage.h:135-197) with batched, device-resident calls.
Torch is used only for device memory and streams. Loading fails loudly if the library has not
been built (``python -m cubiquity_amd.build``); there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import cabi
from .cabi import (CQX_SURFACE_PROPERTIES, MAX_FOOTPRINT_DISABLED, CqxError, GenParams, LaunchCfg,
                   HIT_DTYPE, INTERSECTION_DTYPE, MATERIAL_DTYPE, SPHERE_EDIT_DTYPE, SUBDAG_DTYPE, lib, check)

__all__ = [
    "Materials",
    "traceRayRef",
    "HostVolume", "Context", "DeviceVolume", "findSubDAGs", "intersectVolume", "camera_rays",
    "incoherent_rays", "gen_default_params", "hits_to_intersections", "CqxError", "GenParams",
    "LaunchCfg", "HIT_DTYPE", "INTERSECTION_DTYPE", "SUBDAG_DTYPE", "CQX_SURFACE_PROPERTIES",
    "MAX_FOOTPRINT_DISABLED", "decode_info", "demo_trace", "set_debug_handler", "set_warning_handler",
    "set_progress_handler",
]


_message_hooks = {}  # ctypes callbacks must outlive their registration


def _set_hook(kind, fn, cfunc, wrap):
    setter = getattr(lib(), f"cqx_set_{kind}_handler")
    if fn is None:
        setter(None)
        _message_hooks.pop(kind, None)
        return
    cb = cfunc(wrap(fn))
    setter(ctypes.cast(cb, ctypes.c_void_p))
    _message_hooks[kind] = cb


def set_debug_handler(fn):
    """Cubiquity's set_debug_handler (src/library/base.h:103): fn(message: str) or None."""
    _set_hook("debug", fn, cabi.MESSAGE_FN, lambda f: lambda m: f(m.decode()))


def set_warning_handler(fn):
    """Cubiquity's set_warning_handler (src/library/base.h:104): fn(message: str) or None; every
    error the library raises is reported here as well."""
    _set_hook("warning", fn, cabi.MESSAGE_FN, lambda f: lambda m: f(m.decode()))


def set_progress_handler(fn):
    """Cubiquity's setProgressHandler (src/library/base.h:106-107): fn(task: str, first_step,
    current_step, last_step) or None."""
    _set_hook("progress", fn, cabi.PROGRESS_FN, lambda f: lambda t, a, b, c: f(t.decode(), a, b, c))


def gen_default_params(log2_size, seed, **overrides):
    """Scaled generator defaults for a 2^log2_size volume (cqx_gen_default_params)."""
    p = GenParams()
    check(lib().cqx_gen_default_params(int(log2_size), ctypes.c_uint64(seed), ctypes.byref(p)))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _sphere_edits(edits):
    """(n, 5) centre x, y, z, radius, material -> cqx_sphere_edit records (FP32 centre and radius)."""
    e = np.asarray(edits, dtype=np.float64).reshape(-1, 5)
    arr = np.zeros(len(e), dtype=SPHERE_EDIT_DTYPE)
    arr["centre"] = e[:, :3]
    arr["radius"] = e[:, 3]
    if len(e) and not (np.all(e[:, 4] >= 0) and np.all(e[:, 4] < 2**32)):
        raise ValueError("material out of range")
    arr["material"] = e[:, 4].astype(np.uint64)
    return arr


class HostVolume:
    """Host node store + root: the library's view of a reference ``Volume``."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def generate(cls, params, threads=0):
        h = ctypes.c_void_p()
        check(lib().cqx_hv_generate(ctypes.byref(params), int(threads), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def load_dag(cls, path):
        h = ctypes.c_void_p()
        check(lib().cqx_hv_load_dag(os.fsencode(path), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_dense(cls, grid, lower=(0, 0, 0)):
        """Canonical DAG of a dense uint8 grid indexed [z][y][x] whose voxel (0,0,0) sits at `lower`."""
        grid = np.ascontiguousarray(grid, dtype=np.uint8)
        if grid.ndim != 3:
            raise ValueError("grid must be 3-D, indexed [z][y][x]")
        size = (ctypes.c_int32 * 3)(grid.shape[2], grid.shape[1], grid.shape[0])
        lo = (ctypes.c_int32 * 3)(*[int(v) for v in lower])
        h = ctypes.c_void_p()
        check(lib().cqx_hv_from_dense(grid.ctypes.data_as(ctypes.c_void_p), size, lo, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_nodes(cls, nodes, root):
        nodes = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 8)
        h = ctypes.c_void_p()
        check(lib().cqx_hv_from_nodes(nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                      ctypes.c_uint64(nodes.shape[0]), ctypes.c_uint32(root), ctypes.byref(h)))
        return cls(h)

    def save_dag(self, path):
        check(lib().cqx_hv_save_dag(self._h, os.fsencode(path)))

    def edit_spheres(self, edits, threads=0):
        """Sphere fill/carve edits (fillBrush with SphereBrush semantics). `edits`: (n, 5) array of
        centre x, y, z, radius, material (0 carves); `threads`: 0 all cores, 1 in order on this
        thread. Returns the index of the first appended node."""
        arr = _sphere_edits(edits)
        first = ctypes.c_uint64()
        check(lib().cqx_hv_edit_spheres_mt(self._h, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint64(len(arr)),
                                           ctypes.c_int32(threads), ctypes.byref(first)))
        return int(first.value)

    def checkpoint(self):
        """Volume::checkpoint: later undo() returns to this state."""
        check(lib().cqx_hv_checkpoint(self._h))

    def undo(self):
        check(lib().cqx_hv_undo(self._h))

    def redo(self):
        check(lib().cqx_hv_redo(self._h))

    def compact(self):
        check(lib().cqx_hv_compact(self._h))

    @property
    def node_count(self):
        return int(lib().cqx_hv_node_count(self._h))

    @property
    def root(self):
        return int(lib().cqx_hv_root(self._h))

    @property
    def nodes(self):
        """Read-only numpy view (node_count, 8) of the store, valid while this object lives."""
        n = self.node_count
        ptr = lib().cqx_hv_nodes(self._h)
        arr = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint32)), shape=(n * 8,))
        arr = arr.reshape(n, 8)
        arr.flags.writeable = False
        return arr

    def voxel(self, x, y, z):
        m = ctypes.c_uint32()
        check(lib().cqx_hv_voxel(self._h, int(x), int(y), int(z), ctypes.byref(m)))
        return int(m.value)

    def find_subdags(self):
        return findSubDAGs(self.nodes, self.root)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        # at interpreter exit the module globals may already be gone: the process frees the memory
        if h and cabi is not None and cabi.loaded():
            lib().cqx_hv_free(h)


def findSubDAGs(nodes, root):
    """findSubDAGs (src/library/raytracing.cpp:89-97) over a raw (n, 8) uint32 store."""
    nodes = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 8)
    out = np.zeros(8, dtype=SUBDAG_DTYPE)
    check(lib().cqx_find_subdags(nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                 ctypes.c_uint64(nodes.shape[0]), ctypes.c_uint32(root),
                                 out.ctypes.data_as(ctypes.c_void_p)))
    return out


def camera_rays(eye, target, fov_y_degrees, width, height):
    """Pinhole rays (camera.cpp:12-38 convention), float32 array (height*width, 6)."""
    out = np.empty((width * height, 6), dtype=np.float32)
    e = (ctypes.c_double * 3)(*eye)
    t = (ctypes.c_double * 3)(*target)
    check(lib().cqx_camera_rays(e, t, ctypes.c_double(fov_y_degrees), ctypes.c_uint32(width),
                                ctypes.c_uint32(height), out.ctypes.data_as(ctypes.c_void_p)))
    return out


def incoherent_rays(volume, log2_size, seed, n, threads=0):
    """Uniform origins in empty voxels of [0,N)^3, uniform directions; float32 (n, 6)."""
    out = np.empty((n, 6), dtype=np.float32)
    check(lib().cqx_incoherent_rays(volume._h, int(log2_size), ctypes.c_uint64(seed), ctypes.c_uint64(n),
                                    int(threads), out.ctypes.data_as(ctypes.c_void_p)))
    return out


def hits_to_intersections(rays, hits, surface_properties=True):
    """Expand compact records into RayVolumeIntersection fields (raytracing.h:48-55)."""
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
    hits = np.ascontiguousarray(hits).view(HIT_DTYPE).reshape(-1)
    out = np.zeros(hits.shape[0], dtype=INTERSECTION_DTYPE)
    check(lib().cqx_hits_to_intersections(rays.ctypes.data_as(ctypes.c_void_p), hits.ctypes.data_as(ctypes.c_void_p),
                                          ctypes.c_uint64(hits.shape[0]),
                                          ctypes.c_uint32(CQX_SURFACE_PROPERTIES if surface_properties else 0),
                                          out.ctypes.data_as(ctypes.c_void_p)))
    return out


def decode_info(info):
    """Split the packed ``info`` word of hit records into named numpy fields."""
    info = np.asarray(info, dtype=np.uint32)
    return {
        "hit": (info & 1).astype(bool),
        "slot": (info >> 1) & 7,
        "size_log2": (info >> 4) & 31,
        "normal_code": (info >> 9) & 63,
        "anomaly": ((info >> 15) & 1).astype(bool),
        "material": (info >> 16) & 255,
    }


class DeviceVolume:
    """A node store resident in HBM (cqx_volume_upload)."""

    def __init__(self, ctx, handle, node_count, root):
        self._ctx = ctx
        self._h = handle
        self.node_count = node_count
        self.root = root

    def update(self, volume, first):
        """Append-only delta upload of `volume`'s nodes [first, node_count) (cqx_volume_update)."""
        nodes = volume.nodes
        check(lib().cqx_volume_update(self._ctx._h, self._h, nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                      ctypes.c_uint64(nodes.shape[0]), ctypes.c_uint64(first),
                                      ctypes.c_uint32(volume.root), None))
        self.node_count = nodes.shape[0]
        self.root = volume.root

    def edit_spheres(self, edits):
        """Sphere fill/carve edits on the device copy (cqx_volume_edit_spheres): the store
        HostVolume.edit_spheres(edits, threads > 1) gives, computed on the GPU. `edits`: (n, 5)
        centre x, y, z, radius, material. Returns the index of the first appended node."""
        arr = _sphere_edits(edits)
        first = ctypes.c_uint64()
        check(lib().cqx_volume_edit_spheres(self._ctx._h, self._h, arr.ctypes.data_as(ctypes.c_void_p),
                                            ctypes.c_uint64(len(arr)), ctypes.byref(first)))
        self._refresh()
        return int(first.value)

    def checkpoint(self):
        check(lib().cqx_volume_checkpoint(self._ctx._h, self._h))

    def undo(self):
        check(lib().cqx_volume_undo(self._ctx._h, self._h))
        self._refresh()

    def redo(self):
        check(lib().cqx_volume_redo(self._ctx._h, self._h))
        self._refresh()

    def download(self, first=0, count=None):
        """Nodes [first, first + count) of the device store as a (count, 8) uint32 array."""
        if count is None:
            count = self.node_count - first
        out = np.zeros((count, 8), np.uint32)
        check(lib().cqx_volume_download(self._ctx._h, self._h, ctypes.c_uint64(first), ctypes.c_uint64(count),
                                        out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def _refresh(self):
        n = ctypes.c_uint64()
        r = ctypes.c_uint32()
        check(lib().cqx_volume_info(self._h, None, ctypes.byref(n), ctypes.byref(r), None))
        self.node_count, self.root = int(n.value), int(r.value)

    def info(self):
        nodes = ctypes.c_void_p()
        sub = np.zeros(8, dtype=SUBDAG_DTYPE)
        check(lib().cqx_volume_info(self._h, ctypes.byref(nodes), None, None, sub.ctypes.data_as(ctypes.c_void_p)))
        return nodes.value, sub

    def free(self):
        if self._h and self._ctx._h:
            check(lib().cqx_volume_free(self._ctx._h, self._h))
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """A device context (cqx_create) on one GPU."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().cqx_create(int(device), ctypes.byref(h)))
        self._h = h
        self.device = device

    def upload(self, volume, subdags=None):
        if isinstance(volume, HostVolume):
            nodes, root = volume.nodes, volume.root
        else:
            nodes, root = volume
        nodes = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 8)
        sub_ptr = None
        if subdags is not None:
            subdags = np.ascontiguousarray(subdags, dtype=SUBDAG_DTYPE)
            sub_ptr = subdags.ctypes.data_as(ctypes.c_void_p)
        h = ctypes.c_void_p()
        check(lib().cqx_volume_upload(self._h, nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                      ctypes.c_uint64(nodes.shape[0]), ctypes.c_uint32(root), sub_ptr,
                                      ctypes.byref(h)))
        return DeviceVolume(self, h, nodes.shape[0], root)

    def intersect(self, dvol, rays_ptr, n, hits_ptr, flags=CQX_SURFACE_PROPERTIES,
                  max_footprint=MAX_FOOTPRINT_DISABLED, cfg=None, stream=None):
        """Raw batched call on device pointers (ints)."""
        cfgp = ctypes.byref(cfg) if cfg is not None else None
        check(lib().cqx_intersect(self._h, dvol._h, ctypes.c_void_p(rays_ptr), ctypes.c_uint64(n),
                                  ctypes.c_uint32(flags), ctypes.c_float(max_footprint), ctypes.c_void_p(hits_ptr),
                                  cfgp, ctypes.c_void_p(stream or 0)))

    def intersect_host(self, dvol, rays, flags=CQX_SURFACE_PROPERTIES, max_footprint=MAX_FOOTPRINT_DISABLED):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        hits = np.zeros(rays.shape[0], dtype=HIT_DTYPE)
        check(lib().cqx_intersect_host(self._h, dvol._h, rays.ctypes.data_as(ctypes.c_void_p),
                                       ctypes.c_uint64(rays.shape[0]), ctypes.c_uint32(flags),
                                       ctypes.c_float(max_footprint), hits.ctypes.data_as(ctypes.c_void_p)))
        return hits

    def close(self):
        if self._h:
            check(lib().cqx_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(ctx, dvol, rays, computeSurfaceProperties=True, maxFootprint=MAX_FOOTPRINT_DISABLED, stream=None):
    """Tune the traversal launch on these GPU rays (cqx_autotune); returns (LaunchCfg, ms). The
    choice is cached in the context and used by later auto-configured launches."""
    import torch

    rays = rays.contiguous().view(-1, 6)
    n = rays.shape[0]
    hits = torch.empty((n, 32), dtype=torch.uint8, device=rays.device)
    if stream is None:
        stream = torch.cuda.current_stream(rays.device).cuda_stream
    best = LaunchCfg()
    ms = ctypes.c_float()
    check(lib().cqx_autotune(ctx._h, dvol._h, ctypes.c_void_p(rays.data_ptr()), ctypes.c_uint64(n),
                             ctypes.c_uint32(CQX_SURFACE_PROPERTIES if computeSurfaceProperties else 0),
                             ctypes.c_float(maxFootprint), ctypes.c_void_p(hits.data_ptr()), ctypes.c_void_p(stream),
                             ctypes.byref(best), ctypes.byref(ms)))
    return best, float(ms.value)


def traceRayRef(ctx, dvol, rays, stream=None):
    """The reference's brute-force raycast (cqx_trace_ref) on GPU rays (n, 6) float32; returns a
    uint8 GPU tensor of n 40-byte RayVolumeIntersection records (INTERSECTION_DTYPE)."""
    import torch

    rays = rays.contiguous().view(-1, 6)
    n = rays.shape[0]
    out = torch.empty((n, INTERSECTION_DTYPE.itemsize), dtype=torch.uint8, device=rays.device)
    if stream is None:
        stream = torch.cuda.current_stream(rays.device).cuda_stream
    check(lib().cqx_trace_ref(ctx._h, dvol._h, ctypes.c_void_p(rays.data_ptr()), ctypes.c_uint64(n),
                              ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream)))
    return out


def intersect_batched(ctx, items, cfg=None, stream=None):
    """One persistent launch over many batches. `items`: (DeviceVolume, rays GPU tensor (n, 6),
    hits GPU uint8 tensor (n, 32), computeSurfaceProperties, maxFootprint) tuples."""
    import torch

    arr = (cabi.Batch * max(1, len(items)))()
    for i, (dvol, rays, hits, surf, fp) in enumerate(items):
        n = rays.shape[0]
        if hits.numel() < n * 32:
            raise ValueError(f"hit buffer of batch {i} too small")
        arr[i].volume = dvol._h.value if isinstance(dvol._h, ctypes.c_void_p) else dvol._h
        arr[i].rays = rays.data_ptr()
        arr[i].hits = hits.data_ptr()
        arr[i].n = n
        arr[i].flags = CQX_SURFACE_PROPERTIES if surf else 0
        arr[i].max_footprint = float(fp)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    cfgp = ctypes.byref(cfg) if cfg is not None else None
    check(lib().cqx_intersect_batched(ctx._h, arr, len(items), cfgp, ctypes.c_void_p(stream)))


class Materials:
    """Material metadata (the .toml next to a .dag; Metadata, src/application/base/metadata.h:37-51):
    a structured array of MATERIAL_DTYPE rows indexed by material id, plus optional dimensions."""

    def __init__(self, rows=None, dimensions=None):
        self.rows = np.zeros(0, MATERIAL_DTYPE) if rows is None else np.asarray(rows, MATERIAL_DTYPE)
        self.dimensions = dimensions

    @staticmethod
    def load(path):
        rows = np.zeros(256, MATERIAL_DTYPE)
        count = ctypes.c_uint32()
        dims = (ctypes.c_int32 * 3)()
        has = ctypes.c_uint32()
        check(lib().cqx_materials_load(os.fsencode(path), rows.ctypes.data_as(ctypes.c_void_p), ctypes.byref(count),
                                       dims, ctypes.byref(has)))
        return Materials(rows[: count.value].copy(), tuple(dims) if has.value else None)

    @staticmethod
    def from_tables(albedo, emission=None, names=None):
        """Rows for material ids 0..n from path-tracer tables of ids 1..n (id 0 is empty space)."""
        albedo = np.asarray(albedo, np.float32).reshape(-1, 3)
        rows = np.zeros(albedo.shape[0] + 1, MATERIAL_DTYPE)
        rows[0]["name"] = b"EmptySpace"
        rows[0]["flags"] = cabi.MATERIAL_HAS_NAME
        rows["base_color"][0] = 0.8
        rows["base_color"][1:] = albedo
        rows["flags"][1:] = cabi.MATERIAL_HAS_BASE_COLOR
        if emission is not None:
            rows["emission"][1:] = np.asarray(emission, np.float32).reshape(-1, 3)
            rows["flags"][1:] |= cabi.MATERIAL_HAS_EMISSION
        for i, nm in enumerate(names or []):
            rows[i + 1]["name"] = nm.encode()
            rows[i + 1]["flags"] |= cabi.MATERIAL_HAS_NAME
        return Materials(rows)

    def save(self, path):
        rows = np.ascontiguousarray(self.rows, MATERIAL_DTYPE)
        check(lib().cqx_materials_save(os.fsencode(path), rows.ctypes.data_as(ctypes.c_void_p), len(rows)))

    def colour_table(self):
        """The viewer's 256 x 3 colour table (purple where no material is defined)."""
        rows = np.ascontiguousarray(self.rows, MATERIAL_DTYPE)
        out = np.zeros((256, 3), np.float32)
        check(lib().cqx_materials_colour_table(rows.ctypes.data_as(ctypes.c_void_p), len(rows),
                                               out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def pathtrace_tables(self):
        """(albedo, emission) rows for material ids 1..n, the path tracer's tables."""
        n = max(0, len(self.rows) - 1)
        return (np.ascontiguousarray(self.rows["base_color"][1:], np.float32).reshape(n, 3),
                np.ascontiguousarray(self.rows["emission"][1:], np.float32).reshape(n, 3))


def material_table(n_materials, seed, n_emissive=1, emission=5.0):
    """Albedo U[0.1, 0.9]^3 per material (numpy PCG64(seed)); the last `n_emissive` materials
    emit `emission` on every channel. Rows are material ids 1..n_materials."""
    rng = np.random.Generator(np.random.PCG64(seed))
    albedo = rng.uniform(0.1, 0.9, (n_materials, 3)).astype(np.float32)
    em = np.zeros((n_materials, 3), np.float32)
    if n_emissive:
        em[-n_emissive:] = emission
    return albedo, em


class PathTraceSetup:
    """Parameters of one render (cqx_pt_params); keeps its material tables alive."""

    def __init__(self, camera, width, height, spp, bounces, albedo, emission, sky=(0.8, 0.8, 1.0),
                 bounce_offset=0.01, max_footprint=MAX_FOOTPRINT_DISABLED, seed=0):
        self.albedo = np.ascontiguousarray(albedo, dtype=np.float32).reshape(-1, 3)
        self.emission = np.ascontiguousarray(emission, dtype=np.float32).reshape(-1, 3)
        if self.albedo.shape != self.emission.shape:
            raise ValueError("albedo and emission tables differ in size")
        p = cabi.PtParams()
        p.width, p.height, p.spp, p.bounces = int(width), int(height), int(spp), int(bounces)
        p.eye = (ctypes.c_double * 3)(*camera.eye)
        p.target = (ctypes.c_double * 3)(*camera.target)
        p.fov_y_degrees = float(camera.fov_y)
        p.n_materials = self.albedo.shape[0]
        p.albedo = self.albedo.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        p.emission = self.emission.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        p.sky = (ctypes.c_float * 3)(*sky)
        p.bounce_offset = float(bounce_offset)
        p.max_footprint = float(max_footprint)
        p.seed = int(seed)
        self.params = p
        self.width, self.height, self.spp, self.bounces = p.width, p.height, p.spp, p.bounces
        self.camera = camera
        self.sky = tuple(sky)
        self.bounce_offset = bounce_offset
        self.max_footprint = max_footprint
        self.seed = seed

    def with_tables(self, albedo, emission):
        return PathTraceSetup(self.camera, self.width, self.height, self.spp, self.bounces, albedo, emission,
                              self.sky, self.bounce_offset, self.max_footprint, self.seed)


def pathtrace(ctx, dvol, setup, mode=0, cfg=None, out=None, stream=None):
    """Render (cqx_pathtrace): returns a float32 GPU tensor (height, width, 3), linear RGB."""
    import torch

    if out is None:
        out = torch.empty((setup.height, setup.width, 3), dtype=torch.float32, device="cuda")
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    cfgp = ctypes.byref(cfg) if cfg is not None else None
    check(lib().cqx_pathtrace(ctx._h, dvol._h, ctypes.byref(setup.params), int(mode), cfgp,
                              ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream)))
    return out


def demo_trace(ctx, dvol, rays, colours, max_footprint=0.0035, gamma=0.0, add_noise=True, include_sun=True,
               include_sky=True, seed=0, out=None, accumulate=False, cfg=None, stream=None):
    """The reference CPU tracer's shading (cqx_demo_trace, PathtracingDemo::traceSingleRay,
    src/application/commands/view/pathtracing_demo.cpp:150-194) for one primary ray per pixel.

    ``rays``: float32 GPU tensor (n, 6); ``colours``: (256, 3) material colours (e.g.
    ``Materials.colour_table()``). Returns (or fills / adds to, with ``accumulate``) a float32 GPU
    tensor (n, 3). Per-pixel generator states from ``seed``; gamma 0 is the reference's 1/2.2."""
    import torch

    if not (isinstance(rays, torch.Tensor) and rays.is_cuda and rays.dtype == torch.float32):
        raise TypeError("rays must be a float32 GPU tensor of shape (n, 6)")
    rays = rays.contiguous().view(-1, 6)
    n = rays.shape[0]
    col = np.ascontiguousarray(colours, dtype=np.float32).reshape(256, 3)
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs an output tensor")
        out = torch.empty((n, 3), dtype=torch.float32, device=rays.device)
    if out.numel() < n * 3 or out.dtype != torch.float32:
        raise ValueError("out must be a float32 tensor of n x 3")
    if stream is None:
        stream = torch.cuda.current_stream(rays.device).cuda_stream
    prm = cabi.DemoParams(col.ctypes.data, float(max_footprint), float(gamma), int(bool(add_noise)),
                          int(bool(include_sun)), int(bool(include_sky)), int(bool(accumulate)), int(seed))
    cfgp = ctypes.byref(cfg) if cfg is not None else None
    check(lib().cqx_demo_trace(ctx._h, dvol._h, ctypes.c_void_p(rays.data_ptr()), n, ctypes.byref(prm), cfgp,
                               ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream)))
    return out


def pathtrace_adjoint(ctx, dvol, setup, d_image, cfg=None, stream=None):
    """dLoss/dalbedo and dLoss/demission (n_materials x 3 GPU tensors) for dLoss/dimage `d_image`."""
    import torch

    d_image = d_image.contiguous()
    m = setup.albedo.shape[0]
    da = torch.empty((m, 3), dtype=torch.float32, device="cuda")
    de = torch.empty((m, 3), dtype=torch.float32, device="cuda")
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    cfgp = ctypes.byref(cfg) if cfg is not None else None
    check(lib().cqx_pathtrace_adjoint(ctx._h, dvol._h, ctypes.byref(setup.params), cfgp,
                                      ctypes.c_void_p(d_image.data_ptr()), ctypes.c_void_p(da.data_ptr()),
                                      ctypes.c_void_p(de.data_ptr()), ctypes.c_void_p(stream)))
    return da, de


def intersectVolume(ctx, dvol, rays, computeSurfaceProperties=True, maxFootprint=MAX_FOOTPRINT_DISABLED,
                    cfg=None, stream=None, out=None):
    """Batched intersectVolume (src/library/raytracing.cpp:397-478) on device-resident rays.

    ``rays`` is a torch tensor (n, 6) float32 on the GPU; returns (or fills ``out``) a torch
    uint8 tensor of n 32-byte records (view with ``HIT_DTYPE`` after ``.cpu().numpy()``).
    """
    import torch

    if not (isinstance(rays, torch.Tensor) and rays.is_cuda and rays.dtype == torch.float32):
        raise TypeError("rays must be a float32 GPU tensor of shape (n, 6)")
    rays = rays.contiguous().view(-1, 6)
    n = rays.shape[0]
    if out is None:
        out = torch.empty((n, 32), dtype=torch.uint8, device=rays.device)
    if out.numel() < n * 32:
        raise ValueError("out too small")
    if stream is None:
        stream = torch.cuda.current_stream(rays.device).cuda_stream
    flags = CQX_SURFACE_PROPERTIES if computeSurfaceProperties else 0
    ctx.intersect(dvol, rays.data_ptr(), n, out.data_ptr(), flags, maxFootprint, cfg, stream)
    return out
