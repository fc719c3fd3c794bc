"""Writes tests/golden/raycast_128_golden.npz from the compiled, unmodified reference library
(oracle/_ref/libcubiquity_ref.so, built from the reference sources by oracle/Makefile).

Contents (all produced by the reference's own intersectVolume / traceRayRef / findSubDAGs on the
.dag file written by this library's generator for the 128^3 raycast workload):
  dag_sha256           SHA-256 of the (node_count, 8) uint32 store the fixture was made on
  subdags              findSubDAGs result (padding zeroed: the reference leaves it uninitialised)
  rays                 a subset: 2048 camera rays, 2048 incoherent rays and all special rays
  ref_surf1, ref_surf0 intersectVolume results on `rays` (surface properties on / off)
  ref_lod              intersectVolume results on `rays` with maxFootprint = 0.05
  tref                 traceRayRef hit / material / distance on the first 512 incoherent rays
  camera_rays          all 65,536 camera rays
  camera_ref_sha256    SHA-256 of the reference's records for all camera rays (surface on)

lattice_128_golden.npz (python tests/golden/make_golden.py --lattice), same volume:
  rays                 4096 lattice rays (tests/raysets.py lattice_rays, seed 77) on which the
                       reference terminates (rays with a zero component whose origin lies on a
                       node boundary can loop forever in intersectRayNodeESVO; the oracle's
                       iteration cap finds them beforehand and they are left out)
  ref_surf1, ref_surf0 intersectVolume results (surface properties on / off)
  ref_lod_*            intersectVolume results at maxFootprint 0.05, 0.3 and 2.0

Run: python tests/golden/make_golden.py [--lattice]   (needs the reference checkout to build oracle/_ref)
"""
import hashlib
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import cubiquity_amd as cq  # noqa: E402
from cubiquity_amd import workloads as W  # noqa: E402
from oracle.bindings import Reference, build  # noqa: E402
from tests.raysets import lattice_rays, special_rays  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "raycast_128_golden.npz")
OUT_LATTICE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lattice_128_golden.npz")
LATTICE_LODS = (0.05, 0.3, 2.0)


def lattice():
    from oracle.bindings import Oracle

    build(ref=True)
    ref = Reference()
    orc = Oracle()
    w = W.RAYCAST_128
    hv = cq.HostVolume.generate(w.params())
    rays = lattice_rays(w.side, 4096 + 512, 77)
    sub = orc.find_subdags(hv.nodes, hv.root)
    ok = np.ones(rays.shape[0], bool)
    for surf, fp in [(True, -1.0)] + [(True, f) for f in LATTICE_LODS]:
        h = orc.intersect(hv.nodes, sub, rays, surf, fp, max_iters=1 << 16)
        ok &= ((h["info"] >> 15) & 1) == 0
    rays = rays[ok][:4096]
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "v.dag")
        hv.save_dag(path)
        rv = ref.load(path)
        out = {"dag_sha256": hashlib.sha256(hv.nodes.tobytes()).hexdigest(), "rays": rays,
               "ref_surf1": rv.intersect(rays, True), "ref_surf0": rv.intersect(rays, False)}
        for fp in LATTICE_LODS:
            out[f"ref_lod_{fp}"] = rv.intersect(rays, True, fp)
        np.savez_compressed(OUT_LATTICE, **out)
    print(OUT_LATTICE, os.path.getsize(OUT_LATTICE), "bytes")


def main():
    build(ref=True)
    ref = Reference()
    w = W.RAYCAST_128
    hv = cq.HostVolume.generate(w.params())
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "v.dag")
        hv.save_dag(path)
        rv = ref.load(path)
        c = w.camera
        cam = cq.camera_rays(c.eye, c.target, c.fov_y, c.width, c.height)
        inc = cq.incoherent_rays(hv, w.log2_size, w.incoherent_seed, w.incoherent_rays)
        rng = np.random.Generator(np.random.PCG64(2024))
        ci = np.sort(rng.choice(cam.shape[0], 2048, replace=False))
        rays = np.concatenate([cam[ci], inc[:2048], special_rays(w.side)]).astype(np.float32)
        sub = rv.find_subdags()
        for f in ("pad0", "pad1", "pad2"):
            sub[f] = 0
        tref = rv.trace_ref(inc[:512])
        np.savez_compressed(
            OUT,
            dag_sha256=hashlib.sha256(hv.nodes.tobytes()).hexdigest(),
            subdags=sub,
            rays=rays,
            ref_surf1=rv.intersect(rays, True),
            ref_surf0=rv.intersect(rays, False),
            ref_lod=rv.intersect(rays, True, 0.05),
            tref_rays=inc[:512],
            tref_hit=tref["hit"],
            tref_material=tref["material"],
            tref_distance=tref["distance"],
            camera_rays=cam,
            camera_ref_sha256=hashlib.sha256(rv.intersect(cam, True).tobytes()).hexdigest(),
        )
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    lattice() if "--lattice" in sys.argv else main()
