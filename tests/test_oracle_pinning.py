"""Pins the oracle (the C restatement in oracle/cpu_ref.c) before anything is checked against it.

  * bit-exact against the compiled, unmodified reference (oracle/_ref), live where present and
    through the committed golden fixtures everywhere;
  * against an independent FP64 dense brute force (no reference code path) on small volumes,
    including volumes spanning several root octants;
  * against closed forms: axis-aligned and diagonal rays into a single solid box;
  * the derived outputs (hit voxel, parent node, slot) against the closed-form root descent of
    Volume::voxel (src/library/storage.cpp:481-499).
  * exact-tie lattice rays (tests/raysets.py lattice_rays) against the compiled reference's
    records in tests/golden/lattice_128_golden.npz.
Planted slips (tests/test_oracle_mutations.py: a flipped comparison, operand order, sign, shift
or table entry, each in a copy of the oracle) each fail one of these checks.
"""
import hashlib
import os

import numpy as np
import pytest

from tests.conftest import records_equal

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "raycast_128_golden.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


def _sub(oracle, hv):
    return oracle.find_subdags(hv.nodes, hv.root)


# ---- against the compiled reference -------------------------------------------------------------

@pytest.mark.parametrize("surf,fp", [(True, -1.0), (False, -1.0), (True, 0.05), (True, 0.0035)])
@pytest.mark.parametrize("set_name", ["camera", "incoherent", "special"])
def test_oracle_bitexact_vs_reference(oracle, reference, vol128, dag128, rays128, set_name, surf, fp):
    rays = rays128[set_name]
    rv = reference.load(dag128)
    o = oracle.to_results(rays, oracle.intersect(vol128.nodes, _sub(oracle, vol128), rays, surf, fp), surf)
    r = rv.intersect(rays, surf, fp)
    eq = records_equal(o, r)
    assert eq.all(), f"{(~eq).sum()} of {len(eq)} differ"


def test_trace_ref_vs_reference(oracle, reference, vol128, dag128, rays128):
    rays = rays128["incoherent"][:400]
    rv = reference.load(dag128)
    a = oracle.trace_ref(vol128.nodes, vol128.root, rays)
    b = rv.trace_ref(rays)
    # the reference leaves position/normal of traceRayRef uninitialised; compare what it sets
    for f in ("hit", "material", "distance"):
        assert (a[f] == b[f]).all(), f


def test_subdags_vs_reference(oracle, reference, cq, vol128, dag128):
    rv = reference.load(dag128)
    r = rv.find_subdags()
    o = _sub(oracle, vol128)
    p = vol128.find_subdags()
    for f in ("lower", "height", "node"):  # padding is uninitialised in the reference
        assert (r[f] == o[f]).all() and (p[f] == o[f]).all(), f


# ---- against the golden fixtures (no reference needed) ------------------------------------------

def test_oracle_vs_golden(oracle, vol128, golden):
    assert hashlib.sha256(vol128.nodes.tobytes()).hexdigest() == str(golden["dag_sha256"])
    sub = _sub(oracle, vol128)
    rays = golden["rays"]
    for surf, fp, key in ((True, -1.0, "ref_surf1"), (False, -1.0, "ref_surf0"), (True, 0.05, "ref_lod")):
        o = oracle.to_results(rays, oracle.intersect(vol128.nodes, sub, rays, surf, fp), surf)
        eq = records_equal(o, golden[key])
        assert eq.all(), f"{key}: {(~eq).sum()} differ"
    cam = golden["camera_rays"]
    o = oracle.to_results(cam, oracle.intersect(vol128.nodes, sub, cam, True, threads=8), True)
    assert hashlib.sha256(o.tobytes()).hexdigest() == str(golden["camera_ref_sha256"])
    t = oracle.trace_ref(vol128.nodes, vol128.root, golden["tref_rays"])
    assert (t["hit"] == golden["tref_hit"]).all()
    assert (t["material"] == golden["tref_material"]).all()
    assert (t["distance"] == golden["tref_distance"]).all()
    g = golden["subdags"]
    for f in ("lower", "height", "node"):
        assert (sub[f] == g[f]).all()


def test_oracle_vs_lattice_golden(oracle, vol128, golden):
    from tests import pt_pins

    pt_pins.check_traversal_golden(oracle, vol128, golden, np.load(pt_pins.LATTICE, allow_pickle=False))


def test_oracle_vs_reference_lattice_live(oracle, reference, vol128, dag128):
    """A fresh lattice set (another seed) against the live reference; rays the oracle's cap flags
    (the reference's non-terminating cases) are left out before the reference sees them."""
    from tests.raysets import lattice_rays

    rays = lattice_rays(128, 3000, 78)
    sub = _sub(oracle, vol128)
    h = oracle.intersect(vol128.nodes, sub, rays, True, -1.0, max_iters=1 << 16)
    rays = rays[((h["info"] >> 15) & 1) == 0]
    rv = reference.load(dag128)
    for surf, fp in ((True, -1.0), (False, -1.0), (True, 0.3)):
        o = oracle.to_results(rays, oracle.intersect(vol128.nodes, sub, rays, surf, fp), surf)
        eq = records_equal(o, rv.intersect(rays, surf, fp))
        assert eq.all(), (surf, fp, (~eq).sum())


# ---- against an independent FP64 dense brute force ---------------------------------------------

def _blob_grid(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    z, y, x = np.mgrid[0:n, 0:n, 0:n]
    g = np.zeros((n, n, n), np.uint8)
    for k in range(6):
        c = rng.uniform(0, n, 3)
        r = rng.uniform(n / 10, n / 4)
        g[(x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r] = 1 + k % 5
    g[: n // 5, :, :] = np.where(g[: n // 5] == 0, 7, g[: n // 5])  # a slab (uniform blocks)
    return g


def _brute_rays(n, lower, count, seed):
    """Origins in empty voxels of the box (or outside it), directions uniform."""
    rng = np.random.Generator(np.random.PCG64(seed))
    o = rng.uniform(-0.5, n - 0.5, (count, 3)) + np.asarray(lower, np.float64)
    out_box = rng.uniform(-n, 2 * n, (count // 4, 3)) + np.asarray(lower, np.float64)
    o[: count // 4] = out_box
    d = rng.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


@pytest.mark.parametrize("lower", [(0, 0, 0), (-20, -9, -27), (-48, 5, -1)])
def test_oracle_vs_dense_bruteforce(cq, oracle, lower):
    n = 48
    grid = _blob_grid(n, 70 + lower[0])  # [z][y][x]
    hv = cq.HostVolume.from_dense(grid, lower)
    rays = _brute_rays(n, lower, 3000, 110 + lower[1])
    # keep rays whose origin voxel is empty (inside the box) or that start outside the box
    # heading towards it; the reference does not test exit > 0 when the origin is beyond a
    # SubDAG node and reports geometry behind the origin, which brute force would not
    rel = rays[:, :3].astype(np.float64) - np.asarray(lower)
    v = np.floor(rel + 0.5).astype(int)
    inside = np.all((v >= 0) & (v < n), axis=1)
    empty = np.ones(len(rays), bool)
    empty[inside] = grid[v[inside, 2], v[inside, 1], v[inside, 0]] == 0
    rays = rays[empty]
    sub = oracle.find_subdags(hv.nodes, hv.root)
    hits = oracle.intersect(hv.nodes, sub, rays, True)
    shifted = rays.copy()
    shifted[:, :3] = (rays[:, :3].astype(np.float64) - np.asarray(lower)).astype(np.float32)
    hit, tn, ts, mat, vox = oracle.dense_brute(grid, shifted)
    h = (hits["info"] & 1).astype(bool)
    # exclude near-ties (edge/corner grazing) where the nearest voxel is not unique in FP32, and
    # the reference's hits behind an origin that lies beyond a SubDAG node (negative t)
    ok = ~(np.isfinite(ts) & (ts - tn < 1e-3))
    behind = h & (hits["t"] < 0)
    assert behind.sum() < 0.05 * len(rays)
    ok &= ~behind
    assert ok.sum() > 0.9 * len(rays)
    assert (h[ok] == hit[ok]).all()
    both = ok & h
    assert both.sum() > 200
    m = (hits["info"] >> 16) & 255
    assert (m[both] == mat[both]).all()
    assert (hits["voxel"][both] - np.asarray(lower) == vox[both]).all()
    rel_err = np.abs(hits["t"][both] - tn[both]) / np.maximum(1.0, np.abs(tn[both]))
    assert rel_err.max() < 1e-5


def test_octant_walk_vs_reference(cq, oracle, reference, tmp_path):
    """A volume spanning the origin occupies several root octants (several SubDAGs)."""
    n = 40
    grid = _blob_grid(n, 3)
    hv = cq.HostVolume.from_dense(grid, (-17, -23, -11))
    sub = oracle.find_subdags(hv.nodes, hv.root)
    assert (sub["node"] > 0).sum() == 8
    path = str(tmp_path / "span.dag")
    hv.save_dag(path)
    rv = reference.load(path)
    rays = _brute_rays(n, (-17, -23, -11), 4000, 5)
    for surf, fp in ((True, -1.0), (False, -1.0), (True, 0.2)):
        o = oracle.to_results(rays, oracle.intersect(hv.nodes, sub, rays, surf, fp), surf)
        assert records_equal(o, rv.intersect(rays, surf, fp)).all()


def test_small_and_degenerate_volumes_vs_reference(cq, oracle, reference, tmp_path):
    """Single voxel (a height-0 SubDAG that is a material node), a solid cube (a material
    SubDAG), an empty volume, volumes spanning the origin or in a negative octant."""
    from tests.raysets import special_rays
    from tests.test_gpu_parity import _small_volumes

    for name, (hv, n, lower) in _small_volumes(cq).items():
        path = str(tmp_path / f"{name}.dag")
        hv.save_dag(path)
        rv = reference.load(path)
        sp = special_rays(n)
        sp[:, :3] += np.asarray(lower, np.float32)
        rays = np.concatenate([_brute_rays(n, lower, 4000, 17), sp]).astype(np.float32)
        sub = oracle.find_subdags(hv.nodes, hv.root)
        for surf, fp in ((True, -1.0), (False, -1.0), (True, 0.1)):
            h = oracle.intersect(hv.nodes, sub, rays, surf, fp, max_iters=1 << 16)
            # Rays with a zero direction component whose shifted origin lies exactly on a node
            # boundary of that axis make tChildExit NaN; no axis ever advances and the reference
            # loops forever (raytracing.cpp:259-330). The oracle stops them at the cap and flags
            # them; only the others can be run through the reference.
            an = ((h["info"] >> 15) & 1).astype(bool)
            if an.any():
                zero = (rays[an, 3:] == 0).any(axis=1)
                assert zero.all(), f"{name}: anomaly on a ray without a zero direction component"
                assert (h["iters"][an] == 1 << 16).all()
            ok = ~an
            eq = records_equal(oracle.to_results(rays[ok], h[ok], surf), rv.intersect(rays[ok], surf, fp))
            assert eq.all(), f"{name}: {(~eq).sum()} differ"


def test_deep_sparse_volumes_vs_reference(cq, oracle, reference, tmp_path):
    """SubDAGs 21-22 levels deep (spheres about 2^20 voxels apart): long walks through empty space,
    deep stacks and pops back up, bit-exact against the compiled reference."""
    from tests.raysets import deep_volumes

    for name, (hv, rays) in deep_volumes(cq).items():
        path = str(tmp_path / f"{name}.dag")
        hv.save_dag(path)
        rv = reference.load(path)
        sub = oracle.find_subdags(hv.nodes, hv.root)
        assert max(int(s["height"]) for s in sub if s["node"]) >= 21
        for surf, fp in ((True, -1.0), (False, -1.0), (True, 0.01)):
            h = oracle.intersect(hv.nodes, sub, rays, surf, fp, max_iters=1 << 16)
            assert not ((h["info"] >> 15) & 1).any()
            assert (h["info"] & 1).sum() > len(rays) // 3
            eq = records_equal(oracle.to_results(rays, h, surf), rv.intersect(rays, surf, fp))
            assert eq.all(), f"{name} surf={surf} fp={fp}: {(~eq).sum()} differ"


# ---- closed forms ------------------------------------------------------------------------------

# Direction components of exactly zero are not handled by the reference's octant walk: with
# o' > 0 on such an axis, distToOrigin = -o'/+0 = -inf stays -inf after += FLT_MAX, so that axis
# is stepped again and the walk skips octants (raytracing.cpp:431-473). Near-axis rays use a tiny
# non-zero component instead; it changes none of the distances below.
EPS = 1e-30


def test_zero_direction_component_reference_behaviour(cq, oracle, reference, tmp_path):
    """Pins the reference's (replicated) result for exactly axis-aligned rays: a miss."""
    g = np.zeros((32, 32, 32), np.uint8)
    g[8:24, 8:24, 8:24] = 5
    hv = cq.HostVolume.from_dense(g, (0, 0, 0))
    path = str(tmp_path / "box.dag")
    hv.save_dag(path)
    rays = np.asarray([[-3.25, 12.0, 12.0, 1.0, 0.0, 0.0], [-3.25, 12.0, 12.0, 1.0, EPS, EPS]], np.float32)
    o = oracle.to_results(rays, oracle.intersect(hv.nodes, oracle.find_subdags(hv.nodes, hv.root), rays))
    r = reference.load(path).intersect(rays)
    assert records_equal(o, r).all()
    assert r["hit"].tolist() == [0, 1]


def test_closed_form_box(cq, oracle):
    """A solid box [8,24)^3 of material 5 in a 32^3 grid: near-axis rays enter the face at
    x = 7.5 (voxel 8 spans [7.5, 8.5)), so t = 7.5 - o_x exactly, normal (-1, -0, -0)."""
    g = np.zeros((32, 32, 32), np.uint8)
    g[8:24, 8:24, 8:24] = 5
    hv = cq.HostVolume.from_dense(g, (0, 0, 0))
    sub = oracle.find_subdags(hv.nodes, hv.root)
    ys = np.linspace(7.6, 23.4, 9)
    rays = []
    for y in ys:
        for z in ys:
            rays.append([-3.25, y, z, 1.0, EPS, EPS])
    rays.append([40.0, 12.0, 12.0, -1.0, EPS, EPS])   # from +x: face at 23.5, t = 16.5
    rays.append([12.0, 12.0, -10.0, EPS, EPS, 1.0])   # from -z: face at 7.5, t = 17.5
    rays.append([30.0, 30.0, 30.0, 1.0, 0.0, 0.0])    # misses
    rays = np.asarray(rays, np.float32)
    h = oracle.intersect(hv.nodes, sub, rays, True)
    r = oracle.to_results(rays, h)
    k = len(ys) ** 2
    assert (r["hit"][:k] == 1).all()
    assert (r["distance"][:k] == 7.5 + 3.25).all()
    assert (r["material"][:k] == 5).all()
    assert np.array_equal(np.signbit(r["normal"][:k]), np.tile([True, True, True], (k, 1)))
    assert (r["normal"][:k, 0] == -1.0).all() and (r["normal"][:k, 1:] == 0.0).all()
    assert (h["voxel"][:k, 0] == 8).all()
    assert r["distance"][k] == 16.5 and r["normal"][k, 0] == 1.0 and h["voxel"][k, 0] == 23
    # +0 direction components count as positive: their zero normal components are -0.0
    assert np.signbit(r["normal"][k, 1]) and np.signbit(r["normal"][k, 2])
    assert r["distance"][k + 1] == 17.5 and r["normal"][k + 1, 2] == -1.0 and h["voxel"][k + 1, 2] == 8
    assert r["hit"][k + 2] == 0
    # diagonal into the corner (7.5,7.5,7.5): entry on all three faces, every normal axis set
    d = np.float32(1.0 / np.sqrt(3.0))
    ray = np.asarray([[2.5, 2.5, 2.5, d, d, d]], np.float32)
    hd = oracle.intersect(hv.nodes, sub, ray, True)
    rd = oracle.to_results(ray, hd)
    assert rd["hit"][0] == 1 and tuple(hd["voxel"][0]) == (8, 8, 8)
    assert abs(rd["distance"][0] - 5.0 * np.sqrt(3.0)) < 1e-5


def test_closed_form_miss_and_bounds(cq, oracle):
    g = np.zeros((16, 16, 16), np.uint8)
    g[4, 4, 4] = 9
    hv = cq.HostVolume.from_dense(g, (0, 0, 0))
    sub = oracle.find_subdags(hv.nodes, hv.root)
    rays = np.asarray([[4.0, 4.0, -7.0, EPS, EPS, 1.0],    # hits the single voxel at z face 3.5
                       [4.0, 4.6, -7.0, EPS, EPS, 1.0],    # passes above it (voxel spans y<4.5)
                       [-1e6, 4.0, 4.0, 1.0, EPS, EPS]],   # from far away
                      np.float32)
    r = oracle.to_results(rays, oracle.intersect(hv.nodes, sub, rays))
    assert r["hit"].tolist() == [1, 0, 1]
    assert r["distance"][0] == 10.5 and r["material"][0] == 9
    assert r["distance"][2] == np.float32(1e6) + np.float32(3.5)


# ---- derived outputs: hit voxel, parent node and slot ------------------------------------------

@pytest.mark.parametrize("set_name", ["camera", "incoherent"])
def test_voxel_node_slot_by_descent(oracle, vol128, rays128, set_name):
    rays = rays128[set_name]
    nodes = vol128.nodes
    h = oracle.intersect(nodes, _sub(oracle, vol128), rays, True)
    idx = np.nonzero(h["info"] & 1)[0]
    rng = np.random.Generator(np.random.PCG64(3))
    idx = rng.choice(idx, min(3000, len(idx)), replace=False)
    for i in idx:
        rec = h[i]
        x, y, z = (int(v) for v in rec["voxel"])
        mat = (int(rec["info"]) >> 16) & 255
        k = (int(rec["info"]) >> 4) & 31
        slot = (int(rec["info"]) >> 1) & 7
        assert oracle.voxel(nodes, vol128.root, x, y, z) == mat
        path = oracle.descend(nodes, vol128.root, x, y, z)
        assert path[k + 1] == rec["node"]
        assert nodes[rec["node"], slot] == path[k]
        # the voxel's FP64 slab entry equals the hit distance
        o = rays[i, :3].astype(np.float64)
        d = rays[i, 3:].astype(np.float64)
        lo = (np.array([x, y, z]) - 0.5 - o) / d
        hi = (np.array([x, y, z]) + 0.5 - o) / d
        entry = np.max(np.minimum(lo, hi))
        assert abs(entry - rec["t"]) <= 1e-5 * max(1.0, abs(entry))


def test_iteration_cap(oracle, vol128, rays128):
    rays = rays128["camera"][32768:34768]  # image centre: every ray enters the volume
    h = oracle.intersect(vol128.nodes, _sub(oracle, vol128), rays, True, max_iters=4)
    an = ((h["info"] >> 15) & 1).astype(bool)
    assert an.any()
    assert (h["iters"][an] == 4).all()
    assert (h["t"][an] == 0).all() and ((h["info"][an] & 1) == 0).all()


def test_dense_brute_self_check(oracle):
    g = np.zeros((4, 4, 4), np.uint8)
    g[0, 0, 0] = 3
    rays = np.asarray([[-5.0, 0.0, 0.0, 1.0, 0.0, 0.0], [-5.0, 0.0, 1.0, 1.0, 0.0, 0.0]], np.float32)
    hit, tn, ts, mat, vox = oracle.dense_brute(g, rays)
    assert hit.tolist() == [True, False] and tn[0] == 4.5 and mat[0] == 3
