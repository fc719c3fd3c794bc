"""GPU parity: the HIP traversal (through the C ABI) against the oracle, the golden fixtures made
by the compiled reference, and the compiled reference itself where it is present.

Bar: every field of every 32-byte record bit-identical (t bits, voxel, node, slot, block size,
normal code incl. signed zeros, material, node-entry and iteration counts)."""
import hashlib
import os

import numpy as np
import pytest

from tests.conftest import records_equal

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "raycast_128_golden.npz")


def _gpu(ctx, dvol, rays, surf=True, fp=-1.0, cfg=None):
    import torch

    from cubiquity_amd import HIT_DTYPE, intersectVolume

    r = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32)).cuda()
    out = intersectVolume(ctx, dvol, r, computeSurfaceProperties=surf, maxFootprint=fp, cfg=cfg)
    torch.cuda.synchronize()
    return out.cpu().numpy().view(HIT_DTYPE).reshape(-1)


@pytest.fixture(scope="module")
def dvol128(gpu_ctx, vol128):
    return gpu_ctx.upload(vol128)


def _report(name, eq, rays, g, o):
    if not eq.all():
        i = np.nonzero(~eq)[0][:4]
        return f"{name}: {(~eq).sum()} of {len(eq)} differ; first {i.tolist()}\nrays {rays[i]}\ngpu {g[i]}\noracle {o[i]}"
    return ""


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("surf,fp", [(True, -1.0), (False, -1.0), (True, 0.05), (True, 0.0035)])
@pytest.mark.parametrize("set_name", ["camera", "incoherent", "special"])
def test_gpu_matches_oracle_128(cq, oracle, gpu_ctx, dvol128, vol128, rays128, set_name, surf, fp, variant):
    rays = rays128[set_name]
    sub = oracle.find_subdags(vol128.nodes, vol128.root)
    o = oracle.intersect(vol128.nodes, sub, rays, surf, fp, threads=8)
    g = _gpu(gpu_ctx, dvol128, rays, surf, fp, cq.LaunchCfg(variant=variant))
    eq = records_equal(g, o)
    assert eq.all(), _report(set_name, eq, rays, g, o)


@pytest.mark.parametrize("variant", [2, 3, 4, 5])
@pytest.mark.parametrize("bin_bits", [1, 3, 4])
@pytest.mark.parametrize("set_name", ["camera", "incoherent", "special"])
def test_gpu_binned_matches_oracle_128(cq, oracle, gpu_ctx, dvol128, vol128, rays128, set_name, bin_bits, variant):
    """Rays binned by origin before the persistent launch: the same records, in the caller's order
    (the special set has origins outside the volume, infinite and NaN, which bin to border cells)."""
    rays = rays128[set_name]
    sub = oracle.find_subdags(vol128.nodes, vol128.root)
    o = oracle.intersect(vol128.nodes, sub, rays, True, -1.0, threads=8)
    g = _gpu(gpu_ctx, dvol128, rays, True, -1.0, cq.LaunchCfg(variant=variant, bin_bits=bin_bits))
    eq = records_equal(g, o)
    assert eq.all(), _report(set_name + " binned", eq, rays, g, o)


def test_gpu_binned_ragged_counts(cq, gpu_ctx, dvol128, vol128):
    """Binned launches of 1 ray, a few rays, and counts either side of where the prepass splits the
    rays into two and three workgroup ranges (16384 rays per range at least), against unbinned
    launches; then a smaller batch after a larger one (stale order entries)."""
    rays = cq.incoherent_rays(vol128, 7, 11, 50000)
    full = _gpu(gpu_ctx, dvol128, rays, True, -1.0, cq.LaunchCfg(variant=2))
    for n in (50000, 1, 17, 16383, 16384, 16385, 32767, 32768, 32769, 49151, 49152, 49153):
        g = _gpu(gpu_ctx, dvol128, rays[:n], True, -1.0, cq.LaunchCfg(variant=2, bin_bits=3))
        assert records_equal(g, full[:n]).all(), n


def test_gpu_binned_grid_follows_the_origins(cq, gpu_ctx, dvol128, vol128):
    """The binning grid spans the origins' bounds within the SubDAG box: origins confined to one
    corner of the volume, origins sharing one coordinate (a flat grid axis), and one shared origin
    (every axis falls back to the box) give the records of unbinned launches."""
    rays = cq.incoherent_rays(vol128, 7, 12, 40000)
    corner = rays.copy()
    corner[:, :3] = 1.0 + np.mod(np.abs(corner[:, :3]), 15.0)
    flat = rays.copy()
    flat[:, 1] = 64.0
    point = rays.copy()
    point[:, :3] = np.float32([40.5, 70.25, 90.0])
    for name, r in (("corner", corner), ("flat", flat), ("point", point)):
        r = np.ascontiguousarray(r, dtype=np.float32)
        full = _gpu(gpu_ctx, dvol128, r, True, -1.0, cq.LaunchCfg(variant=2))
        for bits in (2, 4):
            g = _gpu(gpu_ctx, dvol128, r, True, -1.0, cq.LaunchCfg(variant=3, bin_bits=bits))
            assert records_equal(g, full).all(), (name, bits)


def test_gpu_binning_rejects_bad_configs(cq, gpu_ctx, dvol128, rays128):
    import torch

    r = torch.from_numpy(rays128["incoherent"][:64]).cuda()
    for cfg in (cq.LaunchCfg(variant=2, bin_bits=5), cq.LaunchCfg(variant=1, bin_bits=2)):
        with pytest.raises(cq.CqxError):
            cq.intersectVolume(gpu_ctx, dvol128, r, True, cfg=cfg)
    h = torch.zeros((64, 32), dtype=torch.uint8, device="cuda")
    with pytest.raises(cq.CqxError):
        cq.intersect_batched(gpu_ctx, [(dvol128, r, h, True, -1.0)], cfg=cq.LaunchCfg(bin_bits=3))
    torch.cuda.synchronize()


def test_gpu_matches_golden(cq, gpu_ctx, dvol128, vol128):
    """Against the compiled reference's own outputs stored as fixtures (no reference needed)."""
    z = np.load(GOLDEN, allow_pickle=False)
    assert hashlib.sha256(vol128.nodes.tobytes()).hexdigest() == str(z["dag_sha256"])
    rays = z["rays"]
    for surf in (1, 0):
        g = _gpu(gpu_ctx, dvol128, rays, bool(surf))
        res = cq.hits_to_intersections(rays, g, bool(surf))
        ref = z[f"ref_surf{surf}"]
        eq = records_equal(res, ref)
        assert eq.all(), f"surf={surf}: {(~eq).sum()} rays differ from the reference fixture"
    # full 65,536-ray camera set: SHA-256 of the reference records
    cam = z["camera_rays"]
    g = _gpu(gpu_ctx, dvol128, cam, True)
    res = cq.hits_to_intersections(cam, g, True)
    assert hashlib.sha256(res.tobytes()).hexdigest() == str(z["camera_ref_sha256"])


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
def test_gpu_matches_lattice_golden(cq, oracle, gpu_ctx, dvol128, vol128, variant):
    """Lattice rays (exact FP32 ties at child boundaries, +-0 components on node boundaries)
    against the compiled reference's records, surface properties on and off and at three LOD
    footprints, and the full 32-byte records against the oracle."""
    from tests import pt_pins

    lat = np.load(pt_pins.LATTICE, allow_pickle=False)
    assert hashlib.sha256(vol128.nodes.tobytes()).hexdigest() == str(lat["dag_sha256"])
    sub = oracle.find_subdags(vol128.nodes, vol128.root)
    for name, rays, surf, fp, ref in pt_pins.golden_cases(np.load(GOLDEN, allow_pickle=False), lat):
        if not name.startswith("lattice/"):
            continue
        g = _gpu(gpu_ctx, dvol128, rays, surf, fp, cq.LaunchCfg(variant=variant))
        eq = records_equal(cq.hits_to_intersections(rays, g, surf), ref)
        assert eq.all(), f"{name}: {(~eq).sum()} rays differ from the reference fixture"
        o = oracle.intersect(vol128.nodes, sub, rays, surf, fp, threads=8)
        eq = records_equal(g, o)
        assert eq.all(), _report(name, eq, rays, g, o)


def test_gpu_matches_reference_live(cq, reference, gpu_ctx, dvol128, dag128, rays128):
    rv = reference.load(dag128)
    for name in ("camera", "incoherent", "special"):
        rays = rays128[name]
        g = _gpu(gpu_ctx, dvol128, rays, True)
        res = cq.hits_to_intersections(rays, g, True)
        ref = rv.intersect(rays, True)
        eq = records_equal(res, ref)
        assert eq.all(), f"{name}: {(~eq).sum()} differ from the compiled reference"


def _small_volumes(cq):
    from tests.test_oracle_pinning import _blob_grid

    g1 = np.zeros((16, 16, 16), np.uint8)
    g1[4, 4, 4] = 9
    g3 = np.zeros((4, 4, 4), np.uint8)
    return {
        "single_voxel": (cq.HostVolume.from_dense(g1, (0, 0, 0)), 16, (0, 0, 0)),
        "solid_cube": (cq.HostVolume.from_dense(np.full((8, 8, 8), 3, np.uint8), (0, 0, 0)), 8, (0, 0, 0)),
        "empty": (cq.HostVolume.from_dense(g3, (0, 0, 0)), 4, (0, 0, 0)),
        "span_origin": (cq.HostVolume.from_dense(_blob_grid(40, 3), (-17, -23, -11)), 40, (-17, -23, -11)),
        "negative_octant": (cq.HostVolume.from_dense(_blob_grid(32, 8), (-32, -32, -32)), 32, (-32, -32, -32)),
        "gen64_9mat": (cq.HostVolume.generate(cq.gen_default_params(6, 3, n_materials=9)), 64, (0, 0, 0)),
    }


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
def test_gpu_small_and_degenerate_volumes(cq, oracle, gpu_ctx, variant):
    from tests.raysets import special_rays
    from tests.test_oracle_pinning import _brute_rays

    for name, (hv, n, lower) in _small_volumes(cq).items():
        dv = gpu_ctx.upload(hv)
        sp = special_rays(n)
        sp[:, :3] += np.asarray(lower, np.float32)
        rays = np.concatenate([_brute_rays(n, lower, 4000, 17), sp]).astype(np.float32)
        sub = oracle.find_subdags(hv.nodes, hv.root)
        for surf, fp in ((True, -1.0), (False, -1.0), (True, 0.1)):
            # includes rays the reference would never finish (NaN exits); both sides cap them
            o = oracle.intersect(hv.nodes, sub, rays, surf, fp, max_iters=1 << 16)
            g = _gpu(gpu_ctx, dv, rays, surf, fp, cq.LaunchCfg(variant=variant, max_iters=1 << 16))
            eq = records_equal(g, o)
            assert eq.all(), _report(f"{name} surf={surf} fp={fp}", eq, rays, g, o)
        dv.free()


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
def test_gpu_deep_sparse_volumes(cq, oracle, gpu_ctx, variant):
    """SubDAGs 21-22 levels deep (23 stack rows per lane in LDS), every kernel variant."""
    from tests.raysets import deep_volumes

    for name, (hv, rays) in deep_volumes(cq).items():
        dv = gpu_ctx.upload(hv)
        sub = oracle.find_subdags(hv.nodes, hv.root)
        for surf, fp in ((True, -1.0), (False, -1.0), (True, 0.01)):
            o = oracle.intersect(hv.nodes, sub, rays, surf, fp, max_iters=1 << 16)
            g = _gpu(gpu_ctx, dv, rays, surf, fp, cq.LaunchCfg(variant=variant, max_iters=1 << 16))
            eq = records_equal(g, o)
            assert eq.all(), _report(f"{name} surf={surf} fp={fp}", eq, rays, g, o)
        dv.free()


@pytest.mark.parametrize("variant", [2, 3])
@pytest.mark.parametrize("stack_rows", [2, 4])
@pytest.mark.parametrize("bin_bits", [0, 3])
def test_gpu_short_stack_matches_oracle_128(cq, oracle, gpu_ctx, dvol128, vol128, rays128, bin_bits, stack_rows,
                                            variant):
    """Short stacks of 2 and 4 rows per lane on the 8-level SubDAGs of the 128^3 volume: pops past
    the kept rows descend again from the SubDAG root; every record (node-entry and iteration counts
    included) equals the full-stack walk's."""
    sub = oracle.find_subdags(vol128.nodes, vol128.root)
    for set_name in ("camera", "incoherent", "special"):
        rays = rays128[set_name]
        for surf, fp in ((True, -1.0), (True, 0.0035)):
            o = oracle.intersect(vol128.nodes, sub, rays, surf, fp, threads=8)
            cfg = cq.LaunchCfg(variant=variant, stack_rows=stack_rows, bin_bits=bin_bits)
            g = _gpu(gpu_ctx, dvol128, rays, surf, fp, cfg)
            eq = records_equal(g, o)
            assert eq.all(), _report(f"{set_name} fp={fp}", eq, rays, g, o)


@pytest.mark.parametrize("variant", [2, 3])
@pytest.mark.parametrize("stack_rows", [2, 4, 8, 16])
def test_gpu_short_stack_deep_sparse_volumes(cq, oracle, gpu_ctx, variant, stack_rows):
    """SubDAGs 21-22 levels deep walked with 2-16 stack rows per lane instead of 23."""
    from tests.raysets import deep_volumes

    for name, (hv, rays) in deep_volumes(cq).items():
        dv = gpu_ctx.upload(hv)
        sub = oracle.find_subdags(hv.nodes, hv.root)
        for surf, fp in ((True, -1.0), (True, 0.01)):
            o = oracle.intersect(hv.nodes, sub, rays, surf, fp, max_iters=1 << 16)
            cfg = cq.LaunchCfg(variant=variant, stack_rows=stack_rows, max_iters=1 << 16)
            g = _gpu(gpu_ctx, dv, rays, surf, fp, cfg)
            eq = records_equal(g, o)
            assert eq.all(), _report(f"{name} rows={stack_rows} fp={fp}", eq, rays, g, o)
        dv.free()


def test_gpu_short_stack_small_volumes_and_batched(cq, oracle, gpu_ctx):
    """Short stacks on the small and degenerate volumes (one launch each), then all of them in one
    batched launch with a short stack: each batch equals the oracle."""
    import torch

    from tests.raysets import special_rays
    from tests.test_oracle_pinning import _brute_rays

    launch, refs = [], []
    for name, (hv, n, lower) in _small_volumes(cq).items():
        dv = gpu_ctx.upload(hv)
        sp = special_rays(n)
        sp[:, :3] += np.asarray(lower, np.float32)
        rays = np.concatenate([_brute_rays(n, lower, 3000, 23), sp]).astype(np.float32)
        o = oracle.intersect(hv.nodes, oracle.find_subdags(hv.nodes, hv.root), rays, True, -1.0, max_iters=1 << 16)
        g = _gpu(gpu_ctx, dv, rays, True, -1.0, cq.LaunchCfg(variant=2, stack_rows=2, max_iters=1 << 16))
        eq = records_equal(g, o)
        assert eq.all(), _report(name, eq, rays, g, o)
        r = torch.from_numpy(rays).cuda()
        launch.append((dv, r, torch.zeros((len(rays), 32), dtype=torch.uint8, device="cuda"), True, -1.0))
        refs.append(o)
    for variant in (2, 3):
        for _, _, h, _, _ in launch:
            h.zero_()
        cq.intersect_batched(gpu_ctx, launch, cfg=cq.LaunchCfg(variant=variant, stack_rows=4, max_iters=1 << 16))
        torch.cuda.synchronize()
        for k, ((_, r, h, _, _), o) in enumerate(zip(launch, refs)):
            got = h.cpu().numpy().view(cq.HIT_DTYPE).reshape(-1)
            assert records_equal(got, o).all(), f"variant {variant} batch {k}"
    for dv, _, _, _, _ in launch:
        dv.free()


def test_gpu_short_stack_rejects_bad_configs(cq, gpu_ctx, dvol128, rays128):
    import torch

    r = torch.from_numpy(rays128["camera"][:256]).cuda()
    for cfg in (cq.LaunchCfg(variant=2, stack_rows=3), cq.LaunchCfg(variant=2, stack_rows=1),
                cq.LaunchCfg(variant=1, stack_rows=4), cq.LaunchCfg(variant=5, stack_rows=4)):
        with pytest.raises(cq.CqxError):
            cq.intersectVolume(gpu_ctx, dvol128, r, True, cfg=cfg)
    h = torch.zeros((256, 32), dtype=torch.uint8, device="cuda")
    with pytest.raises(cq.CqxError):
        cq.intersect_batched(gpu_ctx, [(dvol128, r, h, True, -1.0)], cfg=cq.LaunchCfg(variant=1, stack_rows=4))


def test_gpu_full_size_1024(cq, oracle, gpu_ctx):
    """The benchmark workload at full size: all 2^24 incoherent rays bit-exact against the
    oracle; both kernel variants agree; hit voxels hold the hit material (Volume::voxel)."""
    import time

    from cubiquity_amd import workloads as W

    w = W.PATHTRACE_1024
    hv = cq.HostVolume.generate(w.params())
    rays = cq.incoherent_rays(hv, w.log2_size, w.incoherent_seed, w.incoherent_rays)
    dv = gpu_ctx.upload(hv)
    t0 = time.time()
    g2 = _gpu(gpu_ctx, dv, rays, True, -1.0, cq.LaunchCfg(variant=2))
    g1 = _gpu(gpu_ctx, dv, rays, True, -1.0, cq.LaunchCfg(variant=1))
    g3 = _gpu(gpu_ctx, dv, rays, True, -1.0, cq.LaunchCfg(variant=3))
    g4 = _gpu(gpu_ctx, dv, rays, True, -1.0, cq.LaunchCfg(variant=4, lds_nodes=1024))
    g5 = _gpu(gpu_ctx, dv, rays, True, -1.0, cq.LaunchCfg(variant=5))
    g2b = _gpu(gpu_ctx, dv, rays, True, -1.0, cq.LaunchCfg(variant=2, bin_bits=3))
    g3b = _gpu(gpu_ctx, dv, rays, True, -1.0, cq.LaunchCfg(variant=3, bin_bits=4))
    g2s = _gpu(gpu_ctx, dv, rays, True, -1.0, cq.LaunchCfg(variant=2, stack_rows=4, bin_bits=3))
    g3s = _gpu(gpu_ctx, dv, rays, True, -1.0, cq.LaunchCfg(variant=3, stack_rows=8))
    print(f"gpu {time.time() - t0:.2f}s", flush=True)
    assert records_equal(g2s, g2).all()
    assert records_equal(g3s, g2).all()
    assert records_equal(g2b, g2).all()
    assert records_equal(g3b, g2).all()
    assert records_equal(g1, g2).all()
    assert records_equal(g3, g2).all()
    assert records_equal(g4, g2).all()
    assert records_equal(g5, g2).all()
    t0 = time.time()
    o = oracle.intersect(hv.nodes, oracle.find_subdags(hv.nodes, hv.root), rays, True, threads=16)
    print(f"oracle {time.time() - t0:.2f}s", flush=True)
    eq = records_equal(g2, o)
    assert eq.all(), _report("1024 incoherent", eq, rays, g2, o)
    assert ((g2["info"] >> 15) & 1).sum() == 0
    hit = np.nonzero(g2["info"] & 1)[0]
    assert 0.3 < len(hit) / len(rays) < 0.95
    rng = np.random.Generator(np.random.PCG64(5))
    for i in rng.choice(hit, 2000, replace=False):
        x, y, z = (int(v) for v in g2["voxel"][i])
        assert hv.voxel(x, y, z) == (int(g2["info"][i]) >> 16) & 255
    dv.free()


def test_gpu_autotuned_launch_is_exact(cq, oracle, gpu_ctx, vol128, rays128):
    """The autotuner's choice is cached and used by auto launches; results stay bit-exact."""
    import torch

    ctx = cq.Context(0)
    dv = ctx.upload(vol128)
    rays = rays128["incoherent"]
    r = torch.from_numpy(rays).cuda()
    best, ms = cq.autotune(ctx, dv, r, True, -1.0)
    assert best.variant in (1, 2, 3, 5) and best.bin_bits in (0, 2, 3, 4) and ms > 0
    assert best.stack_rows in (0, 4, 8) and (best.stack_rows == 0 or best.variant in (2, 3))
    auto = cq.intersectVolume(ctx, dv, r, True).cpu().numpy().view(cq.HIT_DTYPE).reshape(-1)
    o = oracle.intersect(vol128.nodes, oracle.find_subdags(vol128.nodes, vol128.root), rays, True)
    assert records_equal(auto, o).all()
    dv.free()
    ctx.close()


def test_gpu_autotuned_launch_on_deep_volumes(cq, oracle):
    """On SubDAGs 21-22 levels deep the autotuner also weighs short stacks (4, 8, 16 rows): its
    choice, whatever it is, gives the oracle's records through the auto launch."""
    import torch

    from tests.raysets import deep_volumes

    ctx = cq.Context(0)
    for name, (hv, rays) in deep_volumes(cq).items():
        dv = ctx.upload(hv)
        many = np.ascontiguousarray(np.tile(rays, (max(1, (1 << 18) // len(rays)), 1)), dtype=np.float32)
        r = torch.from_numpy(many).cuda()
        best, ms = cq.autotune(ctx, dv, r, True, -1.0)
        assert ms > 0 and best.stack_rows in (0, 4, 8, 16) and (best.stack_rows == 0 or best.variant in (2, 3))
        g = cq.intersectVolume(ctx, dv, torch.from_numpy(rays).cuda(), True).cpu().numpy().view(cq.HIT_DTYPE)
        o = oracle.intersect(hv.nodes, oracle.find_subdags(hv.nodes, hv.root), rays, True, -1.0, max_iters=1 << 22)
        eq = records_equal(g.reshape(-1), o)
        assert eq.all(), _report(f"{name} tuned {best.variant}/{best.stack_rows}", eq, rays, g.reshape(-1), o)
        dv.free()
    ctx.close()


def test_gpu_deterministic(cq, gpu_ctx, dvol128, rays128):
    a = _gpu(gpu_ctx, dvol128, rays128["incoherent"])
    b = _gpu(gpu_ctx, dvol128, rays128["incoherent"])
    assert records_equal(a, b).all()


def test_gpu_iteration_cap_reports_anomaly(cq, oracle, gpu_ctx, dvol128, vol128, rays128):
    """A tiny per-ray iteration cap forced from the host: every capped ray is flagged, none hangs."""
    rays = rays128["camera"][30720:34816]  # image centre: every ray enters the volume
    sub = oracle.find_subdags(vol128.nodes, vol128.root)
    o = oracle.intersect(vol128.nodes, sub, rays, True, -1.0, max_iters=5)
    g = _gpu(gpu_ctx, dvol128, rays, True, -1.0, cq.LaunchCfg(max_iters=5))
    assert ((g["info"] >> 15) & 1).sum() > 0
    assert records_equal(g, o).all()


def test_gpu_empty_and_single(cq, gpu_ctx, dvol128, rays128):
    import torch

    from cubiquity_amd import intersectVolume

    empty = torch.zeros((0, 6), dtype=torch.float32, device="cuda")
    assert intersectVolume(gpu_ctx, dvol128, empty).numel() == 0
    one = _gpu(gpu_ctx, dvol128, rays128["camera"][32896:32897])
    assert one.shape == (1,)


def test_gpu_host_entry_point(cq, oracle, gpu_ctx, dvol128, vol128, rays128):
    rays = rays128["special"]
    g = gpu_ctx.intersect_host(dvol128, rays)
    o = oracle.intersect(vol128.nodes, oracle.find_subdags(vol128.nodes, vol128.root), rays)
    assert records_equal(g, o).all()


def test_gpu_store_beyond_2_27_nodes(cq, oracle, gpu_ctx, vol128, rays128):
    """Stores of more than 2^27 nodes (the 16384^3 workload has 910M) read children with 64-bit
    indexing instead of 32-bit byte offsets. The 128^3 store padded with 2^27 unreachable nodes
    (4.3 GB) takes that path; every kernel variant, binned and not, and traceRayRef must give the
    oracle's records on the unpadded store."""
    import torch

    base = vol128.nodes
    nodes = np.zeros((base.shape[0] + (1 << 27), 8), np.uint32)  # padding nodes hold material 0
    nodes[: base.shape[0]] = base
    dv = gpu_ctx.upload((nodes, vol128.root))
    del nodes
    assert dv.node_count > (1 << 27)
    sub = oracle.find_subdags(base, vol128.root)
    for set_name in ("camera", "incoherent", "special"):
        rays = rays128[set_name]
        o = oracle.intersect(base, sub, rays, True, -1.0, threads=8)
        for cfg in (cq.LaunchCfg(variant=1), cq.LaunchCfg(variant=2), cq.LaunchCfg(variant=3),
                    cq.LaunchCfg(variant=4, lds_nodes=512), cq.LaunchCfg(variant=5),
                    cq.LaunchCfg(variant=3, bin_bits=3)):
            g = _gpu(gpu_ctx, dv, rays, True, -1.0, cfg)
            eq = records_equal(g, o)
            assert eq.all(), _report(f"{set_name} {cfg.variant}/{cfg.bin_bits}", eq, rays, g, o)
    rays = np.concatenate([rays128["camera"][::16], rays128["special"]])
    got = cq.traceRayRef(gpu_ctx, dv, torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32)).cuda())
    torch.cuda.synchronize()
    assert got.cpu().numpy().tobytes() == oracle.trace_ref(base, vol128.root, rays).tobytes()
    dv.free()


def test_gpu_trace_ref_matches_oracle(cq, oracle, gpu_ctx, dvol128, vol128, rays128):
    """cqx_trace_ref (the reference's brute-force raycast) against the oracle's restatement, byte
    for byte, on the 128^3 ray sets and on volumes with 21-22-level SubDAGs."""
    import torch

    from tests.raysets import deep_volumes

    cases = [(vol128, dvol128, np.concatenate([rays128["camera"][::16], rays128["incoherent"], rays128["special"]]))]
    deep = []
    for hv, rays in deep_volumes(cq).values():
        dv = gpu_ctx.upload(hv)
        deep.append(dv)
        cases.append((hv, dv, rays))
    for hv, dv, rays in cases:
        r = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32)).cuda()
        got = cq.traceRayRef(gpu_ctx, dv, r)
        torch.cuda.synchronize()
        ref = oracle.trace_ref(hv.nodes, hv.root, rays)
        assert got.cpu().numpy().tobytes() == ref.tobytes()
        hits = ref["hit"].sum()
        assert 0 < hits < len(rays)
    for dv in deep:
        dv.free()
