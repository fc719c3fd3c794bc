"""Batched traversal: many volumes of uneven depth and ray batches of uneven size (coherent and
incoherent) in one persistent launch. Every batch must equal, bit for bit, its own single-volume
launch and the oracle."""
import numpy as np
import pytest

from tests.conftest import records_equal

pytestmark = pytest.mark.gpu


def _batches(cq, count, seed):
    from cubiquity_amd import workloads as W

    rng = np.random.Generator(np.random.PCG64(seed))
    vols = W.batched_volumes()[:count]
    out = []
    for i, w in enumerate(vols):
        if w.log2_size > 9:  # keep the test volumes small; the bench runs the full set
            w = type(w)(w.name, 6 + i % 4, w.seed, w.overrides, W._camera(6 + i % 4, 64, 64), w.incoherent_seed)
        hv = cq.HostVolume.generate(w.params())
        n = int(round(10 ** rng.uniform(2.0, 4.5)))
        if i % 2 == 0:
            side = int(np.ceil(np.sqrt(n)))
            rays = cq.camera_rays(w.camera.eye, w.camera.target, w.camera.fov_y, side, side)[:n]
        else:
            rays = cq.incoherent_rays(hv, w.log2_size, w.incoherent_seed, n)
        surf = i % 3 != 2
        fp = 0.05 if i % 5 == 4 else -1.0
        out.append((hv, rays, surf, fp))
    return out


def test_batched_matches_single_and_oracle(cq, oracle, gpu_ctx):
    import torch

    items = _batches(cq, 14, 3)
    dvols = [gpu_ctx.upload(hv) for hv, _, _, _ in items]
    launch = []
    for dv, (hv, rays, surf, fp) in zip(dvols, items):
        r = torch.from_numpy(rays).cuda()
        h = torch.zeros((r.shape[0], 32), dtype=torch.uint8, device="cuda")
        launch.append((dv, r, h, surf, fp))
    # an empty batch in the middle is allowed
    empty = torch.zeros((0, 6), dtype=torch.float32, device="cuda")
    launch.insert(5, (dvols[0], empty, torch.zeros((0, 32), dtype=torch.uint8, device="cuda"), True, -1.0))
    cq.intersect_batched(gpu_ctx, launch)
    torch.cuda.synchronize()
    launch.pop(5)
    for k, ((hv, rays, surf, fp), (dv, r, h, _, _)) in enumerate(zip(items, launch)):
        got = h.cpu().numpy().view(cq.HIT_DTYPE).reshape(-1)
        single = cq.intersectVolume(gpu_ctx, dv, r, surf, fp).cpu().numpy().view(cq.HIT_DTYPE).reshape(-1)
        ref = oracle.intersect(hv.nodes, oracle.find_subdags(hv.nodes, hv.root), rays, surf, fp, threads=8)
        assert records_equal(got, single).all(), f"batch {k}"
        assert records_equal(got, ref).all(), f"batch {k}"
    for dv in dvols:
        dv.free()


def test_batched_launch_shapes(cq, oracle, gpu_ctx):
    """Different persistent launch shapes give identical records."""
    import torch

    items = _batches(cq, 6, 9)
    dvols = [gpu_ctx.upload(hv) for hv, _, _, _ in items]
    results = []
    for cfg in (cq.LaunchCfg(), cq.LaunchCfg(waves_per_cu=4, regroup_interval=1, refill_lanes=1),
                cq.LaunchCfg(waves_per_cu=20, regroup_interval=64, refill_lanes=64, block=64)):
        launch = []
        for dv, (hv, rays, surf, fp) in zip(dvols, items):
            r = torch.from_numpy(rays).cuda()
            launch.append((dv, r, torch.zeros((r.shape[0], 32), dtype=torch.uint8, device="cuda"), surf, fp))
        cq.intersect_batched(gpu_ctx, launch, cfg=cfg)
        torch.cuda.synchronize()
        results.append([h.cpu().numpy() for _, _, h, _, _ in launch])
    for a, b in zip(results[0], results[1]):
        assert np.array_equal(a, b)
    for a, b in zip(results[0], results[2]):
        assert np.array_equal(a, b)
    for dv in dvols:
        dv.free()
