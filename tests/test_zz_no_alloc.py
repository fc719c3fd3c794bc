"""cqx_intersect_noalloc (the entry point that allocates nothing: work counters and ray-binning
buffers in the library's static per-device scratch) against cqx_intersect, through the C ABI:
bit-identical records on two calls in a row, the second with fewer rays (the scratch keeps the
first call's contents), unbinned and binned, persistent and one-ray-per-lane; sizes past the
binning limit are refused before anything is launched."""
import ctypes

import numpy as np
import pytest

from tests.conftest import records_equal

pytestmark = pytest.mark.gpu


def _noalloc(cq, ctx, dvol, rays, surf, fp, cfg, out):
    """cqx_intersect_noalloc through the C ABI; returns its status."""
    import torch

    stream = torch.cuda.current_stream().cuda_stream
    flags = cq.cabi.CQX_SURFACE_PROPERTIES if surf else 0
    cfgp = ctypes.byref(cfg) if cfg is not None else None
    return cq.cabi.lib().cqx_intersect_noalloc(ctx._h, dvol._h, ctypes.c_void_p(rays.data_ptr()),
                                               ctypes.c_uint64(rays.shape[0]), flags, ctypes.c_float(fp),
                                               ctypes.c_void_p(out.data_ptr()), cfgp, ctypes.c_void_p(stream))


def _last_error(cq):
    return cq.cabi.lib().cqx_last_error().decode()


@pytest.mark.parametrize("cfg_kw", [dict(variant=2), dict(variant=2, bin_bits=3), dict(variant=3, bin_bits=4),
                                    dict(variant=5, bin_bits=2), dict(variant=1), dict(variant=2, bin_bits=3, stack_rows=4),
                                    None])
def test_noalloc_matches_intersect_twice(cq, gpu_ctx, vol128, cfg_kw):
    import torch

    dv = gpu_ctx.upload(vol128)
    cfg = cq.LaunchCfg(**cfg_kw) if cfg_kw is not None else None
    big = cq.incoherent_rays(vol128, 7, 31, 200000)
    for n in (200000, 37123):  # the second call is smaller: stale counters, cells and order entries
        r = torch.from_numpy(np.ascontiguousarray(big[:n])).cuda()
        want = cq.intersectVolume(gpu_ctx, dv, r, True, -1.0, cfg=cfg)
        got = torch.full((n, 32), 0xAB, dtype=torch.uint8, device="cuda")
        st = _noalloc(cq, gpu_ctx, dv, r, True, -1.0, cfg, got)
        torch.cuda.synchronize()
        assert st == 0, _last_error(cq)
        g = got.cpu().numpy().view(cq.HIT_DTYPE).reshape(-1)
        w = want.cpu().numpy().view(cq.HIT_DTYPE).reshape(-1)
        eq = records_equal(g, w)
        assert eq.all(), f"{cfg_kw} n={n}: {(~eq).sum()} records differ"
    dv.free()


def test_noalloc_refuses_binning_past_its_scratch(cq, gpu_ctx, vol128):
    import torch

    dv = gpu_ctx.upload(vol128)
    n = (1 << 22) + 1
    r = torch.zeros((n, 6), dtype=torch.float32, device="cuda")
    out = torch.zeros((n, 32), dtype=torch.uint8, device="cuda")
    st = _noalloc(cq, gpu_ctx, dv, r, True, -1.0, cq.LaunchCfg(variant=2, bin_bits=2), out)
    torch.cuda.synchronize()
    assert st == -7, st  # CQX_ERR_LIMIT
    assert "bins at most" in _last_error(cq)
    assert int(out.count_nonzero().item()) == 0  # nothing was launched
    dv.free()


def test_noalloc_at_its_binning_limit(cq, gpu_ctx, vol128):
    """Exactly CQX_NOALLOC_MAX_BINNED_RAYS rays binned in the static scratch (its cell histograms,
    order and cells all at their largest) give the records of cqx_intersect."""
    import torch

    dv = gpu_ctx.upload(vol128)
    n = 1 << 22
    r = torch.from_numpy(np.ascontiguousarray(cq.incoherent_rays(vol128, 7, 37, n))).cuda()
    cfg = cq.LaunchCfg(variant=2, bin_bits=4)
    want = cq.intersectVolume(gpu_ctx, dv, r, True, -1.0, cfg=cfg)
    got = torch.full((n, 32), 0xAB, dtype=torch.uint8, device="cuda")
    st = _noalloc(cq, gpu_ctx, dv, r, True, -1.0, cfg, got)
    torch.cuda.synchronize()
    assert st == 0, _last_error(cq)
    assert torch.equal(got, want)
    dv.free()
