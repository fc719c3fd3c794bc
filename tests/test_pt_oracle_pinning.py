"""Pins the path-tracer oracle (oracle/cpu_pt.c) to published known answers, mathematics and
closed forms (tests/pt_pins.py states each check), on the CPU. The GPU renderer is bit-identical
to this oracle (tests/test_pathtrace.py), and the closed forms are also checked on the GPU image
directly (test_gpu_closed_forms below).
"""
import numpy as np
import pytest

from cubiquity_amd import workloads as W
from tests import pt_pins


def test_pcg32_published_stream(oracle):
    pt_pins.check_pcg32(oracle)


def test_sincos2pi_every_uniform_draw(oracle):
    pt_pins.check_sincos(oracle)


def test_bounce_cosine_law(oracle):
    pt_pins.check_bounce(oracle)


def test_camera_ray_pinhole(oracle, cq):
    pt_pins.check_camera(oracle, cq, W)


def test_furnace_and_floor_closed_forms(oracle, cq):
    pt_pins.check_closed_forms(oracle, cq, W)


def test_furnace_paths_never_escape(oracle, cq):
    """With the sky at 1000 any escaping path would dominate the pixel: the furnace image sits on
    e (1 + a + ... + a^B) at every pixel, with every sample (spp 5 at 7 bounces)."""
    hv, s, want = pt_pins.furnace_scene(cq, W, 7, spp=5)
    img = oracle.pathtrace(hv.nodes, hv.root, s.params).astype(np.float64)
    assert np.abs(img - want).max() <= 2e-6 * want.max()


@pytest.mark.gpu
def test_gpu_closed_forms(cq, oracle, gpu_ctx):
    """Both GPU renderers against the furnace and floor closed forms (not through the oracle)."""
    import torch

    for mode in (0, 1):
        def render(hv, s):
            dv = gpu_ctx.upload(hv)
            img = cq.pathtrace(gpu_ctx, dv, s, mode).cpu().numpy()
            torch.cuda.synchronize()
            dv.free()
            return img

        pt_pins.check_closed_forms(oracle, cq, W, render=render)
