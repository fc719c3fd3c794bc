import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle.bindings import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.bindings import Reference

    if not Reference.available():
        pytest.skip("compiled reference (oracle/_ref) not present")
    return Reference()


@pytest.fixture(scope="session")
def cq():
    import cubiquity_amd

    cubiquity_amd.cabi.lib()  # fails loudly when libcqx.so has not been built
    return cubiquity_amd


@pytest.fixture(scope="session")
def vol128(cq):
    from cubiquity_amd import workloads as W

    return cq.HostVolume.generate(W.RAYCAST_128.params())


@pytest.fixture(scope="session")
def dag128(vol128, tmp_path_factory):
    path = tmp_path_factory.mktemp("dag") / "raycast_128.dag"
    vol128.save_dag(str(path))
    return str(path)


@pytest.fixture(scope="session")
def rays128(cq, vol128):
    """Camera (65,536), incoherent (16,384) and special-case rays of the 128^3 workload."""
    from cubiquity_amd import workloads as W
    from tests.raysets import special_rays

    c = W.RAYCAST_128.camera
    cam = cq.camera_rays(c.eye, c.target, c.fov_y, c.width, c.height)
    inc = cq.incoherent_rays(vol128, 7, W.RAYCAST_128.incoherent_seed, W.RAYCAST_128.incoherent_rays)
    spec = special_rays(128)
    return {"camera": cam, "incoherent": inc, "special": spec}


@pytest.fixture(scope="session")
def gpu_ctx(cq):
    ctx = cq.Context(0)
    yield ctx
    ctx.close()


def records_equal(a, b):
    """Per-record bytewise equality of two structured arrays of the same dtype."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return (a.view(np.uint8).reshape(a.shape[0], -1) == b.view(np.uint8).reshape(b.shape[0], -1)).all(axis=1)
