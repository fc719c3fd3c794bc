"""The C++ drop-in API (include/cubiquity_hip.hpp) used from a compiled program, as a reference
caller would use it, checked bit-exactly against the oracle's RayVolumeIntersection records."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import records_equal

EXE = os.path.join(os.path.dirname(__file__), "cpp", "dropin_check")


@pytest.mark.gpu
@pytest.mark.parametrize("surf,fp", [(1, -1.0), (0, -1.0), (1, 0.05)])
def test_cpp_dropin_matches_oracle(oracle, vol128, dag128, rays128, tmp_path, surf, fp):
    assert os.path.exists(EXE), "build() compiles tests/cpp/dropin_check"
    rays = np.concatenate([rays128["camera"][::7], rays128["incoherent"], rays128["special"]]).astype(np.float32)
    rp = tmp_path / "rays.f32"
    op = tmp_path / "out.bin"
    rays.tofile(rp)
    r = subprocess.run([EXE, dag128, str(rp), str(op), str(surf), str(fp)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    from oracle.bindings import RESULT_DTYPE

    got = np.fromfile(op, dtype=RESULT_DTYPE)
    sub = oracle.find_subdags(vol128.nodes, vol128.root)
    ref = oracle.to_results(rays, oracle.intersect(vol128.nodes, sub, rays, bool(surf), fp), bool(surf))
    assert got.shape == ref.shape
    assert records_equal(got, ref).all()


@pytest.mark.gpu
def test_cpp_edit_session(cq, oracle, reference, dag128, rays128, tmp_path):
    """The C++ API's edit session (fillSphere, checkpointed fillSpheres batches uploaded as deltas,
    undo as a root-only update): the GPU results equal the oracle's on the saved store, and that
    store's voxels equal the compiled reference's fillBrush after the same (undone) edits."""
    assert os.path.exists(EXE), "build() compiles tests/cpp/dropin_check"
    from oracle.bindings import RESULT_DTYPE

    rng = np.random.Generator(np.random.PCG64(77))
    n = 41
    edits = np.zeros((n, 5), np.float32)
    edits[:, :3] = rng.uniform(-4.0, 132.0, (n, 3))
    edits[:, 3] = rng.uniform(1.0, 14.0, n)
    edits[:, 4] = rng.integers(0, 10, n)
    rays = np.concatenate([rays128["camera"][::13], rays128["incoherent"][:4000]]).astype(np.float32)
    paths = [str(tmp_path / f) for f in ("e.f32", "r.f32", "hits.bin", "ref.bin", "out.dag")]
    edits.tofile(paths[0])
    rays.tofile(paths[1])
    r = subprocess.run([EXE, "edit", dag128, *paths], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    hv = cq.HostVolume.load_dag(paths[4])
    sub = oracle.find_subdags(hv.nodes, hv.root)
    got = np.fromfile(paths[2], dtype=RESULT_DTYPE)
    ref = oracle.to_results(rays, oracle.intersect(hv.nodes, sub, rays, True), True)
    assert records_equal(got, ref).all()
    assert np.fromfile(paths[3], dtype=RESULT_DTYPE).tobytes() == oracle.trace_ref(hv.nodes, hv.root, rays).tobytes()
    rv = reference.load(dag128)
    half = 1 + (n - 1) // 2
    for e in edits[:half]:  # the undone second batch is not applied
        rv.fill_sphere(e[:3], e[3], int(e[4]))
    lo, size = (-8, -8, -8), (144, 144, 144)
    assert np.array_equal(oracle.dense(hv.nodes, hv.root, lo, size), rv.dense(lo, size))
