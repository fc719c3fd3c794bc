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
