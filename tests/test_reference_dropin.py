"""The batched intersectVolume overload in Cubiquity's own types (include/cubiquity_hip_reference.hpp),
as a Cubiquity maintainer would add it: compiled against the reference library's headers with its
layout checks (static_assert on RayVolumeIntersection and SubDAG sizes and offsets), linked with the
compiled, unmodified reference library (oracle/_ref) into tests/cpp/reference_dropin.cpp's program.

CPU (here, where the reference checkout is): the header compiles against the reference headers.
GPU: the program (built here into oracle/_ref, which travels with the tree) loads a .dag with the
reference's Volume, and every ray's RayVolumeIntersection from the GPU overload equals the
reference's own per-ray intersectVolume, field by field and bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle.bindings import REF_SRC

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "reference_dropin")


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="needs the reference checkout (headers)")
def test_header_compiles_against_reference_headers(tmp_path):
    src = tmp_path / "tu.cpp"
    src.write_text('#include "cubiquity_hip_reference.hpp"\n'
                   "static_assert(sizeof(Cubiquity::RayVolumeIntersection) == 48);\n"
                   "static_assert(offsetof(Cubiquity::RayVolumeIntersection, distance) == 8);\n"
                   "static_assert(offsetof(Cubiquity::RayVolumeIntersection, material) == 16);\n"
                   "static_assert(offsetof(Cubiquity::RayVolumeIntersection, position) == 20);\n"
                   "static_assert(offsetof(Cubiquity::RayVolumeIntersection, normal) == 32);\n")
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", f"-I{REF_SRC}", f"-I{os.path.join(ROOT, 'include')}",
                        str(src)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # and the test program links (built by make -C oracle dropin, which build() runs)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "dropin"], check=True, timeout=300)
    assert os.path.exists(EXE)


@pytest.mark.gpu
@pytest.mark.parametrize("surf,fp", [(1, -1.0), (0, -1.0), (1, 0.05)])
def test_gpu_reference_typed_overload_matches_reference(dag128, rays128, tmp_path, surf, fp):
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/reference_dropin is built only where the reference checkout is")
    from tests import pt_pins

    lattice = np.load(pt_pins.LATTICE, allow_pickle=False)["rays"]
    rays = np.concatenate([rays128["camera"][::5], rays128["incoherent"], rays128["special"], lattice])
    rp = tmp_path / "rays.f32"
    np.ascontiguousarray(rays, np.float32).tofile(rp)
    r = subprocess.run([EXE, dag128, str(rp), str(surf), str(fp)], capture_output=True, text=True, timeout=300)
    print(r.stdout.strip(), flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"{len(rays)} rays" in r.stdout and ", 0 differ" in r.stdout
