"""The device traversal core (cubiquity_amd/csrc/traverse_core.h) compiled for the host and driven
the way the kernels drive it (tests/cpp/core_emulation.cpp): per-ray trace_ray, and 64-lane waves
with the persistent kernel's service-pass policy (deferred SubDAG changes and records, refills
between traversal steps) at several refill / regroup settings, also with the top nodes read from
a copy (the LDS node cache). Every 32-byte record must be
bit-identical to the oracle's, so a logic change to the core is caught on the CPU."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import records_equal

EXE = os.path.join(os.path.dirname(__file__), "cpp", "core_emulation")


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EXE):
        from cubiquity_amd import build

        build.build_cpp_checks()
    assert os.path.exists(EXE)
    return EXE


def _run(emu, tmp_path, cq, hv, rays, mode, surf, fp, max_iters, refill=32, regroup=8):
    sub = cq.findSubDAGs(hv.nodes, hv.root)
    paths = [str(tmp_path / f) for f in ("n.bin", "s.bin", "r.bin", "o.bin")]
    np.ascontiguousarray(hv.nodes).tofile(paths[0])
    np.ascontiguousarray(sub).tofile(paths[1])
    np.ascontiguousarray(rays, dtype=np.float32).tofile(paths[2])
    subprocess.run([emu, mode, *paths, str(int(surf)), repr(float(fp)), str(max_iters), str(refill), str(regroup)],
                   check=True, timeout=300)
    return np.fromfile(paths[3], dtype=cq.HIT_DTYPE)


@pytest.mark.parametrize("mode,refill,regroup", [("lane", 0, 0), ("persistent", 32, 8), ("persistent", 8, 1),
                                                 ("persistent", 64, 3), ("top", 16, 4), ("pair", 16, 8),
                                                 ("short2", 16, 8), ("short4", 32, 4), ("short8", 8, 1)])
def test_core_matches_oracle_128(emu, tmp_path, cq, oracle, vol128, rays128, mode, refill, regroup):
    sub = oracle.find_subdags(vol128.nodes, vol128.root)
    rays = np.concatenate([rays128["camera"][::5], rays128["incoherent"], rays128["special"]])
    for surf, fp in ((True, -1.0), (False, -1.0), (True, 0.05), (True, 0.0035)):
        got = _run(emu, tmp_path, cq, vol128, rays, mode, surf, fp, 1 << 16, refill, regroup)
        ref = oracle.intersect(vol128.nodes, sub, rays, surf, fp, max_iters=1 << 16, threads=8)
        eq = records_equal(got, ref)
        assert eq.all(), f"{mode} surf={surf} fp={fp}: {(~eq).sum()} records differ, first {np.nonzero(~eq)[0][:4]}"


def test_core_iteration_cap(emu, tmp_path, cq, oracle, vol128, rays128):
    rays = rays128["camera"][30720:34816]
    got = _run(emu, tmp_path, cq, vol128, rays, "persistent", True, -1.0, 5)
    ref = oracle.intersect(vol128.nodes, oracle.find_subdags(vol128.nodes, vol128.root), rays, True, -1.0, max_iters=5)
    assert ((got["info"] >> 15) & 1).sum() > 0
    assert records_equal(got, ref).all()


def test_core_small_and_degenerate_volumes(emu, tmp_path, cq, oracle):
    from tests.raysets import special_rays
    from tests.test_gpu_parity import _small_volumes
    from tests.test_oracle_pinning import _brute_rays

    for name, (hv, n, lower) in _small_volumes(cq).items():
        sp = special_rays(n)
        sp[:, :3] += np.asarray(lower, np.float32)
        rays = np.concatenate([_brute_rays(n, lower, 3000, 17), sp]).astype(np.float32)
        sub = oracle.find_subdags(hv.nodes, hv.root)
        for mode in ("lane", "persistent", "top", "pair", "short2", "short4"):
            got = _run(emu, tmp_path, cq, hv, rays, mode, True, -1.0, 1 << 16, 16, 4)
            ref = oracle.intersect(hv.nodes, sub, rays, True, -1.0, max_iters=1 << 16)
            assert records_equal(got, ref).all(), f"{name} {mode}"


@pytest.mark.parametrize("mode", ["lane", "persistent", "pair", "short4", "short8", "short16"])
def test_core_deep_sparse_volumes(emu, tmp_path, cq, oracle, mode):
    """SubDAGs 21-22 levels deep: 23 stack rows per lane, or short stacks of 4-16 rows whose pops
    past them restart from the SubDAG root."""
    from tests.raysets import deep_volumes

    for name, (hv, rays) in deep_volumes(cq).items():
        sub = oracle.find_subdags(hv.nodes, hv.root)
        for surf, fp in ((True, -1.0), (True, 0.01)):
            got = _run(emu, tmp_path, cq, hv, rays, mode, surf, fp, 1 << 16, 16, 8)
            ref = oracle.intersect(hv.nodes, sub, rays, surf, fp, max_iters=1 << 16)
            assert records_equal(got, ref).all(), f"{name} {mode} fp={fp}"


def test_trace_ref_core_matches_oracle(emu, tmp_path, cq, oracle, vol128, rays128):
    """traceRayRef's walk (trace_ref_core.h) on the host against the oracle's restatement, which
    the compiled reference pins: hit, material and FP64 distance bit for bit."""
    from tests.raysets import deep_volumes

    cases = [(vol128, np.concatenate([rays128["camera"][::64], rays128["incoherent"][:2048], rays128["special"]]))]
    cases += [(hv, rays[::4]) for hv, rays in deep_volumes(cq).values()]
    for hv, rays in cases:
        sub = cq.findSubDAGs(hv.nodes, hv.root)
        paths = [str(tmp_path / f) for f in ("n.bin", "s.bin", "r.bin", "o.bin")]
        np.ascontiguousarray(hv.nodes).tofile(paths[0])
        np.ascontiguousarray(sub).tofile(paths[1])
        np.ascontiguousarray(rays, dtype=np.float32).tofile(paths[2])
        subprocess.run([emu, "traceref", *paths, "1", "-1.0", str(hv.root), "0", "0"], check=True, timeout=300)
        got = np.fromfile(paths[3], dtype=cq.INTERSECTION_DTYPE)
        ref = oracle.trace_ref(hv.nodes, hv.root, rays)
        assert got.tobytes() == ref.tobytes()


def test_core_census_of_the_walk_loop(emu, tmp_path, cq, oracle, vol128, rays128):
    """The census mode (the persistent driver with counters): its records equal the oracle's, every
    walk step is one iteration of exactly one path, descents + pops + SubDAG entries are the node
    entries, and the incoherent set keeps every wave step divergent (descend, advance and pop all
    taken in almost every step) -- the fact the persistent kernel's VALU lane utilisation follows."""
    sub = cq.findSubDAGs(vol128.nodes, vol128.root)
    rays = rays128["incoherent"]
    paths = [str(tmp_path / f) for f in ("n.bin", "s.bin", "r.bin", "o.bin")]
    np.ascontiguousarray(vol128.nodes).tofile(paths[0])
    np.ascontiguousarray(sub).tofile(paths[1])
    np.ascontiguousarray(rays, dtype=np.float32).tofile(paths[2])
    r = subprocess.run([emu, "census", *paths, "1", "-1.0", str(1 << 16), "16", "16"], check=True, timeout=300,
                       capture_output=True, text=True)
    words = r.stdout.split()
    assert words[0] == "census"
    c = {k: int(v) for k, v in zip(words[1::2], words[2::2])}
    got = np.fromfile(paths[3], dtype=cq.HIT_DTYPE)
    ref = oracle.intersect(vol128.nodes, oracle.find_subdags(vol128.nodes, vol128.root), rays, True, -1.0,
                           max_iters=1 << 16, threads=8)
    assert records_equal(got, ref).all()
    lanes = sum(c[f"lane_{k}"] for k in "DAPLH")
    assert lanes == c["slots_walk"] == int(ref["iters"].astype(np.int64).sum())
    subdag_entries = int(ref["entries"].astype(np.int64).sum()) - c["lane_D"] - c["lane_P"]
    assert 0 < subdag_entries <= 8 * len(rays)
    assert c["rays"] == len(rays) and c["lane_H"] <= len(rays)
    for k in "DAP":
        assert c[f"wave_{k}"] > 0.95 * c["wave_steps"], c
