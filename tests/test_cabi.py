"""The C ABI (include/*.h) on the CPU: every declared symbol is exported, host entry points
validate their inputs and report errors through status codes, and nothing throws. No GPU compute
is called here."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(cqx_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported(cq):
    L = cq.cabi.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header(cq):
    assert set(declared_functions()) <= set(cq.cabi.SIGNATURES)


def _status(fn, *args):
    return fn(*args)


def test_validation_errors(cq):
    L = cq.cabi.lib()
    base = np.tile(np.arange(256, dtype=np.uint32)[:, None], (1, 8))
    good = np.vstack([base, [[0, 0, 0, 0, 0, 0, 0, 1]]]).astype(np.uint32)
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    assert L.cqx_validate_dag(p(good), len(good), 256) == 0
    bad = good.copy()
    bad[256, 3] = 999
    assert L.cqx_validate_dag(p(bad), len(bad), 256) == -2
    assert b"out of range" in L.cqx_last_error()
    cyc = np.vstack([base, [[0, 0, 0, 0, 0, 0, 0, 257], [256, 0, 0, 0, 0, 0, 0, 1]]]).astype(np.uint32)
    assert L.cqx_validate_dag(p(cyc), len(cyc), 256) == -2
    assert b"cycle" in L.cqx_last_error()
    # a chain of 33 internal nodes is one level deeper than the tree
    chain = [base]
    for k in range(33):
        row = np.zeros((1, 8), np.uint32)
        row[0, 0] = 256 + k + 1 if k < 32 else 1
        chain.append(row)
    deep = np.vstack(chain).astype(np.uint32)
    assert L.cqx_validate_dag(p(deep), len(deep), 256) == -2
    ok = np.vstack(chain[:1] + chain[2:]).astype(np.uint32)  # 32 levels is fine
    ok[256:, 0] = np.where(ok[256:, 0] > 1, ok[256:, 0] - 1, ok[256:, 0])
    assert L.cqx_validate_dag(p(ok), len(ok), 256) == 0
    mat = good.copy()
    mat[5, 2] = 6
    assert L.cqx_validate_dag(p(mat), len(mat), 256) == -2
    assert L.cqx_validate_dag(p(good), len(good), 300) == -2  # root out of range
    assert L.cqx_validate_dag(None, 0, 0) == -1


def test_find_subdags_and_upload_reject_bad_store(cq):
    base = np.tile(np.arange(256, dtype=np.uint32)[:, None], (1, 8))
    bad = np.vstack([base, [[0, 0, 0, 0, 0, 0, 0, 4000]]]).astype(np.uint32)
    with pytest.raises(cq.CqxError) as e:
        cq.findSubDAGs(bad, 256)
    assert e.value.status == -2
    with pytest.raises(cq.CqxError):
        cq.HostVolume.from_nodes(bad, 256)


def test_load_dag_errors(cq, tmp_path):
    with pytest.raises(cq.CqxError) as e:
        cq.HostVolume.load_dag(str(tmp_path / "missing.dag"))
    assert e.value.status == -5
    p = tmp_path / "trunc.dag"
    p.write_bytes(np.asarray([300, 10], np.uint32).tobytes() + b"\0" * 40)
    with pytest.raises(cq.CqxError) as e:
        cq.HostVolume.load_dag(str(p))
    assert e.value.status == -5
    p2 = tmp_path / "badchild.dag"
    p2.write_bytes(np.asarray([256, 1, 0, 0, 0, 0, 0, 0, 0, 77777], np.uint32).tobytes())
    with pytest.raises(cq.CqxError) as e:
        cq.HostVolume.load_dag(str(p2))
    assert e.value.status == -2


def test_generator_argument_errors(cq):
    p = cq.gen_default_params(7, 1)
    p.n_materials = 1
    with pytest.raises(cq.CqxError):
        cq.HostVolume.generate(p)
    with pytest.raises(cq.CqxError):
        cq.gen_default_params(40, 1)


def test_hits_to_intersections_decodes(cq):
    rays = np.asarray([[1.0, 2.0, 3.0, 0.5, -0.25, 0.0]], np.float32)
    h = np.zeros(1, dtype=cq.HIT_DTYPE)
    h["t"] = 3.0
    # hit, normal: x = +1 (code 1), y = -0 (code 2), z = -1 (code 3); material 7
    h["info"] = 1 | (1 << 9) | (2 << 11) | (3 << 13) | (7 << 16)
    r = cq.hits_to_intersections(rays, h, True)[0]
    assert r["hit"] == 1 and r["material"] == 7 and r["distance"] == 3.0
    assert list(r["position"]) == [np.float32(1.0) + np.float32(0.5) * np.float32(3.0), 2.0 - 0.75, 3.0]
    assert r["normal"][0] == 1.0 and r["normal"][1] == 0.0 and np.signbit(r["normal"][1])
    assert r["normal"][2] == -1.0
    r0 = cq.hits_to_intersections(rays, h, False)[0]
    assert r0["material"] == 0 and not np.signbit(r0["normal"]).any()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="only meaningful where no GPU is visible")
def test_create_without_gpu_reports_no_device(cq):
    with pytest.raises(cq.CqxError) as e:
        cq.Context(0)
    assert e.value.status == -6
