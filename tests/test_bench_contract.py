"""bench.py's contract pieces that need no GPU: one GPU only, the fixed synthetic workloads, and
the committed traffic profile the roofline reads."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_edit_frames_are_seeded_by_frame():
    from cubiquity_amd import workloads as W

    a, b, c = W.edit_frame_spheres(0), W.edit_frame_spheres(0), W.edit_frame_spheres(1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.shape == (W.EDIT_FRAME_EDITS, 5)
    assert (a[:, :3] >= 0).all() and (a[:, :3] < 2048).all()
    assert (a[:, 3] >= 2).all() and (a[:, 3] <= 16).all()
    assert set(np.unique(a[:, 4]).astype(int)) <= set(range(17))


def test_traffic_profile_is_consistent():
    d = json.load(open(os.path.join(ROOT, "profiles", "traffic_intersect_incoherent_1024.json")))
    cal = d["calibration"]
    assert abs(cal["fetch_factor"] - cal["copy_bytes"] / 1024 / cal["copy_fetch_size_kib"]) < 1e-4
    assert abs(cal["write_factor"] - cal["copy_bytes"] / 1024 / cal["copy_write_size_kib"]) < 1e-4
    hbm = d["fetch_size_kib_per_launch"] * 1024 * cal["fetch_factor"] + \
        d["write_size_kib_per_launch"] * 1024 * cal["write_factor"]
    assert abs(hbm - d["hbm_bytes_per_launch"]) / hbm < 1e-3
    assert 0 < d["l2_hit_rate"] < 1
    # the records alone are 2^24 x 32 bytes
    assert d["write_size_kib_per_launch"] * 1024 * cal["write_factor"] >= (1 << 24) * 32
