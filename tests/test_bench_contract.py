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
    """Every profiled launch's calibrated HBM bytes follow from its counters; the bench fills
    traffic and the L2 hit rate only for the launch it actually timed."""
    sys.path.insert(0, ROOT)
    import bench

    d = json.load(open(os.path.join(ROOT, "profiles", "traffic_intersect_incoherent_1024.json")))
    assert d["launches"]
    for p in d["launches"]:
        cal = p["calibration"]
        assert abs(cal["fetch_factor"] - cal["copy_bytes"] / 1024 / cal["copy_fetch_size_kib"]) < 1e-4
        assert abs(cal["write_factor"] - cal["copy_bytes"] / 1024 / cal["copy_write_size_kib"]) < 1e-4
        hbm = p["fetch_size_kib_per_launch"] * 1024 * cal["fetch_factor"] + \
            p["write_size_kib_per_launch"] * 1024 * cal["write_factor"]
        assert abs(hbm - p["hbm_bytes_per_launch"]) / hbm < 1e-3
        assert 0 < p["l2_hit_rate"] < 1
        # the records alone are 2^24 x 32 bytes
        assert p["write_size_kib_per_launch"] * 1024 * cal["write_factor"] >= (1 << 24) * 32
        cfg = dict(kv.split("=") for kv in p["launch"].split(","))
        assert p["launch"] == bench.launch_string({k: int(v) for k, v in cfg.items()})
        traffic, l2, prof = bench.traffic_from_profiles(p["launch"])
        assert traffic == p["hbm_bytes_per_launch"] and l2 == p["l2_hit_rate"] and prof
        # the launch the bench would time instead of an unprofiled tuned one parses into a LaunchCfg
        import cubiquity_amd as cq

        c = bench.launch_cfg(cq, p["launch"])
        assert bench.launch_string({k: getattr(c, k) for k in bench.LAUNCH_KEYS + ("stack_rows",)}) == p["launch"]
    assert bench.profiled_launches() == [p["launch"] for p in d["launches"]]
    assert bench.traffic_from_profiles("variant=3,block=128,waves_per_cu=7,regroup_interval=2,refill_lanes=3,"
                                       "lds_nodes=0,bin_bits=0") == (None, None, None)
