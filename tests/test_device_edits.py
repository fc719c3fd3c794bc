"""On-device sphere edits (cqx_volume_edit_spheres), on the GPU:

  * the device store equals the host's region-parallel edit store (HostVolume.edit_spheres with
    8 threads) node for node, with the same root and first appended node, frame after frame (the
    device index kept between calls, the device store grown in place), including spheres reaching
    past the volume (new root octants), near-duplicate spheres across regions, and spheres large
    enough that regions list more nodes than their first slice holds;
  * the voxels equal the compiled reference's fillBrush after the same spheres in order;
  * traversal of the edited device volume is bit-exact against the oracle on the downloaded store;
  * device checkpoint / undo / redo switch roots as the host's do;
  * a batch the device path does not take is refused and leaves the volume unchanged.
"""
import numpy as np
import pytest

from cubiquity_amd import workloads as W
from tests.conftest import records_equal

pytestmark = pytest.mark.gpu


def _frame(log2, f, n, kind):
    e = W.edit_frame_spheres(f, log2_size=log2, n=n)
    if kind == "near_duplicates":
        e[::3, :3] = e[1::3, :3] + 0.5
    elif kind == "past_bounds":
        rng = np.random.Generator(np.random.PCG64(50 + f))
        side = float(1 << log2)
        e[:, :3] = rng.uniform(-0.1 * side, 1.1 * side, (n, 3))
    elif kind == "large":
        rng = np.random.Generator(np.random.PCG64(60 + f))
        e[:, 3] = rng.uniform(10.0, 30.0, n)
    return e


@pytest.mark.parametrize("log2,seed,n,frames,kind", [
    (7, 3, 2000, 3, "near_duplicates"),
    (7, 4, 1200, 2, "past_bounds"),
    (8, 5, 150, 2, "large"),
    (9, 4, 5000, 2, "plain"),
])
def test_device_store_equals_host_store(cq, gpu_ctx, monkeypatch, log2, seed, n, frames, kind):
    if kind == "large":  # first slices of 64 nodes: regions outgrow them and run again, up to 3 times
        monkeypatch.setenv("CQX_EDIT_FIRST_SLICE", "64")
    hv = cq.HostVolume.generate(cq.gen_default_params(log2, seed, n_materials=16))
    dv = gpu_ctx.upload(hv)
    notes = []
    cq.set_debug_handler(notes.append)
    for f in range(frames):
        e = _frame(log2, f, n, kind)
        hf = hv.edit_spheres(e, threads=8)
        df = dv.edit_spheres(e)
        assert df == hf, (f, df, hf)
        assert (dv.node_count, dv.root) == (hv.node_count, hv.root), f
        got = dv.download()
        eq = (got == hv.nodes).all(axis=1)
        assert eq.all(), f"frame {f}: {(~eq).sum()} of {len(eq)} nodes differ, first at {np.nonzero(~eq)[0][:5]}"
    cq.set_debug_handler(None)
    if kind == "large":  # regions listed more nodes than their first slices held, and ran again
        assert any("outgrew their node lists" in m for m in notes), notes
    dv.free()


def test_device_edits_match_reference_fillbrush(cq, oracle, reference, gpu_ctx, tmp_path):
    hv = cq.HostVolume.generate(cq.gen_default_params(6, 4, n_materials=9))
    path = str(tmp_path / "v.dag")
    hv.save_dag(path)
    rv = reference.load(path)
    dv = gpu_ctx.upload(hv)
    rng = np.random.Generator(np.random.PCG64(7))
    e = np.zeros((120, 5))
    e[:, :3] = rng.uniform(-4.0, 68.0, (120, 3))
    e[:, 3] = rng.uniform(0.5, 12.0, 120)
    e[:, 4] = rng.integers(0, 10, 120)
    e[::7, 4] = 0
    dv.edit_spheres(e)
    for s in e:
        rv.fill_sphere(s[:3], s[3], int(s[4]))
    lo, size = (-8, -8, -8), (80, 80, 80)
    ours = oracle.dense(dv.download(), dv.root, lo, size)
    assert np.array_equal(ours, rv.dense(lo, size))
    assert (ours != oracle.dense(hv.nodes, hv.root, lo, size)).sum() > 1000  # the edits did something
    dv.free()


def test_traversal_after_device_edits_and_history(cq, oracle, gpu_ctx):
    import torch

    hv = cq.HostVolume.generate(cq.gen_default_params(8, 6, n_materials=16))
    dv = gpu_ctx.upload(hv)
    rays = cq.incoherent_rays(hv, 8, 17, 30000)
    r = torch.from_numpy(rays).cuda()

    def trace_and_check():
        nodes = dv.download()
        g = cq.intersectVolume(gpu_ctx, dv, r, True).cpu().numpy().view(cq.HIT_DTYPE).reshape(-1)
        o = oracle.intersect(nodes, oracle.find_subdags(nodes, dv.root), rays, True, threads=8)
        assert records_equal(g, o).all()
        return g

    before = trace_and_check()
    root0 = dv.root
    dv.checkpoint()
    dv.edit_spheres(_frame(8, 0, 600, "past_bounds"))
    after = trace_and_check()
    assert not records_equal(before, after).all()
    dv.undo()
    assert dv.root == root0
    assert records_equal(trace_and_check(), before).all()
    dv.redo()
    assert records_equal(trace_and_check(), after).all()
    dv.free()
    torch.cuda.synchronize()


def test_device_edits_refuse_what_the_host_path_must_do(cq, gpu_ctx):
    hv = cq.HostVolume.generate(cq.gen_default_params(6, 2))
    dv = gpu_ctx.upload(hv)
    n0, r0 = dv.node_count, dv.root
    few = _frame(6, 0, 10, "plain")
    far = _frame(6, 1, 40, "plain")
    far[3, :3] = (9.0e6, 1.0, 1.0)
    for e in (few, far):
        with pytest.raises(cq.CqxError):
            dv.edit_spheres(e)
        assert (dv.node_count, dv.root) == (n0, r0)
    with pytest.raises(ValueError):
        dv.edit_spheres([[1, 1, 1, 2, -3]] * 20)
    dv.free()
