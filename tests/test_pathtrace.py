"""Path tracer and path-replay adjoint.

CPU (oracle pinning): the FP64 oracle gradient against central finite differences of the FP64
image (the image is a polynomial in the tables, so FD with h = 1e-3 is accurate); closed forms
(an empty volume shows the sky; emission enters linearly); FP32 oracle image against FP64.
GPU: both renderers (hipGraph wavefront, megakernel) bit-identical to the FP32 oracle image
(stated tolerance: RMSE <= 1e-6 of the mean radiance; observed 0); graph replays interleaved with
eager work all equal; the adjoint within 1e-4 relative of the FP64 oracle gradient and within
1e-3 relative of central finite differences (h = 1e-3) of the GPU loss, and bit-reproducible from
call to call (its sums have a fixed order).
"""
import numpy as np
import pytest

from cubiquity_amd import workloads as W


@pytest.fixture(scope="module")
def scene(cq):
    hv = cq.HostVolume.generate(cq.gen_default_params(7, 1, n_materials=9))
    alb, em = cq.material_table(9, 11)
    setup = cq.PathTraceSetup(W.RAYCAST_128.camera, 96, 64, 4, 4, alb, em)
    return hv, setup


@pytest.fixture(scope="module")
def scene16(cq):
    """A reduced differentiable-rendering case: 64^3, 16 materials, 3 bounces."""
    hv = cq.HostVolume.generate(cq.gen_default_params(6, 5, n_materials=16))
    cam = W._camera(6, 48, 48)
    alb, em = cq.material_table(16, 11)
    tgt_alb, _ = cq.material_table(16, 12)
    setup = cq.PathTraceSetup(cam, 48, 48, 4, 3, alb, em)
    return hv, setup, tgt_alb


def _fd_step(x, h):
    """float32 perturbations and the actual step between them."""
    xp = np.float32(x + h)
    xm = np.float32(x - h)
    return xp, xm, float(xp) - float(xm)


# ---- CPU: pin the oracle ------------------------------------------------------------------------

def test_oracle_fp32_vs_fp64_image(oracle, scene):
    hv, s = scene
    a = oracle.pathtrace(hv.nodes, hv.root, s.params)
    b, _, _ = oracle.pathtrace_f64(hv.nodes, hv.root, s.params)
    assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max()
    assert a.mean() > 0.05  # the scene is lit


def test_oracle_empty_volume_is_sky(cq, oracle):
    hv = cq.HostVolume.from_dense(np.zeros((8, 8, 8), np.uint8))
    alb, em = cq.material_table(3, 1)
    s = cq.PathTraceSetup(W.RAYCAST_128.camera, 16, 12, 2, 3, alb, em, sky=(0.8, 0.7, 0.25))
    img = oracle.pathtrace(hv.nodes, hv.root, s.params)
    assert (img == np.asarray([0.8, 0.7, 0.25], np.float32)).all()


def test_oracle_emission_is_linear(oracle, scene):
    hv, s = scene
    sky0 = s.with_tables(s.albedo, s.emission)
    sky0.params.sky[0] = sky0.params.sky[1] = sky0.params.sky[2] = 0.0
    a, _, _ = oracle.pathtrace_f64(hv.nodes, hv.root, sky0.params)
    e2 = sky0.with_tables(s.albedo, 2 * s.emission)
    e2.params.sky[0] = e2.params.sky[1] = e2.params.sky[2] = 0.0
    b, _, _ = oracle.pathtrace_f64(hv.nodes, hv.root, e2.params)
    assert np.allclose(b, 2 * a, rtol=1e-12, atol=1e-12)
    assert a.max() > 0


def test_oracle_gradient_vs_finite_differences(oracle, scene16):
    hv, s, tgt_alb = scene16
    target, _, _ = oracle.pathtrace_f64(hv.nodes, hv.root, s.with_tables(tgt_alb, s.emission).params)
    img, _, _ = oracle.pathtrace_f64(hv.nodes, hv.root, s.params)
    _, da, de = oracle.pathtrace_f64(hv.nodes, hv.root, s.params, 2 * (img - target))

    def loss(a, e):
        im, _, _ = oracle.pathtrace_f64(hv.nodes, hv.root, s.with_tables(a, e).params)
        return ((im - target) ** 2).sum()

    rng = np.random.Generator(np.random.PCG64(4))
    for table, grad in (("albedo", da), ("emission", de)):
        for m, c in zip(rng.integers(0, 16, 6), rng.integers(0, 3, 6)):
            a, e = s.albedo.copy(), s.emission.copy()
            t = a if table == "albedo" else e
            xp, xm, step = _fd_step(t[m, c], 1e-3)
            t[m, c] = xp
            lp = loss(a, e)
            t[m, c] = xm
            lm = loss(a, e)
            fd = (lp - lm) / step
            scale = max(abs(fd), 1e-3 * np.abs(grad).max())
            assert abs(grad[m, c] - fd) <= 1e-4 * scale, (table, m, c, grad[m, c], fd)


# ---- GPU ------------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def dscene(gpu_ctx, scene):
    hv, s = scene
    return gpu_ctx.upload(hv)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_image_matches_oracle(cq, oracle, gpu_ctx, scene, dscene, mode):
    hv, s = scene
    img = cq.pathtrace(gpu_ctx, dscene, s, mode).cpu().numpy()
    ref = oracle.pathtrace(hv.nodes, hv.root, s.params)
    rmse = float(np.sqrt(((img.astype(np.float64) - ref) ** 2).mean()))
    assert rmse <= 1e-6 * float(np.abs(ref).mean()), rmse
    assert np.array_equal(img, ref), f"{(img != ref).sum()} values differ (rmse {rmse})"


@pytest.mark.gpu
def test_gpu_graph_replays_match_eager(cq, gpu_ctx, scene, dscene):
    """The captured wavefront graph, replayed with eager kernels and syncs in between."""
    import torch

    hv, s = scene
    eager = cq.pathtrace(gpu_ctx, dscene, s, 1).clone()
    for k in range(4):
        img = cq.pathtrace(gpu_ctx, dscene, s, 0).clone()
        torch.cuda.synchronize()
        assert torch.equal(img, eager), f"replay {k}"
        cq.pathtrace(gpu_ctx, dscene, s, 1)  # eager work between replays
        torch.cuda.synchronize()
    # new tables reuse the same graph (they live in a parameter block, not in the graph)
    s2 = s.with_tables(s.albedo * 0.5, s.emission)
    a = cq.pathtrace(gpu_ctx, dscene, s2, 0).clone()
    b = cq.pathtrace(gpu_ctx, dscene, s2, 1).clone()
    assert torch.equal(a, b) and not torch.equal(a, eager)


@pytest.mark.gpu
def test_gpu_adjoint_vs_fp64_and_fd(cq, oracle, gpu_ctx, scene16):
    import torch

    hv, s, tgt_alb = scene16
    dv = gpu_ctx.upload(hv)
    st = s.with_tables(tgt_alb, s.emission)
    target = cq.pathtrace(gpu_ctx, dv, st, 1).double()
    img = cq.pathtrace(gpu_ctx, dv, s, 1).double()
    d_image = (2 * (img - target)).float()
    da, de = cq.pathtrace_adjoint(gpu_ctx, dv, s, d_image)
    da, de = da.cpu().numpy().astype(np.float64), de.cpu().numpy().astype(np.float64)
    # FP64 reference gradient over the same paths, same dLoss/dimage
    _, ra, re = oracle.pathtrace_f64(hv.nodes, hv.root, s.params, d_image.cpu().numpy().astype(np.float64))
    g = np.concatenate([da.ravel(), de.ravel()])
    r = np.concatenate([ra.ravel(), re.ravel()])
    scale = np.maximum(np.abs(r), 1e-3 * np.abs(r).max())
    assert (np.abs(g - r) <= 1e-4 * scale).all(), np.abs(g - r).max()

    def loss(a, e):
        im = cq.pathtrace(gpu_ctx, dv, s.with_tables(a, e), 1).double()
        return float(((im - target) ** 2).sum())

    # central differences of the GPU loss on all 96 parameters; relative error per parameter
    # against max(|fd|, 1e-2 max|fd|) and of the whole gradient vector (L2)
    fds, gs = [], []
    for table, grad in (("albedo", da), ("emission", de)):
        for m in range(16):
            for c in range(3):
                a, e = s.albedo.copy(), s.emission.copy()
                t = a if table == "albedo" else e
                xp, xm, step = _fd_step(t[m, c], 1e-3)
                t[m, c] = xp
                lp = loss(a, e)
                t[m, c] = xm
                lm = loss(a, e)
                fds.append((lp - lm) / step)
                gs.append(grad[m, c])
    fds, gs = np.asarray(fds), np.asarray(gs)
    sc = np.maximum(np.abs(fds), 1e-2 * np.abs(fds).max())
    per = np.abs(gs - fds) / sc
    l2 = np.linalg.norm(gs - fds) / np.linalg.norm(fds)
    print(f"adjoint vs FD: max per-parameter {per.max():.2e}, L2 {l2:.2e}", flush=True)
    assert per.max() <= 1e-3 and l2 <= 1e-3, (per.max(), l2)
    dv.free()
    torch.cuda.synchronize()


@pytest.mark.gpu
def test_gpu_adjoint_is_bit_reproducible(cq, gpu_ctx, scene16):
    """The adjoint's sums have a fixed order (no atomics): repeated calls, with other work on the
    GPU in between, return the same gradient bit for bit."""
    import torch

    hv, s, _ = scene16
    dv = gpu_ctx.upload(hv)
    img = cq.pathtrace(gpu_ctx, dv, s, 0).double()
    d_image = (2 * img - 0.25).float()
    first = [t.cpu().numpy().copy() for t in cq.pathtrace_adjoint(gpu_ctx, dv, s, d_image)]
    for k in range(3):
        cq.pathtrace(gpu_ctx, dv, s.with_tables(s.albedo * 0.9, s.emission), 1)  # other work in between
        again = [t.cpu().numpy() for t in cq.pathtrace_adjoint(gpu_ctx, dv, s, d_image)]
        for a, b in zip(first, again):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"call {k + 2} differs"
    dv.free()
    torch.cuda.synchronize()


@pytest.mark.gpu
@pytest.mark.parametrize("bounces", [0, 7, 8, 17, 64])
def test_gpu_adjoint_at_bounce_limits(cq, oracle, gpu_ctx, scene16, bounces):
    """The adjoint replays the wavefront pass, which records up to bounces + 1 vertex materials
    per path, then sweeps the record in segments of 4 vertices, last first: at 0 bounces (one
    vertex), at 7 and 8 (two full segments, then one more vertex), at 17 (five segments) and at 64
    (the most the path tracer takes) the gradient must equal the FP64 one over the same paths."""
    import torch

    hv, s0, _ = scene16
    s = cq.PathTraceSetup(W._camera(6, 32, 32), 32, 32, 2, bounces, s0.albedo, s0.emission)
    dv = gpu_ctx.upload(hv)
    img = cq.pathtrace(gpu_ctx, dv, s, 0).double()
    d_image = (2 * img).float()  # dLoss/dimage of sum(image^2)
    da, de = cq.pathtrace_adjoint(gpu_ctx, dv, s, d_image)
    da, de = da.cpu().numpy().astype(np.float64), de.cpu().numpy().astype(np.float64)
    _, ra, re = oracle.pathtrace_f64(hv.nodes, hv.root, s.params, d_image.cpu().numpy().astype(np.float64))
    g = np.concatenate([da.ravel(), de.ravel()])
    r = np.concatenate([ra.ravel(), re.ravel()])
    assert np.abs(r).max() > 0
    scale = np.maximum(np.abs(r), 1e-3 * np.abs(r).max())
    assert (np.abs(g - r) <= 1e-4 * scale).all(), (bounces, np.abs(g - r).max())
    dv.free()
    torch.cuda.synchronize()


@pytest.mark.gpu
@pytest.mark.parametrize("spp", [64, 65])
def test_gpu_resolve_both_paths(cq, oracle, gpu_ctx, scene, dscene, spp):
    """k_pt_resolve stages a block's samples in LDS up to 64 spp and reads them directly above;
    both sum each pixel's samples in sample order, so the image equals the oracle's bit for bit."""
    hv, s0 = scene
    s = cq.PathTraceSetup(W.RAYCAST_128.camera, 12, 8, spp, 1, s0.albedo, s0.emission)
    img = cq.pathtrace(gpu_ctx, dscene, s, 0).cpu().numpy()
    ref = oracle.pathtrace(hv.nodes, hv.root, s.params)
    assert np.array_equal(img, ref), f"spp {spp}: {(img != ref).sum()} values differ"


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_odd_image_size(cq, oracle, gpu_ctx, scene, dscene, mode):
    """37 x 23 pixels x 3 spp = 2553 paths: the shading pass's 1024-thread compaction blocks end
    part-full at every depth, and the resolve's last 256-pixel block holds 83 pixels."""
    hv, s0 = scene
    s = cq.PathTraceSetup(W.RAYCAST_128.camera, 37, 23, 3, 4, s0.albedo, s0.emission)
    img = cq.pathtrace(gpu_ctx, dscene, s, mode).cpu().numpy()
    ref = oracle.pathtrace(hv.nodes, hv.root, s.params)
    assert np.array_equal(img, ref), f"mode {mode}: {(img != ref).sum()} values differ"


def _furnace_gradient(setup, alb, em, bounces):
    """d(sum image)/d(albedo, emission) in the closed furnace: every pixel is e (1 + a + ... + a^B)
    per channel, so the gradient is W H (1 + a + ... + a^B) for e and W H e (1 + 2a + ... + B a^(B-1))
    for a."""
    a = alb[0].astype(np.float64)
    e = em[0].astype(np.float64)
    px = setup.params.width * setup.params.height
    de = px * sum(a ** k for k in range(bounces + 1))
    da = px * e * sum(k * a ** (k - 1) for k in range(1, bounces + 1))
    return da, de


@pytest.mark.parametrize("bounces", [8, 17, 64])
def test_oracle_adjoint_furnace_closed_form(cq, oracle, bounces):
    """The FP64 adjoint of the oracle for paths of up to 65 vertices against the furnace's closed
    form (every path bounces until the limit)."""
    from tests.pt_pins import furnace_scene

    hv, s, _ = furnace_scene(cq, W, bounces)
    d_image = np.ones((s.params.height, s.params.width, 3), np.float64)
    _, ra, re = oracle.pathtrace_f64(hv.nodes, hv.root, s.params, d_image)
    da, de = _furnace_gradient(s, s.albedo, s.emission, bounces)
    assert np.allclose(ra[0], da, rtol=1e-9) and np.allclose(re[0], de, rtol=1e-9), (ra, da, re, de)


@pytest.mark.gpu
@pytest.mark.parametrize("bounces", [7, 8, 17, 64])
def test_gpu_adjoint_furnace_deep_paths(cq, oracle, gpu_ctx, bounces):
    """Every furnace path has bounces + 1 vertices, so the adjoint's backward sweep runs 2 to 17
    segments of 4 vertices in every lane: the gradient equals the closed form and the FP64 oracle's
    within 1e-4 relative."""
    import torch

    from tests.pt_pins import furnace_scene

    hv, s, _ = furnace_scene(cq, W, bounces)
    dv = gpu_ctx.upload(hv)
    d_image = torch.ones((s.params.height, s.params.width, 3), dtype=torch.float32, device="cuda")
    ga, ge = cq.pathtrace_adjoint(gpu_ctx, dv, s, d_image)
    ga, ge = ga.cpu().numpy().astype(np.float64), ge.cpu().numpy().astype(np.float64)
    _, ra, re = oracle.pathtrace_f64(hv.nodes, hv.root, s.params, d_image.cpu().numpy().astype(np.float64))
    da, de = _furnace_gradient(s, s.albedo, s.emission, bounces)
    for got, want, ref in ((ga[0], da, ra[0]), (ge[0], de, re[0])):
        assert np.allclose(got, want, rtol=1e-4, atol=0), (bounces, got, want)
        assert np.allclose(got, ref, rtol=1e-4, atol=0), (bounces, got, ref)
    dv.free()
    torch.cuda.synchronize()


@pytest.mark.gpu
def test_gpu_adjoint_many_materials(cq, oracle, gpu_ctx):
    """More materials than the sweep keeps per-lane sums for (40 > 16): the wave groups its lanes by
    material instead; the gradient equals the FP64 one within 1e-4 and repeats bit for bit."""
    import torch

    n = 48  # a ridged floor whose 3-voxel tiles cycle through the 40 materials
    z, y, x = np.mgrid[0:n, 0:n, 0:n]
    grid = np.zeros((n, n, n), np.uint8)
    floor = y < 10 + ((x // 6 + z // 6) % 3)
    grid[floor] = (1 + (x // 3 + 16 * (z // 3)) % 40)[floor]
    hv = cq.HostVolume.from_dense(grid, (0, 0, 0))
    alb, em = cq.material_table(40, 11)
    s = cq.PathTraceSetup(W.Camera((24.0, 40.0, 4.0), (24.0, 8.0, 30.0), 70.0, 32, 32), 32, 32, 2, 3, alb, em)
    dv = gpu_ctx.upload(hv)
    img = cq.pathtrace(gpu_ctx, dv, s, 0).double()
    d_image = (2 * img - 0.1).float()
    da, de = cq.pathtrace_adjoint(gpu_ctx, dv, s, d_image)
    again = cq.pathtrace_adjoint(gpu_ctx, dv, s, d_image)
    assert np.array_equal(da.cpu().numpy().view(np.uint32), again[0].cpu().numpy().view(np.uint32))
    assert np.array_equal(de.cpu().numpy().view(np.uint32), again[1].cpu().numpy().view(np.uint32))
    _, ra, re = oracle.pathtrace_f64(hv.nodes, hv.root, s.params, d_image.cpu().numpy().astype(np.float64))
    g = np.concatenate([da.cpu().numpy().astype(np.float64).ravel(), de.cpu().numpy().astype(np.float64).ravel()])
    r = np.concatenate([ra.ravel(), re.ravel()])
    assert np.count_nonzero(ra[16:]) > 0 and np.count_nonzero(re[16:]) > 0  # materials 17..40 take part
    scale = np.maximum(np.abs(r), 1e-3 * np.abs(r).max())
    assert (np.abs(g - r) <= 1e-4 * scale).all(), np.abs(g - r).max()
    dv.free()
    torch.cuda.synchronize()
