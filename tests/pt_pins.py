"""Checks that pin the path-tracer oracle (oracle/cpu_pt.c) to something outside itself.

Each check takes an oracle (oracle.bindings.Oracle, possibly a planted-slip build) and raises
AssertionError when the oracle disagrees with:

  * pcg32        O'Neill's published pcg32 demo stream (pcg32-demo.c: pcg32_srandom_r(42, 54));
  * sincos       FP64 sin/cos of 2 pi u on every FP32 uniform draw u = k / 2^24;
  * bounce       the cosine-weighted hemisphere law: unit directions, n.d >= 0, E[n.d] = 2/3,
                 E[(n.d)^2] = 1/2, the CDF of n.d (mu^2), uniform azimuth, the origin offset;
  * camera       the FP64 pinhole of src/application/commands/view/camera.cpp:12-38 (forward +
                 right * x - up * y, y down) through the jittered pixel point;
  * furnace      a camera inside a closed shell of one material (albedo a, emission e): every
                 path hits B + 1 times, so every pixel is e (1 + a + ... + a^B);
  * floor        a camera above a slab whose every camera ray hits its top face: a bounce goes
                 up and escapes, so every pixel is e + a sky for B >= 1 and e for B = 0;
  * traversal    the compiled reference's records held in the golden fixture.

Used by tests/test_pt_oracle_pinning.py on the oracle and by tests/test_oracle_mutations.py on
builds with a planted slip (each slip must fail at least one check).
"""
import hashlib
import math
import os

import numpy as np

# pcg32-demo.c (O'Neill, pcg-c-basic), "Round 1" of pcg32_random_r after pcg32_srandom_r(42u, 54u)
PCG32_DEMO_42_54 = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]

SINCOS_MAX_ABS_ERR = 4e-7  # FP32 result against FP64 sin/cos; the oracle measures ~1.8e-7


def check_pcg32(orc):
    got = [int(v) for v in orc.pcg32(42, 54, len(PCG32_DEMO_42_54))]
    assert got == PCG32_DEMO_42_54, [hex(v) for v in got]
    # the uniform draw is the top 24 bits scaled by 2^-24: exactly representable, in [0, 1)
    u = orc.pcg_uniform(42, 54, len(PCG32_DEMO_42_54)).astype(np.float64)
    assert (u == np.asarray([v >> 8 for v in PCG32_DEMO_42_54], np.float64) / 2.0**24).all(), u


def check_sincos(orc):
    k = np.arange(1 << 24, dtype=np.float64)
    u = (k / 2.0**24).astype(np.float32)
    c, s = orc.sincos2pi(u)
    th = 2.0 * math.pi * k / 2.0**24
    ec = np.abs(c.astype(np.float64) - np.cos(th)).max()
    es = np.abs(s.astype(np.float64) - np.sin(th)).max()
    assert ec <= SINCOS_MAX_ABS_ERR and es <= SINCOS_MAX_ABS_ERR, (ec, es)


# face normals as the traversal reports them (components in {-1, 0, 1}, several at edges/corners)
BOUNCE_NORMALS = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
                  (1, 1, 0), (0, -1, 1), (-1, 0, -1), (1, -1, 1), (-1, -1, -1)]


def _tangent_basis(n):
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    e1 = np.cross(n, a)
    e1 /= np.linalg.norm(e1)
    return e1, np.cross(n, e1)


def check_bounce(orc, count=1 << 17, normals=BOUNCE_NORMALS):
    pos = np.array([3.25, -7.5, 11.0], np.float32)
    for i, nrm in enumerate(normals):
        nf = np.asarray(nrm, np.float32)
        rays = orc.bounce_rays(nf, pos, 0.01, 1000 + i, 7 + i, count)
        assert rays.shape[0] == count, (nrm, rays.shape)
        # origin pos + n * offset with the face normal as given (FP32)
        want_o = pos + nf * np.float32(0.01)
        assert (rays[:, :3] == want_o).all(), (nrm, rays[0, :3], want_o)
        d = rays[:, 3:].astype(np.float64)
        n = nf.astype(np.float64) / np.linalg.norm(nf)
        assert np.abs(np.linalg.norm(d, axis=1) - 1.0).max() <= 2e-6, (nrm, "not unit length")
        mu = d @ n
        assert mu.min() >= -1e-6, (nrm, "below the surface", mu.min())
        # cosine-weighted: mu has density 2 mu on [0, 1]; E[mu] = 2/3 (var 1/18), E[mu^2] = 1/2 (var 1/12)
        for m, want, var in ((mu.mean(), 2.0 / 3.0, 1.0 / 18.0), ((mu * mu).mean(), 0.5, 1.0 / 12.0)):
            assert abs(m - want) <= 5.0 * math.sqrt(var / count), (nrm, m, want)
        srt = np.sort(np.clip(mu, 0.0, 1.0))
        ks = np.max(np.abs(srt * srt - (np.arange(1, count + 1) - 0.5) / count))
        assert ks <= 1.95 / math.sqrt(count), (nrm, "CDF of n.d", ks)
        # azimuth uniform: chi-square over 32 sectors (p ~ 1e-4 cut), tangential mean ~ 0
        e1, e2 = _tangent_basis(n)
        phi = np.arctan2(d @ e2, d @ e1)
        hist, _ = np.histogram(phi, bins=32, range=(-math.pi, math.pi))
        expect = count / 32.0
        chi2 = float(((hist - expect) ** 2 / expect).sum())
        assert chi2 < 70.0, (nrm, "azimuth", chi2)
        tmean = np.abs([(d @ e1).mean(), (d @ e2).mean()]).max()
        assert tmean <= 5.0 * math.sqrt(0.25 / count), (nrm, "tangential mean", tmean)


def _pinhole64(eye, target, fov_y, width, height, px, py):
    """camera.cpp:12-38 geometry in FP64 at the image-plane point (px, py) (pixel units, y down)."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.array([-f[2], 0.0, f[0]])
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    scale = 2.0 * math.tan(math.radians(fov_y) / 2.0)
    aspect = width / height
    d = f + r * ((px - width / 2.0) / width * aspect * scale) - u * ((py - height / 2.0) / height * scale)
    return d / np.linalg.norm(d), f, r, u


def check_camera(orc, cq, W):
    cam = W.RAYCAST_128.camera
    width, height, spp = 40, 30, 3
    alb, em = cq.material_table(2, 1)
    s = cq.PathTraceSetup(cam, width, height, spp, 1, alb, em, seed=5)
    rng = np.random.Generator(np.random.PCG64(3))
    paths = list(rng.integers(0, width * height * spp, 200)) + [0, width * height * spp - 1]
    for path in paths:
        path = int(path)
        ray, jit = orc.pt_camera_ray(s.params, path)
        pixel, sample = path // spp, path % spp
        # the jitter is the path stream's first two draws: initstate (seed << 32) + pixel, stream sample
        want = orc.pcg_uniform((5 << 32) + pixel, sample, 2)
        assert (jit == want).all(), (path, jit, want)
        x, y = pixel % width, pixel // width
        d64, f, r, u = _pinhole64(cam.eye, cam.target, cam.fov_y, width, height, x + float(jit[0]),
                                  y + float(jit[1]))
        assert (ray[:3] == np.asarray(cam.eye, np.float32)).all()
        assert np.abs(ray[3:].astype(np.float64) - d64).max() <= 2e-6, (path, ray[3:], d64)
        # orientation: right of centre looks right, below centre looks down (y inverted)
        dd = ray[3:].astype(np.float64)
        assert (dd @ r > 0) == (x + jit[0] > width / 2.0) or abs(x + jit[0] - width / 2.0) < 1e-3
        assert (dd @ u < 0) == (y + jit[1] > height / 2.0) or abs(y + jit[1] - height / 2.0) < 1e-3


def furnace_scene(cq, W, bounces, spp=2):
    n = 24
    g = np.ones((n, n, n), np.uint8)
    g[2:n - 2, 2:n - 2, 2:n - 2] = 0  # a closed shell two voxels thick
    hv = cq.HostVolume.from_dense(g)
    cam = W.Camera((11.7, 12.3, 11.1), (13.0, 10.0, 30.0), 60.0, 16, 12)
    alb = np.array([[0.5, 0.7, 0.9]], np.float32)
    em = np.array([[0.3, 0.2, 0.1]], np.float32)
    s = cq.PathTraceSetup(cam, 16, 12, spp, bounces, alb, em, sky=(1000.0, 1000.0, 1000.0), seed=9)
    want = em[0].astype(np.float64) * sum(alb[0].astype(np.float64) ** k for k in range(bounces + 1))
    return hv, s, want


def floor_scene(cq, W, bounces, spp=2):
    n = 32
    g = np.zeros((n, n, n), np.uint8)
    g[:, 0:4, :] = 1  # [z][y][x]: a slab under y = 3.5
    hv = cq.HostVolume.from_dense(g)
    cam = W.Camera((16.0, 10.0, 16.0), (16.2, 0.0, 17.0), 60.0, 16, 12)
    alb = np.array([[0.25, 0.5, 0.75]], np.float32)
    em = np.array([[0.125, 0.0, 0.5]], np.float32)
    sky = np.array([0.8, 0.8, 1.0], np.float32)
    s = cq.PathTraceSetup(cam, 16, 12, spp, bounces, alb, em, sky=tuple(sky), seed=4)
    e, a = em[0].astype(np.float64), alb[0].astype(np.float64)
    want = e + a * sky.astype(np.float64) if bounces >= 1 else e
    return hv, s, want


def check_closed_forms(orc, cq, W, render=None):
    """Furnace and floor images against their closed forms. `render(hv, setup)` defaults to the
    oracle's FP32 image; the GPU tests pass the GPU renderer."""
    if render is None:
        render = lambda hv, s: orc.pathtrace(hv.nodes, hv.root, s.params)  # noqa: E731
    for make in (furnace_scene, floor_scene):
        for b in (0, 1, 2, 4, 7):
            hv, s, want = make(cq, W, b)
            img = np.asarray(render(hv, s), np.float64)
            err = np.abs(img - want).max() / want.max()
            assert err <= 2e-6, (make.__name__, b, err, img.reshape(-1, 3)[0], want)


def golden_cases(golden, lattice):
    """(name, rays, surface properties, max footprint, reference records) of both fixtures."""
    rays = golden["rays"]
    out = [("golden/" + key, rays, surf, fp, golden[key])
           for surf, fp, key in ((True, -1.0, "ref_surf1"), (False, -1.0, "ref_surf0"), (True, 0.05, "ref_lod"))]
    rays = lattice["rays"]
    out += [("lattice/ref_surf1", rays, True, -1.0, lattice["ref_surf1"]),
            ("lattice/ref_surf0", rays, False, -1.0, lattice["ref_surf0"])]
    out += [(f"lattice/{k}", rays, True, float(k[len("ref_lod_"):]), lattice[k])
            for k in lattice.files if k.startswith("ref_lod_")]
    return out


def check_traversal_golden(orc, vol128, golden, lattice):
    """The oracle's traversal records against the compiled reference's, held in the golden
    fixtures (bit for bit)."""
    for z in (golden, lattice):
        assert hashlib.sha256(vol128.nodes.tobytes()).hexdigest() == str(z["dag_sha256"])
    sub = orc.find_subdags(vol128.nodes, vol128.root)
    for name, rays, surf, fp, ref in golden_cases(golden, lattice):
        # the cap only matters to a broken build (the fixtures' rays take far fewer iterations)
        o = orc.to_results(rays, orc.intersect(vol128.nodes, sub, rays, surf, fp, max_iters=1 << 14), surf)
        ga = np.ascontiguousarray(ref)
        eq = (o.view(np.uint8).reshape(len(o), -1) == ga.view(np.uint8).reshape(len(ga), -1)).all(axis=1)
        assert eq.all(), f"{name}: {(~eq).sum()} differ"


GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "raycast_128_golden.npz")
LATTICE = os.path.join(os.path.dirname(__file__), "golden", "lattice_128_golden.npz")
