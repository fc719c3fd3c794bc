"""Special-case ray sets for parity tests: axis-aligned rays (+-0 components), rays through voxel
edges and corners, grazing rays along faces, rays starting outside and pointing away, rays
starting inside the volume. All deterministic."""
import numpy as np


def special_rays(n):
    """About 600 rays for a volume occupying [0, n)^3 (voxel v spans [v-0.5, v+0.5))."""
    rays = []
    c = n / 2.0
    far = 2.0 * n
    # axis-aligned from outside along each axis, both directions, with +0 and -0 zeros
    for axis in range(3):
        for sgn in (1.0, -1.0):
            for zero in (0.0, -0.0):
                for k in range(8):
                    u = (k + 0.5) * n / 8.0
                    o = [u, c + 0.25 * k, u]
                    o[axis] = -far if sgn > 0 else n + far
                    d = [zero, zero, zero]
                    d[axis] = sgn
                    rays.append(o + d)
    # through voxel corners and edges exactly (origins on integer + 0.5 lattice)
    for k in range(64):
        o = [-10.5 + k % 4, n + 10.5 - (k // 4) % 4, -10.5 + (k // 16)]
        t = [c + 0.5 * (k % 7) - 1.5, c * 0.5 + 0.5 * (k % 5), c - 0.5 * (k % 3)]
        d = np.array(t) - np.array(o)
        d = d / np.linalg.norm(d)
        rays.append(o + list(d))
    # diagonal rays (equal |components|) hit edges/corners simultaneously
    for k in range(32):
        s = [1.0 if (k >> b) & 1 else -1.0 for b in range(3)]
        o = [c - s[0] * (far + (k % 4) * 0.5), c - s[1] * far, c - s[2] * (far + (k // 8) * 0.5)]
        inv = 1.0 / np.sqrt(3.0)
        rays.append(o + [s[0] * inv, s[1] * inv, s[2] * inv])
    # grazing rays in a face plane (origin exactly on a voxel boundary)
    for k in range(64):
        y = float(k % 16) + 0.5
        o = [-5.0, y, (k // 16) * 7.5 + 0.5]
        d = [1.0, 0.0, 0.0] if k % 2 == 0 else [0.8, 0.0, 0.6]
        rays.append(o + d)
    # starting outside the volume and pointing away (reference returns hits behind the origin
    # when the origin lies in the outer half-voxel shell)
    for k in range(64):
        o = [n - 0.5 + 0.1 * (k % 4), c + (k % 7), c - (k % 5)]
        a = 0.3 * (k % 9) - 1.2
        d = np.array([1.0, np.sin(a), np.cos(a) * 0.5])
        d = d / np.linalg.norm(d)
        rays.append(o + list(d))
    # starting inside the volume in random directions (includes origins inside solid voxels)
    rng = np.random.Generator(np.random.PCG64(1234))
    for k in range(256):
        o = rng.uniform(0.0, n, 3) - 0.5
        d = rng.normal(size=3)
        d = d / np.linalg.norm(d)
        if k % 16 == 0:
            d[k % 3] = 0.0
            if np.linalg.norm(d) == 0:
                d[(k + 1) % 3] = 1.0
            d = d / np.linalg.norm(d)
        rays.append(list(o) + list(d))
    # tiny direction components (near-axis) and huge origins
    for k in range(32):
        eps = 10.0 ** (-(k % 8) - 1)
        o = [-100.0 - k, c + 0.3, c - 0.7]
        d = np.array([1.0, eps, -eps * (k % 3)])
        d = d / np.linalg.norm(d)
        rays.append(o + list(d))
    return np.asarray(rays, dtype=np.float32)


def deep_volumes(cq):
    """Sparse volumes whose SubDAGs are 20+ levels deep: spheres about 2^20 voxels apart (edits on an
    empty store), in one octant and across octants. Returns {name: (host volume, rays)}; the rays
    start near one sphere or far outside and are aimed at the spheres, so that walks cross large
    empty regions, descend deep stacks and pop back up."""
    import numpy as np

    cases = {
        "deep_one_octant": [(3, 5, 7, 4, 2), (1 << 20, 1 << 19, 9, 6, 3), (700000, 12, 1 << 20, 5, 4)],
        "deep_all_octants": [(-(1 << 18), 7, -(1 << 21), 6, 5), (1 << 20, -(1 << 19), 3, 3, 6), (5, 5, 5, 4, 7),
                             (-1000, -(1 << 20), 1 << 19, 5, 8)],
    }
    out = {}
    for k, (name, spheres) in enumerate(cases.items()):
        hv = cq.HostVolume.from_dense(np.zeros((4, 4, 4), np.uint8), (0, 0, 0))
        hv.edit_spheres([[x, y, z, r, m] for x, y, z, r, m in spheres])
        rng = np.random.Generator(np.random.PCG64(900 + k))
        rays = []
        for (x, y, z, r, _) in spheres:
            c = np.asarray([x, y, z], np.float64)
            for (x2, y2, z2, _, _) in spheres:  # from near each sphere towards every sphere
                o = c + rng.uniform(-3 * r, 3 * r, (60, 3))
                t = np.asarray([x2, y2, z2], np.float64) + rng.uniform(-r, r, (60, 3))
                d = t - o
                rays.append(np.concatenate([o, d / np.linalg.norm(d, axis=1, keepdims=True)], axis=1))
            o = c + rng.normal(size=(200, 3)) * 4000.0  # from far away, aimed near the sphere
            d = c + rng.uniform(-r, r, (200, 3)) - o
            rays.append(np.concatenate([o, d / np.linalg.norm(d, axis=1, keepdims=True)], axis=1))
        d = rng.normal(size=(400, 3))  # random directions from the origin region
        rays.append(np.concatenate([rng.uniform(-8, 8, (400, 3)), d / np.linalg.norm(d, axis=1, keepdims=True)], 1))
        out[name] = (hv, np.concatenate(rays).astype(np.float32))
    return out


def lattice_rays(n, count, seed):
    """Rays whose traversal arithmetic is exact in FP32: origins on the voxel-corner lattice
    (integer - 0.5, so the reference's shifted origin o + 0.5 is an integer) and direction
    components in {0, +-1/4, +-1/2, +-1, +-2}, signed towards the volume's centre (about one ray in
    ten has a zero component, with +0 or -0). Child boundaries, entry and exit times are then
    dyadic, so entries on a child's centre plane, edge and corner crossings and equal axis
    distances (the `<` / `<=` ties of findFirstChild, the sibling advance and the octant walk) occur
    on most rays, and zero components on a node boundary give the NaN operands whose order the
    min/max helpers keep. Not normalised (intersectVolume does not need unit directions)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    o = rng.integers(-n // 2, n + n // 2, (count, 3)).astype(np.float64) - 0.5
    d = rng.choice(np.array([0.25, 0.5, 1.0, 2.0]), (count, 3))
    d = d * np.sign(n / 2.0 - o + 0.01)
    zero = rng.random(count) < 0.1
    axis = rng.integers(0, 3, count)
    neg = rng.random(count) < 0.5
    d[zero, axis[zero]] = np.where(neg[zero], -0.0, 0.0)
    return np.concatenate([o, d], 1).astype(np.float32)
