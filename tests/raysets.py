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
