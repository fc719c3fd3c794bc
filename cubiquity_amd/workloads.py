"""The fixed synthetic workloads (volumes, cameras, ray sets) of the benchmark and the tests.

Version 1 of the definitions; changing any value changes every fixture and number derived from
it, so values are only ever added, never edited.

  raycast_128     128^3 volume (seed 1): fBm heightfield (4 octaves, gain 0.5) + 32 spheres r in
                  [4,16], materials {1 ground, 2 rock below y=40, 3 spheres}; 256x256 primary rays
                  from (64,150,-100) towards (64,40,64), 60 deg vertical FOV; plus incoherent rays
  pathtrace_1024  1024^3 volume (seed 1): 6 octaves, 2048 spheres r in [8,32], 9 materials
                  (8 diffuse + 1 emissive); 1280x720 camera; the headline incoherent set is
                  2^24 rays with seed 7
  batched         64 volumes, seeds 100..163, sides 2^(6 + i mod 6)
  diffrender_512  512^3 volume (seed 5), 16 materials, 512x512 camera
  large_16384     16384^3 volume (seed 9), 1M spheres, 255 materials
"""
from dataclasses import dataclass, field

from . import gen_default_params


@dataclass(frozen=True)
class Camera:
    eye: tuple
    target: tuple
    fov_y: float
    width: int
    height: int


@dataclass(frozen=True)
class Workload:
    name: str
    log2_size: int
    seed: int
    overrides: dict = field(default_factory=dict)
    camera: Camera = None
    incoherent_seed: int = 7
    incoherent_rays: int = 0

    @property
    def side(self):
        return 1 << self.log2_size

    def params(self):
        return gen_default_params(self.log2_size, self.seed, **self.overrides)


def _camera(log2_size, width, height):
    s = (1 << log2_size) / 128.0
    return Camera((64.0 * s, 150.0 * s, -100.0 * s), (64.0 * s, 40.0 * s, 64.0 * s), 60.0, width, height)


RAYCAST_128 = Workload("raycast_128", 7, 1, {}, _camera(7, 256, 256), 7, 16384)
PATHTRACE_1024 = Workload("pathtrace_1024", 10, 1, {"n_materials": 9}, _camera(10, 1280, 720), 7, 1 << 24)
DIFFRENDER_512 = Workload("diffrender_512", 9, 5, {"n_materials": 16}, _camera(9, 512, 512), 7, 0)
LARGE_16384 = Workload("large_16384", 14, 9, {"n_materials": 255, "n_spheres": 1 << 20}, None, 7, 1 << 26)


# The headline volume family at larger sides (same generator, seed 1, 9 materials; default sphere
# counts and radii scale with the side): 4096^3 holds 11.6M nodes (0.37 GB), 8192^3 57.4M nodes
# (1.84 GB), past the 256 MB MALL, so traversal there reads its nodes from HBM.
def scaled_headline(log2_size):
    return Workload(f"pathtrace_family_{1 << log2_size}", log2_size, 1, {"n_materials": 9}, None, 7, 1 << 24)


# Edit frames (a scaled form of the 16384^3 / 1M-sphere / 10k-edits-per-frame workload, whose
# volume does not fit a 64 GB build host): a 2048^3 volume, then per frame 1000 sphere edits
# (seeded by frame) applied on the host and uploaded as an append-only delta, then 2^24
# incoherent rays traced against the edited volume.
EDIT_2048 = Workload("edit_2048", 11, 21, {"n_materials": 16}, None, 23, 1 << 24)
EDIT_FRAME_EDITS = 1000


# The full-size edit workload: LARGE_16384 (16384^3, 1M spheres, 255 materials; ~800M nodes,
# a ~26 GB store, whose host build needs ~70 GB), per frame 10k sphere edits seeded by frame, then
# 64M incoherent rays.
LARGE_FRAME_EDITS = 10000


def edit_frame_spheres(frame, log2_size=11, n_materials=16, n=EDIT_FRAME_EDITS):
    """(n, 5) edits of one frame: centre U[0, N)^3, radius U[2, 16] * (N / 2048)^(1/3), material
    U{0..n_materials} (0 carves), numpy PCG64(7000 + frame)."""
    import numpy as np

    rng = np.random.Generator(np.random.PCG64(7000 + frame))
    side = float(1 << log2_size)
    e = np.zeros((n, 5))
    e[:, :3] = rng.uniform(0.0, side, (n, 3))
    e[:, 3] = rng.uniform(2.0, 16.0, n) * (side / 2048.0) ** (1.0 / 3.0)
    e[:, 4] = rng.integers(0, n_materials + 1, n)
    return e


def batched_volumes():
    """The 64 volumes of the batched workload: seed 100+i, side 2^(6 + i mod 6)."""
    return [Workload(f"batched_{i}", 6 + i % 6, 100 + i, {}, _camera(6 + i % 6, 256, 256), 1000 + i, 0)
            for i in range(64)]


def batched_sizes(seed=42):
    """64 batch sizes log-uniform in [1e3, 1e6] (PCG-free numpy draw, fixed seed)."""
    import numpy as np

    rng = np.random.Generator(np.random.PCG64(seed))
    return [int(round(10 ** u)) for u in rng.uniform(3.0, 6.0, 64)]


ALL = {w.name: w for w in (RAYCAST_128, PATHTRACE_1024, DIFFRENDER_512, LARGE_16384, EDIT_2048)}
