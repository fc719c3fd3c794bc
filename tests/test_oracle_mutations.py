"""Planted slips in the oracle, on the CPU: each build below changes one line of the oracle's C
sources (a dropped term, a wrong sign, constant, comparison or index: mistakes a restatement could
plausibly carry) and must fail at least one of the pinning checks of tests/pt_pins.py. A control
build of the unchanged sources, made the same way, passes all of them. The mutants are built from
copies in a temporary directory; the oracle's own files are never changed.
"""
import concurrent.futures
import os
import shutil
import subprocess

import numpy as np
import pytest

from cubiquity_amd import workloads as W
from tests import pt_pins

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
CFLAGS = ["-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-pthread", "-shared"]

# name -> (file, original text, planted text, which checks apply)
PT = "pt"
TRAV = "traversal"
MUTANTS = {
    # path tracer (cpu_pt.c)
    "cosine_radius_u1": ("cpu_pt.c", "float r = sqrtf(u1), c, s;", "float r = u1, c, s;", PT),
    "cosine_height_swapped": ("cpu_pt.c", "lz = sqrtf(1.0f - u1);", "lz = sqrtf(u1);", PT),
    "dropped_albedo": ("cpu_pt.c", "b[a] = b[a] * c->albedo[m * 3 + a];", "b[a] = b[a];", PT),
    "one_bounce_short": ("cpu_pt.c", "if (depth >= p->bounces) break;", "if (depth + 1 >= p->bounces) break;", PT),
    "emission_unweighted": ("cpu_pt.c", "out->L[a] += b[a] * c->emission[m * 3 + a];",
                            "out->L[a] += c->emission[m * 3 + a];", PT),
    "sky_unweighted": ("cpu_pt.c", "out->L[a] += b[a] * p->sky[a];", "out->L[a] += p->sky[a];", PT),
    "pcg_multiplier": ("cpu_pt.c", "6364136223846793005ULL", "6364136223846793007ULL", PT),
    "pcg_rotation": ("cpu_pt.c", "uint32_t rot = (uint32_t)(old >> 59u);", "uint32_t rot = (uint32_t)(old >> 58u);", PT),
    "pcg_stream_shift": ("cpu_pt.c", "pcg_t r = { 0, (initseq << 1u) | 1u };", "pcg_t r = { 0, initseq | 1u };", PT),
    "uniform_scale": ("cpu_pt.c", "(1.0f / 16777216.0f)", "(1.0f / 16777215.0f)", PT),
    "sincos_coefficient": ("cpu_pt.c", "8.33333333e-3f", "8.33333333e-4f", PT),
    "sincos_quadrant": ("cpu_pt.c", "case 1: *c = -sp; *s = cp; break;", "case 1: *c = sp; *s = cp; break;", PT),
    "duff_basis_sign": ("cpu_pt.c", "float a = -1.0f / (sg + nz);", "float a = 1.0f / (sg + nz);", PT),
    "duff_basis_index": ("cpu_pt.c", "float b0 = b, b1 = sg + ny * ny * a, b2 = -ny;",
                         "float b0 = b, b1 = sg + nx * ny * a, b2 = -ny;", PT),
    "bounce_offset_sign": ("cpu_pt.c", "ray[k] = pos[k] + n[k] * p->bounce_offset;",
                           "ray[k] = pos[k] - n[k] * p->bounce_offset;", PT),
    "camera_jitter_swapped": ("cpu_pt.c", "float jx = pcg_u(rng), jy = pcg_u(rng);",
                              "float jy = pcg_u(rng), jx = pcg_u(rng);", PT),
    "camera_y_not_inverted": ("cpu_pt.c", "- c->up[a] * cy;", "+ c->up[a] * cy;", PT),
    "camera_pixel_centre": ("cpu_pt.c", "float cx = (((float)x + jx) - c->half_w) * c->kx;",
                            "float cx = (((float)x + 0.5f) - c->half_w) * c->kx;", PT),
    # traversal (cpu_ref.c)
    "advance_strict": ("cpu_ref.c", "if (ct1[a] <= tcx) flips |= 1u << a;", "if (ct1[a] < tcx) flips |= 1u << a;", TRAV),
    "first_child_nonstrict": ("cpu_ref.c", "if (tm < entry) id |= 1u << a;", "if (tm <= entry) id |= 1u << a;", TRAV),
    "normal_sign": ("cpu_ref.c", "float n = (float)(tce == ct0[a]) * -sign[a];",
                    "float n = (float)(tce == ct0[a]) * sign[a];", TRAV),
    "origin_shift": ("cpu_ref.c", "o[a] = (r[a] + 0.5f) * sign[a];", "o[a] = (r[a] - 0.5f) * sign[a];", TRAV),
    "octant_tie": ("cpu_ref.c", "if (dist[1] <= m) { step += 2;", "if (dist[1] < m) { step += 2;", TRAV),
    "nearest_material_order": ("cpu_ref.c", "(0x76534210u | 0x88888888u)", "(0x76543210u | 0x88888888u)", TRAV),
    "max_operand_order": ("cpu_ref.c", "static inline float smax(float a, float b) { return (a < b) ? b : a; }",
                          "static inline float smax(float a, float b) { return (b < a) ? a : b; }", TRAV),
    "min_operand_order": ("cpu_ref.c", "static inline float smin(float a, float b) { return (b < a) ? b : a; }",
                          "static inline float smin(float a, float b) { return (a < b) ? a : b; }", TRAV),
    "pop_height": ("cpu_ref.c", "height = msb + 1;", "height = msb + 2;", TRAV),
}


def _build(tmp, name, mutation):
    d = os.path.join(tmp, name)
    os.makedirs(d)
    for f in os.listdir(ORACLE_DIR):
        if f.endswith(".c") or f.endswith(".h"):
            shutil.copy(os.path.join(ORACLE_DIR, f), d)
    if mutation is not None:
        fname, old, new, _ = mutation
        path = os.path.join(d, fname)
        src = open(path).read()
        assert src.count(old) == 1, f"{name}: the line to mutate is not in {fname} exactly once"
        open(path, "w").write(src.replace(old, new))
    srcs = sorted(os.path.join(d, f) for f in os.listdir(d) if f.startswith("cpu_") and f.endswith(".c"))
    lib = os.path.join(d, "liboracle.so")
    subprocess.run(["gcc"] + CFLAGS + ["-o", lib] + srcs + ["-lm"], check=True, capture_output=True)
    return lib


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("mutants"))
    jobs = {"control": None, **MUTANTS}
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        futs = {k: ex.submit(_build, tmp, k, m) for k, m in jobs.items()}
        return {k: f.result() for k, f in futs.items()}


@pytest.fixture(scope="module")
def golden():
    return np.load(pt_pins.GOLDEN, allow_pickle=False), np.load(pt_pins.LATTICE, allow_pickle=False)


def _checks(kind, orc, cq, vol128, golden):
    if kind == PT:
        return [
            ("pcg32", lambda: pt_pins.check_pcg32(orc)),
            ("sincos", lambda: pt_pins.check_sincos(orc)),
            ("bounce", lambda: pt_pins.check_bounce(orc, count=1 << 15)),
            ("camera", lambda: pt_pins.check_camera(orc, cq, W)),
            ("closed_forms", lambda: pt_pins.check_closed_forms(orc, cq, W)),
        ]
    return [("traversal_golden", lambda: pt_pins.check_traversal_golden(orc, vol128, *golden))]


def test_control_build_passes_every_check(builds, cq, vol128, golden):
    from oracle.bindings import Oracle

    orc = Oracle(builds["control"])
    for kind in (PT, TRAV):
        for name, fn in _checks(kind, orc, cq, vol128, golden):
            fn()


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_slip_is_caught(builds, cq, vol128, golden, name):
    from oracle.bindings import Oracle

    orc = Oracle(builds[name])
    caught = []
    for check, fn in _checks(MUTANTS[name][3], orc, cq, vol128, golden):
        try:
            fn()
        except AssertionError:
            caught.append(check)
    assert caught, f"planted slip {name} passed every check"
