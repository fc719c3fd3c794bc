"""Material metadata (.toml next to a .dag): the reader against an independent TOML parser
(tomli) on the files the reference's Metadata::save writes and on hand-edited variants, the
defaults of Material (0.8 grey base colour) and of the viewer's colour table (purple), and errors."""
import numpy as np
import pytest

tomli = pytest.importorskip("tomli")


def _expect_from_toml(text):
    """Materials as tomli reads them, with Material's defaults applied."""
    doc = tomli.loads(text)
    out = []
    for t in doc.get("materials", []):
        out.append((t.get("name"), np.float32(t.get("base_color", [0.8, 0.8, 0.8])),
                    np.float32(t.get("emission", [0.0, 0.0, 0.0]))))
    return out, doc.get("dimensions")


def _check_against_tomli(cq, path):
    m = cq.Materials.load(path)
    want, dims = _expect_from_toml(open(path, encoding="utf-8").read())
    assert len(m.rows) == len(want)
    for row, (name, bc, em) in zip(m.rows, want):
        assert (row["name"].decode() if row["flags"] & 1 else None) == name
        assert np.array_equal(row["base_color"], bc)
        assert np.array_equal(row["emission"], em)
    assert m.dimensions == (tuple(dims) if dims is not None else None)
    return m


def test_round_trip_reference_layout(cq, tmp_path):
    rng = np.random.Generator(np.random.PCG64(4))
    alb = rng.uniform(0, 1, (255, 3)).astype(np.float32)
    em = np.zeros((255, 3), np.float32)
    em[-3:] = 5.0
    m = cq.Materials.from_tables(alb, em, names=[f"m{i}" for i in range(255)])
    p = tmp_path / "v.toml"
    m.save(p)
    text = p.read_text()
    assert "[[materials]] # Material index 255" in text and "base_color = [" in text
    back = _check_against_tomli(cq, p)
    assert len(back.rows) == 256
    # four-decimal colours, as Metadata::save writes them
    assert np.allclose(back.rows["base_color"][1:], alb, atol=5.1e-5)
    a2, e2 = back.pathtrace_tables()
    assert a2.shape == (255, 3) and np.array_equal(e2[-1], [5.0, 5.0, 5.0])


HAND_EDITED = r'''# header comment
title = "ignored"   # top-level key the reader does not use
dimensions = [ 128, 64,
               32 ]  # several lines

[other.table]
base_color = [9, 9, 9]  # not a material

[[materials]]
name = 'EmptySpace'   # literal string, no base_color: Material::DefaultColor

[[ materials ]]
"name" = "Rock \"granite\"\tgrey"
base_color = [0.5, 1, 2.5e-1,]   # trailing comma, integer element
emission = [
  0.0,  # red
  1_000.0,
  +0.25e+1,
]

[[materials]]
base_color = [0.1, 0.2, 0.3]
extra = { ignored = true }
'''


def test_hand_edited_file(cq, tmp_path):
    p = tmp_path / "h.toml"
    p.write_text(HAND_EDITED)
    m = _check_against_tomli(cq, p)
    assert m.dimensions == (128, 64, 32)
    assert len(m.rows) == 3
    assert m.rows[1]["name"] == b'Rock "granite"\tgrey'
    assert np.array_equal(m.rows[0]["base_color"], np.float32([0.8, 0.8, 0.8]))
    assert m.rows[2]["flags"] == 2  # base_color only


def test_colour_table_defaults(cq, tmp_path):
    p = tmp_path / "c.toml"
    p.write_text("[[materials]]\nname = \"EmptySpace\"\n\n[[materials]]\nbase_color = [0.25, 0.5, 0.75]\n")
    t = cq.Materials.load(p).colour_table()
    assert t.shape == (256, 3)
    assert np.array_equal(t[0], np.float32([0.8, 0.8, 0.8])) and np.array_equal(t[1], [0.25, 0.5, 0.75])
    assert np.array_equal(t[2:], np.tile(np.float32([1, 0, 1]), (254, 1)))


@pytest.mark.parametrize("text,line", [
    ("[[materials]]\nbase_color = [0.1, 0.2]\n", 2),
    ("[[materials]]\nbase_color = [0.1, 0.2, 0.3\n", 3),
    ("[[materials]\nname = \"x\"\n", 1),
    ("[[materials]]\nname = 5\n", 2),
    ("[[materials]]\nname = \"unterminated\n", 2),
    ("dimensions = [1.5, 2, 3]\n", 1),
    ("[[materials]]\nbase_color = [0.1, 0.2, 0.3] junk\n", 2),
    ("[[materials]]\nname = \"" + "x" * 70 + "\"\n", 2),
    ("[[materials]]\n" * 257, 257),
])
def test_malformed_files_fail_with_the_line(cq, tmp_path, text, line):
    p = tmp_path / "bad.toml"
    p.write_text(text)
    with pytest.raises(cq.CqxError) as e:
        cq.Materials.load(p)
    assert e.value.status == -8
    assert f":{line}:" in str(e.value)


def test_missing_file_and_bad_names(cq, tmp_path):
    with pytest.raises(cq.CqxError) as e:
        cq.Materials.load(tmp_path / "none.toml")
    assert e.value.status == -5
    m = cq.Materials.from_tables(np.zeros((1, 3)), names=['bad"name'])
    with pytest.raises(cq.CqxError):
        m.save(tmp_path / "x.toml")


def test_random_materials_round_trip(cq, tmp_path):
    rng = np.random.Generator(np.random.PCG64(9))
    for k in range(20):
        n = int(rng.integers(0, 40))
        rows = np.zeros(n, cq.MATERIAL_DTYPE)
        for i in range(n):
            f = int(rng.integers(0, 8))
            rows[i]["flags"] = f
            if f & 1:
                rows[i]["name"] = "".join(chr(int(c)) for c in rng.integers(32, 127, int(rng.integers(0, 20)))
                                          if chr(int(c)) not in '"\\').encode()
            rows[i]["base_color"] = rng.uniform(-10, 10, 3)
            rows[i]["emission"] = rng.uniform(0, 100, 3)
        p = tmp_path / f"r{k}.toml"
        cq.Materials(rows).save(p)
        _check_against_tomli(cq, p)


TOML_SCENE = """# materials for the 128^3 raycast volume (ids 0..3), in Metadata::save's layout, hand-edited
dimensions = [128, 128, 128]

[[materials]] # Material index 0
name = "EmptySpace"

[[materials]] # Material index 1
name = "Ground"
base_color = [0.4000, 0.6500, 0.2500]

[[materials]] # Material index 2
name = "Rock"
base_color = [0.5500, 0.5000, 0.4500]
emission = [0.0, 0.0, 0.0]

[[materials]] # Material index 3
name = "Lamp"
base_color = [0.9000, 0.8500, 0.3000]
emission = [4.0, 3.5, 1.25]
"""


@pytest.mark.gpu
def test_gpu_renders_with_tables_from_toml(cq, oracle, gpu_ctx, vol128, tmp_path):
    """A .toml written by hand is read (cqx_materials_load), turned into the viewer's colour table
    and the path tracer's albedo/emission tables, and both GPU renderers use them: the
    reference-shading frame and the path-traced image equal the oracle's with the same tables."""
    import torch

    from cubiquity_amd import workloads as W

    p = tmp_path / "scene.toml"
    p.write_text(TOML_SCENE)
    m = _check_against_tomli(cq, p)
    col = m.colour_table()
    alb, em = m.pathtrace_tables()
    assert alb.shape == (3, 3) and np.array_equal(em[2], np.float32([4.0, 3.5, 1.25]))
    assert np.array_equal(col[3], np.float32([0.9, 0.85, 0.3]))
    dv = gpu_ctx.upload(vol128)
    sub = oracle.find_subdags(vol128.nodes, vol128.root)
    c = W.RAYCAST_128.camera
    rays = cq.camera_rays(c.eye, c.target, c.fov_y, 96, 72)
    g = cq.demo_trace(gpu_ctx, dv, torch.from_numpy(rays).cuda(), col, gamma=1.0, seed=3).cpu().numpy()
    o = oracle.demo_trace(vol128.nodes, sub, rays, col, gamma=1.0, seed=3)
    assert np.array_equal(g, o)
    s = cq.PathTraceSetup(c, 80, 60, 4, 3, alb, em)
    for mode in (0, 1):
        img = cq.pathtrace(gpu_ctx, dv, s, mode).cpu().numpy()
        assert np.array_equal(img, oracle.pathtrace(vol128.nodes, vol128.root, s.params)), mode
    torch.cuda.synchronize()
    dv.free()
