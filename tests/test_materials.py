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
