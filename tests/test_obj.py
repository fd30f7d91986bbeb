"""Wavefront OBJ reader (x3d_obj_load, the reference's wavefront_obj::load_geometry,
resource.cpp:137-255) on the host — no GPU needed.

- The reference's Cornell box files (testdata/cornellbox, the meshes tst_renderer.cpp loads) give,
  triangle by triangle, the same corner positions and normals as the fixture meshes the
  reference's BVH fixtures were built from, bit for bit.
- Every vertex keeps its file position and order; positions are the decimal values of the file
  rounded once to float32.
- The reference's pattern rules: extra fields on v / vt lines, lines with leading blanks, faces
  with fewer than three corners, corner field counts, 1-based and out-of-range indices, the last
  corner naming a vertex setting its normal, and the errors (named by line).
"""

import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PARTS = ["left_wall", "right_wall", "back_wall", "ceiling", "floor", "left_sphere", "right_sphere", "light"]


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("part", PARTS)
def test_cornell_meshes_match_fixture(x3d, part):
    v, n, uv, t = x3d.load_obj(os.path.join(GOLD, "cornellbox", part + ".obj"))
    z = np.load(os.path.join(GOLD, "cornell_meshes.npz"))
    gv, gn, gt = z["verts_" + part], z["normals_" + part], z["tris_" + part]
    assert t.shape == gt.shape
    # per triangle and corner: identical positions and normals (vertex numbering may differ)
    assert np.array_equal(_bits(v[t]), _bits(gv[gt]))
    assert np.array_equal(_bits(n[t]), _bits(gn[gt]))
    # positions: the file's "v" lines in order, each decimal rounded once to float32
    lines = [ln.split() for ln in open(os.path.join(GOLD, "cornellbox", part + ".obj"))]
    want = np.array([[np.float32(x) for x in p[1:4]] for p in lines if p and p[0] == "v"], np.float32)
    assert np.array_equal(_bits(v), _bits(want))
    assert uv.shape == (v.shape[0], 2)


def _write(tmp_path, text, name="m.obj"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_pattern_rules(x3d, tmp_path):
    text = "\n".join([
        "# comment",
        "v 0 0 0",
        "v 1 0 0 7",          # four fields: the greedy groups read x from the first, y z from the last two
        "v 0 1 0",
        "  v 9 9 9",          # leading blank: not a vertex line
        "v 0 0",              # two fields: no match, skipped
        "vt 0.5 0.25 0.75",   # u from the first field, v from the last
        "vt 1 1",
        "vn 0 0 1",
        "vn 1 0 0",
        "f 1/1/1 2/1/1 3/2/1",
        "f 1/1/2 3/2/2",      # two corners: skipped
        "f 3//2 2//2 1//2",   # no texture index; vertex 1, 2, 3 normals now (1, 0, 0)
        "o ignored", "s 1", "usemtl x",
    ]) + "\n"
    v, n, uv, t = x3d.load_obj(_write(tmp_path, text))
    assert np.array_equal(v, np.array([[0, 0, 0], [1, 0, 7], [0, 1, 0]], np.float32))
    assert np.array_equal(t, np.array([[0, 1, 2], [2, 1, 0]], np.uint32))
    assert np.array_equal(n, np.array([[1, 0, 0]] * 3, np.float32))  # last corner naming each vertex wins
    assert np.array_equal(uv, np.array([[0.5, 0.75], [0.5, 0.75], [1, 1]], np.float32))


@pytest.mark.parametrize("bad, what", [
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1/1/1 2/1/1 3/1/1 4/1/1\n", "line 5"),   # quad
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n", "line 4"),                             # one field per corner
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1// 2// 3//\n", "line 4"),                       # trailing empty field dropped
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf //1 2// 3//\n", "line 4"),                       # vertex index missing
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1// 2// 4//\n", "line 4"),                       # vertex out of range
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0// 1// 2//\n", "line 4"),                       # 0 is not a 1-based index
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1//1 2// 3//\n", "line 4"),                      # normal out of range
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/1/ 2// 3//\n", "line 4"),                      # texcoord out of range
    ("v a 0 0\n", "line 1"),                                                        # not a number
    ("# no vertices\n", "no vertex"),
])
def test_errors(x3d, tmp_path, bad, what):
    with pytest.raises(x3d.X3DError) as e:
        x3d.load_obj(_write(tmp_path, bad))
    assert what in str(e.value)


def test_missing_file(x3d, tmp_path):
    with pytest.raises(x3d.X3DError):
        x3d.load_obj(str(tmp_path / "absent.obj"))
