"""OBJ input and output (tina/assimp/obj.py behaviour) against hand-worked
expectations: polygon fans, index forms, material ids, orientation, scaling,
flat normals and a write/read round trip."""
import io

import numpy as np
import pytest

from taichi_three_amd import assimp

SRC = """# a unit square, a pentagon and a triangle with mixed index forms
mtllib scene.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0.5 2 0
vt 0 0
vt 1 0
vt 1 1
vn 0 0 1
vn 0 0 -1
o ignored record
usemtl red
f 1/1/1 2/2/1 3/3/1 4/1/1
usemtl blue
f 1//2 2//2 3//2 5//2 4//2
usemtl red
f 3/2 5/3 4/1
"""


def test_polygons_fan_around_their_first_corner():
    o = assimp.readobj_text(SRC)
    assert o['v'].shape == (5, 3) and o['vt'].shape == (3, 2) and o['vn'].shape == (2, 3)
    fv = o['f'][:, :, 0].tolist()
    # quad (0, 1, 2), (2, 3, 0); pentagon (0, 1, 2), (0, 2, 4), (0, 4, 3); triangle
    assert fv == [[0, 1, 2], [2, 3, 0], [0, 1, 2], [0, 2, 4], [0, 4, 3], [2, 4, 3]]
    # 'v//vn' has no texcoord: it reads as index 0; 'v/vt' has no normal: also 0
    assert o['f'][2, :, 1].tolist() == [0, 0, 0] and o['f'][2, :, 2].tolist() == [1, 1, 1]
    assert o['f'][5, :, 1].tolist() == [1, 2, 0] and o['f'][5, :, 2].tolist() == [0, 0, 0]
    assert o['mtllib'] == b'scene.mtl'
    assert o['usemtl'] == [[0, b'red'], [2, b'blue'], [5, b'red']]
    assert o['f'].dtype == np.int32 and o['v'].dtype == np.float32


def test_quadok_keeps_polygons_and_simple_returns_positions():
    o = assimp.readobj_text('v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n', quadok=True)
    assert o['f'].shape == (1, 4, 3) and o['f'][0, :, 0].tolist() == [0, 1, 2, 3]
    v, f = assimp.readobj_text('v 0 0 0\nv 1 0 0\nv 0 1 0\nf 3 1 2\n', simple=True)
    assert v.shape == (3, 3) and f.tolist() == [[2, 0, 1]]


def test_empty_file_reads_one_zero_row_each():
    o = assimp.readobj(io.BytesIO(b'# nothing\n'))
    assert o['v'].tolist() == [[0, 0, 0]] and o['vt'].tolist() == [[0, 0]]
    assert o['vn'].tolist() == [[0, 0, 0]] and o['f'].shape == (1, 3, 3) and not o['f'].any()


def test_material_ids_and_parts():
    o = assimp.readobj_text(SRC)
    assert assimp.objmtlids(o).tolist() == [1, 1, 2, 2, 2, 1]
    parts = assimp.objunpackmtls(o)
    assert sorted(parts) == [b'blue', b'red']
    assert parts[b'red']['f'][:, :, 0].tolist() == [[0, 1, 2], [2, 3, 0], [2, 4, 3]]
    assert len(parts[b'blue']['f']) == 3


def test_orientation_swizzles_negates_and_reverses():
    base = assimp.readobj_text(SRC)
    o = assimp.readobj_text(SRC, orient='xzy')
    assert np.array_equal(o['v'], base['v'][:, [0, 2, 1]]) and np.array_equal(o['vn'], base['vn'][:, [0, 2, 1]])
    o = assimp.readobj_text(SRC, orient='xYz')
    assert np.array_equal(o['v'][:, 1], -base['v'][:, 1]) and np.array_equal(o['v'][:, 0], base['v'][:, 0])
    o = assimp.readobj_text(SRC, orient='-xyz')
    assert np.array_equal(o['f'], base['f'][:, ::-1, :])


def test_autoscale_and_scale():
    o = assimp.readobj_text(SRC, scale='auto')
    assert np.allclose(o['v'].mean(axis=0), 0, atol=1e-7) and np.isclose(np.abs(o['v']).max(), 1.0)
    o = assimp.readobj_text(SRC, scale=3.0)
    assert np.array_equal(o['v'], assimp.readobj_text(SRC)['v'] * 3)


def test_flat_normals_follow_the_reference_cross_order():
    o = assimp.readobj_text('v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n')
    assimp.objmknorm(o)
    # (c - a) x (b - a) for a counter-clockwise triangle in the xy plane points to -z
    assert np.allclose(o['vn'], [[0, 0, -1]]) and o['f'][0, :, 2].tolist() == [0, 0, 0]


def test_write_read_round_trip_is_exact():
    rng = np.random.default_rng(0)
    obj = {'v': rng.normal(size=(40, 3)).astype(np.float32), 'vt': rng.random((40, 2)).astype(np.float32),
           'vn': rng.normal(size=(40, 3)).astype(np.float32), 'f': rng.integers(0, 40, (25, 3, 3)).astype(np.int32)}
    buf = io.StringIO()
    assimp.writeobj(buf, obj)
    back = assimp.readobj_text(buf.getvalue())
    for k in ('v', 'vt', 'vn', 'f'):
        assert np.array_equal(back[k], obj[k]), k
    buf = io.StringIO()
    assimp.writeobj(buf, {'v': obj['v'], 'f': obj['f'][:, :, 0]})
    assert np.array_equal(assimp.readobj_text(buf.getvalue())['f'][:, :, 0], obj['f'][:, :, 0])


def test_degenerate_faces_raise():
    with pytest.raises(ValueError):
        assimp.readobj_text('v 0 0 0\nv 1 0 0\nf 1 2\n')
