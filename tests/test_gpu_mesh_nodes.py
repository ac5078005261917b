"""Mesh nodes upstream of set_object against NumPy constructions of the same
faces, and the rendered ids against the oracle raster of those faces:
MeshFlipCulling, MeshNoCulling, MeshFlipNormal, MeshGrid (faces, tex, the
central-difference normals of a height field) and the export helpers."""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu


def _ids(mesh, W=160, H=120, **opts):
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd.util.matrix import lookat, perspective
    scene = tina.Scene(W, H, **opts)
    scene.add_object(mesh)
    scene.set_camera(lookat([0, 0, 0], [0.3, 0.4, 3.0]), perspective(50, W / H, 0.1, 20))
    scene.render()
    torch.cuda.synchronize()
    return scene, scene.engine.ids.to_numpy().T


def test_culling_nodes(gpu):
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    print('case: MeshFlipCulling / MeshNoCulling / MeshFlipNormal', flush=True)
    sphere = tina.PrimitiveMesh.sphere(lons=12, lats=8)
    v = sphere.face_buffers()['verts'].cpu().numpy()
    n = sphere.face_buffers()['norms'].cpu().numpy()
    flip = tina.MeshFlipCulling(sphere)
    assert np.array_equal(flip.face_buffers()['verts'].cpu().numpy(), v[:, ::-1])
    scene, ids = _ids(flip)
    cam = orc.camera(scene.engine.W2V, scene.engine.V2W)
    assert np.array_equal(ids, orc.raster(160, 120, cam, np.ascontiguousarray(v[:, ::-1]), flags=_abi.CULL | _abi.CLIP)[1])
    both = tina.MeshNoCulling(sphere)
    fb = both.face_buffers()
    exp = np.stack([v, v[:, ::-1]], axis=1).reshape(-1, 3, 3)
    assert fb['nfaces'] == 2 * len(v) == both.get_nfaces()
    assert np.array_equal(fb['verts'].cpu().numpy(), exp)
    assert np.array_equal(fb['norms'].cpu().numpy(), np.stack([n, -n[:, ::-1]], axis=1).reshape(-1, 3, 3))
    scene, ids = _ids(both)
    assert np.array_equal(ids, orc.raster(160, 120, cam, exp, flags=_abi.CULL | _abi.CLIP)[1])
    assert np.array_equal(tina.MeshFlipNormal(sphere).face_buffers()['norms'].cpu().numpy(), -n)


def test_mesh_grid(gpu):
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    print('case: MeshGrid 9x9 height field', flush=True)
    g = tina.MeshGrid(9)
    p = g.pos.to_numpy()
    i, j = np.meshgrid(np.arange(9), np.arange(9), indexing='ij')
    assert np.allclose(p[..., 0], i / 8 * 2 - 1) and np.allclose(p[..., 1], j / 8 * 2 - 1) and np.all(p[..., 2] == 0)
    assert np.allclose(g.tex.to_numpy(), np.stack([i / 8, j / 8], axis=2))
    p[..., 2] = 0.3 * np.sin(2 * p[..., 0]) * np.cos(3 * p[..., 1])
    g.pos.from_numpy(p.astype(np.float32))
    g.pre_compute()
    # central differences, clamped at the border (grid.py:26-35), in float64
    P = p.astype(np.float64)
    ip, im = np.minimum(i + 1, 8), np.maximum(i - 1, 0)
    jp, jm = np.minimum(j + 1, 8), np.maximum(j - 1, 0)
    dx, dy = P[ip, j] - P[im, j], P[i, jp] - P[i, jm]
    nn = np.cross(dx, dy)
    nn /= np.linalg.norm(nn, axis=2, keepdims=True)
    assert np.abs(g.nrm.to_numpy() - nn).max() <= 1e-6
    fb = g.face_buffers()
    assert fb['nfaces'] == 2 * 8 * 8 == g.get_nfaces()
    a, b, c, d = p[:-1, :-1], p[1:, :-1], p[1:, 1:], p[:-1, 1:]
    exp = np.stack([np.stack([a, b, c], 2), np.stack([a, c, d], 2)], 2).reshape(-1, 3, 3)
    assert np.array_equal(fb['verts'].cpu().numpy(), exp.astype(np.float32))
    scene, ids = _ids(g, smoothing=True, culling=False)
    cam = orc.camera(scene.engine.W2V, scene.engine.V2W)
    assert (ids >= 0).sum() > 1000
    assert np.array_equal(ids, orc.raster(160, 120, cam, exp.astype(np.float32), flags=_abi.CLIP)[1])
    quads = tina.MeshGrid(5, as_quad=True)
    assert quads.get_npolygon() == 4 and quads.face_buffers()['verts'].shape == (16, 4, 3)
    assert tina.MeshToWire(quads).face_buffers()['verts'].shape == (64, 2, 3)


def test_export_helpers(gpu):
    import taichi_three_amd as tina
    sphere = tina.PrimitiveMesh.sphere(lons=6, lats=4)
    obj = tina.export_simple_mesh(sphere)
    assert np.array_equal(obj['fv'], sphere.face_buffers()['verts'].cpu().numpy())
    conn = tina.simple_mesh_to_connective(obj)
    assert np.array_equal(conn['v'][conn['f']], obj['fv']) and np.array_equal(conn['vn'][conn['f']], obj['fn'])
    cm = tina.ConnectiveMesh(conn['v'], conn['f'])
    back = tina.export_connective_mesh(cm)
    assert np.array_equal(back['v'][back['f']], obj['fv'])
