"""The mesh nodes upstream of set_object as device passes: MeshTransform
(tina/mesh/trans.py:29-40), MeshFlatNormal and MeshSmoothNormal
(tina/mesh/norm.py:5-55) bit for bit against the oracle's serial restatements;
scenes built from them render like the oracle fed the untransformed mesh (the
oracle applies the transform and the normals itself)."""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu


def _trans(seed, projective=False):
    rng = np.random.default_rng(seed)
    M = np.eye(4)
    M[:3, :3] = rng.normal(size=(3, 3)) * 0.4 + np.eye(3)
    M[:3, 3] = rng.uniform(-0.3, 0.3, 3)
    if projective:
        M[3] = [0.03, -0.02, 0.05, 1.1]
    return M


def _tn(M):
    return np.transpose(np.linalg.inv(M))[:3, :3].astype(np.float32)


def _np(t):
    return t.detach().cpu().numpy()


def _random_indexed(seed, nv=3000, nf=6000, unused=7):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1, 1, (nv, 3)).astype(np.float32)
    f = np.stack([rng.choice(nv - unused, 3, replace=False) for _ in range(nf)]).astype(np.int32)
    f[5] = [f[5, 0], f[5, 0], f[5, 1]]           # a degenerate face: its normal is NaN, as in the reference
    return v, f


@pytest.mark.parametrize('projective', [False, True])
def test_mesh_transform_bit_exact(gpu, projective):
    import taichi_three_amd as tina
    rng = np.random.default_rng(5)
    print(f'case: MeshTransform projective={projective}', flush=True)
    verts = rng.uniform(-1, 1, (20000, 3, 3)).astype(np.float32)
    norms = rng.normal(size=(20000, 3, 3)).astype(np.float32)
    m = tina.SimpleMesh()
    m.set_face_verts(verts)
    m.set_face_norms(norms)
    M = _trans(1, projective)
    fb = tina.MeshTransform(m, M).face_buffers()
    ov, on = orc.mesh_transform(M.astype(np.float32), _tn(M), verts, norms)
    assert np.array_equal(_np(fb['verts']), ov)
    assert np.array_equal(_np(fb['norms']), on)


def test_flat_normals_bit_exact(gpu):
    import taichi_three_amd as tina
    rng = np.random.default_rng(6)
    print('case: MeshFlatNormal', flush=True)
    verts = rng.uniform(-1, 1, (30000, 3, 3)).astype(np.float32)
    verts[3] = verts[3, 0]                              # zero area: NaN normal
    m = tina.SimpleMesh()
    m.set_face_verts(verts)
    got = _np(tina.MeshFlatNormal(m).face_buffers()['norms'])
    exp = np.repeat(orc.face_normals(verts)[:, None, :], 3, axis=1)
    assert np.array_equal(got, exp, equal_nan=True)
    assert np.isnan(got[3]).all()


@pytest.mark.parametrize('which', ['icosphere', 'random'])
def test_smooth_normals_bit_exact(gpu, which):
    import taichi_three_amd as tina
    print(f'case: MeshSmoothNormal {which}', flush=True)
    if which == 'icosphere':
        V, F = tina.PrimitiveMesh.icosphere(24)
        v, f = V.astype(np.float32), F.astype(np.int32)
    else:
        v, f = _random_indexed(7)
    conn = tina.ConnectiveMesh(v, f)
    node = tina.MeshSmoothNormal(conn)
    got = _np(node.norm)
    exp = orc.smooth_normals(v[f], f, len(v))
    assert np.array_equal(got, exp, equal_nan=True)
    fb = node.face_buffers()
    assert np.array_equal(_np(fb['norms']), exp[f], equal_nan=True)
    # the cached node keeps its normals until update_normal; a new topology version recomputes them
    v2 = (v * np.float32(1.5)).astype(np.float32)
    conn.set_mesh(v2, f)
    fb2 = node.face_buffers()
    assert np.array_equal(_np(fb2['norms']), orc.smooth_normals(v2[f], f, len(v))[f], equal_nan=True)


@pytest.mark.parametrize('order', ['transform_of_smooth', 'smooth_of_transform'])
def test_scene_through_mesh_nodes_matches_oracle(gpu, order):
    """An icosphere through MeshSmoothNormal and MeshTransform (both orders) in a
    smooth-shaded Scene: ids and depth bit for bit against the oracle's raster of
    the transformed faces, colour within 1e-5 of the binary32 oracle shading with
    the oracle's own normals."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd.matr import lower_material
    print(f'case: scene through mesh nodes, {order}', flush=True)
    V, F = tina.PrimitiveMesh.icosphere(10)
    v, f = V.astype(np.float32), F.astype(np.int32)
    M = _trans(3)
    conn = tina.ConnectiveMesh(v, f)
    if order == 'transform_of_smooth':
        mesh = tina.MeshTransform(tina.MeshSmoothNormal(conn), M)
        nv = orc.smooth_normals(v[f], f, len(v))
        tv, tn = orc.mesh_transform(M.astype(np.float32), _tn(M), v[f], nv[f])
    else:
        mesh = tina.MeshSmoothNormal(tina.MeshTransform(conn, M))
        tv, _ = orc.mesh_transform(M.astype(np.float32), _tn(M), v[f])
        tn = orc.smooth_normals(tv, f, len(v))[f]
    W, H = 320, 240
    scene = tina.Scene(W, H, smoothing=True, maxfaces=len(f))
    scene.engine.set_camera(tina.lookat([0, 0, 0], [0.6, 0.4, 3.2]), tina.perspective(45, W / H, 0.1, 20))
    material = tina.Classic(color=(0.7, 0.5, 0.3), shineness=32)
    scene.add_object(mesh, material)
    scene.render()
    torch.cuda.synchronize()
    ids = scene.engine.ids.to_numpy().T
    depth = scene.engine.depth.to_numpy().T
    img = scene.img.to_numpy().transpose(1, 0, 2)
    cam = orc.camera(scene.engine.W2V, scene.engine.V2W)
    od, oi = orc.raster(W, H, cam, tv)
    assert (oi >= 0).sum() > 10000
    assert np.array_equal(ids, oi)
    assert np.array_equal(depth, od)
    lo = lower_material(material, dict(attrs={}))
    o32 = orc.shade(W, H, cam, orc.HostMesh(tv, tn), lo.struct, scene.lighting.to_struct(), oi,
                    _abi.SMOOTH | _abi.TONEMAP | _abi.BG)
    assert np.abs(img - o32).max() <= 1e-5
