"""An OBJ mesh (tina.readobj + tina.MeshModel) renders exactly like the same
faces given as a face soup, and its ids and depths equal the oracle's."""
import io

import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu


def test_obj_mesh_renders_like_its_face_soup(gpu):
    import torch
    import taichi_three_amd as tina
    print('case: obj icosphere(12) round trip', flush=True)
    V, F = tina.PrimitiveMesh.icosphere(12)
    V = np.asarray(V, np.float32)
    N = (V / np.linalg.norm(V, axis=1, keepdims=True)).astype(np.float32)
    F = np.asarray(F, np.int64)
    buf = io.StringIO()
    tina.writeobj(buf, {'v': V, 'vn': N, 'vt': np.zeros((1, 2), np.float32),
                        'f': np.stack([F, np.zeros_like(F), F], axis=2)})
    obj = tina.readobj(io.BytesIO(buf.getvalue().encode()))
    W, H = 320, 240
    imgs, ids = [], []
    for mesh in (tina.MeshModel(obj), None):
        scene = tina.Scene((W, H), smoothing=True)
        if mesh is None:
            mesh = tina.SimpleMesh(maxfaces=len(F))
            mesh.set_face_verts(V[F])
            mesh.set_face_norms(N[F])
        scene.add_object(mesh, tina.Classic())
        scene.render()
        torch.cuda.synchronize()
        imgs.append(scene.img.to_numpy().copy())
        ids.append(scene.engine.ids.to_numpy().T.copy())
        dep = scene.engine.depth.to_numpy().T.copy()
        cam = orc.camera(scene.engine.W2V, scene.engine.V2W)
    assert (ids[0] >= 0).sum() > 1000
    assert np.array_equal(ids[0], ids[1]) and np.array_equal(imgs[0], imgs[1])
    od, oi = orc.raster(W, H, cam, V[F], flags=3)
    assert np.array_equal(ids[1], oi) and np.array_equal(dep, od)
