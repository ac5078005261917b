"""Scenes loaded with Scene.load_gltf render like the oracle: one MeshModel +
MeshTransform object per primitive with its PBR material (a base-colour texture
in cornell.gltf), ids and depth bit-exact, colour within 1e-5 of the binary32
oracle."""
import os

import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu

ASSETS = os.path.join(os.path.dirname(__file__), 'golden', 'assets')
COLOR_TOL_F32 = 1e-5


def _device_tensors(node, out, seen):
    """Every device tensor (Param values, textures) reachable from a material tree."""
    if id(node) in seen:
        return
    seen.add(id(node))
    t = getattr(node, 'tensor', None)
    if t is not None and hasattr(t, 'data_ptr'):
        out[t.data_ptr()] = t
    vals = node.values() if isinstance(node, dict) else node if isinstance(node, (list, tuple)) else \
        vars(node).values() if hasattr(node, '__dict__') else ()
    for v in vals:
        if isinstance(v, (dict, list, tuple)) or hasattr(v, '__dict__'):
            _device_tensors(v, out, seen)


@pytest.mark.parametrize('name', ['sphere', 'cornell', 'caustics'])
def test_gltf_scene_matches_oracle(gpu, name):
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd.matr import lower_material
    from taichi_three_amd.util.matrix import lookat, perspective
    W, H = 480, 360
    print(f'case: gltf {name}', flush=True)
    scene = tina.Scene(W, H, smoothing=True, texturing=True)
    scene.load_gltf(os.path.join(ASSETS, f'{name}.gltf'))
    scene.set_camera(lookat([0, 1, 0], [1.5, 3.0, 9.0]), perspective(40, W / H, 0.5, 60))
    scene.render()
    torch.cuda.synchronize()
    ids = scene.engine.ids.to_numpy().T
    depth = scene.engine.depth.to_numpy().T
    img = scene.img.to_numpy().transpose(1, 0, 2)

    eng = scene.engine
    cam = orc.camera(eng.W2V, eng.V2W)
    od = np.full((H, W), 2 ** 30, np.int32)
    oi = np.full((H, W), -1, np.int32)
    base = 0
    parts = []
    for obj, info in scene.objects.items():
        fb = obj.face_buffers()
        host = {k: (None if fb[k] is None else fb[k].cpu().numpy()) for k in ('verts', 'norms', 'coors', 'colors')}
        od, oi = orc.raster(W, H, cam, host['verts'], id_base=base, flags=_abi.CULL | _abi.CLIP, depth=od, occup=oi)
        parts.append((base, host, info.material, fb))
        base += len(host['verts'])
    assert np.array_equal(ids, oi), f'{(ids != oi).sum()} id mismatches'
    assert np.array_equal(depth, od)
    assert (oi >= 0).sum() > 0.02 * W * H
    lt = scene.lighting.to_struct()
    exp = np.zeros((H, W, 3), np.float32)
    for k, (b, host, material, fb) in enumerate(parts):
        lo = lower_material(material, fb)
        tensors, keep = {}, []
        _device_tensors(material, tensors, set())
        for i in range(lo.struct.nparams):
            p = lo.struct.params[i]
            if p.kind in (_abi.P_PARAM, _abi.P_TEXTURE, _abi.P_FACE) and p.ptr in tensors:
                h = np.ascontiguousarray(tensors[p.ptr].cpu().numpy())
                keep.append(h)
                p.ptr = h.ctypes.data
        flags = _abi.SMOOTH | _abi.TEX | _abi.TONEMAP | (_abi.BG if k == 0 else 0)
        flags |= _abi.VCOLOR if host['colors'] is not None else 0
        mesh = orc.HostMesh(host['verts'], host['norms'], host['coors'], host['colors'], id_base=b)
        o32 = orc.shade(W, H, cam, mesh, lo.struct, lt, oi, flags)
        mine = ((oi >= b) & (oi < b + len(host['verts'])))[..., None]
        exp = np.where(mine | ((oi < 0)[..., None] & (k == 0)), o32, exp)
    assert np.abs(img - exp).max() <= COLOR_TOL_F32
