"""Pins the differentiable-shading oracle (oracle/diffshade.py): with roughness
and metallic constant over the mesh it must equal the C oracle's binary64
shade of the PBR preset under the same point lights, per-vertex colours and
smooth normals."""
import numpy as np
import pytest

from oracle import diffshade as ds
from oracle import orc


def diff_case(freq=7, res=64, seed=0, const_rm=None):
    """A small icosphere scene: (W, H, W2V, V2W, verts, faces, params, lights, amb)."""
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    from taichi_three_amd.mesh import PrimitiveMesh
    rng = np.random.default_rng(seed)
    verts, faces = PrimitiveMesh.icosphere(freq)
    nv = len(verts)
    view, proj = wl.default_camera(1.0)
    W2V, V2W = tina.camera_matrices(view, proj)
    normals = verts + 0.1 * rng.normal(size=verts.shape)
    colors = rng.uniform(0.1, 0.9, size=(nv, 3))
    if const_rm is None:
        rough, metal = rng.uniform(0.2, 0.9, nv), rng.uniform(0, 1, nv)
    else:
        rough, metal = np.full(nv, const_rm[0]), np.full(nv, const_rm[1])
    lp = rng.normal(size=(4, 3))
    lp = 4.0 * lp / np.linalg.norm(lp, axis=1, keepdims=True)
    lights = [(tuple(p), (1.0, 0.9, 0.8)) for p in lp]
    inten = rng.uniform(0.5, 2.0, 4)
    amb = (0.1, 0.1, 0.1)
    p = dict(normals=normals.astype(np.float32), colors=colors.astype(np.float32),
             rough=rough.astype(np.float32), metal=metal.astype(np.float32), intens=inten.astype(np.float32))
    return res, res, W2V, V2W, verts.astype(np.float32), faces, p, lights, amb


@pytest.mark.parametrize('r,m', [(0.3, 0.0), (0.7, 1.0), (0.5, 0.25)])
def test_diffshade_oracle_equals_the_pbr_oracle(r, m):
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd.matr import lower_material
    W, H, W2V, V2W, verts, faces, p, lights, amb = diff_case(const_rm=(r, m))
    cam = orc.camera(W2V, V2W)
    soup = verts[faces]
    _, ids = orc.raster(W, H, cam, soup)
    assert (ids >= 0).sum() > 500
    mine = ds.render(W, H, W2V, V2W, verts, faces, ids, p['normals'], p['colors'], p['rough'], p['metal'],
                     lights, p['intens'], amb)
    lo = lower_material(tina.PBR(basecolor='color', metallic=float(np.float32(m)), roughness=float(np.float32(r))))
    lt = _abi.Lights()
    lt.n = len(lights)
    for i, ((lp, lc), I) in enumerate(zip(lights, p['intens'])):
        for k in range(3):
            lt.dirs[i][k] = lp[k]
            lt.colors[i][k] = float(np.float32(I) * np.float32(lc[k]))
        lt.dirs[i][3] = 1.0
    for k in range(3):
        lt.ambient[k] = amb[k]
    mesh = orc.HostMesh(soup, p['normals'][faces], None, p['colors'][faces])
    ref = orc.shade(W, H, cam, mesh, lo.struct, lt, ids, _abi.SMOOTH | _abi.VCOLOR, f64=True)
    err = float(np.abs(mine - ref).max())
    assert err <= 1e-5 * max(1.0, float(np.abs(ref).max())), err
