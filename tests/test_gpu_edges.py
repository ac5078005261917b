"""Edge cases of the HIP path against the oracle: degenerate screens (1 x 1,
one row, one column), a screen too large for the binned path's LDS tables (the
global-atomic path), empty scenes and meshes with no faces, with and without
image-based lighting, and a batch holding an empty scene."""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu


def _raster_vs_oracle(W, H, verts, culling=True):
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    view, proj = wl.default_camera(W / H)
    W2V, V2W = tina.camera_matrices(view, proj)
    eng = tina.Engine(W, H, maxfaces=max(1, len(verts)))
    eng.set_camera(view, proj)
    r = tina.TriangleRaster(eng, maxfaces=max(1, len(verts)), culling=culling)
    m = tina.SimpleMesh(maxfaces=max(1, len(verts)))
    m.set_face_verts(verts)
    eng.clear_depth()
    r.set_object(m)
    r.render_occup()
    ids, depth = eng.resolve()
    torch.cuda.synchronize()
    od, oi = orc.raster(W, H, orc.camera(W2V, V2W), verts, flags=(1 if culling else 0) | 2)
    return ids.cpu().numpy(), depth.cpu().numpy(), oi, od


@pytest.mark.parametrize('W,H', [(1, 1), (1, 700), (700, 1), (17, 3), (3840, 2160), (4000, 2600)])
def test_degenerate_and_oversized_screens(gpu, W, H):
    """(3840, 2160) has 32,400 tiles, which the binned path holds only with the
    160 KB LDS of a gfx950 workgroup; (4000, 2600) has 40,750, more than its LDS
    tables hold, so it runs the global-atomic path."""
    from taichi_three_amd import workloads as wl
    rng = np.random.default_rng(W * 7 + H)
    verts = wl.triangle_soup(rng, 20000, np.log(0.05), 0.8)
    print(f'case: screen {W}x{H}', flush=True)
    for culling in (True, False):
        ids, depth, oi, od = _raster_vs_oracle(W, H, verts, culling)
        assert np.array_equal(ids, oi), (W, H, culling, int((ids != oi).sum()))
        assert np.array_equal(depth, od), (W, H, culling)


def test_empty_scene_and_empty_mesh(gpu):
    """A Scene with no objects paints the background; a mesh with no faces adds
    nothing; ids are all -1 and depth 2**30."""
    import torch
    import taichi_three_amd as tina
    print('case: empty scene', flush=True)
    scene = tina.Scene(64, 48, bgcolor=(0.2, 0.3, 0.4), tonemap=False)
    scene.render()
    torch.cuda.synchronize()
    img = scene.img.to_numpy()
    assert np.allclose(img, np.array([0.2, 0.3, 0.4], np.float32)[None, None], atol=0)
    m = tina.SimpleMesh(maxfaces=4)
    m.set_face_verts(np.zeros((0, 3, 3), np.float32))
    scene.add_object(m)
    scene.render()
    torch.cuda.synchronize()
    assert (scene.engine.ids.to_numpy() == -1).all()
    assert (scene.engine.depth.to_numpy() == 2 ** 30).all()
    assert np.allclose(scene.img.to_numpy(), np.array([0.2, 0.3, 0.4], np.float32)[None, None], atol=0)


def test_empty_ibl_scene_is_the_sky(gpu):
    """With image-based lighting and nothing drawn, every pixel is the
    tone-mapped environment along its view ray (Engine.render_background,
    tina/core/engine.py:41-49), as the binary32 oracle computes it."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd import workloads as wl
    from taichi_three_amd.skybox import make_ibl_struct
    print('case: empty ibl scene', flush=True)
    sky = tina.Skybox.procedural(128, 64)
    scene = tina.Scene(96, 64, ibl=True, skybox=sky, ibl_precision=1)
    view, proj = wl.default_camera(96 / 64)
    scene.set_camera(view, proj)
    scene.render()
    torch.cuda.synchronize()
    img = scene.img.to_numpy().transpose(1, 0, 2)
    W2V, V2W = tina.camera_matrices(view, proj)
    ibl, keep = scene.lighting.ibltab.host_copy()
    ids = -np.ones((64, 96), np.int32)
    mat = tina.matr.lower_material(tina.Diffuse())
    ref = orc.shade(96, 64, orc.camera(W2V, V2W), orc.HostMesh(np.zeros((1, 3, 3), np.float32)), mat.struct,
                    _abi.Lights(), ids, _abi.IBL | _abi.BG | _abi.TONEMAP | _abi.NOLIGHT, ibl=ibl)
    err = float(np.abs(img - ref).max())
    assert err <= 1e-5, err


def test_batch_with_an_empty_scene(gpu):
    import torch
    from taichi_three_amd import workloads as wl
    from taichi_three_amd.batch import BatchRenderer
    print('case: batch with an empty scene', flush=True)
    ws = wl.batch_views(n=3, seed=5, faces=(1000, 5000))
    empty = dict(ws[0])
    for k in ('verts', 'colors'):
        empty[k] = ws[0][k][:0]
    empty['attrs'] = {k: v[:0] for k, v in ws[0]['attrs'].items()}
    plan = BatchRenderer([ws[1], empty, ws[2]])
    plan.render()
    torch.cuda.synchronize()
    ids, depth = plan.resolve(1)
    assert (ids.cpu().numpy() == -1).all()
    assert (plan.resolve(0)[0].cpu().numpy() >= 0).any()
    plan.close()
