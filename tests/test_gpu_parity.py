"""HIP path vs the CPU oracle: bit-exact id/depth buffers, colour within 1e-5 of
the binary32 oracle and 1e-4 of the binary64 oracle (tolerances stated per test)."""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu

COLOR_TOL_F32 = 1e-5   # abs, tone-mapped colour vs the binary32 oracle
COLOR_TOL_F64 = 1e-4   # abs, tone-mapped colour vs the binary64 oracle


def _render(w, **opts):
    import torch
    from taichi_three_amd import workloads as wl
    scene, mesh = wl.build_scene(w, **opts)
    scene.render()
    torch.cuda.synchronize()
    ids = scene.engine.ids.to_numpy().T
    depth = scene.engine.depth.to_numpy().T
    img = scene.img.to_numpy().transpose(1, 0, 2)
    return scene, ids, depth, img


def _oracle(w, flags_raster=3, tonemap=True):
    from taichi_three_amd import workloads as wl, _abi
    from taichi_three_amd.matr import lower_material
    W, H = w.res
    cam = orc.camera(w.W2V, w.V2W)
    depth, ids = orc.raster(W, H, cam, w.verts, flags=flags_raster)
    lo = lower_material(wl.make_material(w.material), dict(attrs=w.attrs))
    flags = _abi.VCOLOR | _abi.BG | (_abi.TONEMAP if tonemap else 0) | (_abi.SMOOTH if w.smoothing else 0)
    mesh = orc.HostMesh(w.verts, w.norms, None, w.colors)
    lt = wl.make_lighting(w).to_struct()
    img32 = orc.shade(W, H, cam, mesh, lo.struct, lt, ids, flags)
    img64 = orc.shade(W, H, cam, mesh, lo.struct, lt, ids, flags, f64=True)
    return ids, depth, img32, img64


def test_plain_512_bit_exact(gpu):
    from taichi_three_amd import workloads as wl
    w = wl.plain_512()
    print('case: plain_512 seed 0', flush=True)
    _, ids, depth, img = _render(w)
    oids, odepth, o32, o64 = _oracle(w)
    assert (ids >= 0).sum() > 50000
    assert np.array_equal(ids, oids), f'{(ids != oids).sum()} id mismatches'
    assert np.array_equal(depth, odepth)
    assert np.abs(img - o32).max() <= COLOR_TOL_F32
    assert np.abs(img - o64).max() <= COLOR_TOL_F64


def test_plain_512_classic_no_tonemap(gpu):
    from taichi_three_amd import workloads as wl
    w = wl.plain_512(seed=3)
    w['material'] = 'classic'
    print('case: plain_512 classic seed 3', flush=True)
    _, ids, depth, img = _render(w, tonemap=False)
    oids, odepth, o32, o64 = _oracle(w, tonemap=False)
    assert np.array_equal(ids, oids)
    # before tone mapping the colour is unbounded: relative tolerance
    assert np.all(np.abs(img - o32) <= COLOR_TOL_F32 * np.maximum(1, np.abs(o32)))
    assert np.all(np.abs(img - o64) <= COLOR_TOL_F64 * np.maximum(1, np.abs(o64)))


@pytest.mark.parametrize('seed', [0, 1, 2, 3])
def test_soup_small_pbr(gpu, seed):
    """A 200k-face soup of the 1080p generator at 640x360: PBR with per-face
    roughness/metallic and four point lights."""
    from taichi_three_amd import workloads as wl
    w = wl.soup_1080p(seed=seed, res=(640, 360), nfaces=200_000)
    print(f'case: soup 640x360 200k seed {seed}', flush=True)
    _, ids, depth, img = _render(w)
    oids, odepth, o32, o64 = _oracle(w)
    assert np.array_equal(ids, oids), f'{(ids != oids).sum()} id mismatches'
    assert np.array_equal(depth, odepth)
    assert np.abs(img - o32).max() <= COLOR_TOL_F32
    assert np.abs(img - o64).max() <= COLOR_TOL_F64


@pytest.mark.parametrize('edge_mu,seed', [(np.log(0.3), 5), (np.log(1.5), 6), (np.log(0.02), 7)])
def test_raster_sizes_and_culling(gpu, edge_mu, seed):
    """Small, big and giant faces (all three raster paths), culling on and off."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    rng = np.random.default_rng(seed)
    verts = wl.triangle_soup(rng, 3000, edge_mu, 0.6)
    W, H = 333, 257
    view, proj = wl.default_camera(W / H)
    W2V, V2W = tina.camera_matrices(view, proj)
    cam = orc.camera(W2V, V2W)
    for culling in (True, False):
        print(f'case: sizes mu={edge_mu:.2f} seed={seed} culling={culling}', flush=True)
        eng = tina.Engine(W, H, maxfaces=len(verts))
        eng.set_camera(view, proj)
        r = tina.TriangleRaster(eng, maxfaces=len(verts), culling=culling)
        m = tina.SimpleMesh(maxfaces=len(verts))
        m.set_face_verts(verts)
        eng.clear_depth()
        r.set_object(m)
        r.render_occup()
        ids, depth = eng.resolve()
        torch.cuda.synchronize()
        od, oi = orc.raster(W, H, cam, verts, flags=(1 if culling else 0) | 2)
        assert np.array_equal(ids.cpu().numpy(), oi)
        assert np.array_equal(depth.cpu().numpy(), od)


def test_multi_object_ties_and_probe(gpu):
    """Two objects with coincident faces: the earlier object wins every tie; the
    probe reports local face indices of the winners."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    rng = np.random.default_rng(11)
    a = wl.triangle_soup(rng, 500, np.log(0.4), 0.5)
    b = np.concatenate([a[::-1][:250], wl.triangle_soup(rng, 300, np.log(0.4), 0.5)])  # duplicates of a
    scene = tina.Scene(256, 200, maxfaces=1000)
    view, proj = wl.default_camera(256 / 200)
    scene.engine.set_camera(view, proj)
    probe = tina.ProbeShader(scene.res)
    scene.post_shaders.append(probe)
    ma, mb = tina.SimpleMesh(), tina.SimpleMesh()
    ma.set_face_verts(a)
    mb.set_face_verts(b)
    scene.add_object(ma)
    scene.add_object(mb, tina.Classic(color=(0.3, 0.6, 0.9)))
    scene.render()
    torch.cuda.synchronize()
    ids = scene.engine.ids.to_numpy().T
    cam = orc.camera(scene.engine.W2V, scene.engine.V2W)
    od, oi = orc.raster(256, 200, cam, a, id_base=0)
    od, oi = orc.raster(256, 200, cam, b, id_base=len(a), depth=od, occup=oi)
    assert np.array_equal(ids, oi)
    assert (oi >= len(a)).sum() > 100 and ((oi >= 0) & (oi < len(a))).sum() > 100
    el = probe.elmid.to_numpy().T
    exp = np.where(oi >= len(a), oi - len(a), oi)
    assert np.array_equal(el, exp)


def test_deterministic_repeat(gpu):
    import torch
    from taichi_three_amd import workloads as wl
    w = wl.soup_1080p(seed=9, res=(480, 270), nfaces=100_000)
    scene, _ = wl.build_scene(w)
    out = []
    for _ in range(3):
        scene.render()
        torch.cuda.synchronize()
        out.append((scene.engine.ids.to_numpy().copy(), scene.img.to_numpy().copy()))
    for i in range(1, 3):
        assert np.array_equal(out[0][0], out[i][0])
        assert np.array_equal(out[0][1], out[i][1])


def test_plain_512_matches_golden_fixture(gpu):
    """The HIP frame against the committed fixture tests/golden/plain_512.npz."""
    import os
    from conftest import ROOT
    from taichi_three_amd import workloads as wl
    g = np.load(os.path.join(ROOT, 'tests', 'golden', 'plain_512.npz'))
    _, ids, depth, img = _render(wl.plain_512())
    assert np.array_equal(ids, g['ids']) and np.array_equal(depth, g['depth'])
    assert np.abs(img[192:320, 192:320] - g['img_crop']).max() <= COLOR_TOL_F32


def test_errors_raise(gpu):
    import taichi_three_amd as tina
    eng = tina.Engine(64, 64, maxfaces=10)
    r = tina.TriangleRaster(eng, maxfaces=10)
    m = tina.SimpleMesh()
    m.set_face_verts(np.zeros((20, 3, 3), dtype=np.float32))
    with pytest.raises(ValueError):
        r.set_object(m)
    from taichi_three_amd import _abi
    with pytest.raises(ValueError):
        _abi.check(eng._lib.tina_raster(eng._ctx, _abi.ptr(m.verts), 20, 0, 3, None), eng._ctx)
