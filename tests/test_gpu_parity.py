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
    print(f'colour max |gpu - binary32| = {np.abs(img - o32).max():.2e}, '
          f'|gpu - binary64| = {np.abs(img - o64).max():.2e}', flush=True)
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
    print(f'colour max |gpu - binary32| = {np.abs(img - o32).max():.2e}, '
          f'|gpu - binary64| = {np.abs(img - o64).max():.2e}', flush=True)
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


def test_raster_thin_and_screen_filling_faces(gpu):
    """Bboxes one pixel wide or tall (the queue pass's row/column split), faces larger
    than the screen with a vertex inside the NDC cube, and faces crossing w = 0
    (no near clipping in the reference: culled or clipped as it decides)."""
    import torch
    import taichi_three_amd as tina
    W, H = 160, 1200
    M = np.diag([1.0, 1.0, -1.0, 1.0])
    xs = np.linspace(-0.9, 0.9, 40)
    faces = []
    for i, x in enumerate(xs):
        dx = (i % 3) * 0.004 + 0.0005                      # 0.04 to 0.7 px wide
        faces.append([[x, -0.95, 0.1 * (i % 5)], [x + dx, -0.95, 0.0], [x, 0.95, 0.2]])
        faces.append([[-0.95, x, 0.3], [0.95, x, 0.1], [-0.95, x + dx / 8, -0.1]])
    faces.append([[-3, -3, 0.5], [3, -3, 0.5], [0, 0.5, 0.5]])   # larger than the screen
    faces.append([[-0.5, -0.5, 0.2], [0.5, -0.5, 0.2], [0.0, 0.5, -1.5]])  # one vertex beyond z = -1
    verts = np.array(faces, dtype=np.float32)
    rng = np.random.default_rng(3)
    persp = np.array([[1.0, 0, 0, 0], [0, 1.0, 0, 0], [0, 0, -1.0, -0.1], [0, 0, -1.0, 0.5]])
    for name, W2V in (('identity', M), ('projective', persp)):
        for culling in (True, False):
            print(f'case: thin faces {name} culling={culling}', flush=True)
            eng = tina.Engine(W, H, maxfaces=len(verts))
            eng.W2V = W2V.astype(np.float32)
            eng.V2W = np.linalg.inv(W2V).astype(np.float32)
            eng._push_camera()
            r = tina.TriangleRaster(eng, maxfaces=len(verts), culling=culling)
            m = tina.SimpleMesh()
            m.set_face_verts(verts[rng.permutation(len(verts))] if name == 'projective' else verts)
            eng.clear_depth()
            r.set_object(m)
            r.render_occup()
            ids, depth = eng.resolve()
            torch.cuda.synchronize()
            cam = orc.camera(eng.W2V, eng.V2W)
            od, oi = orc.raster(W, H, cam, m.verts.cpu().numpy(), flags=(1 if culling else 0) | 2)
            assert np.array_equal(ids.cpu().numpy(), oi)
            assert np.array_equal(depth.cpu().numpy(), od)


def test_raster_slivers_and_subpixel_faces(gpu):
    """Faces whose pixel set is decided by rounding: near-collinear slivers (the
    third vertex 1e-6 to 1e-2 px off the opposite edge), sub-pixel triangles,
    and triangles with vertices on pixel centres (edges through centres) -- the
    cases where the binned path's shrunken per-face bboxes must still hold every
    pixel the reference's coverage test accepts.  Bit-exact ids and depths."""
    import torch
    import taichi_three_amd as tina
    W, H = 640, 480
    rng = np.random.default_rng(77)
    n = 30000
    c = rng.uniform([-0.95, -0.95], [0.95, 0.95], (n, 2))
    d = rng.normal(0, 0.03, (n, 2))
    t = rng.uniform(0.1, 0.9, (n, 1))
    perp = np.stack([-d[:, 1], d[:, 0]], 1) / np.linalg.norm(d, axis=1, keepdims=True)
    off = (10.0 ** rng.uniform(-6, -2, (n, 1))) * (2.0 / W) * rng.choice([-1, 1], (n, 1))
    sliver = np.stack([c, c + d, c + t * d + off * perp], 1)
    tiny = c[:, None, :] + rng.normal(0, 0.6 / W, (n, 3, 2))
    grid = (np.floor((c[:, None, :] * 0.5 + 0.5) * [W, H]) + rng.integers(-4, 5, (n, 3, 2)) + 0.5) / [W, H] * 2 - 1
    xy = np.concatenate([sliver, tiny, grid])
    z = rng.uniform(-0.9, 0.9, xy.shape[:2] + (1,))
    verts = np.concatenate([xy, z], -1).astype(np.float32)
    M = np.diag([1.0, 1.0, -1.0, 1.0]).astype(np.float32)
    for culling in (True, False):
        print(f'case: slivers and sub-pixel faces culling={culling}', flush=True)
        eng = tina.Engine(W, H, maxfaces=len(verts))
        eng.W2V = M
        eng.V2W = np.linalg.inv(M).astype(np.float32)
        eng._push_camera()
        r = tina.TriangleRaster(eng, maxfaces=len(verts), culling=culling)
        m = tina.SimpleMesh(maxfaces=len(verts))
        m.set_face_verts(verts)
        eng.clear_depth()
        r.set_object(m)
        r.render_occup()
        ids, depth = eng.resolve()
        torch.cuda.synchronize()
        cam = orc.camera(eng.W2V, eng.V2W)
        od, oi = orc.raster(W, H, cam, verts, flags=(1 if culling else 0) | 2)
        assert (oi >= 0).sum() > 10000
        assert np.array_equal(ids.cpu().numpy(), oi), f'{(ids.cpu().numpy() != oi).sum()} id mismatches'
        assert np.array_equal(depth.cpu().numpy(), od)


@pytest.mark.parametrize('seed', [31, 32])
def test_raster_camera_inside_the_cloud(gpu, seed):
    """Camera in the middle of the soup: faces next to the eye (1/w far from 1,
    huge and tiny bboxes, vertices behind the camera) exercise the binned path's
    non-well-conditioned faces, the giant queue and the clip rule together."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    rng = np.random.default_rng(seed)
    verts = wl.triangle_soup(rng, 60000, np.log(0.05), 1.0)
    W, H = 640, 480
    view = tina.lookat(pos=(0.1, 0.05, 0.0), back=(0.0, 0.0, 0.3))
    proj = tina.perspective(90, W / H, 0.01, 50)
    W2V, V2W = tina.camera_matrices(view, proj)
    cam = orc.camera(W2V, V2W)
    for culling in (True, False):
        print(f'case: camera inside the cloud seed={seed} culling={culling}', flush=True)
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
        assert (oi >= 0).sum() > 1000
        assert np.array_equal(ids.cpu().numpy(), oi), f'{(ids.cpu().numpy() != oi).sum()} id mismatches'
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


def test_giant_faces_over_a_loaded_frame_repeatedly(gpu):
    """A soup of small faces, then an object of giant faces (bboxes over many
    tiles, some crossing the soup in depth) drawn into the same keys, three
    frames in a row: the tiles take each giant's part inside them whether they
    start from empty keys or load the first object's, and the counters the
    raster calls alternate between stay right from frame to frame."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    rng = np.random.default_rng(23)
    small = wl.triangle_soup(rng, 20000, np.log(0.03), 0.5)
    big = (rng.uniform(-1.5, 1.5, (40, 3, 3)) * np.array([1.0, 1.0, 0.6])).astype(np.float32)
    W, H = 480, 360
    scene = tina.Scene(W, H, maxfaces=len(small) + len(big))
    view, proj = wl.default_camera(W / H)
    scene.engine.set_camera(view, proj)
    ms, mb = tina.SimpleMesh(), tina.SimpleMesh()
    ms.set_face_verts(small)
    mb.set_face_verts(big)
    scene.add_object(ms)
    scene.add_object(mb, tina.Classic(color=(0.3, 0.6, 0.9)))
    cam = orc.camera(scene.engine.W2V, scene.engine.V2W)
    od, oi = orc.raster(W, H, cam, small, id_base=0)
    od, oi = orc.raster(W, H, cam, big, id_base=len(small), depth=od, occup=oi)
    assert (oi >= len(small)).sum() > 20000 and ((oi >= 0) & (oi < len(small))).sum() > 5000
    for frame in range(3):
        print(f'case: giant faces over a loaded frame, frame {frame}', flush=True)
        scene.render()
        torch.cuda.synchronize()
        assert np.array_equal(scene.engine.ids.to_numpy().T, oi)
        assert np.array_equal(scene.engine.depth.to_numpy().T, od)


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


def test_generic_material_program_and_nodes(gpu):
    """A material tree no fixed kernel matches (the interpreter path), with Param,
    Texture (texturing=True), Input('normal') and an Emission lobe."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl, _abi
    rng = np.random.default_rng(21)
    W, H = 200, 150
    verts = wl.triangle_soup(rng, 4000, np.log(0.15), 0.5)
    coors = rng.uniform(0, 1, size=(len(verts), 3, 2)).astype(np.float32)
    tex = rng.uniform(0, 1, size=(16, 8, 3)).astype(np.float32)
    rough = tina.Param(initial=0.3)
    mat = tina.MixMaterial(tina.Lambert() * tina.Texture(tex),
                           tina.CookTorrance(roughness=rough, fresnel=0.5) + tina.Phong(shineness=8) * 'normal',
                           0.5) + tina.Emission() * 0.05
    scene = tina.Scene(W, H, texturing=True, maxfaces=len(verts))
    view, proj = wl.default_camera(W / H)
    scene.engine.set_camera(view, proj)
    mesh = tina.SimpleMesh()
    mesh.set_face_verts(verts)
    mesh.set_face_coors(coors)
    scene.add_object(mesh, mat)
    cam = orc.camera(scene.engine.W2V, scene.engine.V2W)
    for r in (0.3, 0.7):
        rough.value[None] = r
        print(f'case: generic material roughness {r}', flush=True)
        scene.render()
        torch.cuda.synchronize()
        ids = scene.engine.ids.to_numpy().T
        img = scene.img.to_numpy().transpose(1, 0, 2)
        _, oi = orc.raster(W, H, cam, verts)
        assert np.array_equal(ids, oi)
        # the oracle reads Param / Texture from host copies of the same values
        lo = tina.matr.lower_material(mat, dict(attrs={}))
        host = []
        for i in range(lo.struct.nparams):
            p = lo.struct.params[i]
            if p.kind in (_abi.P_PARAM, _abi.P_TEXTURE):
                src = rough.tensor if p.kind == _abi.P_PARAM else mat.mat1.mat1.params['factor'].tensor
                h = np.ascontiguousarray(src.cpu().numpy())
                host.append(h)
                p.ptr = h.ctypes.data
        flags = _abi.TEX | _abi.TONEMAP | _abi.BG
        o32 = orc.shade(W, H, cam, orc.HostMesh(verts, None, coors, None), lo.struct,
                        scene.lighting.to_struct(), oi, flags)
        assert np.abs(img - o32).max() <= COLOR_TOL_F32


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


@pytest.mark.parametrize('seed', list(range(8)))
def test_raster_random_scenes(gpu, seed):
    """Random soups, cameras and screens (binned path: bbox shrinking, giant
    faces, strips of tiles) against the oracle, bit for bit, culling on and off."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    rng = np.random.default_rng(1000 + seed)
    W, H = int(rng.integers(64, 1400)), int(rng.integers(64, 900))
    n = int(rng.integers(2000, 60000))
    verts = wl.triangle_soup(rng, n, np.log(rng.uniform(0.005, 0.2)), rng.uniform(0.3, 1.2))
    d = rng.normal(size=3)
    view = tina.lookat(pos=tuple(rng.uniform(-0.3, 0.3, 3)), back=tuple(rng.uniform(1.5, 4.0) * d / np.linalg.norm(d)))
    proj = tina.perspective(float(rng.uniform(30, 100)), W / H, 0.05, 500)
    W2V, V2W = tina.camera_matrices(view, proj)
    cam = orc.camera(W2V, V2W)
    for culling in (True, False):
        print(f'case: random scene seed={seed} {W}x{H} n={n} culling={culling}', flush=True)
        eng = tina.Engine(W, H, maxfaces=n)
        eng.set_camera(view, proj)
        r = tina.TriangleRaster(eng, maxfaces=n, culling=culling)
        m = tina.SimpleMesh(maxfaces=n)
        m.set_face_verts(verts)
        eng.clear_depth()
        r.set_object(m)
        r.render_occup()
        ids, depth = eng.resolve()
        torch.cuda.synchronize()
        od, oi = orc.raster(W, H, cam, verts, flags=(1 if culling else 0) | 2)
        assert np.array_equal(ids.cpu().numpy(), oi), f'{(ids.cpu().numpy() != oi).sum()} id mismatches'
        assert np.array_equal(depth.cpu().numpy(), od)


@pytest.mark.parametrize('seed', [0, 1, 2, 3])
def test_raster_random_large_calls(gpu, seed):
    """Calls of 1.1-1.5 M faces (and 3 M: two rounds of binning groups, one
    group per CU in each).
    Random screens, cameras (some inside the cloud, so the clip rule acts), face
    sizes from sub-pixel to giant; ids and depth bit for bit against the
    oracle's parallel raster, culling on and off."""
    import os
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    rng = np.random.default_rng(2000 + seed)
    W, H = int(rng.integers(300, 1400)), int(rng.integers(200, 900))
    n = int(rng.integers(1_100_000, 1_500_000)) if seed < 3 else 3_000_000
    verts = wl.triangle_soup(rng, n, np.log(rng.uniform(0.004, 0.02)), rng.uniform(0.5, 1.2))
    big = rng.choice(n, 200, replace=False)  # a few faces over many tiles
    verts[big] = wl.triangle_soup(rng, 200, np.log(0.8), 0.3)
    d = rng.normal(size=3)
    dist = float((0.4, 1.6, 3.2, 2.2)[seed] * rng.uniform(0.9, 1.1))  # inside the cloud, at its edge, outside
    view = tina.lookat(pos=tuple(rng.uniform(-0.2, 0.2, 3)), back=tuple(dist * d / np.linalg.norm(d)))
    proj = tina.perspective(float(rng.uniform(40, 90)), W / H, 0.05, 500)
    W2V, V2W = tina.camera_matrices(view, proj)
    cam = orc.camera(W2V, V2W)
    nthreads = max(1, len(os.sched_getaffinity(0)))
    eng = tina.Engine(W, H, maxfaces=n)
    eng.set_camera(view, proj)
    m = tina.SimpleMesh(maxfaces=n)
    m.set_face_verts(verts)
    for culling in (True, False):
        print(f'case: large call seed={seed} {W}x{H} n={n} dist={dist:.2f} culling={culling}', flush=True)
        r = tina.TriangleRaster(eng, maxfaces=n, culling=culling)
        eng.clear_depth()
        r.set_object(m)
        r.render_occup()
        ids, depth = eng.resolve()
        torch.cuda.synchronize()
        keys = orc.raster_omp(W, H, cam, verts, flags=(1 if culling else 0) | 2, nthreads=nthreads)
        od, oi = orc.unpack(keys)
        assert (oi >= 0).sum() > 0.1 * W * H
        assert np.array_equal(ids.cpu().numpy(), oi), f'{(ids.cpu().numpy() != oi).sum()} id mismatches'
        assert np.array_equal(depth.cpu().numpy(), od)


def test_raster_binning_group_boundaries(gpu):
    """Face counts at the edges of the binning-group sizes (groups of at least
    512 faces, whole rounds of one group per CU, at most 8,192 faces a group at
    this screen): 1, 31-33, 255-257 and 511-513 faces, 32 and 512 faces a CU
    and one face more or less, and around a full round of 8,192-face groups.  Every call renders the
    same scene's first n faces; ids and depth bit for bit against the oracle."""
    import os
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    W, H = 640, 360
    counts = sorted({1, 31, 32, 33, 255, 256, 257, 511, 512, 513, 32 * ncu - 1, 32 * ncu, 32 * ncu + 1,
                     512 * ncu - 1, 512 * ncu, 512 * ncu + 1, 8192 * ncu - 1, 8192 * ncu, 8192 * ncu + 1})
    n = counts[-1]
    rng = np.random.default_rng(77)
    verts = wl.triangle_soup(rng, n, np.log(0.01), 0.7)
    view = tina.lookat(back=(0.3, 0.4, 3.0))
    proj = tina.perspective(60, W / H, 0.05, 500)
    W2V, V2W = tina.camera_matrices(view, proj)
    cam = orc.camera(W2V, V2W)
    nthreads = max(1, len(os.sched_getaffinity(0)))
    eng = tina.Engine(W, H, maxfaces=n)
    eng.set_camera(view, proj)
    r = tina.TriangleRaster(eng, maxfaces=n)
    for k in counts:
        print(f'case: binning boundary {W}x{H} n={k} ({ncu} CUs)', flush=True)
        m = tina.SimpleMesh(maxfaces=k)
        m.set_face_verts(verts[:k])
        eng.clear_depth()
        r.set_object(m)
        r.render_occup()
        ids, depth = eng.resolve()
        torch.cuda.synchronize()
        od, oi = orc.unpack(orc.raster_omp(W, H, cam, verts[:k], flags=3, nthreads=nthreads))
        assert k < 255 or (oi >= 0).sum() > 20  # faces of ~0.5 px: 255 of them cover ~70 pixels
        assert np.array_equal(ids.cpu().numpy(), oi), f'n={k}: {(ids.cpu().numpy() != oi).sum()} id mismatches'
        assert np.array_equal(depth.cpu().numpy(), od), f'n={k}'
