"""Wireframe raster on the HIP path vs the CPU oracle: depth and ids bit-exact,
the image exactly the line colour where a wire won and the background elsewhere
(both through ACES, within 1e-6), alone and in scenes with triangles and
particles in either order."""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu

COLOR_TOL_F32 = 1e-5
COLOR_TOL_F64 = 1e-4
LINE = np.array([0.9, 0.6, 0.0], dtype=np.float32)   # WireframeRaster's default linecolor


def _camera(W, H, eye=(0.4, 0.3, 3.2), fov=45):
    from taichi_three_amd.util.matrix import lookat, perspective
    return lookat([0, 0, 0], list(eye)), perspective(fov, W / H, 0.1, 50)


def _wires(n, seed, W, H, extreme=True):
    rng = np.random.default_rng(seed)
    w = rng.uniform(-1.2, 1.2, (n, 2, 3)).astype(np.float32)
    if extreme:
        from taichi_three_amd.util.matrix import camera_matrices
        W2V, V2W = camera_matrices(*_camera(W, H))
        V2W = np.float64(V2W)

        def world(ndc):
            h = V2W @ np.append(ndc, 1.0)
            return h[:3] / h[3]
        eye = np.array([0.4, 0.3, 3.2])
        fwd = -eye / np.linalg.norm(eye)
        right = np.cross(fwd, [0, 1, 0])
        right /= np.linalg.norm(right)
        extra = np.array([
            [[0.2, 0.1, 0.3], [0.2, 0.1, 0.3]],                                # zero length on screen
            [eye + right * 1.0 + fwd * 2e-5, [0.1, -0.2, 0.0]],                # w ~ 2e-5: ~10^7 px long
            [eye - right * 0.5 + fwd * 1e-3, eye + right * 0.5 + fwd * 1e-3],  # in front of the near plane
            [world([-0.9, -0.9, 0.5]), world([0.9, 0.9, 0.5])],                # the screen diagonal
            [world([0.3, -0.7, 0.2]), world([0.3, 0.7, 0.2])],                 # vertical
        ], dtype=np.float32)
        w = np.concatenate([w, extra])
    return w


def _oracle(scene, wires, flags, id_base=0, depth=None, occ=None):
    eng = scene.engine
    W, H = (int(x) for x in scene.res)
    cam = orc.camera(eng.W2V, eng.V2W, eng.bias)
    return orc.raster_wires(W, H, cam, wires, id_base=id_base, flags=flags, depth=depth, occup=occ)


def _run(scene):
    import torch
    scene.render()
    torch.cuda.synchronize()
    eng = scene.engine
    return eng.ids.to_numpy().T, eng.depth.to_numpy().T, scene.img.to_numpy().transpose(1, 0, 2)


@pytest.mark.parametrize('case', [(640, 360, 20000, 0, False), (1920, 1080, 50000, 1, True), (333, 517, 3000, 2, False)])
def test_wire_soup_bit_exact(gpu, case):
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    W, H, n, seed, clipping = case
    print(f'case: wires {W}x{H} n={n} seed {seed} clipping={clipping}', flush=True)
    wires = _wires(n, seed, W, H)
    bg = np.array([0.2, 0.3, 0.4], dtype=np.float32)
    scene = tina.Scene(W, H, clipping=clipping, bgcolor=bg)
    mesh = tina.SimpleMesh(npolygon=2)
    mesh.set_face_verts(wires)
    scene.add_object(mesh)
    scene.set_camera(*_camera(W, H))
    ids, depth, img = _run(scene)
    od, oi = _oracle(scene, wires, _abi.CLIP if clipping else 0)
    assert (oi >= 0).sum() > 0.05 * W * H
    assert np.array_equal(ids, oi), f'{(ids != oi).sum()} id mismatches'
    assert np.array_equal(depth, od)
    exp = np.where((oi >= 0)[..., None], orc.aces(LINE[None])[0], orc.aces(bg[None])[0])
    assert np.abs(img - exp).max() <= 1e-6
    if not clipping:
        assert np.any(oi == n + 1), 'the 10^7-pixel wire crosses the screen'


def test_mesh_to_wire_edges(gpu):
    import taichi_three_amd as tina
    sphere = tina.PrimitiveMesh.sphere(lons=8, lats=6)
    wire = tina.MeshToWire(sphere)
    fb = wire.face_buffers()
    v = sphere.face_buffers()['verts'].cpu().numpy()
    exp = np.stack([v, v[:, [1, 2, 0]]], axis=2).reshape(-1, 2, 3)
    assert wire.get_npolygon() == 2 and wire.get_nfaces() == 3 * len(v) == fb['nfaces']
    assert np.array_equal(fb['verts'].cpu().numpy(), exp)


@pytest.mark.parametrize('wire_first', [False, True])
def test_triangles_wires_particles_in_one_scene(gpu, wire_first):
    """The three rasters share one depth|id buffer: ids in object order, each
    object's colour pass over the pixels it won."""
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd.matr import lower_material
    W, H = 480, 320
    print(f'case: triangles + wires + particles wire_first={wire_first}', flush=True)
    scene = tina.Scene(W, H, maxpars=2000, linecolor=(0.1, 0.8, 0.3))
    sphere = tina.PrimitiveMesh.sphere()
    cage = tina.MeshToWire(tina.MeshTransform(tina.PrimitiveMesh.sphere(lons=12, lats=8)))
    cage.set_transform(np.diag([1.15, 1.15, 1.15, 1.0]))
    rng = np.random.default_rng(4)
    pos = rng.uniform(-1.4, 1.4, (2000, 3)).astype(np.float32)
    rad = rng.uniform(0.01, 0.05, 2000).astype(np.float32)
    col = rng.uniform(0.1, 1.0, (2000, 3)).astype(np.float32)
    pars = tina.SimpleParticles(2000)
    pars.set_particles(pos)
    pars.set_particle_radii(rad)
    pars.set_particle_colors(col)
    mat_tri, mat_par = tina.Classic(), tina.Diffuse()
    objs = [(cage, None), (sphere, mat_tri), (pars, mat_par)] if wire_first else \
        [(sphere, mat_tri), (cage, None), (pars, mat_par)]
    for o, m in objs:
        scene.add_object(o, m)
    scene.set_camera(*_camera(W, H))
    ids, depth, img = _run(scene)

    eng = scene.engine
    cam = orc.camera(eng.W2V, eng.V2W, eng.bias)
    tri = sphere.face_buffers()['verts'].cpu().numpy()
    wv = cage.face_buffers()['verts'].cpu().numpy()
    depth_o = np.full((H, W), 2 ** 30, np.int32)
    occ = np.full((H, W), -1, np.int32)
    base = 0
    ranges = {}
    for o, _ in objs:
        if o is sphere:
            depth_o, occ = orc.raster(W, H, cam, tri, id_base=base, flags=_abi.CULL | _abi.CLIP, depth=depth_o,
                                      occup=occ)
            ranges['tri'] = (base, len(tri))
            base += len(tri)
        elif o is cage:
            depth_o, occ = orc.raster_wires(W, H, cam, wv, id_base=base, flags=0, depth=depth_o, occup=occ)  # wires default to clipping=False
            ranges['wire'] = (base, len(wv))
            base += len(wv)
        else:
            depth_o, occ = orc.raster_particles(W, H, cam, pos, rad, id_base=base, flags=_abi.CLIP, depth=depth_o,
                                                occup=occ)
            ranges['par'] = (base, len(pos))
            base += len(pos)
    assert np.array_equal(ids, occ), f'{(ids != occ).sum()} id mismatches'
    assert np.array_equal(depth, depth_o)
    lt = scene.lighting.to_struct()

    def inr(key):
        b, n = ranges[key]
        return (occ >= b) & (occ < b + n)
    assert inr('tri').any() and inr('wire').any() and inr('par').any()
    tm = _abi.TONEMAP
    lo_t = lower_material(mat_tri, dict(attrs={}))
    t32 = orc.shade(W, H, cam, orc.HostMesh(tri, id_base=ranges['tri'][0]), lo_t.struct, lt, occ, _abi.BG | tm)
    lo_p = lower_material(mat_par, dict(attrs={}))
    pm = orc.particle_mesh(pos, rad, col, id_base=ranges['par'][0])
    p32 = orc.shade(W, H, cam, pm, lo_p.struct, lt, occ, _abi.PARTICLES | _abi.VCOLOR | tm)
    exp = np.where(inr('par')[..., None], p32, t32)
    exp = np.where(inr('wire')[..., None], orc.aces(np.array([[0.1, 0.8, 0.3]], np.float32))[0], exp)
    nan = np.isnan(img)
    assert np.array_equal(nan, np.isnan(exp))
    assert np.abs(img[~nan] - exp[~nan]).max() <= COLOR_TOL_F32


@pytest.mark.parametrize('seed', range(6))
def test_random_particle_and_wire_scenes(gpu, seed):
    """Random screens, perspective or orthographic cameras, sub-pixel biases and
    clipping choices with a particle cloud and a wire soup in one scene: ids and
    depth bit for bit against the oracle, particle colour within 1e-5 of binary32."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd.matr import lower_material
    rng = np.random.default_rng(500 + seed)
    W, H = int(rng.integers(48, 1300)), int(rng.integers(48, 800))
    npar, nw = int(rng.integers(500, 30000)), int(rng.integers(100, 8000))
    spread = rng.uniform(0.5, 2.0)
    pos = rng.uniform(-spread, spread, (npar, 3)).astype(np.float32)
    rad = np.exp(rng.uniform(np.log(0.002), np.log(0.15), npar)).astype(np.float32)
    col = rng.uniform(0.1, 1, (npar, 3)).astype(np.float32)
    wires = rng.uniform(-spread, spread, (nw, 2, 3)).astype(np.float32)
    d = rng.normal(size=3)
    view = tina.lookat(pos=tuple(rng.uniform(-0.3, 0.3, 3)), back=tuple(rng.uniform(1.0, 4.0) * d / np.linalg.norm(d)))
    ortho = bool(rng.integers(0, 2))
    proj = tina.orthogonal(spread * 1.5, W / H) if ortho else tina.perspective(float(rng.uniform(30, 100)), W / H, 0.05, 50)
    clipping = seed % 2 == 1
    bias = rng.uniform(0, 1, 2).astype(np.float32)
    print(f'case: random particles+wires seed={seed} {W}x{H} npar={npar} nwires={nw} ortho={ortho} '
          f'clipping={clipping}', flush=True)
    scene = tina.Scene(W, H, maxpars=npar, clipping=clipping)
    pars = tina.SimpleParticles(npar)
    pars.set_particles(pos)
    pars.set_particle_radii(rad)
    pars.set_particle_colors(col)
    mesh = tina.SimpleMesh(nw, npolygon=2)
    mesh.set_face_verts(wires)
    mat = tina.Diffuse()
    scene.add_object(pars, mat)
    scene.add_object(mesh)
    scene.set_camera(view, proj)
    scene.engine.bias = bias
    scene.engine._push_camera()
    ids, depth, img = _run(scene)
    eng = scene.engine
    cam = orc.camera(eng.W2V, eng.V2W, eng.bias)
    fl = _abi.CLIP if clipping else 0
    od, oi = orc.raster_particles(W, H, cam, pos, rad, flags=fl)
    od, oi = orc.raster_wires(W, H, cam, wires, id_base=npar, flags=fl, depth=od, occup=oi)
    assert np.array_equal(ids, oi), f'{(ids != oi).sum()} id mismatches'
    assert np.array_equal(depth, od)
    lo = lower_material(mat, dict(attrs={}))
    p32 = orc.shade(W, H, cam, orc.particle_mesh(pos, rad, col), lo.struct, scene.lighting.to_struct(), oi,
                    _abi.PARTICLES | _abi.VCOLOR | _abi.BG | _abi.TONEMAP)
    exp = np.where(((oi >= npar))[..., None], orc.aces(LINE[None])[0], p32)
    nan = np.isnan(img)
    assert np.array_equal(nan, np.isnan(exp))
    assert np.abs(img[~nan] - exp[~nan]).max() <= COLOR_TOL_F32


def test_boundary_rejects_bad_arguments(gpu):
    """The particle, wire, paint, buffer and SSAO entry points reject malformed
    calls with TINA_ERR_INVALID (raised as ValueError) before launching anything."""
    import ctypes
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    eng = tina.Engine(64, 48)
    lib, ctx = eng._lib, eng._ctx
    pos = torch.zeros((4, 3), device='cuda')
    rad = torch.ones(4, device='cuda')
    img = torch.zeros((48, 64, 3), device='cuda')
    col = (ctypes.c_float * 3)(1, 1, 1)
    m = _abi.Mesh()
    m.verts, m.norms, m.nfaces = _abi.ptr(pos), 0, 4
    calls = [
        lambda: lib.tina_raster_particles(ctx, _abi.ptr(pos), _abi.ptr(rad), 4, 0, _abi.CULL, None),   # flag
        lambda: lib.tina_raster_particles(ctx, _abi.ptr(pos), None, 4, 0, 0, None),                    # radii
        lambda: lib.tina_raster_particles(ctx, _abi.ptr(pos), _abi.ptr(rad), 4, 0xFFFFFFFE, 0, None),  # ids
        lambda: lib.tina_raster_wires(ctx, _abi.ptr(pos), -1, 0, 0, None),
        lambda: lib.tina_raster_wires(ctx, None, 2, 0, 0, None),
        lambda: lib.tina_paint_ids(ctx, 0, 4, ctypes.byref(col), None, _abi.BG, _abi.ptr(img), None),  # bg missing
        lambda: lib.tina_paint_ids(ctx, 0, 4, ctypes.byref(col), None, _abi.SMOOTH, _abi.ptr(img), None),
        lambda: lib.tina_gbuffer(ctx, ctypes.byref(m), 0, 12, 3, None, ctypes.c_float(1), _abi.ptr(img), None),
        lambda: lib.tina_gbuffer(ctx, ctypes.byref(m), 0, 9, 3, None, ctypes.c_float(0), _abi.ptr(img), None),
        lambda: lib.tina_gbuffer(ctx, ctypes.byref(m), _abi.PARTICLES, 0, 3, None, ctypes.c_float(1), _abi.ptr(img),
                                 None),                                                                  # no radii
        lambda: lib.tina_ssao(ctx, _abi.ptr(img), _abi.ptr(pos), 0, _abi.ptr(rad), 2, ctypes.c_float(0.2),
                              ctypes.c_float(0), ctypes.c_float(1), _abi.ptr(rad), None),                # nsamples
        lambda: lib.tina_ssao_apply(_abi.ptr(img), 0, 48, _abi.ptr(rad), 4, 1, None),
    ]
    for k, call in enumerate(calls):
        with pytest.raises(ValueError):
            _abi.check(call(), ctx if k < len(calls) - 1 else None)
    # an empty call is a no-op, not an error
    _abi.check(lib.tina_raster_particles(ctx, None, None, 0, 0, 0, None), ctx)
    _abi.check(lib.tina_raster_wires(ctx, None, 0, 0, 0, None), ctx)
