"""Particle raster on the HIP path vs the CPU oracle: ids and depth bit-exact,
colour within 1e-5 of the binary32 oracle everywhere (tone-mapped), and within
the binary32 conditioning bound of the reference's formula of the binary64
oracle.

That bound: the shaded normal is d - Z sqrt(1 - |d|^2) with d = (Pl - c) / R
(tina/core/particle.py:134-141).  Pl - c cancels, so its binary32 error is
about eps32 times the coordinates' size, i.e. 1e-6 here, and d is off by
1e-6 / R; the view-axis component sqrt(1 - |d|^2) then moves by that over
sqrt(1 - |d|^2).  For Lambert shading (colour Lipschitz in the normal with a
constant near 1) the per-pixel binary64 tolerance is therefore
1e-4 + 1e-5 / (R sqrt(1 - |d|^2)) (measured on the oracle pair itself:
|f32 - f64| R sqrt(1 - |d|^2) <= 2.1e-6 over 10^5 pixels).  Specular lobes
multiply the normal's error by their sharpness, so Classic and PBR particles
are held to the binary32 oracle only; their binary64 agreement on well-sized
spheres is pinned on the CPU (tests/test_particles_oracle.py).  A pixel whose
1 - |d|^2 rounds below 0 is NaN in both binary32 images, as in the reference.
"""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu

COLOR_TOL_F32 = 1e-5
COLOR_TOL_F64 = 1e-4


def _cloud(n, seed, r_lo=0.005, r_hi=0.03, spread=1.0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    rad = rng.uniform(r_lo, r_hi, n).astype(np.float32)
    col = rng.uniform(0.1, 1.0, (n, 3)).astype(np.float32)
    return pos, rad, col


def _camera(W, H, eye=(0.4, 0.3, 3.2), fov=45):
    from taichi_three_amd.util.matrix import lookat, perspective
    return lookat([0, 0, 0], list(eye)), perspective(fov, W / H, 0.1, 50)


def _scene(W, H, pos, rad, col, material=None, **opts):
    import taichi_three_amd as tina
    scene = tina.Scene(W, H, maxpars=max(len(pos), 1), **opts)
    pars = tina.SimpleParticles(max(len(pos), 1))
    pars.set_particles(pos)
    if len(pos):
        pars.set_particle_radii(rad)
        pars.set_particle_colors(col)
    scene.add_object(pars, material)
    scene.set_camera(*_camera(W, H))
    return scene, pars


def _run(scene):
    import torch
    scene.render()
    torch.cuda.synchronize()
    eng = scene.engine
    return eng.ids.to_numpy().T, eng.depth.to_numpy().T, scene.img.to_numpy().transpose(1, 0, 2)


def _oracle_particles(scene, pos, rad, col, material, flags_raster, shade_flags, id_base=0, depth=None, occ=None):
    from taichi_three_amd.matr import lower_material
    eng = scene.engine
    W, H = (int(x) for x in scene.res)
    cam = orc.camera(eng.W2V, eng.V2W, eng.bias)
    depth, occ = orc.raster_particles(W, H, cam, pos, rad, id_base=id_base, flags=flags_raster, depth=depth, occup=occ)
    lo = lower_material(material, dict(attrs={}))
    lt = scene.lighting.to_struct()
    mesh = orc.particle_mesh(pos, rad, col, id_base=id_base)
    img32 = orc.shade(W, H, cam, mesh, lo.struct, lt, occ, shade_flags)
    img64 = orc.shade(W, H, cam, mesh, lo.struct, lt, occ, shade_flags, f64=True)
    return depth, occ, img32, img64


def _f64_tol(scene, occ, pos, rad, id_base=0):
    """[H, W] binary64 tolerance: 1e-4, plus 1e-5 / (R sqrt(1 - |d|^2)) on the pixels of
    particles [id_base, +len(pos)), d = (Pl - c) / R evaluated in binary64 (particle.py:134-140)."""
    eng = scene.engine
    H, W = occ.shape
    W2V, V2W = np.float64(eng.W2V), np.float64(eng.V2W)
    i = occ.astype(np.int64) - id_base
    mine = (i >= 0) & (i < len(pos))
    Y, X = np.nonzero(mine)
    c, R = np.float64(pos[i[mine]]), np.float64(rad[i[mine]])
    h = np.concatenate([c, np.ones((len(c), 1))], axis=1) @ W2V.T
    zv = h[:, 2] / h[:, 3]
    u = (X + np.float64(eng.bias[0])) / W * 2 - 1
    v = (Y + np.float64(eng.bias[1])) / H * 2 - 1
    P = np.stack([u, v, zv, np.ones_like(u)], axis=1) @ V2W.T
    d = (P[:, :3] / P[:, 3:] - c) / R[:, None]
    q = np.maximum(1 - (d * d).sum(1), 1e-30)
    tol = np.full(occ.shape, COLOR_TOL_F64)
    tol[Y, X] += 1e-5 / (R * np.sqrt(q))
    return tol


def _check_colour(img, o32, o64=None, tol64=None):
    nan = np.isnan(img)
    assert np.array_equal(nan, np.isnan(o32)), f'{(nan != np.isnan(o32)).sum()} NaN mismatches'
    ok = ~nan
    assert np.abs(img[ok] - o32[ok]).max(initial=0) <= COLOR_TOL_F32
    if o64 is None:
        return int(nan.any(axis=-1).sum()), None
    ok &= np.isfinite(o64)
    excess = np.where(ok, np.abs(img - o64) - tol64[..., None], 0)
    assert excess.max() <= 0, f'binary64 bound exceeded by {excess.max()}'
    return int(nan.any(axis=-1).sum()), int((tol64 > 1e-3).sum())


@pytest.mark.parametrize('case', [(640, 360, 20000, 0, 'diffuse'), (640, 360, 20000, 1, 'classic'),
                                  (1920, 1080, 100000, 2, 'pbr'), (333, 517, 5000, 3, 'diffuse')])
def test_particle_cloud_bit_exact(gpu, case):
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    W, H, n, seed, mname = case
    print(f'case: particles {W}x{H} n={n} seed {seed} {mname}', flush=True)
    material = {'diffuse': tina.Diffuse(), 'classic': tina.Classic(),
                'pbr': tina.PBR(roughness=0.35, metallic=0.2)}[mname]
    pos, rad, col = _cloud(n, seed)
    scene, _ = _scene(W, H, pos, rad, col, material)
    ids, depth, img = _run(scene)
    od, oi, o32, o64 = _oracle_particles(scene, pos, rad, col, material, _abi.CLIP,
                                         _abi.PARTICLES | _abi.VCOLOR | _abi.BG | _abi.TONEMAP)
    assert (oi >= 0).sum() > 0.2 * W * H
    assert np.array_equal(ids, oi), f'{(ids != oi).sum()} id mismatches'
    assert np.array_equal(depth, od)
    if mname == 'diffuse':
        nan_px, loose_px = _check_colour(img, o32, o64, _f64_tol(scene, oi, pos, rad))
    else:
        nan_px, loose_px = _check_colour(img, o32)
    print(f'  covered {(oi >= 0).sum()} px, binary64 bound above 1e-3 on {loose_px} px, NaN {nan_px} px', flush=True)


def test_particles_and_triangles_share_the_depth_buffer(gpu):
    """A triangle mesh then a particle set in one scene: one depth|id buffer,
    particle ids after the mesh's faces, each object shaded over the pixels it won."""
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd.matr import lower_material
    W, H = 480, 320
    print('case: mesh + particles 480x320', flush=True)
    scene = tina.Scene(W, H, maxpars=3000)
    mesh = tina.PrimitiveMesh.sphere()
    mat_mesh, mat_par = tina.Classic(), tina.Diffuse()
    scene.add_object(mesh, mat_mesh)
    pos, rad, col = _cloud(3000, 11, 0.01, 0.06, spread=1.3)
    pars = tina.SimpleParticles(3000)
    pars.set_particles(pos)
    pars.set_particle_radii(rad)
    pars.set_particle_colors(col)
    scene.add_object(pars, mat_par)
    scene.set_camera(*_camera(W, H))
    ids, depth, img = _run(scene)

    eng = scene.engine
    cam = orc.camera(eng.W2V, eng.V2W, eng.bias)
    fb = mesh.face_buffers()
    verts = fb['verts'].cpu().numpy()
    T = len(verts)
    od, oi = orc.raster(W, H, cam, verts, flags=_abi.CULL | _abi.CLIP)
    lt = scene.lighting.to_struct()
    lo_m = lower_material(mat_mesh, dict(attrs={}))
    hm = orc.HostMesh(verts, None, None, None)
    m32 = orc.shade(W, H, cam, hm, lo_m.struct, lt, oi, _abi.BG | _abi.TONEMAP)
    m64 = orc.shade(W, H, cam, hm, lo_m.struct, lt, oi, _abi.BG | _abi.TONEMAP, f64=True)
    od, oi, p32, p64 = _oracle_particles(scene, pos, rad, col, mat_par, _abi.CLIP,
                                         _abi.PARTICLES | _abi.VCOLOR | _abi.TONEMAP, id_base=T, depth=od, occ=oi)
    assert (oi < T).sum() > 0 and (oi >= T).sum() > 0
    assert np.array_equal(ids, oi), f'{(ids != oi).sum()} id mismatches'
    assert np.array_equal(depth, od)
    # the mesh's launch paints the background and its faces; the particle pixels come from the second launch
    is_par = (oi >= T)[..., None]
    o32 = np.where(is_par, p32, m32)
    o64 = np.where(is_par, p64, m64)
    _check_colour(img, o32, o64, _f64_tol(scene, oi, pos, rad, id_base=T))


def test_unclipped_jittered_and_edge_cases(gpu):
    """clipping=False (a particle behind the near plane still draws), a sub-pixel
    bias, twins (the earlier wins), a particle bigger than the screen, particles
    off the screen and the empty set."""
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    W, H = 400, 300
    pos, rad, col = _cloud(4000, 21, 0.01, 0.05, spread=2.5)
    # twins, a sphere around the whole view, one off the screen, one between the eye and the near plane
    eye = np.array([0.4, 0.3, 3.2])
    inside_near = eye * (1 - 0.05 / np.linalg.norm(eye))
    extra = np.array([[0.1, 0.1, 0.0], [0.1, 0.1, 0.0], [0.0, 0.0, -0.6], [9.0, 0.0, 0.0], inside_near],
                     dtype=np.float32)
    pos = np.concatenate([pos, extra])
    rad = np.concatenate([rad, np.array([0.2, 0.2, 3.0, 0.5, 0.01], dtype=np.float32)])
    col = np.concatenate([col, np.ones((5, 3), np.float32)])
    for clipping in (True, False):
        print(f'case: edge particles clipping={clipping}', flush=True)
        material = tina.Diffuse()
        scene, _ = _scene(W, H, pos, rad, col, material, clipping=clipping)
        scene.engine.bias = np.array([0.3125, 0.8125], dtype=np.float32)
        scene.engine._push_camera()
        ids, depth, img = _run(scene)
        od, oi, o32, o64 = _oracle_particles(scene, pos, rad, col, material, _abi.CLIP if clipping else 0,
                                             _abi.PARTICLES | _abi.VCOLOR | _abi.BG | _abi.TONEMAP)
        assert np.array_equal(ids, oi), f'{(ids != oi).sum()} id mismatches'
        assert np.array_equal(depth, od)
        assert not np.any(oi == len(pos) - 4)   # the second twin never wins
        assert np.any(oi == len(pos) - 1) != clipping   # z_view < -1: drawn only without clipping
        _check_colour(img, o32, o64, _f64_tol(scene, oi, pos, rad))
    print('case: empty particle set', flush=True)
    scene, _ = _scene(W, H, pos[:0], rad[:0], col[:0], tina.Diffuse(), bgcolor=(0.2, 0.3, 0.4))
    ids, depth, img = _run(scene)
    assert np.all(ids == -1) and np.all(depth == 2 ** 30)
    assert np.allclose(img, orc.aces(np.array([[0.2, 0.3, 0.4]], np.float32))[0], atol=1e-6)


def test_transformed_particles_and_standalone_raster(gpu):
    """ParsTransform (centres through a matrix, radii scaled) through the scene, and
    ParticleRaster driven directly (set_particles, render_occup, occup)."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    W, H = 320, 240
    print('case: ParsTransform 320x240', flush=True)
    pos, rad, col = _cloud(2000, 5, 0.01, 0.04)
    M = np.array([[0.8, -0.3, 0.1, 0.2], [0.3, 0.9, 0.0, -0.1], [0.0, 0.2, 1.1, 0.3], [0.0, 0.0, 0.0, 1.0]],
                 dtype=np.float32)
    scene = tina.Scene(W, H, maxpars=2000)
    base = tina.SimpleParticles(2000)
    base.set_particles(pos)
    base.set_particle_radii(rad)
    base.set_particle_colors(col)
    tp = tina.ParsTransform(base)
    tp.set_transform(M, 1.5)
    material = tina.Diffuse()
    scene.add_object(tp, material)
    scene.set_camera(*_camera(W, H))
    ids, depth, img = _run(scene)
    # the transform in binary32, in mapply's order (tina/common.py:169-183)
    h = [np.full(len(pos), M[i, 3], np.float32) for i in range(4)]
    for i in range(4):
        for j in range(3):
            h[i] = (h[i] + M[i, j] * pos[:, j]).astype(np.float32)
    tpos = np.stack([h[0] / h[3], h[1] / h[3], h[2] / h[3]], axis=1).astype(np.float32)
    trad = (np.float32(1.5) * rad).astype(np.float32)
    od, oi, o32, o64 = _oracle_particles(scene, tpos, trad, col, material, _abi.CLIP,
                                         _abi.PARTICLES | _abi.VCOLOR | _abi.BG | _abi.TONEMAP)
    assert np.array_equal(ids, oi), f'{(ids != oi).sum()} id mismatches'
    assert np.array_equal(depth, od)
    _check_colour(img, o32, o64, _f64_tol(scene, oi, tpos, trad))

    print('case: standalone ParticleRaster', flush=True)
    eng = tina.Engine(W, H)
    eng.set_camera(*_camera(W, H))
    r = tina.ParticleRaster(eng, maxpars=2000)
    eng.clear_depth()
    r.set_particles(pos)
    r.set_particle_radii(rad)
    r.render_occup()
    torch.cuda.synchronize()
    cam = orc.camera(eng.W2V, eng.V2W, eng.bias)
    od, oi = orc.raster_particles(W, H, cam, pos, rad, flags=_abi.CLIP)
    assert np.array_equal(r.occup.to_numpy().T, oi)
    small = tina.ParticleRaster(eng, maxpars=10)
    small.set_particles(pos)                 # keeps the first maxpars (particle.py:47)
    assert small.npars == 10
