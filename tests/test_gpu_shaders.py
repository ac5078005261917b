"""The auxiliary shaders (NormalShader, ViewNormalShader, TexcoordShader,
ColorShader, PositionShader, DepthShader, ConstShader, ViewdirShader,
SimpleShader, ChessboardShader) passed to render_color on their own: the
fragment values equal the binary32 oracle's (a Lamp material of Input(name)
shades to exactly that value), the derived ones equal their formulas on those
values, each written only where the object won."""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu


def _img(H, W, c):
    import torch
    from taichi_three_amd.field import Field
    return Field(torch.zeros((H, W, c) if c > 1 else (H, W), dtype=torch.float32, device='cuda'))


def _np(field):
    return field.storage.cpu().numpy()


def test_auxiliary_shaders(gpu):
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd.matr import lower_material
    W, H = 240, 180
    print('case: auxiliary shaders on a textured, coloured sphere and particles 240x180', flush=True)
    eng = tina.Engine(W, H)
    eng.set_camera(tina.lookat([0, 0, 0], [0.4, 0.5, 3.0]), tina.perspective(50, W / H, 0.1, 20))
    sphere = tina.PrimitiveMesh.sphere(lons=16, lats=12)
    v = sphere.face_buffers()['verts'].cpu().numpy()
    rng = np.random.default_rng(6)
    col = rng.uniform(0, 1, v.shape).astype(np.float32)
    sphere.set_face_colors(col)
    r = tina.TriangleRaster(eng, smoothing=True, texturing=True)
    eng.clear_depth()
    r.set_object(sphere)
    r.render_occup()
    pos = rng.uniform(-1.3, 1.3, (300, 3)).astype(np.float32)
    rad = rng.uniform(0.03, 0.08, 300).astype(np.float32)
    pcol = rng.uniform(0, 1, (300, 3)).astype(np.float32)
    pr = tina.ParticleRaster(eng, maxpars=300)
    pr.set_particles(pos)
    pr.set_particle_radii(rad)
    pr.set_particle_colors(pcol)
    pr.render_occup()
    torch.cuda.synchronize()
    ids = eng.ids.to_numpy().T
    depth = eng.depth.to_numpy().T
    cam = orc.camera(eng.W2V, eng.V2W, eng.bias)
    lt = tina.Lighting().to_struct()
    fb = sphere.face_buffers()
    tri = orc.HostMesh(v, fb['norms'].cpu().numpy(), fb['coors'].cpu().numpy(), col)
    par = orc.particle_mesh(pos, rad, pcol, id_base=len(v))
    mine_t = (ids >= 0) & (ids < len(v))
    mine_p = ids >= len(v)
    assert mine_t.sum() > 3000 and mine_p.sum() > 300

    def oracle(name, mesh, flags):
        lo = lower_material(tina.Lamp(name), dict(attrs={}))
        return orc.shade(W, H, cam, mesh, lo.struct, lt, ids, flags)

    tflags = _abi.SMOOTH | _abi.TEX | _abi.VCOLOR
    pflags = _abi.PARTICLES | _abi.VCOLOR
    for cls, name in ((tina.NormalShader, 'normal'), (tina.PositionShader, 'pos'), (tina.ColorShader, 'color'),
                      (tina.TexcoordShader, 'texcoord')):
        ch = 2 if name == 'texcoord' else 3
        img = _img(H, W, ch)
        r.render_color(cls(img))
        pr.render_color(cls(img))
        got = _np(img)
        exp = oracle(name, tri, tflags) + oracle(name, par, pflags)
        assert np.abs(got - exp[..., :ch]).max() <= 1e-6, name
        assert np.all(got[~(mine_t | mine_p)] == 0)

    nimg = _img(H, W, 3)
    r.render_color(tina.NormalShader(nimg))
    pr.render_color(tina.NormalShader(nimg))
    n = _np(nimg)
    f32 = np.float32
    M = np.asarray(eng.W2V, f32)
    vn = [((M[i, 3] * f32(0) + M[i, 0] * n[..., 0]).astype(f32) + M[i, 1] * n[..., 1]).astype(f32) + M[i, 2] * n[..., 2]
          for i in range(3)]
    vn = np.stack([x.astype(f32) for x in vn], axis=-1)
    vn = vn * (f32(1) / np.sqrt(((vn[..., 0] ** 2 + vn[..., 1] ** 2) + vn[..., 2] ** 2).astype(f32)))[..., None]
    img = _img(H, W, 3)
    r.render_color(tina.ViewNormalShader(img))
    pr.render_color(tina.ViewNormalShader(img))
    covered = mine_t | mine_p
    assert np.abs(_np(img)[covered] - vn[covered]).max() <= 1e-6

    # view ray through each pixel centre in float64 (calc_viewdir, shader.py:81-93)
    X, Y = np.meshgrid(np.arange(W), np.arange(H))
    u = (X + 0.5) / W * 2 - 1
    w = (Y + 0.5) / H * 2 - 1
    Vi = np.linalg.inv(np.float64(eng.W2V))
    a = np.stack([u, w, -np.ones_like(u), np.ones_like(u)], -1) @ Vi.T
    b = np.stack([u, w, np.ones_like(u), np.ones_like(u)], -1) @ Vi.T
    d = b[..., :3] / b[..., 3:] - a[..., :3] / a[..., 3:]
    viewdir = -d / np.linalg.norm(d, axis=-1, keepdims=True)
    img = _img(H, W, 3)
    r.render_color(tina.ViewdirShader(img))
    assert np.abs(_np(img)[mine_t] - (viewdir * 0.5 + 0.5)[mine_t]).max() <= 1e-5
    img = _img(H, W, 1)
    r.render_color(tina.SimpleShader(img))
    assert np.abs(_np(img)[mine_t] - np.abs((n * viewdir).sum(-1))[mine_t]).max() <= 1e-5

    img = _img(H, W, 1)
    r.render_color(tina.DepthShader(img))
    pr.render_color(tina.DepthShader(img))
    assert np.array_equal(_np(img)[covered], depth[covered].astype(np.float32))
    img = _img(H, W, 3)
    r.render_color(tina.ConstShader(img, (0.25, 0.5, 0.75)))
    assert np.all(_np(img)[mine_t] == np.float32([0.25, 0.5, 0.75])) and np.all(_np(img)[~mine_t] == 0)
    img = _img(H, W, 1)
    r.render_color(tina.ChessboardShader(img, size=8))
    chess = np.where(((X + 0.5) // 8 + (Y + 0.5) // 8) % 2 == 1, 0.9, 0.4).astype(np.float32)
    assert np.array_equal(_np(img)[mine_t], chess[mine_t])


def test_probe_records_ids_and_texcoords(gpu):
    """ProbeShader (tina/probe.py:18-20) in a ShaderGroup with the colour shader:
    elmid is the local face index and texcoord the interpolated texcoord of every
    pixel the textured sphere won (the binary32 oracle's fragment value), -1 and
    0 elsewhere; an untextured second object leaves texcoord zero on its pixels."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd.matr import lower_material
    W, H = 200, 150
    print('case: probe ids and texcoords 200x150', flush=True)
    eng = tina.Engine(W, H)
    eng.set_camera(tina.lookat([0, 0, 0], [0.3, 0.4, 3.0]), tina.perspective(50, W / H, 0.1, 20))
    sphere = tina.PrimitiveMesh.sphere(lons=20, lats=10)
    fb = sphere.face_buffers()
    v = fb['verts'].cpu().numpy()
    rng = np.random.default_rng(8)
    soup = np.concatenate([v[:40] + 0.01, rng.uniform(-1.2, 1.2, (60, 3, 3))]).astype(np.float32)
    flat = tina.SimpleMesh()
    flat.set_face_verts(soup)
    r1 = tina.TriangleRaster(eng, smoothing=True, texturing=True)
    r2 = tina.TriangleRaster(eng)
    img = _img(H, W, 3)
    probe = tina.ProbeShader((W, H))
    lt = tina.Lighting()
    group = tina.ShaderGroup([tina.Shader(img, lt, tina.Diffuse(color=(0.5, 0.5, 0.5))), probe])
    eng.clear_depth()
    probe.clear_buffer()
    r1.set_object(sphere)
    r1.render_occup()
    r2.set_object(flat)
    r2.render_occup()
    r1.render_color(group)
    el1 = _np(probe.elmid).copy()
    tc1 = _np(probe.texcoord).copy()
    r2.render_color(group)
    torch.cuda.synchronize()
    ids = eng.ids.to_numpy().T
    cam = orc.camera(eng.W2V, eng.V2W, eng.bias)
    od, oi = orc.raster(W, H, cam, v)
    od, oi = orc.raster(W, H, cam, soup, id_base=len(v), depth=od, occup=oi)
    assert np.array_equal(ids, oi)
    on1 = (oi >= 0) & (oi < len(v))
    on2 = oi >= len(v)
    assert on1.sum() > 2000 and on2.sum() > 300
    lo = lower_material(tina.Lamp('texcoord'), dict(attrs={}))
    tri = orc.HostMesh(v, fb['norms'].cpu().numpy(), fb['coors'].cpu().numpy(), None)
    exp = orc.shade(W, H, cam, tri, lo.struct, tina.Lighting().to_struct(), oi, _abi.SMOOTH | _abi.TEX)[..., :2]
    # after the sphere's pass: its pixels hold its ids and texcoords, the rest is untouched
    assert np.array_equal(el1, np.where(on1, oi, -1))
    assert np.abs(tc1 - exp).max() <= 1e-6
    assert np.all(tc1[~on1] == 0) and np.abs(tc1[on1]).max() > 0.5
    # after the soup's pass: its pixels hold its local ids and zero texcoords
    el = _np(probe.elmid)
    tc = _np(probe.texcoord)
    assert np.array_equal(el, np.where(on2, oi - len(v), np.where(on1, oi, -1)))
    assert np.abs(tc[on1] - exp[on1]).max() <= 1e-6
    assert np.all(tc[~on1] == 0)
