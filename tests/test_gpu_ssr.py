"""Screen-space reflections (Scene(ssr=True), tina/postp/ssr.py:6-103) on the
GPU: the reflection buffer equals the oracle's SSR on the same depth, id,
normal and image buffers, the applied image equals the oracle's apply of it,
and the mirror-floor scene reflects the wall as the float64 construction says."""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu


def _capture(scene):
    """Wraps scene.ssr.apply to keep the image before it and the reflection buffer."""
    got = {}
    apply = scene.ssr.apply

    def wrapped(image):
        got['pre'] = image.storage.detach().cpu().numpy().copy()
        got['buf'] = scene.ssr.img.detach().cpu().numpy().copy()
        apply(image)
    scene.ssr.apply = wrapped
    return got


def test_ssr_matches_oracle_on_the_same_buffers(gpu):
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    from taichi_three_amd.matr import lower_material
    print('case: ssr scene 320x240: sphere (Blinn), PBR soup, mirror-ish floor, particles', flush=True)
    W, H = 320, 240
    rng = np.random.default_rng(12)
    scene = tina.Scene(W, H, ssr=True, tonemap=False, smoothing=True, maxfaces=20000)
    scene.engine.set_camera(tina.lookat([0, -0.2, 0], [0.3, 1.2, 3.4]), tina.perspective(50, W / H, 0.1, 30))
    sph = tina.PrimitiveMesh.sphere(lons=24, lats=12, rad=0.6)
    sph_col = rng.uniform(0, 1, sph.face_buffers()['verts'].shape).astype(np.float32)
    sph.set_face_colors(sph_col)
    soup = wl.triangle_soup(rng, 3000, np.log(0.08), 0.4) * np.float32(1.3)
    soup[:, :, 1] += np.float32(0.2)
    soup_n = wl.face_normals(soup)
    soup_c = rng.uniform(0, 1, soup.shape).astype(np.float32)
    rough = rng.uniform(0.05, 1.0, len(soup)).astype(np.float32)
    metal = (rng.random(len(soup)) < 0.5).astype(np.float32)
    m2 = tina.SimpleMesh()
    m2.set_face_verts(soup)
    m2.set_face_norms(soup_n)
    m2.set_face_colors(soup_c)
    m2.set_face_attr('roughness', rough)
    m2.set_face_attr('metallic', metal)
    # the clip rule drops faces with every vertex off screen: the floor is a fine grid
    from test_ssr_oracle import _grid
    floor_v = _grid((-3, -0.7, 2), (6, 0, 0), (0, 0, -5), 16)
    floor_n = np.broadcast_to(np.array([0, 1, 0], np.float32), floor_v.shape).copy()
    fl = tina.SimpleMesh()
    fl.set_face_verts(floor_v)
    fl.set_face_norms(floor_n)
    pars = tina.SimpleParticles(200)
    ppos = rng.uniform(-1.2, 1.2, (200, 3)).astype(np.float32)
    pars.set_particles(ppos)
    pars.set_particle_radii(np.full(200, 0.05, np.float32))
    mats = [tina.Blinn(color='color', shineness=32),
            tina.PBR(basecolor='color', metallic=tina.FaceAttr('metallic'), roughness=tina.FaceAttr('roughness')),
            tina.PBR(basecolor=(0.6, 0.6, 0.7), metallic=0.9, roughness=0.1),
            tina.Diffuse(color=(0.8, 0.3, 0.2))]
    for obj, m in zip((sph, m2, fl, pars), mats):
        scene.add_object(obj, m)
    got = _capture(scene)
    scene.render()
    torch.cuda.synchronize()
    ids = scene.engine.ids.to_numpy().T.copy()
    depth = scene.engine.depth.to_numpy().T.copy()
    norm = scene.norm_buffer.storage.cpu().numpy()
    final = scene.image.storage.cpu().numpy()
    cam = orc.camera(scene.engine.W2V, scene.engine.V2W)
    sv = sph.face_buffers()['verts'].cpu().numpy()
    host = [(orc.HostMesh(sv, None, None, sph_col), {}),
            (orc.HostMesh(soup, soup_n, None, soup_c), {'roughness': rough, 'metallic': metal}),
            (orc.HostMesh(floor_v, floor_n), {}),
            (orc.particle_mesh(ppos, 0.05), {})]
    out = np.zeros((H, W, 4), np.float32)
    base = 0
    for (hm, attrs), m in zip(host, mats):
        hm.struct.id_base = base
        hm.struct.nfaces = len(hm.arrays[0])
        lo = lower_material(m, dict(attrs=attrs))
        orc.ssr(W, H, cam, hm, lo.struct, depth, ids, norm, got['pre'], out=out)
        base += hm.struct.nfaces
    for lo_id, hi_id in ((0, len(sv)), (len(sv), len(sv) + len(soup)),
                         (len(sv) + len(soup), len(sv) + len(soup) + len(floor_v))):
        assert ((ids >= lo_id) & (ids < hi_id)).sum() > 500
    hits = out[..., 3] > 0
    assert hits.sum() > 2000
    d = np.abs(got['buf'] - out).max(axis=2)
    bad = d > 1e-5
    print(f'ssr buffer: max diff {d.max():.2e}, {int(bad.sum())} pixels above 1e-5, {int(hits.sum())} with hits',
          flush=True)
    assert bad.sum() <= 1e-3 * W * H
    assert np.abs(final - orc.ssr_apply(got['pre'], got['buf'])).max() <= 1e-6


def test_mirror_floor_on_the_gpu(gpu):
    """The float64 construction of the oracle test, rendered by the GPU path: a
    mirror floor (no lights, so its own colour is 0) in front of an emissive wall."""
    import torch
    import taichi_three_amd as tina
    from test_ssr_oracle import _scene, _classify
    print('case: ssr mirror floor 160x120', flush=True)
    W, H, W2V, V2W, floor, wall = _scene()
    C = np.array([0.2, 0.5, 0.8], np.float32)
    scene = tina.Scene(W, H, ssr=True, tonemap=False, smoothing=True, maxfaces=len(floor))
    scene.engine.W2V, scene.engine.V2W = W2V.astype(np.float32), V2W.astype(np.float32)
    scene.engine._push_camera()
    fl, wl_ = tina.SimpleMesh(), tina.SimpleMesh()
    fl.set_face_verts(floor)
    fl.set_face_norms(np.broadcast_to(np.array([0, 1, 0], np.float32), floor.shape).copy())
    wl_.set_face_verts(wall)
    wl_.set_face_norms(np.broadcast_to(np.array([0, 0, 1], np.float32), wall.shape).copy())
    scene.add_object(fl, tina.CookTorrance(roughness=0.0, fresnel=1.0))
    scene.add_object(wl_, tina.Lamp(color=tuple(C)))
    scene.lighting.clear_lights()
    scene.lighting.set_ambient_light((0, 0, 0))
    got = _capture(scene)
    scene.render()
    torch.cuda.synchronize()
    final = scene.image.storage.cpu().numpy()
    ids = scene.engine.ids.to_numpy().T
    on_floor, on_wall = (ids >= 0) & (ids < len(floor)), ids >= len(floor)
    cls = _classify(W, H, W2V, V2W)

    def box_all(mask):
        ok = np.zeros_like(mask)
        for y in range(2, H - 1):
            for x in range(2, W - 1):
                ok[y, x] = mask[y - 2:y + 2, x - 2:x + 2].all()
        return ok
    hit, miss, wall_in = box_all(cls == 1), box_all(cls == 0), box_all(on_wall)
    assert hit.sum() > 300 and miss.sum() > 200 and wall_in.sum() > 500
    assert np.all(got['pre'][on_floor] == 0)
    assert np.allclose(got['buf'][hit & on_floor][:, 3], 1.0)
    assert np.allclose(final[hit], C, rtol=1e-4, atol=0)
    assert np.all(final[miss] == 0)
    assert np.array_equal(final[wall_in], np.broadcast_to(C, (int(wall_in.sum()), 3)))


def test_ssr_with_textured_roughness(gpu):
    """texturing=True: the texcoord pre-shader fills the coordinate buffer and the
    floor's roughness is a Texture sampled there (tina/postp/ssr.py:55-80 reads
    the material at the pixel's texcoord); the reflection buffer equals the
    oracle's on the same depth, id, normal, texcoord and image buffers."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd.matr import lower_material
    from test_ssr_oracle import _grid
    print('case: ssr with a textured roughness 240x180', flush=True)
    W, H = 240, 180
    rng = np.random.default_rng(31)
    scene = tina.Scene(W, H, ssr=True, tonemap=False, texturing=True, maxfaces=4000)
    scene.engine.set_camera(tina.lookat([0, -0.2, 0], [0.2, 1.0, 3.0]), tina.perspective(50, W / H, 0.1, 30))
    floor_v = _grid((-3, -0.7, 2), (6, 0, 0), (0, 0, -5), 16)
    floor_n = np.broadcast_to(np.array([0, 1, 0], np.float32), floor_v.shape).copy()
    floor_c = rng.uniform(0, 1, floor_v.shape[:2] + (2,)).astype(np.float32)
    wall_v = _grid((-3, -0.7, -1.5), (6, 0, 0), (0, 1.5, 0), 12)
    wall_n = np.broadcast_to(np.array([0, 0, 1], np.float32), wall_v.shape).copy()
    wall_c = np.zeros(wall_v.shape[:2] + (2,), np.float32)
    fl, wa = tina.SimpleMesh(), tina.SimpleMesh()
    fl.set_face_verts(floor_v)
    fl.set_face_norms(floor_n)
    fl.set_face_coors(floor_c)
    wa.set_face_verts(wall_v)
    wa.set_face_norms(wall_n)
    wa.set_face_coors(wall_c)
    tex = tina.Texture(rng.uniform(0.02, 0.6, (8, 8)).astype(np.float32))
    floor_mat = tina.CookTorrance(roughness=tex, fresnel=0.9)
    wall_mat = tina.Lamp(color=(0.9, 0.6, 0.3))
    scene.add_object(fl, floor_mat)
    scene.add_object(wa, wall_mat)
    got = _capture(scene)
    scene.render()
    torch.cuda.synchronize()
    ids = scene.engine.ids.to_numpy().T.copy()
    depth = scene.engine.depth.to_numpy().T.copy()
    norm = scene.norm_buffer.storage.cpu().numpy()
    coor = scene.coor_buffer.storage.cpu().numpy()
    cam = orc.camera(scene.engine.W2V, scene.engine.V2W)
    out = np.zeros((H, W, 4), np.float32)
    host = []
    base = 0
    for hm, m in ((orc.HostMesh(floor_v, floor_n, floor_c), floor_mat),
                  (orc.HostMesh(wall_v, wall_n, wall_c), wall_mat)):
        hm.struct.id_base = base
        hm.struct.nfaces = len(hm.arrays[0])
        lo = lower_material(m, dict(attrs={}))
        # the oracle samples the texture from a host copy of the same values
        for i in range(lo.struct.nparams):
            p = lo.struct.params[i]
            if p.kind == _abi.P_TEXTURE:
                h = np.ascontiguousarray(tex.tensor.cpu().numpy())
                host.append(h)
                p.ptr = h.ctypes.data
        orc.ssr(W, H, cam, hm, lo.struct, depth, ids, norm, got['pre'], coor=coor, out=out)
        base += hm.struct.nfaces
    on_floor = (ids >= 0) & (ids < len(floor_v))
    assert on_floor.sum() > 5000 and (ids >= len(floor_v)).sum() > 2000
    assert np.abs(coor[on_floor]).max() > 0
    assert (out[..., 3] > 0).sum() > 1000
    d = np.abs(got['buf'] - out).max(axis=2)
    print(f'ssr textured: max diff {d.max():.2e}, {int((d > 1e-5).sum())} pixels above 1e-5', flush=True)
    assert (d > 1e-5).sum() <= 1e-3 * W * H
