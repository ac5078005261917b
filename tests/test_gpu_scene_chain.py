"""Every raster and every post pass in one scene: a smooth-shaded mesh, a wire
cage and particles with ssao, blooming and fxaa.  The frame equals the
reference's chain (tina/scene/raster.py:168-214) applied by the oracle to the
same scene's linear frame: fxaa(aces(bloom(ssao_apply(linear, ao)))), with the
occlusion the scene computed (itself checked against the oracle elsewhere)."""
import numpy as np
import pytest

from oracle import orc
from oracle import postp as opp

pytestmark = pytest.mark.gpu


def _objects():
    import taichi_three_amd as tina
    sphere = tina.PrimitiveMesh.sphere()
    cage = tina.MeshToWire(tina.MeshTransform(tina.PrimitiveMesh.sphere(lons=10, lats=6)))
    cage.set_transform(np.diag([1.3, 1.3, 1.3, 1.0]))
    rng = np.random.default_rng(8)
    pars = tina.SimpleParticles(600)
    pars.set_particles(rng.uniform(-1.8, 1.8, (600, 3)).astype(np.float32))
    pars.set_particle_radii(rng.uniform(0.04, 0.1, 600).astype(np.float32))
    pars.set_particle_colors(rng.uniform(0.2, 4.0, (600, 3)).astype(np.float32))   # some above the bloom threshold
    return [(sphere, tina.PBR(roughness=0.3, metallic=0.5)), (cage, None), (pars, tina.Diffuse())]


def _render(**opts):
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd.util.matrix import lookat, perspective
    W, H = 320, 240
    scene = tina.Scene(W, H, smoothing=True, maxpars=600, linecolor=(2.0, 1.5, 0.2), **opts)
    for o, m in _objects():
        scene.add_object(o, m)
    scene.set_camera(lookat([0, 0, 0], [0.5, 0.7, 3.6]), perspective(50, W / H, 0.1, 50))
    scene.render()
    torch.cuda.synchronize()
    return scene, scene.img.to_numpy().transpose(1, 0, 2).copy()


def test_all_passes_follow_the_reference_chain(gpu):
    print('case: mesh + wires + particles with ssao, blooming, fxaa at 320x240', flush=True)
    _, lin = _render(tonemap=False)
    scene, img = _render(ssao=True, blooming=True, fxaa=True)
    ao = scene.ssao.img.cpu().numpy()
    assert (ao > 0).sum() > 100 and (lin > 1).any()
    bl = scene.blooming
    want = opp.fxaa(orc.aces(opp.bloom(opp.ssao_apply(lin, ao, 4, box=True), gwei=bl.gwei.cpu().numpy(),
                                       radius=bl.radius)))
    nan = np.isnan(img)
    assert np.array_equal(nan, np.isnan(want))
    assert np.abs(img[~nan] - want[~nan]).max() <= 1e-5


def test_all_passes_with_taa_average(gpu):
    import torch
    print('case: the same scene with taa, 3 frames', flush=True)
    scene, _ = _render(ssao=True, blooming=True, fxaa=True, taa=True)
    frames = [scene.image.to_numpy().copy()]
    for _ in range(2):
        scene.render()
        torch.cuda.synchronize()
        frames.append(scene.image.to_numpy().copy())
    acc = scene.img.to_numpy()
    assert scene.accum.count == 3 and np.isfinite(acc).all()
    assert np.abs(acc - np.mean(frames, axis=0)).max() <= 1e-5


def test_normals_under_the_wire_cage_follow_the_reference(gpu):
    """The normal buffer and the occlusion of the mesh + wire cage + particles
    scene against the oracle's per-object construction: in the reference each
    object's render_color runs the NormalShader on the pixels the object wins at
    its turn, and a wire drawn later only blends its colour, so a wire pixel keeps
    the normal of the surface it was drawn over (raster.py:183-187,
    wireframe.py:88-95, shader.py:9-18)."""
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd.matr import lower_material
    print('case: normals under a wire cage, ssao 320x240', flush=True)
    scene, _ = _render(ssao=True)
    eng = scene.engine
    W, H = int(scene.res[0]), int(scene.res[1])
    cam = orc.camera(eng.W2V, eng.V2W, eng.bias)
    lt = scene.lighting.to_struct()
    lamp = lower_material(tina.Lamp('normal'), dict(attrs={})).struct
    sphere, cage, pars = list(scene.objects)
    od = np.full((H, W), 2 ** 30, np.int32)
    oi = np.full((H, W), -1, np.int32)
    norm = np.zeros((H, W, 3), np.float32)
    fb = sphere.face_buffers()
    sv, sn = fb['verts'].cpu().numpy(), fb['norms'].cpu().numpy()
    od, oi = orc.raster(W, H, cam, sv, id_base=0, flags=_abi.CULL | _abi.CLIP, depth=od, occup=oi)
    norm = orc.shade(W, H, cam, orc.HostMesh(sv, sn), lamp, lt, oi, _abi.SMOOTH)
    base = len(sv)
    wv = cage.face_buffers()['verts'].cpu().numpy()
    wflags = _abi.CLIP if scene.wireframe_raster.clipping else 0
    od, oi = orc.raster_wires(W, H, cam, wv, id_base=base, flags=wflags, depth=od, occup=oi)
    wire_px = (oi >= base) & (oi < base + len(wv))
    base += len(wv)
    pb = pars.particle_buffers()
    pp, pr = pb['pos'].cpu().numpy(), pb['radius'].cpu().numpy()
    od, oi = orc.raster_particles(W, H, cam, pp, pr, id_base=base, flags=_abi.CLIP, depth=od, occup=oi)
    mine = oi >= base
    pn = orc.shade(W, H, cam, orc.particle_mesh(pp, pr, id_base=base), lamp, lt, oi, _abi.PARTICLES)
    norm = np.where(mine[..., None], pn, norm)
    assert np.array_equal(eng.ids.to_numpy().T, oi)
    got = scene.norm_buffer.storage.cpu().numpy()
    over = wire_px & (np.abs(norm).sum(axis=2) > 0)
    assert over.sum() > 50                   # wire pixels drawn over the sphere
    assert np.abs(got - norm).max() <= 1e-6
    ao = opp.ssao(eng.W2V, eng.V2W, eng.bias, od, norm, scene.ssao.samples.cpu().numpy(),
                  scene.ssao.rotations.cpu().numpy(), radius=scene.ssao.radius, thresh=scene.ssao.thresh,
                  factor=scene.ssao.factor)
    assert np.abs(scene.ssao.img.cpu().numpy() - ao).max() <= 1e-6
