"""Scene(ssao=True) on the HIP path vs the oracle: the normal buffer the
NormalShader pass writes equals the binary32 oracle's shading normal (a Lamp
material of Input('normal') shaded by the oracle, no tone mapping), the
occlusion equals oracle/postp.ssao on the same depth, normals, samples and
rotations, and the final image equals ACES(ssao_apply(linear image)) of the
oracle within 1e-5."""
import numpy as np
import pytest

from oracle import orc
from oracle import postp as opp

pytestmark = pytest.mark.gpu

COLOR_TOL_F32 = 1e-5


def _scene(W, H, taa=False):
    import taichi_three_amd as tina
    from taichi_three_amd.util.matrix import lookat, perspective
    scene = tina.Scene(W, H, ssao=True, smoothing=True, taa=taa, maxpars=400)
    sphere = tina.PrimitiveMesh.sphere()                  # per-vertex normals of the UV sphere
    floor = tina.MeshTransform(tina.PrimitiveMesh.sphere(lons=4, lats=2))
    floor.set_transform(np.diag([3.0, 0.05, 3.0, 1.0]) @ np.array(
        [[1, 0, 0, 0], [0, 1, 0, -21.0], [0, 0, 1, 0], [0, 0, 0, 1]], dtype=float))
    rng = np.random.default_rng(2)
    pars = tina.SimpleParticles(400)
    pars.set_particles(rng.uniform(-1.6, 1.6, (400, 3)).astype(np.float32) * [1, 0.3, 1] + [0, -0.7, 0])
    pars.set_particle_radii(rng.uniform(0.05, 0.12, 400).astype(np.float32))
    scene.add_object(sphere, tina.Classic())
    scene.add_object(floor, tina.Diffuse())
    scene.add_object(pars, tina.Diffuse())
    scene.set_camera(lookat([0, -0.3, 0], [0.6, 1.4, 3.4]), perspective(45, W / H, 0.1, 50))
    return scene, (sphere, floor, pars)


def test_ssao_scene_matches_oracle(gpu):
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd.matr import lower_material
    W, H = 400, 300
    print('case: ssao sphere + floor + particles 400x300', flush=True)
    scene, (sphere, floor, pars) = _scene(W, H)
    scene.render()
    torch.cuda.synchronize()
    eng = scene.engine
    ids = eng.ids.to_numpy().T
    depth = eng.depth.to_numpy().T
    img = scene.img.to_numpy().transpose(1, 0, 2)
    norm = scene.norm_buffer.storage.cpu().numpy()
    ao = scene.ssao.img.cpu().numpy()
    cam = orc.camera(eng.W2V, eng.V2W, eng.bias)
    lt = scene.lighting.to_struct()

    # ids and depth of the three objects, in order
    od = np.full((H, W), 2 ** 30, np.int32)
    oi = np.full((H, W), -1, np.int32)
    meshes = []
    base = 0
    for obj in (sphere, floor):
        fb = obj.face_buffers()
        v, n = fb['verts'].cpu().numpy(), fb['norms'].cpu().numpy()
        od, oi = orc.raster(W, H, cam, v, id_base=base, flags=_abi.CULL | _abi.CLIP, depth=od, occup=oi)
        meshes.append((orc.HostMesh(v, n, None, None, id_base=base), _abi.SMOOTH))
        base += len(v)
    pb = pars.particle_buffers()
    pp, pr, pc = (pb[k].cpu().numpy() for k in ('pos', 'radius', 'colors'))
    od, oi = orc.raster_particles(W, H, cam, pp, pr, id_base=base, flags=_abi.CLIP, depth=od, occup=oi)
    meshes.append((orc.particle_mesh(pp, pr, pc, id_base=base), _abi.PARTICLES | _abi.VCOLOR))
    assert np.array_equal(ids, oi) and np.array_equal(depth, od)

    # the normal buffer: a Lamp(Input('normal')) material shades to the shading normal
    lamp = lower_material(tina.Lamp('normal'), dict(attrs={}))
    exp_n = np.zeros((H, W, 3), np.float32)
    for m, fl in meshes:
        exp_n += orc.shade(W, H, cam, m, lamp.struct, lt, oi, fl & ~_abi.VCOLOR)
    assert np.abs(norm - exp_n).max() <= 1e-6

    # the occlusion from the same inputs
    samples = scene.ssao.samples.cpu().numpy()
    rot = scene.ssao.rotations.cpu().numpy()
    exp_ao = opp.ssao(eng.W2V, eng.V2W, eng.bias, depth, norm, samples, rot, radius=scene.ssao.radius,
                      thresh=scene.ssao.thresh, factor=scene.ssao.factor)
    assert np.abs(ao - exp_ao).max() <= 1e-6
    assert (ao > 0).sum() > 500                      # the sphere's foot and the particles occlude the floor

    # the image: linear shading, occlusion, then ACES (raster.py:189-203)
    lin = np.zeros((H, W, 3), np.float32)
    for k, ((m, fl), mat) in enumerate(zip(meshes, (tina.Classic(), tina.Diffuse(), tina.Diffuse()))):
        lo = lower_material(mat, dict(attrs={}))
        part = orc.shade(W, H, cam, m, lo.struct, lt, oi, fl | (_abi.BG if k == 0 else 0))
        mine = (oi >= m.struct.id_base) & (oi < m.struct.id_base + m.struct.nfaces)
        lin = np.where((mine | ((oi < 0) & (k == 0)))[..., None], part, lin)
    exp = orc.aces(opp.ssao_apply(lin, ao, 4, box=True))
    nan = np.isnan(img)
    assert np.array_equal(nan, np.isnan(exp))
    assert np.abs(img[~nan] - exp[~nan]).max() <= COLOR_TOL_F32


def test_ssao_with_taa_runs_and_averages(gpu):
    import torch
    W, H = 200, 150
    print('case: ssao + taa 200x150, 4 frames', flush=True)
    scene, _ = _scene(W, H, taa=True)
    for _ in range(4):
        scene.render()
    torch.cuda.synchronize()
    ao = scene.ssao.img.cpu().numpy()
    img = scene.img.to_numpy()
    assert scene.ssao.nsamples == 16 and np.all((ao >= 0) & (ao <= 1)) and np.isfinite(img).all()
