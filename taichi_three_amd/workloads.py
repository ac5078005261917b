"""Synthetic scenes of the benchmark workloads (seeded NumPy generators).

  plain_512     one 512x512 frame: UV sphere (64x32 segments, r = 1, 3,968 faces)
                + 200 random equilateral triangles (centres U[-1,1]^3, edge
                lognormal(ln 0.05, 0.5), uniform random orientation); smooth
                normals (analytic on the sphere, the face normal on the soup);
                one directional light; Blinn-Phong m = 32 over per-vertex albedo
                U[0,1]^3; camera at distance 3, fov 60.  Seed 0.
  soup_1080p    1920x1080, 2,000,000-face soup (centres U[-1,1]^3, edge
                lognormal(ln 0.01, 0.7), random orientation); face normals;
                PBR with per-face roughness U[0.05,1] and metallic Bernoulli(0.3),
                per-vertex albedo U[0,1]^3; four point lights on a radius-4 sphere
                with intensity U[0.5,2]; camera at distance 3, fov 60, aspect 16/9.  Seed 1.
  batch_views   64 scenes of the 512x512 workload's triangle and material
                generators, 1k-500k faces (log-uniform), resolutions from
                {256^2, 512^2, 1024^2, 1920x1080}, cameras on the radius-3 sphere.  Seed 2.

Arrays are float32 [T, 3, 3] (face, corner, xyz) like TriangleRaster.verts.
"""
import numpy as np

from .mesh import PrimitiveMesh
from .util.matrix import camera_matrices, lookat, perspective


def random_rotations(rng, n):
    """Uniform random rotation matrices from uniform unit quaternions (Shoemake)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def triangle_soup(rng, n, mu, sigma, lo=-1.0, hi=1.0):
    """n equilateral triangles: centre U[lo,hi]^3, edge lognormal(mu, sigma), random rotation.
    Returns float32 [n, 3, 3]."""
    centres = rng.uniform(lo, hi, size=(n, 3))
    edge = rng.lognormal(mu, sigma, size=n)
    ang = np.array([0.0, 2 * np.pi / 3, 4 * np.pi / 3])
    base = np.stack([np.cos(ang), np.sin(ang), np.zeros(3)], axis=1) / np.sqrt(3.0)  # circumradius 1/sqrt3
    R = random_rotations(rng, n)
    tri = np.einsum('nij,kj->nki', R, base) * edge[:, None, None] + centres[:, None, :]
    return tri.astype(np.float32)


def face_normals(v):
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return np.repeat(n[:, None, :], 3, axis=1).astype(np.float32)


def default_camera(aspect=1.0, dist=3.0):
    view = lookat(back=(0, 0, dist))
    proj = perspective(60, aspect, 0.05, 500)
    return view, proj


class Workload(dict):
    """verts, norms, colors, attrs, camera (W2V, V2W), lights, res, material name."""
    __getattr__ = dict.__getitem__


def plain_512(seed=0, res=512, nrand=200):
    rng = np.random.default_rng(seed)
    sph = PrimitiveMesh.sphere_faces(lons=64, lats=32, rad=1).astype(np.float32)
    sv, sn = sph[:, :, 0], sph[:, :, 1]
    rv = triangle_soup(rng, nrand, np.log(0.05), 0.5)
    rn = face_normals(rv)
    verts = np.concatenate([sv, rv]).astype(np.float32)
    norms = np.concatenate([sn, rn]).astype(np.float32)
    colors = rng.uniform(0, 1, size=verts.shape).astype(np.float32)
    view, proj = default_camera(1.0)
    W2V, V2W = camera_matrices(view, proj)
    return Workload(name='plain_512', res=(res, res), verts=verts, norms=norms, colors=colors, attrs={},
                    view=view, proj=proj, W2V=W2V, V2W=V2W, smoothing=True,
                    lights=[dict(dir=(1, 2, 3), color=(0.9, 0.9, 0.9))], ambient=(0.1, 0.1, 0.1),
                    material='blinn')


def soup_1080p(seed=1, res=(1920, 1080), nfaces=2_000_000):
    rng = np.random.default_rng(seed)
    verts = triangle_soup(rng, nfaces, np.log(0.01), 0.7)
    colors = rng.uniform(0, 1, size=verts.shape).astype(np.float32)
    rough = rng.uniform(0.05, 1.0, size=nfaces).astype(np.float32)
    metal = (rng.random(nfaces) < 0.3).astype(np.float32)
    lp = rng.normal(size=(4, 3))
    lp = 4.0 * lp / np.linalg.norm(lp, axis=1, keepdims=True)
    li = rng.uniform(0.5, 2.0, size=4)
    view, proj = default_camera(res[0] / res[1])
    W2V, V2W = camera_matrices(view, proj)
    return Workload(name='soup_1080p', res=tuple(res), verts=verts, norms=None, colors=colors,
                    attrs={'roughness': rough, 'metallic': metal}, view=view, proj=proj, W2V=W2V, V2W=V2W,
                    smoothing=False,
                    lights=[dict(pos=tuple(p), color=(float(i),) * 3) for p, i in zip(lp, li)],
                    ambient=(0.1, 0.1, 0.1), material='pbr_face')


def ibl_1080p(seed=3, res=(1920, 1080), nfaces=1_000_000, sky=(512, 256), precision=32):
    """The 1080p soup generator under image-based lighting: PBR with per-face
    roughness/metallic lit by the procedural HDR sky (SkyboxLighting) plus one
    directional light along the brightest sun."""
    from .skybox import SUNS
    w = soup_1080p(seed=seed, res=res, nfaces=nfaces)
    sun = np.asarray(SUNS[0][0], dtype=np.float64)
    w.update(name='ibl_1080p', ibl=True, sky=tuple(sky), ibl_precision=int(precision),
             lights=[dict(dir=tuple(sun / np.linalg.norm(sun)), color=(2.0, 2.0, 2.0))], ambient=(0.0, 0.0, 0.0))
    return w


BATCH_RESOLUTIONS = ((256, 256), (512, 512), (1024, 1024), (1920, 1080))


def batch_views(n=64, seed=2, faces=(1000, 500_000), resolutions=BATCH_RESOLUTIONS):
    """n independent scenes with the triangle and material generators of the
    512x512 workload: face counts log-uniform in `faces`, equilateral triangles
    (centres U[-1,1]^3, edge lognormal(ln 0.05, 0.5), random orientation) with
    their face normal at every corner (smooth shading), per-vertex albedo
    U[0,1]^3, Blinn-Phong m = 32 under one directional light (1, 2, 3) of
    colour 0.9 and ambient 0.1; resolution drawn uniformly from `resolutions`,
    camera on the radius-3 sphere looking at the origin (up +y)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        T = int(round(np.exp(rng.uniform(np.log(faces[0]), np.log(faces[1])))))
        res = tuple(resolutions[int(rng.integers(len(resolutions)))])
        srng = np.random.default_rng(int(rng.integers(2 ** 31)))
        verts = triangle_soup(srng, T, np.log(0.05), 0.5)
        norms = face_normals(verts)
        colors = srng.uniform(0, 1, size=verts.shape).astype(np.float32)
        d = rng.normal(size=3)
        d = 3.0 * d / np.linalg.norm(d)
        view = lookat(pos=(0, 0, 0), back=tuple(d))
        proj = perspective(60, res[0] / res[1], 0.05, 500)
        W2V, V2W = camera_matrices(view, proj)
        out.append(Workload(name='batch_view', res=res, verts=verts, norms=norms, colors=colors, attrs={},
                            view=view, proj=proj, W2V=W2V, V2W=V2W, smoothing=True,
                            lights=[dict(dir=(1, 2, 3), color=(0.9, 0.9, 0.9))], ambient=(0.1, 0.1, 0.1),
                            material='blinn'))
    return out


def make_material(name):
    from . import matr
    if name == 'blinn':
        return matr.Blinn(color='color', shineness=32)
    if name == 'classic':
        return matr.Classic(color='color', shineness=32)
    if name == 'pbr_face':
        return matr.PBR(basecolor='color', metallic=matr.FaceAttr('metallic'),
                        roughness=matr.FaceAttr('roughness'))
    raise ValueError(name)


def make_lighting(w):
    from .lighting import Lighting
    lt = Lighting()
    for l in w.lights:
        lt.add_light(**l)
    lt.set_ambient_light(w.ambient)
    return lt


def build_scene(w, **options):
    """A tina.Scene holding workload w (device upload happens here, once)."""
    from .scene import Scene
    from .mesh import SimpleMesh
    W, H = w.res
    T = len(w.verts)
    if w.get('ibl'):
        from .skybox import Skybox
        options = dict(options, ibl=True, skybox=Skybox.procedural(*w.sky), ibl_precision=w.ibl_precision)
    scene = Scene(W, H, smoothing=w.smoothing, maxfaces=T, **options)
    scene.engine.set_camera(w.view, w.proj)
    scene.lighting.clear_lights()
    for l in w.lights:
        scene.lighting.add_light(**l)
    scene.lighting.set_ambient_light(w.ambient)
    mesh = SimpleMesh(maxfaces=T)
    mesh.set_face_verts(w.verts)
    if w.norms is not None:
        mesh.set_face_norms(w.norms)
    if w.colors is not None:
        mesh.set_face_colors(w.colors)
    for k, v in w.attrs.items():
        mesh.set_face_attr(k, v)
    scene.add_object(mesh, make_material(w.material))
    return scene, mesh
