"""Pins the particle part of the CPU oracle (orc_raster_particles and the
TINA_PARTICLES branch of orc_shade) against what the geometry fixes:

  coverage  - orthographic camera with a non-square screen: the covered pixels
              are exactly the pixel centres within the radius of the centre
              (float64 distance in the world), the depth is int(z_view * 2**30)
              for view depths chosen exact in binary32
            - perspective camera: every pixel whose ray meets the plane of the
              centre's view depth inside the sphere's radius (float64 ray
              casting with numpy's own matrix inverse) is covered, away from
              a thin rim
            - ties to the earlier particle, the nearer particle wins, clipping
              of z_view outside [-1, 1]
  shading   - under the orthographic camera the shaded point is the sphere's
              front point behind the pixel, pos = c + R n with
              n = (d_x, d_y, sqrt(1 - |d|^2)) in camera axes: the binary64
              image equals orc_shade_point at that closed-form point
"""
import numpy as np
import pytest

from oracle import orc
from taichi_three_amd import _abi
from taichi_three_amd.util.matrix import camera_matrices, lookat, perspective

W, H = 80, 56
SX, SY, SZ = 0.5, 0.625, -0.5   # view = (SX x, SY y, SZ z): the camera looks down world -z


def _ortho_cam(bias=(0.5, 0.5)):
    W2V = np.diag([SX, SY, SZ, 1.0]).astype(np.float32)
    V2W = np.diag([1 / SX, 1 / SY, 1 / SZ, 1.0]).astype(np.float32)
    return orc.camera(W2V, V2W, bias), W2V, V2W


def _pixel_world_ortho(V2W, bias=(0.5, 0.5)):
    X, Y = np.meshgrid(np.arange(W), np.arange(H))
    u = (X + bias[0]) / W * 2 - 1
    v = (Y + bias[1]) / H * 2 - 1
    return u * np.float64(V2W[0, 0]), v * np.float64(V2W[1, 1])


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_orthographic_disc_is_exact(seed):
    rng = np.random.default_rng(seed)
    cam, W2V, V2W = _ortho_cam()
    xw, yw = _pixel_world_ortho(V2W)
    for _ in range(6):
        c = np.array([rng.uniform(-1.8, 1.8), rng.uniform(-1.4, 1.4), -0.5], dtype=np.float32)  # z_view = 0.25
        R = np.float32(rng.uniform(0.05, 0.6))
        depth, occ = orc.raster_particles(W, H, cam, c[None], [R])
        d2 = (xw - np.float64(c[0])) ** 2 + (yw - np.float64(c[1])) ** 2
        inside, rim = d2 <= np.float64(R) ** 2, np.abs(d2 - np.float64(R) ** 2) <= 1e-5 * np.float64(R) ** 2
        got = occ == 0
        assert np.array_equal(got[~rim], inside[~rim])
        assert np.all(depth[got] == 2 ** 28)          # 0.25 * 2**30
        assert np.all(depth[~got] == 2 ** 30) and np.all(occ[~got] == -1)


def test_perspective_disc_matches_ray_casting():
    view = lookat([0, 0, 0], [0.3, 0.2, 3.0])
    proj = perspective(50, W / H, 0.5, 20)
    W2V, V2W = camera_matrices(view, proj)
    cam = orc.camera(W2V, V2W)
    rng = np.random.default_rng(7)
    M = np.asarray(W2V, dtype=np.float32).astype(np.float64)
    Minv = np.linalg.inv(M)
    X, Y = np.meshgrid(np.arange(W), np.arange(H))
    u, v = (X + 0.5) / W * 2 - 1, (Y + 0.5) / H * 2 - 1
    checked = 0
    for _ in range(8):
        c = rng.uniform(-1.0, 1.0, 3).astype(np.float32)
        R = np.float32(rng.uniform(0.05, 0.3))
        depth, occ = orc.raster_particles(W, H, cam, c[None], [R])
        h = M @ np.append(np.float64(c), 1.0)
        zv = h[2] / h[3]
        P = np.stack([u, v, np.full_like(u, zv), np.ones_like(u)], axis=-1) @ Minv.T
        Pl = P[..., :3] / P[..., 3:]
        d2 = ((Pl - np.float64(c)) ** 2).sum(-1)
        r2 = np.float64(R) ** 2
        inside = d2 <= r2 * (1 - 1e-4)
        outside = d2 >= r2 * (1 + 1e-4)
        got = occ == 0
        # the reference's pixel box (particle.py:89-104) reaches at least the disc's extent along
        # the view axes for this camera: no pixel inside is lost
        assert np.all(got[inside]) and not np.any(got[outside])
        assert np.all(np.abs(depth[got].astype(np.int64) - zv * 2 ** 30) <= 512)  # binary32 z_view
        checked += int(inside.sum())
    assert checked > 200


def test_ties_go_to_the_earlier_particle_and_the_nearer_one_wins():
    cam, _, _ = _ortho_cam()
    c = np.array([[0.0, 0.0, -0.5], [0.0, 0.0, -0.5], [0.3, 0.0, -1.0]], dtype=np.float32)  # z_view .25 .25 .5
    depth, occ = orc.raster_particles(W, H, cam, c, [0.4, 0.4, 0.4], id_base=10)
    assert set(np.unique(occ)) == {-1, 10, 12}       # the twin (id 11) never shows
    nearer = orc.raster_particles(W, H, cam, c[::-1].copy(), [0.4, 0.4, 0.4], id_base=0)[1]
    # reversed order: the far particle is now first, the near twins follow; the near ones still win
    assert np.array_equal(occ == 12, nearer == 0) and np.array_equal(occ == 10, nearer == 1)


def test_clipping_of_view_depth():
    cam, _, _ = _ortho_cam()
    near = np.array([[0.0, 0.0, 2.4]], dtype=np.float32)    # z_view = -1.2
    far = np.array([[0.0, 0.0, -2.4]], dtype=np.float32)    # z_view = 1.2
    for c in (near, far):
        assert np.all(orc.raster_particles(W, H, cam, c, [0.5], flags=_abi.CLIP)[1] == -1)
    d, o = orc.raster_particles(W, H, cam, near, [0.5], flags=0)
    assert np.any(o == 0) and np.all(d[o == 0] == np.int32(np.float32(-1.2) * np.float32(2 ** 30)))
    # unclipped but beyond maxdepth: never beats the cleared depth (particle.py:114)
    assert np.all(orc.raster_particles(W, H, cam, far, [0.5], flags=0)[1] == -1)


def test_jittered_bias_moves_the_pixel_centres():
    bias = (0.125, 0.875)
    cam, _, V2W = _ortho_cam(bias)
    xw, yw = _pixel_world_ortho(V2W, bias)
    c = np.array([0.31, -0.22, -0.5], dtype=np.float32)
    R = np.float32(0.45)
    occ = orc.raster_particles(W, H, cam, c[None], [R])[1]
    d2 = (xw - np.float64(c[0])) ** 2 + (yw - np.float64(c[1])) ** 2
    rim = np.abs(d2 - np.float64(R) ** 2) <= 1e-5 * np.float64(R) ** 2
    assert np.array_equal((occ == 0)[~rim], (d2 <= np.float64(R) ** 2)[~rim])


def test_shading_is_the_front_point_of_the_sphere():
    import taichi_three_amd as tina
    from taichi_three_amd.lighting import Lighting
    from taichi_three_amd.matr import lower_material
    cam, W2V, V2W = _ortho_cam()
    xw, yw = _pixel_world_ortho(V2W)
    c = np.array([[0.2, -0.1, -0.5], [-1.1, 0.6, -0.5]], dtype=np.float32)
    R = np.array([0.7, 0.3], dtype=np.float32)
    col = np.array([[1.0, 0.5, 0.25], [0.2, 0.9, 0.4]], dtype=np.float32)
    depth, occ = orc.raster_particles(W, H, cam, c, R)
    lt = Lighting()
    lt.add_light(dir=[1, 2, 3], color=[0.9, 0.9, 0.9])
    lt.add_light(pos=[0.5, 0.5, 2.0], color=[0.6, 0.6, 0.6])
    lt.set_ambient_light([0.1, 0.1, 0.1])
    lts = lt.to_struct()
    mat = lower_material(tina.Classic(shineness=12, specular=0.6), dict(attrs={})).struct  # colour = Input('color')
    mesh = orc.particle_mesh(c, R, col)
    flags = _abi.PARTICLES | _abi.VCOLOR
    img64 = orc.shade(W, H, cam, mesh, mat, lts, occ, flags, f64=True)
    img32 = orc.shade(W, H, cam, mesh, mat, lts, occ, flags)
    viewdir = np.array([0.0, 0.0, 1.0])              # towards the camera (it looks down -z)
    n_checked = 0
    for Y, X in zip(*np.nonzero(occ >= 0)):
        i = occ[Y, X]
        d = np.array([xw[Y, X] - c[i, 0], yw[Y, X] - c[i, 1]]) / np.float64(R[i])
        q = 1 - d @ d
        if q < 1e-6:
            continue                                 # the rim, where sqrt(1 - |d|^2) loses its digits
        n = np.array([d[0], d[1], np.sqrt(q)])
        pos = np.float64(c[i]) + np.float64(R[i]) * n
        ref = orc.shade_point(mat, lts, pos, n, viewdir, slots=_slots_with_color(mat, col[i]), f64=True)
        assert np.allclose(img64[Y, X], ref, rtol=0, atol=1e-6), (X, Y)
        assert np.allclose(img32[Y, X], ref, rtol=0, atol=1e-4), (X, Y)
        n_checked += 1
    assert n_checked > 300


def _slots_with_color(mat, color):
    """Per-slot values for orc_shade_point: the material's constants, with the
    Input('color') slot bound to the particle's colour."""
    vals = np.zeros((mat.nparams, 3), dtype=np.float64)
    for k in range(mat.nparams):
        p = mat.params[k]
        if p.kind == _abi.P_COLOR:
            vals[k] = np.float64(np.asarray(color, dtype=np.float32))
        elif p.kind == _abi.P_CONST:
            vals[k] = [p.value[0]] * 3 if p.dim == 1 else [p.value[0], p.value[1], p.value[2]]
        else:
            raise AssertionError(f'slot kind {p.kind} not expected in this material')
    return vals
