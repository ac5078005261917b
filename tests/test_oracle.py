"""Pins the CPU oracle against things the mathematics fixes, not against a
retyped copy of its own formula:

  coverage/depth  - float64 ray casting through each pixel centre (a different
                    method: 3-D ray/triangle intersection) on perspective scenes
                  - exact rational edge tests (fractions.Fraction) under the
                    identity and orthographic cameras
                  - the docs/triangle.py triangle: every pixel, exact
                  - tie, culling and clipping rules of tina/core/triangle.py:95-125
  shading         - closed forms: Lambert, Phong at the mirror direction,
                    Tina's Cook-Torrance at normal incidence, ACES values
                  - binary32 vs binary64 agreement on the 512x512 workload
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import orc
from taichi_three_amd import _abi
from taichi_three_amd import workloads as wl
from taichi_three_amd.util.matrix import camera_matrices, lookat, ortho, perspective


def _cam(view, proj):
    W2V, V2W = camera_matrices(view, proj)
    return orc.camera(W2V, V2W), W2V


# ---------------------------------------------------------------- coverage

def raycast_f64(W, H, view, proj, verts, culling=True, clipping=True):
    """Per pixel: cast the ray through the pixel centre in float64, intersect every
    face (Moller-Trumbore), keep the nearest NDC depth.  Returns (id, ndc_z,
    margin) where margin is the smallest |barycentric| of the winning hit and the
    depth gap to the runner-up, used to skip pixels float32 cannot decide."""
    M = np.asarray(proj, np.float64) @ np.asarray(view, np.float64)
    Minv = np.linalg.inv(M)
    v = verts.astype(np.float64)
    hom = np.concatenate([v, np.ones(v.shape[:2] + (1,))], axis=2) @ M.T
    ndc = hom[..., :3] / hom[..., 3:]
    facing = ((ndc[:, 1, 0] - ndc[:, 0, 0]) * (ndc[:, 2, 1] - ndc[:, 0, 1])
              - (ndc[:, 1, 1] - ndc[:, 0, 1]) * (ndc[:, 2, 0] - ndc[:, 0, 0]))
    keep = np.ones(len(v), bool)
    if culling:
        keep &= facing > 0
    if clipping:
        inside = np.all((ndc >= -1) & (ndc <= 1), axis=2)
        keep &= inside.any(axis=1)
    ids = np.full((H, W), -1)
    z = np.full((H, W), np.inf)
    margin = np.full((H, W), np.inf)
    xs = (np.arange(W) + 0.5) / W * 2 - 1
    ys = (np.arange(H) + 0.5) / H * 2 - 1
    X, Y = np.meshgrid(xs, ys)
    p0 = np.stack([X, Y, -np.ones_like(X), np.ones_like(X)], -1) @ Minv.T
    p1 = np.stack([X, Y, np.ones_like(X), np.ones_like(X)], -1) @ Minv.T
    o = p0[..., :3] / p0[..., 3:]
    d = p1[..., :3] / p1[..., 3:] - o
    second = np.full((H, W), np.inf)
    for f in np.nonzero(keep)[0]:
        a, b, c = v[f]
        e1, e2 = b - a, c - a
        pv = np.cross(d, e2)
        det = pv @ e1
        with np.errstate(divide='ignore', invalid='ignore'):
            inv = 1.0 / det
            tv = o - a
            u = (tv * pv).sum(-1) * inv
            qv = np.cross(tv, e1)
            vv = (d * qv).sum(-1) * inv
            t = (qv @ e2) * inv
        w = 1 - u - vv
        hit = (u >= 0) & (vv >= 0) & (w >= 0) & np.isfinite(t)
        # Tina's depth (triangle.py:121): the vertices' NDC z blended with the
        # perspective-correct weights, which are the 3-D barycentrics of the hit
        zz = w * ndc[f, 0, 2] + u * ndc[f, 1, 2] + vv * ndc[f, 2, 2]
        m = np.minimum(np.minimum(np.abs(u), np.abs(vv)), np.abs(w))
        better = hit & (zz < z)
        second = np.where(better, z, np.where(hit, np.minimum(second, zz), second))
        # a pixel whose edge decision is close for any candidate is ambiguous
        near_edge = (m < 2e-4) & (zz < 1.5)
        margin = np.where(near_edge, 0.0, margin)
        ids = np.where(better, f, ids)
        z = np.where(better, zz, z)
    with np.errstate(invalid='ignore'):  # inf - inf where no face is hit
        gap = second - z
    margin = np.minimum(margin, np.where(np.isfinite(gap), gap * 1e3, np.inf))
    return ids, z, margin


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_coverage_matches_raycast(seed):
    rng = np.random.default_rng(seed)
    W, H = 96, 64
    verts = wl.triangle_soup(rng, 60, np.log(0.5), 0.6)
    view = lookat(back=(0.3, -0.2, 3.0))
    proj = perspective(60, W / H, 0.05, 500)
    cam, _ = _cam(view, proj)
    depth, occ = orc.raster(W, H, cam, verts)
    ids, z, margin = raycast_f64(W, H, view, proj, verts)
    ok = margin > 1e-3
    assert ok.sum() > 0.9 * W * H
    assert (occ[ok] >= 0).sum() > 500
    assert np.array_equal(occ[ok], ids[ok])
    cov = ok & (ids >= 0)
    assert np.abs(depth[cov] / 2.0 ** 30 - z[cov]).max() < 2e-6


def test_coverage_without_culling_matches_raycast():
    rng = np.random.default_rng(7)
    W, H = 80, 80
    verts = wl.triangle_soup(rng, 50, np.log(0.6), 0.5)
    view, proj = lookat(back=(0, 0, 3)), perspective(60, 1, 0.05, 500)
    cam, _ = _cam(view, proj)
    _, occ = orc.raster(W, H, cam, verts, flags=_abi.CLIP)
    ids, _, margin = raycast_f64(W, H, view, proj, verts, culling=False)
    ok = margin > 1e-3
    assert np.array_equal(occ[ok], ids[ok])
    _, occ_c = orc.raster(W, H, cam, verts, flags=_abi.CLIP | _abi.CULL)
    assert (occ >= 0).sum() > (occ_c >= 0).sum()   # back faces appear only without culling


def _frac_cover(px, py, a, b, c):
    """Inclusive exact edge test at the pixel centre (affine screen, wscale = 1)."""
    P = (Fraction(2 * px + 1, 2), Fraction(2 * py + 1, 2))

    def cross(o, p, q):
        return (p[0] - o[0]) * (q[1] - o[1]) - (p[1] - o[1]) * (q[0] - o[0])

    n = cross(a, b, c)
    if n == 0:
        return False
    l0, l1, l2 = cross(b, c, P) / n, cross(c, a, P) / n, cross(a, b, P) / n
    return l0 >= 0 and l1 >= 0 and l2 >= 0


def test_docs_triangle_identity_camera_exact():
    """docs/triangle.py:18-22 under the default camera W2V = diag(1, 1, -1, 1)."""
    W = H = 64
    verts = np.array([[[-1, -1, 0], [1, -1, 0], [0, 1, 0]]], dtype=np.float32)
    M = np.diag([1.0, 1.0, -1.0, 1.0])
    cam = orc.camera(M, M)
    depth, occ = orc.raster(W, H, cam, verts)
    a, b, c = (Fraction(0), Fraction(0)), (Fraction(W), Fraction(0)), (Fraction(W // 2), Fraction(H))
    exp = np.array([[_frac_cover(x, y, a, b, c) for x in range(W)] for y in range(H)])
    assert np.array_equal(occ >= 0, exp)
    assert exp.sum() == 64 * 64 // 2  # the area: no centre (x+1/2, y+1/2) lies on a slanted edge
    assert np.all(depth[exp] == 0)


@pytest.mark.parametrize('seed', [3, 4])
def test_orthographic_exact_edges(seed):
    """Orthographic camera: w = 1 so the weights are affine; coverage must equal the
    exact rational test except within float32 reach of an edge."""
    rng = np.random.default_rng(seed)
    W, H = 48, 40
    verts = wl.triangle_soup(rng, 30, np.log(0.8), 0.4)
    view = lookat(back=(0, 0, 3))
    proj = ortho(-1.2, 1.2, -1.0, 1.0, 0.1, 10)
    cam, W2V = _cam(view, proj)
    for f in range(len(verts)):
        _, occ = orc.raster(W, H, cam, verts[f:f + 1], flags=_abi.CLIP)
        hom = np.concatenate([verts[f].astype(np.float64), np.ones((3, 1))], 1) @ W2V.astype(np.float64).T
        sc = [((h[0] / h[3]) * 0.5 + 0.5) * W for h in hom], [((h[1] / h[3]) * 0.5 + 0.5) * H for h in hom]
        A, B, C = [(Fraction(sc[0][k]), Fraction(sc[1][k])) for k in range(3)]
        for y in range(H):
            for x in range(W):
                e = _frac_cover(x, y, A, B, C)
                # distance of the centre to the nearest edge, in pixels
                P = np.array([x + 0.5, y + 0.5])
                V = [np.array([float(A[0]), float(A[1])]), np.array([float(B[0]), float(B[1])]),
                     np.array([float(C[0]), float(C[1])])]
                def cross2(u, v):
                    return u[0] * v[1] - u[1] * v[0]
                dist = min(abs(cross2(V[(k + 1) % 3] - V[k], P - V[k])) / np.linalg.norm(V[(k + 1) % 3] - V[k])
                           for k in range(3))
                if dist > 1e-3:
                    assert (occ[y, x] >= 0) == e, (f, x, y)


def test_tie_goes_to_lowest_id_and_earlier_object():
    rng = np.random.default_rng(5)
    W, H = 64, 64
    tri = wl.triangle_soup(rng, 20, np.log(0.7), 0.3)
    verts = np.concatenate([tri, tri])        # exact duplicates: equal depths
    cam, _ = _cam(lookat(back=(0, 0, 3)), perspective(60, 1, 0.05, 500))
    _, occ = orc.raster(W, H, cam, verts, flags=_abi.CLIP)
    assert occ.max() < 20 and (occ >= 0).sum() > 200
    # two objects: the second's duplicates never win (serial strict '<')
    d, o = orc.raster(W, H, cam, tri, id_base=0, flags=_abi.CLIP)
    d, o = orc.raster(W, H, cam, tri, id_base=100, flags=_abi.CLIP, depth=d, occup=o)
    assert o.max() < 20
    # the packed-key parallel form agrees
    keys = orc.raster_omp(W, H, cam, verts, flags=_abi.CLIP, nthreads=4)
    assert np.array_equal(orc.unpack(keys)[1], occ)


def test_permutation_invariance():
    rng = np.random.default_rng(9)
    W, H = 72, 56
    verts = wl.triangle_soup(rng, 80, np.log(0.4), 0.5)
    cam, _ = _cam(lookat(back=(0, 0, 3)), perspective(60, W / H, 0.05, 500))
    d0, o0 = orc.raster(W, H, cam, verts)
    perm = rng.permutation(len(verts))
    d1, o1 = orc.raster(W, H, cam, verts[perm])
    assert np.array_equal(d0, d1)
    assert np.array_equal(o0, np.where(o1 >= 0, perm[np.maximum(o1, 0)], -1))


def test_clip_rule_and_culling():
    W, H = 32, 32
    M = np.diag([1.0, 1.0, -1.0, 1.0])
    cam = orc.camera(M, M)
    big = np.array([[[-5, -5, 0], [5, -5, 0], [0, 5, 0]]], dtype=np.float32)   # all vertices outside NDC
    _, occ = orc.raster(W, H, cam, big)
    assert (occ >= 0).sum() == 0                                                  # triangle.py:100-104
    _, occ = orc.raster(W, H, cam, big, flags=_abi.CULL)
    assert (occ >= 0).sum() == W * H                                              # no clipping: drawn
    one_in = big.copy()
    one_in[0, 2] = [0, 0.5, 0]
    _, occ = orc.raster(W, H, cam, one_in)
    assert (occ >= 0).sum() > 0
    cw = big[:, ::-1].copy()
    _, occ = orc.raster(W, H, cam, cw, flags=_abi.CULL)
    assert (occ >= 0).sum() == 0                                                  # back face culled
    _, occ = orc.raster(W, H, cam, cw, flags=0)
    assert (occ >= 0).sum() == W * H


def test_degenerate_and_nan_faces_cover_nothing():
    W, H = 16, 16
    M = np.diag([1.0, 1.0, -1.0, 1.0])
    cam = orc.camera(M, M)
    v = np.array([[[0, 0, 0], [0.5, 0.5, 0], [1, 1, 0]],            # collinear: n = 0
                  [[0, 0, 0], [np.nan, 0, 0], [0, 0.5, 0]]], dtype=np.float32)
    _, occ = orc.raster(W, H, cam, v, flags=0)
    assert (occ >= 0).sum() == 0


# ---------------------------------------------------------------- shading

def f32(x):
    """Inputs reach the oracle as binary32 (lights, Const slots): round them the same way."""
    return np.float64(np.float32(x))


def _mat(material, attrs=None):
    from taichi_three_amd.matr import lower_material
    return lower_material(material, dict(attrs=attrs or {})).struct


def _lights(dirs, colors, ambient=(0, 0, 0)):
    from taichi_three_amd.lighting import Lighting
    lt = Lighting()
    for d, c in zip(dirs, colors):
        if len(d) == 4 and d[3] == 1:
            lt.add_light(pos=d[:3], color=c)
        else:
            lt.add_light(dir=d[:3], color=c)
    lt.set_ambient_light(ambient)
    return lt.to_struct()


def test_lambert_closed_form():
    import taichi_three_amd as tina
    color = np.array([0.2, 0.5, 0.8])
    mat = _mat(tina.Diffuse(color=tuple(color)))
    n = np.array([0.0, 0.0, 1.0])
    L = np.array([1.0, 0.0, 1.0]) / np.sqrt(2)
    lt = _lights([L], [(0.7, 0.7, 0.7)], ambient=(0.1, 0.1, 0.1))
    got = orc.shade_point(mat, lt, [0, 0, 0], n, [0, 0, 1])
    c, L32 = f32(color), f32(L)        # the light table and Const slots hold binary32
    d = np.linalg.norm(L32)            # lighting.py:91-95 renormalises by |dir| and divides by d**2
    exp = f32(0.1) * c + (L32[2] / d) * (f32(0.7) / d ** 2) * c / np.pi
    assert np.allclose(got, exp, rtol=0, atol=1e-12)
    # light from below contributes nothing
    lt = _lights([-L], [(0.7, 0.7, 0.7)], ambient=(0, 0, 0))
    assert np.all(orc.shade_point(mat, lt, [0, 0, 0], n, [0, 0, 1]) == 0)


def test_point_light_inverse_square():
    import taichi_three_amd as tina
    mat = _mat(tina.Diffuse(color=(1.0, 1.0, 1.0)))
    lt = _lights([(0, 0, 2.0, 1)], [(1.0, 1.0, 1.0)])
    got = orc.shade_point(mat, lt, [0, 0, 0], [0, 0, 1], [0, 0, 1])
    assert np.allclose(got, 1.0 / 4.0 / np.pi, atol=1e-12)


def test_classic_phong_mirror_direction():
    import taichi_three_amd as tina
    color = np.array([0.3, 0.4, 0.5])
    mat = _mat(tina.Classic(color=tuple(color), shineness=32, specular=0.4))
    n = np.array([0.0, 0.0, 1.0])
    lt = _lights([n], [(1.0, 1.0, 1.0)], ambient=(0.1, 0.1, 0.1))
    got = orc.shade_point(mat, lt, [0, 0, 0], n, n)   # L = V = N: V.R = 1
    c, s = f32(color), f32(0.4)
    exp = f32(0.1) * ((1 - s) * c + s) + ((1 - s) * c / np.pi + s * (32 + 2) / 2)
    assert np.allclose(got, exp, atol=1e-12)
    # 60 degrees off the mirror direction: (cos 60)^32 * 17
    V = np.array([np.sin(np.pi / 3), 0, np.cos(np.pi / 3)])
    got = orc.shade_point(mat, _lights([n], [(1.0, 1.0, 1.0)]), [0, 0, 0], n, V)
    assert np.allclose(got, (1 - s) * c / np.pi + s * 0.5 ** 32 * 17, atol=1e-12)


def test_cooktorrance_normal_incidence_closed_form():
    """N = L = V: D = 1/alpha2, G = 1, F = f0; Tina divides by lerp(alpha2, 1, 4pi)."""
    import taichi_three_amd as tina
    for r in (0.1, 0.5, 1.0):
        f0 = np.array([0.04, 0.3, 0.9])
        mat = _mat(tina.CookTorrance(roughness=r, fresnel=tuple(f0)))
        n = np.array([0.0, 0.0, 1.0])
        got = orc.shade_point(mat, _lights([n], [(1.0, 1.0, 1.0)]), [0, 0, 0], n, n)
        r32, f032 = f32(r), f32(f0)
        a2 = max(1e-6, r32 * r32)
        exp = f032 * (1 / a2) / ((1 - a2) + 4 * np.pi * a2)
        assert np.allclose(got, exp, rtol=1e-12, atol=0)


def test_pbr_mixes_with_fresnel_factor():
    import taichi_three_amd as tina
    base = np.array([0.8, 0.2, 0.1])
    m, s, r = 0.6, 0.5, 0.3
    mat = _mat(tina.PBR(basecolor=tuple(base), metallic=m, roughness=r, specular=s))
    n = np.array([0.0, 0.0, 1.0])
    got = orc.shade_point(mat, _lights([n], [(1.0, 1.0, 1.0)], ambient=(0.2, 0.2, 0.2)), [0, 0, 0], n, n)
    base, m, s, r = f32(base), f32(m), f32(s), f32(r)
    f0 = m * base + (1 - m) * 0.16 * s * s
    a2 = r * r
    ct = f0 * (1 / a2) / ((1 - a2) + 4 * np.pi * a2)
    exp = f32(0.2) * ((1 - f0) * base + f0) + ((1 - f0) * base / np.pi + f0 * ct)
    assert np.allclose(got, exp, rtol=1e-12)


def test_aces_values():
    x = np.array([0.0, 1.0, 0.5, 10.0, 1e6])
    got = orc.aces(np.repeat(x[:, None], 3, 1), f64=True)[:, 0]
    exp = x * (2.51 * x + 0.03) / (x * (2.43 * x + 0.59) + 0.14)
    assert np.allclose(got, exp, rtol=1e-15)
    assert abs(got[1] - 2.54 / 3.16) < 1e-15 and got[0] == 0
    assert abs(got[-1] - 2.51 / 2.43) < 1e-6


def test_f32_and_f64_oracles_agree_on_plain_512():
    from taichi_three_amd.matr import lower_material
    w = wl.plain_512()
    W, H = w.res
    cam = orc.camera(w.W2V, w.V2W)
    _, occ = orc.raster(W, H, cam, w.verts)
    lo = lower_material(wl.make_material(w.material))
    mesh = orc.HostMesh(w.verts, w.norms, None, w.colors)
    flags = _abi.SMOOTH | _abi.VCOLOR | _abi.TONEMAP | _abi.BG
    lt = wl.make_lighting(w).to_struct()
    a = orc.shade(W, H, cam, mesh, lo.struct, lt, occ, flags)
    b = orc.shade(W, H, cam, mesh, lo.struct, lt, occ, flags, f64=True)
    assert np.abs(a - b).max() < 1e-4


def test_oracle_reproduces_golden_fixture():
    """The committed plain_512 fixture (written by tests/golden/make_golden.py)."""
    import os
    from conftest import ROOT
    from taichi_three_amd.matr import lower_material
    g = np.load(os.path.join(ROOT, 'tests', 'golden', 'plain_512.npz'))
    w = wl.plain_512()
    W, H = w.res
    cam = orc.camera(w.W2V, w.V2W)
    depth, occ = orc.raster(W, H, cam, w.verts)
    assert np.array_equal(occ, g['ids']) and np.array_equal(depth, g['depth'])
    lo = lower_material(wl.make_material(w.material))
    img = orc.shade(W, H, cam, orc.HostMesh(w.verts, w.norms, None, w.colors), lo.struct,
                    wl.make_lighting(w).to_struct(), occ, _abi.SMOOTH | _abi.VCOLOR | _abi.TONEMAP | _abi.BG)
    assert np.abs(img[192:320, 192:320] - g['img_crop']).max() <= 1e-6


def test_lut_corner_fixture_shape():
    """assets/lut.jpg corner texels (tests/golden/lut_corner.json): A near 0, B near 1
    at VoH -> 0, the values Tina's bilerp(lut, (VoH, roughness)) can reach."""
    import json
    import os
    from conftest import ROOT
    d = json.load(open(os.path.join(ROOT, 'tests', 'golden', 'lut_corner.json')))
    c = np.array(d['corner'])
    assert c.shape == (3, 3, 2)
    assert np.all(c[..., 0] < 0.05) and np.all(c[..., 1] > 0.9)


# ---------------------------------------------------------------- IBL

def test_lut_corner_table_equals_fixture():
    """The LUT corner the package ships equals the texels read from assets/lut.jpg."""
    import json
    import os
    from conftest import ROOT
    from taichi_three_amd.skybox import LUT_CORNER
    d = json.load(open(os.path.join(ROOT, 'tests', 'golden', 'lut_corner.json')))
    assert np.array_equal(LUT_CORNER, np.array(d['corner']))


def test_skybox_sample_inverts_unmapcoor():
    """Skybox.sample(unmapcoor(texel)) returns that texel: spheremap and
    unspheremap (tina/skybox.py:23-35) are inverses on the texel grid."""
    from taichi_three_amd.skybox import unmapcoor
    w, h = 48, 24
    env = np.random.default_rng(0).random((w, h, 3)).astype(np.float32)
    d = unmapcoor(w, h)
    for x in range(w):
        for y in range(1, h - 1):
            assert np.abs(orc.skybox_sample(env, d[x, y]) - env[x, y]).max() <= 1e-9


def test_bakes_of_a_constant_environment_are_that_constant():
    env = np.full((40, 20, 3), 0.7, np.float32)
    assert np.abs(orc.bake_lambert(env, 10, 5, 64) - 0.7).max() <= 1e-7
    assert np.abs(orc.bake_ggx(env, 10, 5, 0.2, 64) - 0.7).max() <= 1e-7


def _linear_env(w, h, a):
    """L(d) = 1 + a.d, tabulated on the equirect grid; exact along y under bilerp."""
    from taichi_three_amd.skybox import unmapcoor
    return (1 + unmapcoor(w, h) @ a)[..., None].repeat(3, -1).astype(np.float32)


def test_lambert_bake_of_a_linear_environment():
    """Cosine-weighted hemisphere mean of 1 + a.d around n is 1 + (2/3) a.n."""
    from taichi_three_amd.skybox import unmapcoor
    a = np.array([0.3, 0.5, 0.2])
    out = orc.bake_lambert(_linear_env(256, 128, a), 12, 6, 4096)
    n = unmapcoor(12, 6)
    assert np.abs(out[..., 0] - (1 + 2 / 3 * (n @ a))).max() <= 2e-4


def test_ggx_bake_of_a_linear_environment():
    """Around n the azimuths average out, so the u-weighted mean of 1 + a.d is
    1 + (a.n) sum(u^2) / sum(u) over the GGX cosines u of the sample set."""
    from taichi_three_amd.skybox import unmapcoor
    a = np.array([0.3, 0.5, 0.2])
    ns, r = 512, np.float32(0.65)
    out = orc.bake_ggx(_linear_env(256, 128, a), 12, 6, r, ns)
    a2 = float(r) * float(r)
    u0 = (np.arange(ns) + 0.5) / ns
    u = np.sqrt((1 - u0) / (1 - u0 * (1 - a2)))
    n = unmapcoor(12, 6)
    assert np.abs(out[..., 0] - (1 + (n @ a) * (u * u).sum() / u.sum())).max() <= 1e-3


def _host_ibl(c_spec, c_diff=0.25):
    """Host IBL tables with constant maps: spec level i holds c_spec[i]."""
    from taichi_three_amd.skybox import make_ibl_struct
    spec = [np.full((8, 4, 3), c, np.float32) for c in c_spec]
    diff = np.full((6, 3, 3), c_diff, np.float32)
    return make_ibl_struct(diff, spec), (spec, diff)


def test_cooktorrance_ibl_level_lerp_and_lut_corner():
    """CookTorrance.sample_ibl (material.py:255-277) at normal incidence: the
    reflected ray is n, VoH = 1, so AB = bilerp(lut, (1, r)) = lerp over the
    lut[1][0..1] texels, and the specular maps lerp between the two levels that
    bracket r: spec(r) * (f0 A + B)."""
    import taichi_three_amd as tina
    from taichi_three_amd.matr import lower_material
    from taichi_three_amd.skybox import LUT_CORNER, ROUGH_LEVELS
    levels = [1.0, 2.0, 3.0, 5.0, 7.0, 11.0, 13.0]
    ibl, keep = _host_ibl(levels)
    n = np.array([0.0, 0.0, 1.0])
    for r in (0.05, 0.1, 0.3, 0.5, 0.9):
        f0 = 0.6
        lo = lower_material(tina.CookTorrance(roughness=r, fresnel=f0))
        got = orc.shade_point(lo.struct, _abi.Lights(), np.zeros(3), n, n, flags=_abi.IBL | _abi.NOLIGHT, ibl=ibl)
        k = int(np.searchsorted(ROUGH_LEVELS, r))
        hi, lo_r = ROUGH_LEVELS[k], ([0] + ROUGH_LEVELS)[k]
        fac = (hi - r) / (hi - lo_r)
        spec = levels[k + 1] * (1 - fac) + levels[k] * fac
        A = LUT_CORNER[1, 0, 0] * (1 - r) + LUT_CORNER[1, 1, 0] * r
        B = LUT_CORNER[1, 0, 1] * (1 - r) + LUT_CORNER[1, 1, 1] * r
        assert np.allclose(got, spec * (f0 * A + B), rtol=1e-6), (r, got, spec * (f0 * A + B))


def test_lambert_ibl_reads_the_irradiance_map():
    """Lambert.sample_ibl (material.py:440-442) is diff.sample(n): a constant
    irradiance map c gives albedo * c, with no ambient term under SkyboxLighting."""
    import taichi_three_amd as tina
    from taichi_three_amd.matr import lower_material
    ibl, keep = _host_ibl([1.0] * 7, c_diff=0.25)
    lo = lower_material(tina.Lambert() * tina.Const((0.2, 0.4, 0.8)))
    lt = _abi.Lights()
    lt.ambient[0] = lt.ambient[1] = lt.ambient[2] = 5.0
    got = orc.shade_point(lo.struct, lt, np.zeros(3), np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0]),
                          flags=_abi.IBL | _abi.NOLIGHT, ibl=ibl)
    assert np.allclose(got, 0.25 * np.array([0.2, 0.4, 0.8]), rtol=1e-12)
