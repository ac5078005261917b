"""Pins of the oracle's specular lobes away from normal incidence.

Cook-Torrance (tina/matr/material.py:322-362) at geometries where each of its
three factors reduces to a number worked out by hand, plus the normalisation of
its distribution term by quadrature; the Blinn-Phong lobe (extension) at a known
N.H and through its directional albedo, whose closed form is integrated below.
Each check is one a plausible slip (the Smith k taken as roughness/2 or alpha2,
a wrong Schlick exponent, a dropped square in D, a wrong Blinn normalisation)
fails.
"""
import math

import numpy as np

from oracle import orc


def f32(x):
    return np.float64(np.float32(x))


def _mat(material):
    from taichi_three_amd.matr import lower_material
    return lower_material(material, dict(attrs={})).struct


def _lights(dirs, colors, ambient=(0, 0, 0)):
    from taichi_three_amd.lighting import Lighting
    lt = Lighting()
    for d, c in zip(dirs, colors):
        lt.add_light(dir=tuple(float(x) for x in d), color=c)
    lt.set_ambient_light(ambient)
    return lt.to_struct()


def _light_dir(lt, i=0):
    """The light direction as the oracle reads it (normalised, stored in binary32)."""
    return np.array([lt.dirs[i][0], lt.dirs[i][1], lt.dirs[i][2]], dtype=np.float64)


N = np.array([0.0, 0.0, 1.0])


def test_cooktorrance_roughness_1_off_the_normal():
    """roughness 1: alpha2 = 1, so D = 1 whatever N.H is, k = 1/2 and
    lerp(1, 1, 4 pi) = 4 pi.  L = N, V 60 degrees off it: N.L = 1, N.V = 1/2,
    V.H = cos 30 deg.  brdf = F / ((1/2 * 1/2 + 1/2) * (1 * 1/2 + 1/2)) / (4 pi)
    with F = f0 + (1 - f0) (1 - cos 30 deg)^5."""
    import taichi_three_amd as tina
    f0 = np.array([0.04, 0.5, 0.95])
    mat = _mat(tina.CookTorrance(roughness=1.0, fresnel=tuple(f0)))
    lt = _lights([N], [(1.0, 1.0, 1.0)])
    V = np.array([math.sin(math.pi / 3), 0.0, math.cos(math.pi / 3)])
    got = orc.shade_point(mat, lt, [0, 0, 0], N, V)
    F = f32(f0) + (1 - f32(f0)) * (1 - math.cos(math.pi / 6)) ** 5
    exp = 1.0 * F / (0.75 * 1.0) / (4 * math.pi)          # N.L = 1 multiplies the brdf
    assert np.allclose(got, exp, rtol=1e-6, atol=0), (got, exp)
    # the same lobe with the light and the viewer swapped (reciprocity)
    lt2 = _lights([V], [(1.0, 1.0, 1.0)])
    got2 = orc.shade_point(mat, lt2, [0, 0, 0], N, N)
    assert np.allclose(got2, 0.5 * exp, rtol=1e-6, atol=0)   # now N.L = 1/2


def test_cooktorrance_half_roughness_at_120_degrees():
    """roughness 1/2, L and V 60 degrees either side of N: H = N, so N.H = 1,
    D = 1/alpha2 = 4; N.L = N.V = 1/2, k = alpha2/2 = 1/8, so
    G = 1 / (15/16)^2 / (3/4 + pi); V.H = 1/2, so F = f0 + (1 - f0)/32."""
    import taichi_three_amd as tina
    f0 = np.array([0.02, 0.25, 0.9])
    mat = _mat(tina.CookTorrance(roughness=0.5, fresnel=tuple(f0)))
    s = math.sin(math.pi / 3)
    L = np.array([-s, 0.0, 0.5])
    V = np.array([s, 0.0, 0.5])
    lt = _lights([L], [(1.0, 1.0, 1.0)])
    got = orc.shade_point(mat, lt, [0, 0, 0], N, V)
    F = f32(f0) + (1 - f32(f0)) / 32
    G = 1 / (15 / 16) ** 2 / (0.75 + math.pi)
    exp = 0.5 * F * G * 4.0
    assert np.allclose(got, exp, rtol=1e-6, atol=0), (got, exp)


def test_pbr_metallic_1_takes_the_albedo_as_f0():
    """PBR with metallic 1: f0 = basecolor, and the mix factor is f0, so the
    result is (1 - f0) base / pi + f0 * CT(f0) at the 120-degree geometry."""
    import taichi_three_amd as tina
    base = np.array([0.9, 0.6, 0.2])
    mat = _mat(tina.PBR(basecolor=tuple(base), metallic=1.0, roughness=0.5, specular=0.5))
    s = math.sin(math.pi / 3)
    lt = _lights([np.array([-s, 0.0, 0.5])], [(1.0, 1.0, 1.0)])
    got = orc.shade_point(mat, lt, [0, 0, 0], N, np.array([s, 0.0, 0.5]))
    f0 = f32(base)
    ct = (f0 + (1 - f0) / 32) * (1 / (15 / 16) ** 2 / (0.75 + math.pi)) * 4.0
    exp = 0.5 * ((1 - f0) * f0 / math.pi + f0 * ct)
    assert np.allclose(got, exp, rtol=1e-6, atol=0), (got, exp)


def test_ggx_distribution_normalises_to_pi():
    """Tina's D is pi times the GGX distribution, so over the hemisphere of half
    vectors D(h) (n.h) integrates to pi for every roughness.  With L = V the half
    vector is V, V.H clamps to 1 - 1e-10 (F = f0 to double precision) and
    G = 1 / (c k + 1 - k)^2 / lerp(alpha2, 1, 4 pi) with c = N.V, which the two
    tests above pin; dividing the brdf by F G leaves D(c)."""
    import taichi_three_amd as tina
    xs, ws = np.polynomial.legendre.leggauss(400)
    c_nodes = 0.5 * (xs + 1.0)
    for r in (0.2, 0.5, 0.8, 1.0):
        f0 = 0.5
        mat = _mat(tina.CookTorrance(roughness=r, fresnel=(f0, f0, f0)))
        a2 = max(1e-6, f32(r) * f32(r))
        k = a2 / 2
        acc = 0.0
        for c, wgt in zip(c_nodes, ws):
            lt = _lights([(math.sqrt(1 - c * c), 0.0, c)], [(1.0, 1.0, 1.0)])
            L = _light_dir(lt)
            cl = float(L[2])
            got = orc.shade_point(mat, lt, [0, 0, 0], N, L)[0] / cl   # remove the N.L factor
            G = 1 / (cl * k + 1 - k) ** 2 / ((1 - a2) + 4 * math.pi * a2)
            D = got / (f32(f0) * G)
            acc += 0.5 * wgt * D * cl
        assert abs(2 * math.pi * acc - math.pi) <= 1e-5 * math.pi, (r, 2 * math.pi * acc)


def _blinn_albedo(m):
    """Directional albedo of max(0, N.H)^m (m + 8)/8 seen along the normal:
    2 pi (m + 8)/8 * 4 [2 c^(m+4)/(m+4) - c^(m+2)/(m+2)] from c = 1/sqrt2 to 1
    (c = cos(theta/2) = N.H when V = N and L is theta off the normal)."""
    def prim(c):
        return 2 * c ** (m + 4) / (m + 4) - c ** (m + 2) / (m + 2)
    return 2 * math.pi * (m + 8) / 8 * 4 * (prim(1.0) - prim(1 / math.sqrt(2)))


def test_blinn_lobe_at_a_known_half_angle():
    """V = N, L 60 degrees off it: H is 30 degrees off N, so the lobe is
    cos(30 deg)^m (m + 8)/8 and the light adds cos(60 deg) times that."""
    import taichi_three_amd as tina
    for m in (8.0, 32.0):
        mat = _mat(tina.BlinnPhong(shineness=m))
        L = np.array([math.sin(math.pi / 3), 0.0, 0.5])
        got = orc.shade_point(mat, _lights([L], [(1.0, 1.0, 1.0)]), [0, 0, 0], N, N)
        exp = 0.5 * math.cos(math.pi / 6) ** m * (m + 8) / 8
        assert np.allclose(got, exp, rtol=1e-6, atol=0), (m, got, exp)


def test_blinn_directional_albedo_by_quadrature():
    """Integrating the oracle's lit value over every light direction of the
    hemisphere (V = N, so only the polar angle matters) gives the albedo's
    closed form."""
    import taichi_three_amd as tina
    xs, ws = np.polynomial.legendre.leggauss(200)
    th = 0.25 * math.pi * (xs + 1.0)
    for m in (8.0, 32.0, 128.0):
        mat = _mat(tina.BlinnPhong(shineness=m))
        acc = 0.0
        for t, wgt in zip(th, ws):
            lt = _lights([(math.sin(t), 0.0, math.cos(t))], [(1.0, 1.0, 1.0)])
            v = orc.shade_point(mat, lt, [0, 0, 0], N, N)[0]       # cos(theta) * brdf
            acc += 0.25 * math.pi * wgt * v * math.sin(t)
        got = 2 * math.pi * acc
        assert abs(got - _blinn_albedo(m)) <= 1e-5 * _blinn_albedo(m), (m, got, _blinn_albedo(m))
