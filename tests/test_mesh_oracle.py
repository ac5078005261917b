"""Pins of the oracle's mesh-node restatements (MeshTransform, MeshFlatNormal,
MeshSmoothNormal; tina/mesh/trans.py:29-40, tina/mesh/norm.py:5-55) against
float64 mathematics and exact cases."""
import numpy as np

from oracle import orc


def _f64_transform(M, p):
    h = np.concatenate([p.astype(np.float64), np.ones((len(p), 1))], axis=1) @ M.astype(np.float64).T
    return h[:, :3] / h[:, 3:]


def test_transform_matches_float64_projective():
    """A projective transform (w != 1): the binary32 result within a few ulps of
    the float64 one; a transposed matrix would be far off."""
    rng = np.random.default_rng(1)
    M = np.eye(4) + 0.3 * rng.normal(size=(4, 4))
    M[3] = [0.05, -0.04, 0.1, 1.2]
    M32 = M.astype(np.float32)
    p = rng.uniform(-2, 2, (5000, 3)).astype(np.float32)
    got, _ = orc.mesh_transform(M32, np.eye(3), p)
    exp = _f64_transform(M32, p)
    assert np.allclose(got, exp, rtol=2e-6, atol=2e-6)
    bad, _ = orc.mesh_transform(M32.T.copy(), np.eye(3), p)
    assert not np.allclose(bad, exp, rtol=1e-3, atol=1e-3)


def test_transform_exact_for_dyadic_affine():
    """Scale by powers of two and translate by dyadic offsets: every operation is
    exact, so the result is the exact affine image."""
    M = np.diag([2.0, 0.5, 4.0, 1.0])
    M[:3, 3] = [0.25, -1.5, 3.0]
    p = (np.arange(300, dtype=np.float32).reshape(100, 3) / 8).astype(np.float32)
    got, _ = orc.mesh_transform(M, np.eye(3), p)
    exp = (p.astype(np.float64) * np.array([2.0, 0.5, 4.0]) + np.array([0.25, -1.5, 3.0])).astype(np.float32)
    assert np.array_equal(got, exp)


def test_transform_normals_by_inverse_transpose():
    """Normals go through the upper-left 3x3 of transpose(inv(trans)): a normal of
    a plane stays perpendicular to the transformed plane's tangents."""
    rng = np.random.default_rng(2)
    A = np.eye(4)
    A[:3, :3] = rng.normal(size=(3, 3)) + 2 * np.eye(3)
    A[:3, 3] = [0.5, -0.2, 0.1]
    tn = np.transpose(np.linalg.inv(A))[:3, :3]
    n = rng.normal(size=(200, 3)).astype(np.float32)
    t = np.cross(n, rng.normal(size=(200, 3))).astype(np.float32)       # tangents, n . t = 0
    _, got = orc.mesh_transform(A, tn, np.zeros((1, 3), np.float32), n)
    exp = n.astype(np.float64) @ tn.astype(np.float32).astype(np.float64).T
    assert np.allclose(got, exp, rtol=1e-6, atol=1e-6)
    tt = t.astype(np.float64) @ A[:3, :3].T
    assert np.abs((got.astype(np.float64) * tt).sum(1)).max() <= 1e-5 * np.abs(got).max() * np.abs(tt).max()


def test_face_normals_exact_and_float64():
    tri = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]],        # counter-clockwise in xy: +z
                    [[0, 0, 0], [0, 1, 0], [1, 0, 0]],        # clockwise: -z
                    [[0, 0, 0], [0, 0, 2], [0, 3, 0]]], np.float32)   # -x
    n = orc.face_normals(tri)
    assert np.array_equal(n, np.array([[0, 0, 1], [0, 0, -1], [-1, 0, 0]], np.float32))
    rng = np.random.default_rng(3)
    soup = rng.uniform(-1, 1, (2000, 3, 3)).astype(np.float32)
    n = orc.face_normals(soup)
    v = soup.astype(np.float64)
    c = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    exp = c / np.linalg.norm(c, axis=1, keepdims=True)
    assert np.allclose(n, exp, atol=1e-5)


def _icosahedron():
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int32)
    return v.astype(np.float32), f


def test_smooth_normals_of_the_icosahedron_are_radial():
    """By symmetry every vertex normal of the regular icosahedron is the vertex's
    direction from the centre."""
    v, f = _icosahedron()
    n = orc.smooth_normals(v[f], f, len(v))
    exp = v.astype(np.float64) / np.linalg.norm(v, axis=1, keepdims=True)
    assert np.allclose(n, exp, atol=1e-6)


def test_smooth_normals_match_float64_accumulation():
    """A random indexed mesh: sum of the faces' unit normals per vertex,
    normalized, in float64; vertices of no face are 0/0 (NaN) as in the reference."""
    rng = np.random.default_rng(4)
    nv = 400
    v = rng.uniform(-1, 1, (nv, 3)).astype(np.float32)
    f = np.stack([rng.choice(nv - 10, 3, replace=False) for _ in range(1500)]).astype(np.int32)
    n = orc.smooth_normals(v[f], f, nv)
    vd = v.astype(np.float64)[f]
    c = np.cross(vd[:, 1] - vd[:, 0], vd[:, 2] - vd[:, 0])
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    acc = np.zeros((nv, 3))
    for k in range(3):
        np.add.at(acc, f[:, k], c)
    used = np.linalg.norm(acc, axis=1) > 1e-3
    exp = acc[used] / np.linalg.norm(acc[used], axis=1, keepdims=True)
    assert np.allclose(n[used], exp, atol=1e-4)
    assert np.isnan(n[nv - 10:]).all()


def test_flat_mesh_smooth_normals_exact():
    """A planar grid in z = 0, every face counter-clockwise: each face normal is
    exactly (0, 0, 1), so every sum is (0, 0, k) and every vertex normal exactly +z."""
    xs = np.arange(5, dtype=np.float32)
    v = np.array([[x, y, 0] for y in xs for x in xs], np.float32)
    f = []
    for j in range(4):
        for i in range(4):
            a = j * 5 + i
            f += [[a, a + 1, a + 6], [a, a + 6, a + 5]]
    f = np.array(f, np.int32)
    n = orc.smooth_normals(v[f], f, len(v))
    assert np.array_equal(n, np.tile(np.array([0, 0, 1], np.float32), (len(v), 1)))
