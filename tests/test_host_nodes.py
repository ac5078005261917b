"""Host-side input nodes on CPU tensors (no GPU): SimpleParticles' count and
truncation rules (tina/pars/simple.py:39-55), ParsTransform's binary32 mapply
order (tina/common.py:169-183, tina/pars/trans.py:21-30), export_simple_pars,
and MeshGrid's faces and border-clamped normals (tina/mesh/grid.py:26-60)."""
import numpy as np
import pytest
import torch

import taichi_three_amd as tina

CPU = torch.device('cpu')


def test_simple_particles_counts_and_truncation():
    p = tina.SimpleParticles(5, radius=0.25, device=CPU)
    assert p.get_npars() == 0
    x = np.arange(24, dtype=np.float32).reshape(8, 3)
    p.set_particles(x)
    assert p.get_npars() == 5 and np.array_equal(p.verts.to_numpy(), x[:5])
    assert np.all(p.sizes.to_numpy() == np.float32(0.25)) and np.all(p.colors.to_numpy() == 1)
    p.set_particle_radii(np.linspace(0.1, 0.9, 7).astype(np.float32))
    assert np.allclose(p.sizes.to_numpy(), np.linspace(0.1, 0.9, 7)[:5])
    with pytest.raises(ValueError):
        p.set_particle_colors(np.ones((3, 3), np.float32))     # fewer colours than particles
    with pytest.raises(ValueError):
        p.set_particles(np.ones((4, 2), np.float32))


def test_pars_transform_is_binary32_mapply():
    rng = np.random.default_rng(3)
    x = rng.uniform(-2, 2, (50, 3)).astype(np.float32)
    p = tina.SimpleParticles(64, radius=0.05, device=CPU)
    p.set_particles(x)
    t = tina.ParsTransform(p)
    M = rng.uniform(-1, 1, (4, 4)).astype(np.float32)
    M[3] = [0.1, -0.2, 0.05, 1.5]
    t.set_transform(M, 2.0)
    fb = t.particle_buffers()
    f32 = np.float32
    h = []
    for i in range(4):
        a = np.full(len(x), M[i, 3], f32)
        for j in range(3):
            a = (a + (M[i, j] * x[:, j]).astype(f32)).astype(f32)
        h.append(a)
    exp = np.stack([h[0] / h[3], h[1] / h[3], h[2] / h[3]], axis=1).astype(f32)
    assert np.array_equal(fb['pos'].numpy(), exp)
    assert np.array_equal(fb['radius'].numpy(), np.full(50, f32(0.1)))
    out = tina.export_simple_pars(t)
    assert np.array_equal(out['v'], exp) and out['vr'].shape == (50,) and out['vc'].shape == (50, 3)


def test_mesh_grid_faces_and_normals():
    g = tina.MeshGrid((6, 6), device=CPU)
    p = g.pos.to_numpy().copy()
    p[..., 2] = (p[..., 0] ** 2 - p[..., 1]).astype(np.float32)
    g.pos.from_numpy(p)
    g.pre_compute()
    n = g.nrm.to_numpy()
    assert np.allclose(np.linalg.norm(n, axis=2), 1, atol=1e-6)
    # interior: dx = (2h, 0, z(i+1)-z(i-1)), dy = (0, 2h, z(j+1)-z(j-1)), n = dx x dy
    h = np.float64(2 / 5)
    i, j = 2, 3
    dzx = np.float64(p[i + 1, j, 2]) - np.float64(p[i - 1, j, 2])
    dzy = np.float64(p[i, j + 1, 2]) - np.float64(p[i, j - 1, 2])
    e = np.cross([2 * h, 0, dzx], [0, 2 * h, dzy])
    assert np.allclose(n[i, j], e / np.linalg.norm(e), atol=1e-6)
    # a border row uses a one-sided difference (the clamped neighbour is the point itself)
    dzx0 = np.float64(p[1, j, 2]) - np.float64(p[0, j, 2])
    e0 = np.cross([h, 0, dzx0], [0, 2 * h, np.float64(p[0, j + 1, 2]) - np.float64(p[0, j - 1, 2])])
    assert np.allclose(n[0, j], e0 / np.linalg.norm(e0), atol=1e-6)
    fb = g.face_buffers()
    v = fb['verts'].numpy()
    assert v.shape == (50, 3, 3)
    # cell (1, 2) = cell number 1 * 5 + 2 = 7: faces 14 (a, b, c) and 15 (a, c, d)
    a, b, c, d = p[1, 2], p[2, 2], p[2, 3], p[1, 3]
    assert np.array_equal(v[14], np.stack([a, b, c])) and np.array_equal(v[15], np.stack([a, c, d]))
