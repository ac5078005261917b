"""Checkpoints of the trained state (the role of tina/shield.py:7-57, which
pickles Taichi fields): a surrogate saved mid-training and loaded again resumes
with the same weights, moments and step, so further training continues bit for
bit as if uninterrupted; IBL tables loaded from a file equal the baked ones and
shade the same frame."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lighting(precision=4):
    import taichi_three_amd as tina
    from taichi_three_amd.lighting import SkyboxLighting
    from taichi_three_amd.skybox import SUNS
    sun = np.asarray(SUNS[0][0]) / np.linalg.norm(SUNS[0][0])
    lt = SkyboxLighting(tina.Skybox.procedural(128, 64), precision=precision)
    lt.add_light(dir=tuple(sun), color=(2.0, 2.0, 2.0))
    return lt


def test_surrogate_resumes_bit_for_bit(gpu, tmp_path):
    import torch
    from taichi_three_amd.neural import NeuralShading
    print('case: surrogate checkpoint', flush=True)
    lt = _lighting()
    a = NeuralShading(lt, seed=5, threshold=0.02)
    X, Y = a.samples(8192, seed=3)
    for s in range(6):
        a.train_step(X, Y, (s % 4) * 2048, 2048, 1e-3)
    a.pack()
    path = tmp_path / 'nn.npz'
    a.save(path)
    b = NeuralShading.load(path, lt)
    assert b.step == a.step == 6 and b.threshold == 0.02
    for k in ('params', 'm1', 'm2', 'packed'):
        assert torch.equal(getattr(a, k), getattr(b, k)), k
    for s in range(6, 10):
        a.train_step(X, Y, (s % 4) * 2048, 2048, 1e-3)
        b.train_step(X, Y, (s % 4) * 2048, 2048, 1e-3)
    torch.cuda.synchronize()
    assert torch.equal(a.params, b.params) and torch.equal(a.m2, b.m2)
    bad = tmp_path / 'bad.npz'
    np.savez(bad, params=np.zeros(5, np.float32), m1=np.zeros(5, np.float32), m2=np.zeros(5, np.float32),
             step=np.array(1), specular=np.array(0.5), threshold=np.array(0.01))
    with pytest.raises(ValueError):
        NeuralShading.load(bad, lt)


def test_ibl_tables_load_without_baking(gpu, tmp_path):
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    from taichi_three_amd.lighting import SkyboxLighting
    from taichi_three_amd.skybox import IBLTables
    print('case: IBL tables checkpoint', flush=True)
    lt = _lighting()
    path = tmp_path / 'ibl.npz'
    lt.ibltab.save(path)
    tab = IBLTables.load(path, lt.skybox)
    assert torch.equal(tab.diff, lt.ibltab.diff)
    assert len(tab.spec) == len(lt.ibltab.spec)
    for x, y in zip(tab.spec, lt.ibltab.spec):
        assert torch.equal(x, y)
    assert tab.spec_samples == lt.ibltab.spec_samples
    with pytest.raises(ValueError):
        IBLTables.load(path, tina.Skybox.procedural(256, 128))
    # the same frame through the baked and the loaded tables
    rng = np.random.default_rng(4)
    soup = wl.triangle_soup(rng, 3000, np.log(0.1), 0.5)
    imgs = []
    for ibltab in (lt.ibltab, tab):
        lt2 = SkyboxLighting(lt.skybox, ibltab=ibltab)
        lt2.add_light(dir=(0.3, 0.8, 0.5), color=(1.0, 1.0, 1.0))
        scene = tina.Scene(160, 120, ibl=False, maxfaces=len(soup))
        scene.lighting = lt2
        m = tina.SimpleMesh()
        m.set_face_verts(soup)
        scene.add_object(m, tina.PBR(basecolor=(0.7, 0.6, 0.5), metallic=0.3, roughness=0.4))
        scene.render()
        torch.cuda.synchronize()
        imgs.append(scene.img.to_numpy())
    assert np.array_equal(imgs[0], imgs[1])
    assert np.ptp(imgs[0]) > 0
