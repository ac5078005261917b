"""Blooming and FXAA on the HIP path against the post-processing oracle, and
Scene(blooming=True, fxaa=True) against the oracle chain on the linear frame."""
import numpy as np
import pytest

from oracle import orc, postp

pytestmark = pytest.mark.gpu
f32 = np.float32


def _field(arr):
    import torch
    from taichi_three_amd.field import Field
    return Field(torch.as_tensor(np.ascontiguousarray(arr), device='cuda'))


@pytest.mark.parametrize('W,H', [(97, 64), (256, 144)])
def test_bloom_and_fxaa_match_the_oracle(gpu, W, H):
    import torch
    import taichi_three_amd as tina
    print(f'case: bloom + fxaa {W}x{H}', flush=True)
    rng = np.random.default_rng(W + H)
    img = (rng.random((H, W, 3)) ** 3 * 4).astype(f32)
    img[H // 3:H // 2, W // 4:W // 2] = 6.0  # a bright block with hard edges
    bl = tina.Blooming((W, H))
    f = _field(img)
    bl.apply(f)
    torch.cuda.synchronize()
    got = f.storage.cpu().numpy()
    want = postp.bloom(img, gwei=bl.gwei.cpu().numpy(), radius=bl.radius)
    assert np.abs(got - want).max() <= 1e-6 * max(1.0, float(np.abs(want).max()))
    ldr = np.clip(got / (1 + got), 0, 1).astype(f32)
    fx = tina.FXAA((W, H))
    g = _field(ldr)
    fx.apply(g)
    torch.cuda.synchronize()
    got2 = g.storage.cpu().numpy()
    want2 = postp.fxaa(ldr)
    assert np.abs(got2 - want2).max() <= 1e-6
    assert (got2 != ldr).any()


def test_scene_post_chain(gpu):
    """Scene(blooming=True, fxaa=True) = fxaa(aces(bloom(linear frame)))."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    print('case: scene blooming + fxaa', flush=True)
    w = wl.plain_512(seed=4)
    imgs = []
    for opts in (dict(tonemap=False), dict(blooming=True, fxaa=True)):
        scene, mesh = wl.build_scene(w, **opts)
        scene.render()
        torch.cuda.synchronize()
        imgs.append(scene.img.to_numpy().transpose(1, 0, 2).copy())
        bl = scene.blooming
    lin, post = imgs
    want = postp.fxaa(orc.aces(postp.bloom(lin, gwei=bl.gwei.cpu().numpy(), radius=bl.radius)))
    assert np.abs(post - want).max() <= 1e-5
