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


@pytest.mark.parametrize('W,H,radius', [(97, 64, None), (256, 144, None), (1920, 1080, None), (601, 333, 90),
                                         (500, 300, 200)])
def test_bloom_and_fxaa_match_the_oracle(gpu, W, H, radius):
    """radius None is the reference's min(res) // 16 (67 at 1920x1080); 90 takes the
    16-column vertical tile, 200 the direct (unstaged) vertical pass."""
    import torch
    import taichi_three_amd as tina
    print(f'case: bloom + fxaa {W}x{H} radius {radius}', flush=True)
    rng = np.random.default_rng(W + H)
    img = (rng.random((H, W, 3)) ** 3 * 4).astype(f32)
    img[H // 3:H // 2, W // 4:W // 2] = 6.0  # a bright block with hard edges
    bl = tina.Blooming((W, H))
    if radius is not None:
        bl.radius = radius
        bl.init_gwei()
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


def test_taa_is_the_running_mean_of_the_jittered_frames(gpu):
    """Scene(taa=True): frame k is rendered with a random pixel-centre bias (the
    first with the centre), and img is the running mean acc = acc (1 - 1/k) +
    frame / k.  Rendering the same biases without TAA and averaging the same way
    in binary32 gives the same image; clear() restarts the mean."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import workloads as wl
    print('case: taa running mean', flush=True)
    w = wl.plain_512(seed=5)
    taa, _ = wl.build_scene(w, taa=True)
    plain, _ = wl.build_scene(w)
    np.random.seed(3)
    acc = None
    biases = []
    for k in range(1, 5):
        taa.render()
        biases.append(taa.engine.bias.copy())
        plain.engine.bias = biases[-1]
        plain.engine._push_camera()
        plain.render()
        torch.cuda.synchronize()
        frame = plain.img.to_numpy()
        inv = f32(1.0) / f32(k)
        acc = frame.copy() * f32(0) if acc is None else acc
        acc = (acc * (f32(1) - inv)).astype(f32)
        acc = (acc + frame * inv).astype(f32)
    assert np.array_equal(biases[0], np.array([0.5, 0.5], f32)) and not np.array_equal(biases[1], biases[0])
    got = taa.img.to_numpy()
    assert np.abs(got - acc).max() <= 1e-6
    taa.clear()
    taa.render()
    torch.cuda.synchronize()
    assert np.array_equal(taa.engine.bias, np.array([0.5, 0.5], f32))
