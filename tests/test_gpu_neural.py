"""The learned shading surrogate on the HIP path: training pairs vs the analytic
oracle, float32 gradients and the Adam step vs the NumPy oracle, the bf16 MFMA
network vs the oracle's bf16 arithmetic, and a trained network against the
analytic frame it replaces."""
import math

import numpy as np
import pytest

from oracle import nn, orc

pytestmark = pytest.mark.gpu


def _lighting(precision=4, sky=(128, 64)):
    import taichi_three_amd as tina
    from taichi_three_amd.lighting import SkyboxLighting
    from taichi_three_amd.skybox import SUNS
    sun = np.asarray(SUNS[0][0]) / np.linalg.norm(SUNS[0][0])
    lt = SkyboxLighting(tina.Skybox.procedural(*sky), precision=precision)
    lt.add_light(dir=tuple(sun), color=(2.0, 2.0, 2.0))
    return lt


def test_samples_match_the_analytic_oracle(gpu):
    """k_nn_samples targets = the binary32 oracle's PBR + IBL + sun at the same
    inputs (recovered from the features; rough surfaces, where the recovered
    view direction's rounding cannot move the colour past the tolerance)."""
    import torch
    import taichi_three_amd as tina
    from taichi_three_amd import _abi
    from taichi_three_amd.matr import lower_material
    from taichi_three_amd.neural import NeuralShading
    print('case: nn samples vs oracle', flush=True)
    lt = _lighting()
    ns = NeuralShading(lt)
    X, Y = ns.samples(64, seed=7, roughness=(0.3, 1.0), p_metal=-1.0)
    torch.cuda.synchronize()
    X, Y = X.cpu().numpy().astype(np.float64), Y.cpu().numpy()
    ibl, keep = lt.ibltab.host_copy()
    lts = lt.to_struct()
    for i in range(64):
        n, R = X[i, 0:3], X[i, 3:6]
        V = 2 * (n @ R) * n - R
        r, m, alb = math.exp(X[i, 7]), X[i, 8], X[i, 9:12]
        lo = lower_material(tina.PBR(basecolor=tina.Const(tuple(alb)), metallic=float(m), roughness=float(r)))
        c = orc.shade_point(lo.struct, lts, np.zeros(3), n, V, flags=_abi.IBL, ibl=ibl, f64=False)
        c = orc.aces(c.astype(np.float32).reshape(1, 3)).reshape(3)
        assert np.abs(c - Y[i, :3]).max() <= 1e-4, (i, c, Y[i])
    assert np.all(Y[:, 3] == 0)


def test_gradient_matches_the_numpy_oracle(gpu):
    """tina_mlp_train_step with lr = 0: the float32 batch gradient equals the
    float64 oracle's to float32 accumulation accuracy."""
    import torch
    from taichi_three_amd.neural import NeuralShading
    print('case: nn gradient vs oracle', flush=True)
    ns = NeuralShading(_lighting(), seed=3)
    X, Y = ns.samples(1000, seed=11)
    g = ns.gradient(X, Y, first=100, batch=700).cpu().numpy().astype(np.float64)
    p = ns.params.cpu().numpy().astype(np.float64)
    Xn, Yn = X.cpu().numpy()[100:800].astype(np.float64), Y.cpu().numpy()[100:800].astype(np.float64)
    ref = nn.backward(p, Xn, Yn)
    for name, shape, off in nn.layout():
        a, b = g[off:off + int(np.prod(shape))], ref[off:off + int(np.prod(shape))]
        scale = np.abs(b).max() + 1e-12
        assert np.abs(a - b).max() <= 1e-4 * scale, (name, np.abs(a - b).max(), scale)


def test_adam_step_matches_the_numpy_oracle(gpu):
    import torch
    from taichi_three_amd.neural import NeuralShading
    print('case: nn adam vs oracle', flush=True)
    ns = NeuralShading(_lighting(), seed=4)
    X, Y = ns.samples(512, seed=12)
    p0 = ns.params.cpu().numpy().astype(np.float64)
    g = ns.gradient(X, Y).cpu().numpy().astype(np.float64)
    ns.train_step(X, Y, 0, 512, 1e-3)
    p1 = ns.params.cpu().numpy().astype(np.float64)
    ref, _, _ = nn.adam(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    assert np.abs(p1 - ref).max() <= 1e-6


def test_mfma_network_matches_the_bf16_oracle(gpu):
    """k_mlp_pack + k_mlp_rows (the MFMA layout and the permuted K order) vs
    the oracle's bf16 arithmetic, for random weights and biases: 1000 rows (a
    ragged last wave) and 300,000 rows (every persistent wave loops)."""
    import torch
    from taichi_three_amd.neural import NeuralShading
    print('case: nn mfma vs oracle', flush=True)
    ns = NeuralShading(_lighting(), seed=5)
    rng = np.random.default_rng(5)
    p = nn.init_params(5) + 0.05 * rng.normal(size=nn.NPARAMS)
    ns.params.copy_(torch.from_numpy(p.astype(np.float32)))
    ns.pack()
    for n in (1000, 300_000):
        X, _ = ns.samples(n, seed=13)
        O = ns.infer(X).cpu().numpy()
        ref = nn.infer_bf16(p.astype(np.float32), X.cpu().numpy())
        scale = np.abs(ref).max()
        assert np.abs(O - ref).max() <= 2e-2 * scale, (n, np.abs(O - ref).max() / scale)
        assert np.abs(O - ref).mean() <= 2e-3 * scale, n


def test_analytic_rows_reproduce_the_sample_targets(gpu):
    """tina_nn_analytic_rows recovers V and the roughness from the features and
    shades them: on rough surfaces it reproduces k_nn_samples' targets, which the
    oracle pins above; an empty call writes nothing."""
    import torch
    from taichi_three_amd.neural import NeuralShading
    print('case: nn analytic rows vs samples', flush=True)
    ns = NeuralShading(_lighting(), seed=6)
    X, Y = ns.samples(100_003, seed=17, roughness=(0.3, 1.0), p_metal=-1.0)
    O = ns.analytic(X)
    err = (O[:, :3] - Y[:, :3]).abs()
    assert float(err.max()) <= 1e-4, float(err.max())
    assert bool((O[:, 3] == 0).all())
    assert ns.analytic(X[:0]).shape == (0, 4)
    assert ns.infer(X[:0]).shape == (0, 4)


def _psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return 10 * math.log10(1.0 / max(mse, 1e-20))


def test_trained_surrogate_reaches_40db_with_fallback(gpu):
    """Train on analytic pairs, render the IBL soup both ways: with the error
    fallback the frame is >= 40 dB PSNR against the analytic frame, the
    fallback pixels are bit-identical to it, and the trained network gains
    >= 15 dB over the same network untrained."""
    import torch
    from taichi_three_amd import workloads as wl
    from taichi_three_amd.neural import NeuralShading
    print('case: nn trained frame 480x270', flush=True)
    w = wl.ibl_1080p(seed=5, res=(480, 270), nfaces=100_000, sky=(512, 256), precision=8)
    scene, mesh = wl.build_scene(w)
    scene.render()
    ref = scene.img.to_numpy().copy()
    ns = NeuralShading(scene.lighting, seed=1)
    scene.neural = ns
    scene.render()
    untrained = scene.img.to_numpy().copy()
    ns.train(nsamples=1 << 21, steps=4000, batch=8192)
    scene.render()
    torch.cuda.synchronize()
    img = scene.img.to_numpy()
    nfall = ns.fallback_count(scene.engine)
    p_tr, p_un = _psnr(img, ref), _psnr(untrained, ref)
    print(f'psnr trained {p_tr:.2f} dB untrained {p_un:.2f} dB fallback {nfall}', flush=True)
    assert p_tr >= 40.0
    assert p_tr - p_un >= 15.0
    assert nfall < 480 * 270
