"""Pins the surrogate's NumPy oracle (oracle/nn.py) against finite differences,
closed forms and torch's own bfloat16 conversion."""
import numpy as np
import pytest

from oracle import nn


def _batch(seed, B=8):
    rng = np.random.default_rng(seed)
    n = rng.normal(size=(B, 3)); n /= np.linalg.norm(n, axis=1, keepdims=True)
    V = rng.normal(size=(B, 3)); V /= np.linalg.norm(V, axis=1, keepdims=True)
    V = np.where((V * n).sum(1, keepdims=True) < 0, -V, V)
    L = np.array([0.4, 0.5, -0.6]); L /= np.linalg.norm(L)
    X = nn.features(n, V, L, rng.uniform(0.05, 1, B), rng.random(B), rng.random((B, 3)), dtype=np.float64)
    Y = rng.random((B, 4))
    return X, Y


def test_layout_counts_the_spec_network():
    """12-64-64-64-3 has 9,347 parameters; the error row adds 64 + 1."""
    assert nn.NPARAMS == 9347 + 65


@pytest.mark.parametrize('seed', [0, 1])
def test_backward_matches_central_differences(seed):
    """Trunk and colour rows against d(colour loss)/dp; the error row (W4[3],
    b4[3]) against d(loss)/dp -- the error term trains that row only."""
    X, Y = _batch(seed)
    p = nn.init_params(seed) * 0.5 + 0.01 * np.random.default_rng(seed).normal(size=nn.NPARAMS)
    g = nn.backward(p, X, Y)
    rng = np.random.default_rng(100 + seed)
    lay = {name: (off, int(np.prod(shape))) for name, shape, off in nn.layout()}
    err_idx = [lay['W4'][0] + 3 * 64 + k for k in range(0, 64, 7)] + [lay['b4'][0] + 3]
    trunk_idx = list(rng.choice(lay['W4'][0] + 3 * 64, 60, replace=False)) + \
        [lay['b4'][0] + k for k in range(3)]
    h = 1e-6
    for idx, part in [(i, 'colour') for i in trunk_idx] + [(i, 'all') for i in err_idx]:
        e = np.zeros(nn.NPARAMS); e[idx] = h
        fd = (nn.loss(p + e, X, Y, part) - nn.loss(p - e, X, Y, part)) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-6 + 1e-4 * abs(fd), (idx, part, fd, g[idx])


def test_first_adam_step_moves_by_lr_times_sign():
    rng = np.random.default_rng(3)
    p, g = rng.normal(size=50), rng.normal(size=50)
    p1, m, v = nn.adam(p, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    assert np.allclose(p1, p - 1e-3 * np.sign(g), rtol=0, atol=1e-8)


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.normal(size=10000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 10000),
                        np.array([0.0, -0.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, np.inf, -np.inf, 3.4e38],
                                 np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(nn.bf16(x), ref)


def test_features_mirror_view_about_the_normal():
    """R = reflect(-V, n) is a unit vector with R.n = V.n and R + V parallel to n."""
    X, _ = _batch(5, B=64)
    n, R = X[:, 0:3], X[:, 3:6]
    rng = np.random.default_rng(5)
    assert np.allclose(np.linalg.norm(R, axis=1), 1.0, atol=1e-12)
    V = 2 * (R * n).sum(1, keepdims=True) * n - R
    assert np.allclose(np.linalg.norm(V, axis=1), 1.0, atol=1e-12)
    assert np.allclose(np.cross(R + V, n), 0.0, atol=1e-12)


def test_bf16_inference_tracks_the_float64_network():
    X, _ = _batch(6, B=256)
    p = nn.init_params(6)
    o64, _ = nn.forward(p, X)
    o16 = nn.infer_bf16(p, X.astype(np.float32))
    assert np.abs(np.maximum(o64[:, :3], 0) - o16[:, :3]).max() <= 0.05 * np.abs(o64).max()
