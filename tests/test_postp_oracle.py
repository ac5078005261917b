"""The post-processing oracle (oracle/postp.py) against closed forms."""
import numpy as np

from oracle import postp

f32 = np.float32


def test_fxaa_leaves_a_flat_image_alone():
    img = np.full((20, 30, 3), 0.4, f32)
    assert np.array_equal(postp.fxaa(img), img)


def test_fxaa_on_a_vertical_step_edge():
    """Left half black, right half white (luminance L = 0.9999).  At the black
    pixel beside the edge: n + e + s + w = L, the diagonals 2L, so
    filt = |(2 L + 2 L) / 12 - 0| / L = 1/3; at the white one
    |(6 L + 2 L) / 12 - L| / L = 1/3 too.  smoothstep(1/3) = 7/27, squared
    49/729.  vert > hori (a vertical edge), so each pixel is re-sampled along x
    by 49/729 towards the side whose luminance differs most: the black pixel
    reads 49/729 of white, the white one 1 - 49/729."""
    W, H = 16, 8
    img = np.zeros((H, W, 3), f32)
    img[:, W // 2:] = 1.0
    out = postp.fxaa(img)
    b = 49.0 / 729.0
    x0, x1 = W // 2 - 1, W // 2
    assert np.allclose(out[3, x0], b, atol=2e-6) and np.allclose(out[3, x1], 1 - b, atol=2e-6)
    assert np.array_equal(out[:, :x0], img[:, :x0]) and np.array_equal(out[:, x1 + 1:], img[:, x1 + 1:])


def test_bloom_weights_are_a_normalised_gaussian():
    for r in (1, 4, 67):
        w = postp.gauss_weights(r).astype(np.float64)
        assert abs(w[0] + 2 * w[1:].sum() - 1.0) < 1e-6
        ref = np.exp(-(np.arange(r + 1) / r) ** 2)
        assert np.allclose(w, ref / (ref[0] + 2 * ref[1:].sum()), rtol=1e-6)


def test_bloom_below_threshold_is_identity():
    rng = np.random.default_rng(0)
    img = rng.random((24, 40, 3)).astype(f32) * 0.99
    assert np.array_equal(postp.bloom(img, radius=3), img)


def test_bloom_of_a_flat_bright_image_adds_the_filtered_value():
    c = 3.0
    img = np.full((32, 48, 3), c, f32)
    out = postp.bloom(img, radius=2)
    add = 1.0 * (1 - 1 / (1 + 0.25 * (c - 1.0)))
    assert np.allclose(out - img, add, rtol=2e-6)


def test_bloom_of_one_bright_block_is_the_separable_kernel():
    """A 2x2 block of value v at half-res pixel (hx, hy): the blurred half-res
    image is f(v) g_|dx| g_|dy| away from the borders."""
    W, H, r = 64, 48, 3
    img = np.zeros((H, W, 3), f32)
    hx, hy = 10, 9
    img[2 * hy:2 * hy + 2, 2 * hx:2 * hx + 2] = 5.0
    out = postp.bloom(img, radius=r)
    g = postp.gauss_weights(r).astype(np.float64)
    fv = 1 - 1 / (1 + 0.25 * 4.0)
    # full-res pixel (2 hx + 2 dx, 2 hy + 2 dy) samples half-res (hx + dx, hy + dy) exactly
    for dx in range(-r, r + 1):
        for dy in range(-r, r + 1):
            want = fv * g[abs(dx)] * g[abs(dy)]
            y, x = 2 * (hy + dy), 2 * (hx + dx)
            got = out[y, x, 0] - img[y, x, 0]
            assert abs(got - want) < 1e-6 * max(1.0, want), (dx, dy, got, want)
