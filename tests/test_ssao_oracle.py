"""Pins the SSAO restatement (oracle/postp.py: ssao, ssao_apply) against closed
forms under a scaling camera (view = (x/2, y/2, -z/2)) with normals facing the
camera:

  - a flat depth buffer occludes nothing (every sample rises towards the camera)
  - a one-pixel pit of depth D_far in a floor at D_near: only the pit is
    occluded, by smoothstep(v_r / (D_far - D_near)) per sample that lands off
    the pit pixel, v_r = radius / 2 (the view depth of radius along the view
    ray), counted in float64 with the rotation tile's matrix [[c, s], [-c, s]]
    (ssao.py:80-81); factor and thresh scale and shift it before the clamp
  - apply: a single occluded pixel darkens the 4 x 4 box of pixels that read
    it (offsets -2..1) by ao / 16; at the corner the clamped reads count it
    nine times; the TAA form multiplies by 1 - ao
"""
import numpy as np

from oracle import postp as opp

W, H = 40, 32
W2V = np.diag([0.5, 0.5, -0.5, 1.0]).astype(np.float32)
V2W = np.diag([2.0, 2.0, -2.0, 1.0]).astype(np.float32)
BIAS = (0.5, 0.5)


def _inputs(seed=0, n=64):
    from taichi_three_amd.postp import SSAO
    rng = np.random.default_rng(seed)
    samples = SSAO.make_samples(rng, n)
    t = rng.random((4, 4)) * 2 * np.pi
    rot = np.stack([np.cos(t), np.sin(t)], axis=2).astype(np.float32)
    norm = np.zeros((H, W, 3), np.float32)
    norm[..., 2] = 1                                   # facing the camera (world +z)
    return samples, rot, norm


def test_flat_floor_is_not_occluded():
    samples, rot, norm = _inputs()
    depth = np.full((H, W), 2 ** 28, np.int32)         # view depth 0.25
    ao = opp.ssao(W2V, V2W, BIAS, depth, norm, samples, rot, radius=0.2)
    assert np.all(ao == 0)


def test_one_pixel_pit():
    samples, rot, norm = _inputs(3)
    near, far = 2 ** 28, 2 ** 29                       # view depths 0.25 and 0.5
    cx, cy = 17, 13
    depth = np.full((H, W), near, np.int32)
    depth[cy, cx] = far
    radius = 0.2
    # float64 landing pixel of every sample taken at the pit
    n = np.array([0.0, 0.0, 1.0])
    bitan = np.cross(n, [233.0, 666.0, 512.0])
    bitan /= np.linalg.norm(bitan)
    tan = np.cross(bitan, n)
    c, s = rot[cx % 4, cy % 4].astype(np.float64)
    pos = np.array([((cx + 0.5) / W * 2 - 1) * 2, ((cy + 0.5) / H * 2 - 1) * 2])
    off_pit, ambiguous = 0, 0
    for sm in samples.astype(np.float64):
        sx, sy = c * sm[0] + s * sm[1], -c * sm[0] + s * sm[1]
        w = pos + radius * (tan[:2] * sx + bitan[:2] * sy)
        D = (w * 0.5 * 0.5 + 0.5) * [W, H]
        ambiguous += np.any(np.abs(D - np.round(D)) < 1e-4)
        off_pit += (int(np.floor(D[0])), int(np.floor(D[1]))) != (cx, cy)
    assert ambiguous == 0 and 0 < off_pit < len(samples)
    g = 0.4 ** 2 * (3 - 2 * 0.4)                       # smoothstep((radius / 2) / (0.5 - 0.25))
    for factor, thresh in ((1.0, 0.0), (2.0, 0.05)):
        ao = opp.ssao(W2V, V2W, BIAS, depth, norm, samples, rot, radius=radius, thresh=thresh, factor=factor)
        exp = np.clip(factor * (g * off_pit / len(samples) - thresh), 0, 1)
        assert abs(ao[cy, cx] - exp) <= 1e-5, (ao[cy, cx], exp)
        ao[cy, cx] = 0
        assert np.all(ao == 0)


def test_apply_box_and_taa_forms():
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 2, (H, W, 3)).astype(np.float32)
    ao = np.zeros((H, W), np.float32)
    ao[10, 20] = 0.8
    out = opp.ssao_apply(img, ao, 4, box=True)
    hit = np.zeros((H, W), bool)
    hit[9:13, 19:23] = True                            # rows/cols y0 - 1 .. y0 + 2
    assert np.allclose(out[hit], img[hit] * (1 - 0.8 / 16), rtol=1e-6)
    assert np.array_equal(out[~hit], img[~hit])
    corner = np.zeros((H, W), np.float32)
    corner[0, 0] = 0.8
    out = opp.ssao_apply(img, corner, 4, box=True)
    assert np.isclose(out[0, 0, 0], img[0, 0, 0] * (1 - 9 * 0.8 / 16), rtol=1e-6)
    assert np.allclose(opp.ssao_apply(img, corner, 4, box=False)[0, 0], img[0, 0] * 0.2, rtol=1e-6)
