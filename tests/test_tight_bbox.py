"""The binned raster's per-face bbox shrinking (tight_bbox in
taichi_three_amd/csrc/bin_kernels.h), restated in float64 here, against brute
force in binary32: for random faces -- small, sub-pixel, near-collinear slivers,
vertices on pixel centres, partly off screen -- every pixel of the reference's
bbox (tina/core/triangle.py:108-109) that the cheap rejection keeps lies inside
the shrunken bbox.  Pixels outside it are rejected by that test, which implies
the reference's coverage test fails there, so skipping them changes no output.
The device code itself is pinned by the bit-exact GPU parity tests."""
import math

import numpy as np

f32 = np.float32
U, TINY = 2.0 ** -24, 2.0 ** -40


def _setup(a, b, c, W, H):
    """Screen-space part of setup_face (triangle.py:106-112) in binary32."""
    ax, ay, bx, by, cx, cy = (f32(v) for v in (a[0], a[1], b[0], b[1], c[0], c[1]))
    s = dict(bx=bx, by=by, cx=cx, cy=cy,
             botx=int(max(math.floor(min(ax, bx, cx)), 0)), boty=int(max(math.floor(min(ay, by, cy)), 0)),
             topx=int(min(math.ceil(max(ax, bx, cx)), W - 1)), topy=int(min(math.ceil(max(ay, by, cy)), H - 1)))
    n = f32(f32(f32(bx - ax) * f32(cy - ay)) - f32(f32(by - ay) * f32(cx - ax)))
    with np.errstate(all='ignore'):
        s.update(bcnx=f32((bx - cx) / n), bcny=f32((by - cy) / n), canx=f32((cx - ax) / n), cany=f32((cy - ay) / n))
    return s


def _well_conditioned(s):
    pos = [abs(float(s[k])) for k in ('bx', 'by', 'cx', 'cy')]
    coef = [abs(float(s[k])) for k in ('bcnx', 'bcny', 'canx', 'cany')]
    if any(math.isnan(v) for v in pos + coef):
        return False
    return max(pos) <= 2 ** 20 and max(coef) <= 2 ** 40


def _tight(s, bias):
    bx, by, cx, cy = (float(s[k]) for k in ('bx', 'by', 'cx', 'cy'))
    bcnx, bcny, canx, cany = (float(s[k]) for k in ('bcnx', 'bcny', 'canx', 'cany'))
    full = (s['botx'], s['boty'], s['topx'], s['topy'])
    px0, px1, py0, py1 = s['botx'] + bias[0], s['topx'] + bias[0], s['boty'] + bias[1], s['topy'] + bias[1]
    rx, ry = max(abs(px0), abs(px1)) * U, max(abs(py0), abs(py1)) * U
    dbx, dby = max(abs(px0 - bx), abs(px1 - bx)) + rx + 1, max(abs(py0 - by), abs(py1 - by)) + ry + 1
    dcx, dcy = max(abs(px0 - cx), abs(px1 - cx)) + rx + 1, max(abs(py0 - cy), abs(py1 - cy)) + ry + 1
    m_a, m_b = dbx * abs(bcny) + dby * abs(bcnx), dcx * abs(cany) + dcy * abs(canx)
    e_a, e_b = TINY + 8 * U * m_a, TINY + 8 * U * m_b
    e_t = TINY + e_a + e_b + 4 * U * (1 + m_a + m_b)
    det = bcnx * cany - bcny * canx
    if not abs(det) > (abs(bcnx * cany) + abs(bcny * canx)) * 2 ** -20:
        return full
    inv = 1.0 / det
    A0, B0 = bx * bcny - by * bcnx, cx * cany - cy * canx
    xs, ys, mag = [], [], 0.0
    for a, b in ((-e_a, -e_b), (1 + e_t + e_b, -e_b), (-e_a, 1 + e_t + e_a)):
        ta, tb = a + A0, b + B0
        X, Y = (bcnx * tb - canx * ta) * inv, (bcny * tb - cany * ta) * inv
        xs.append(X)
        ys.append(Y)
        mag = max(mag, (abs(bcnx * tb) + abs(canx * ta) + abs(bcny * tb) + abs(cany * ta)) * abs(inv) + abs(X) + abs(Y))
    d = 2 ** -30 * (mag + 1)
    return (max(math.ceil(min(xs) - d - bias[0]), s['botx']), max(math.ceil(min(ys) - d - bias[1]), s['boty']),
            min(math.floor(max(xs) + d - bias[0]), s['topx']), min(math.floor(max(ys) + d - bias[1]), s['topy']))


def _cheap_survivors(s, bias):
    """Pixels of the reference bbox the cheap rejection keeps (pixel_rejects), binary32."""
    X, Y = np.meshgrid(np.arange(s['botx'], s['topx'] + 1), np.arange(s['boty'], s['topy'] + 1))
    px, py = X.astype(f32) + f32(bias[0]), Y.astype(f32) + f32(bias[1])
    with np.errstate(all='ignore'):
        wbc = (px - s['bx']) * s['bcny'] - (py - s['by']) * s['bcnx']
        wca = (px - s['cx']) * s['cany'] - (py - s['cy']) * s['canx']
        t = (f32(1) - wbc) - wca
        keep = ~(np.fmin(np.fmin(wbc, wca), t) < f32(-TINY))
    return X[keep], Y[keep]


def _faces(rng, W, H, n):
    for i in range(n):
        kind = i % 5
        if kind == 0:    # small
            yield rng.uniform(-20, [W + 20, H + 20]) + rng.normal(0, 3, (3, 2))
        elif kind == 1:  # near-collinear sliver
            c0, d = rng.uniform(0, [W, H]), rng.normal(0, 40, 2)
            yield np.stack([c0, c0 + d, c0 + d * rng.uniform(0.2, 0.8) + rng.normal(0, 1e-3, 2)])
        elif kind == 2:  # sub-pixel
            yield rng.uniform(0, [W, H]) + rng.normal(0, 0.3, (3, 2))
        elif kind == 3:  # vertices on pixel centres
            yield np.floor(rng.uniform(0, [W, H])) + 0.5 + np.round(rng.normal(0, 6, (3, 2)))
        else:            # large, partly off screen
            yield rng.uniform(-200, [W + 200, H + 200]) + rng.normal(0, 60, (3, 2))


def test_shrunken_bbox_holds_every_cheap_survivor():
    rng = np.random.default_rng(5)
    W, H = 1920, 1080
    checked = shrunk = 0
    for bias in ((0.5, 0.5), (0.25, 0.75)):
        for p in _faces(rng, W, H, 6000):
            s = _setup(p[0], p[1], p[2], W, H)
            if s['topx'] < s['botx'] or s['topy'] < s['boty'] or not _well_conditioned(s):
                continue
            if (s['topx'] - s['botx'] + 1) * (s['topy'] - s['boty'] + 1) > 20000:
                continue
            lx, ly, hx, hy = _tight(s, bias)
            shrunk += (lx, ly, hx, hy) != (s['botx'], s['boty'], s['topx'], s['topy'])
            X, Y = _cheap_survivors(s, bias)
            if len(X):
                checked += 1
                assert lx <= X.min() and X.max() <= hx and ly <= Y.min() and Y.max() <= hy, (p.tolist(), bias)
    assert checked > 3000 and shrunk > 3000
