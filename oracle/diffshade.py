"""NumPy float64 restatement of the differentiable shading pass -- TEST
INFRASTRUCTURE ONLY (imported by tests/ as a reference, never by the product).

The shading is Tina's PBR material (tina/matr/material.py:322-362, 387-396,
701-706) under point lights (tina/core/lighting.py:84-98) with every material
input interpolated per vertex: albedo, smooth normal, roughness, metallic; the
lights carry a scalar intensity each.  Per covered pixel, with the perspective-
correct weights w of its face (tina/core/triangle.py:140-151):

  n = normalize(sum w N), a = sum w A, r = sum w R, m = sum w M
  f0 = m a + (1 - m) 0.16 specular^2
  C = amb * ((1 - f0) a + f0)
    + sum_l [n.L_l > 0] (n.L_l) (I_l c_l / d_l^2) ((1 - f0) a / pi + f0 CT(r, f0, n, L_l, V))

Pinned (tests/test_diff_oracle.py) against the C oracle's binary64 shade of the
same material with constant roughness and metallic, and by finite differences.
"""
import numpy as np

EPS = 1e-10
INV_PI = 1.0 / np.pi


def weights(W, H, W2V, verts, faces, ids, bias=(0.5, 0.5)):
    """Perspective-correct weights (float64) of every covered pixel's face.
    Returns (pix [P] flat indices, face [P], w [P][3])."""
    M = np.asarray(W2V, np.float64).reshape(4, 4)
    pix = np.nonzero(ids.reshape(-1) >= 0)[0]
    f = ids.reshape(-1)[pix].astype(np.int64)
    P = verts[faces[f]].astype(np.float64)                       # [P][3][3]
    hom = np.concatenate([P, np.ones(P.shape[:2] + (1,))], axis=2) @ M.T
    wv = hom[..., 3]
    ndc = hom[..., :3] / wv[..., None]
    res = np.array([W, H], np.float64)
    sc = (ndc[..., :2] * 0.5 + 0.5) * res                          # [P][3][2]
    a, b, c = sc[:, 0], sc[:, 1], sc[:, 2]
    n = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    bcn = (b - c) / n[:, None]
    can = (c - a) / n[:, None]
    px = np.stack([pix % W + bias[0], pix // W + bias[1]], axis=1).astype(np.float64)
    w_bc = (px[:, 0] - b[:, 0]) * bcn[:, 1] - (px[:, 1] - b[:, 1]) * bcn[:, 0]
    w_ca = (px[:, 0] - c[:, 0]) * can[:, 1] - (px[:, 1] - c[:, 1]) * can[:, 0]
    t = 1.0 - w_bc - w_ca
    w = np.stack([w_bc / wv[:, 0], w_ca / wv[:, 1], t / wv[:, 2]], axis=1)
    w = w / w.sum(axis=1, keepdims=True)
    return pix, f, w


def view_dirs(W, H, V2W, pix, bias=(0.5, 0.5)):
    """calc_viewdir (tina/core/shader.py:81-93) in float64: toward the camera."""
    M = np.asarray(V2W, np.float64).reshape(4, 4)
    x = (pix % W + bias[0]) / W * 2 - 1
    y = (pix // W + bias[1]) / H * 2 - 1

    def app(z):
        h = np.stack([x, y, np.full_like(x, z), np.ones_like(x)], axis=1) @ M.T
        return h[:, :3] / h[:, 3:4]
    d = app(1.0) - app(-1.0)
    return -d / np.linalg.norm(d, axis=1, keepdims=True)


def shade(pos, n_raw, a, r, m, V, lights, intens, amb, specular=0.5):
    """Colour of each pixel (float64).  lights: [(pos(3), colour(3))]; intens [L]."""
    n = n_raw / np.sqrt((n_raw * n_raw).sum(1, keepdims=True))
    f0 = m[:, None] * a + (1 - m[:, None]) * (0.16 * specular * specular)
    C = np.asarray(amb)[None, :] * ((1 - f0) * a + f0)
    r2 = np.maximum(1e-6, r * r)
    for (lp, lc), I in zip(lights, intens):
        L = np.asarray(lp)[None, :] - pos
        d = np.sqrt((L * L).sum(1))
        L = L / d[:, None]
        cos = (n * L).sum(1)
        h = L + V
        h = h / np.sqrt((h * h).sum(1, keepdims=True))
        NoH = np.maximum(EPS, (h * n).sum(1))
        NoL = np.maximum(EPS, (L * n).sum(1))
        NoV = np.maximum(EPS, (V * n).sum(1))
        VoH = np.minimum(1 - EPS, np.maximum(EPS, (h * V).sum(1)))
        den = 1 - NoH * NoH * (1 - r2)
        ndf = r2 / (den * den)
        k = r2 / 2
        vdf = 1 / ((NoV * k + 1 - k) * (NoL * k + 1 - k)) / ((1 - r2) + 4 * np.pi * r2)
        F = f0 + (1 - f0) * ((1 - VoH) ** 5)[:, None]
        brdf = (1 - f0) * a * INV_PI + f0 * (ndf * vdf)[:, None] * F
        lit = (cos > 0)[:, None] * cos[:, None] * (I * np.asarray(lc)[None, :] / (d * d)[:, None]) * brdf
        C = C + lit
    return C


def render(W, H, W2V, V2W, verts, faces, ids, normals, colors, rough, metal, lights, intens, amb,
           specular=0.5):
    """Image [H][W][3] (zero where nothing is covered)."""
    pix, f, w = weights(W, H, W2V, verts, faces, ids)
    fv = faces[f]
    interp = lambda A: np.einsum('pk,pk...->p...', w, np.asarray(A, np.float64)[fv])
    pos = interp(verts)
    C = shade(pos, interp(normals), interp(colors), interp(rough), interp(metal), view_dirs(W, H, V2W, pix),
              lights, intens, amb, specular)
    img = np.zeros((H * W, 3))
    img[pix] = C
    return img.reshape(H, W, 3)
