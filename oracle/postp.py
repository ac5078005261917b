"""NumPy restatement of Tina's image post-processing -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ as the reference they compare with; the product never imports it.  Binary32
throughout, with the reference's grouping:

  bloom(img, thresh, factor, scale, gwei, radius)   Blooming.apply
      (tina/postp/blooming.py:45-68)
  fxaa(img, abs_thresh, rel_thresh, factor)         FXAA.apply
      (tina/postp/fxaa.py:26-68)
  ssao(...), ssao_apply(img, ao, nrot, box)          SSAO.render_at / apply
      (tina/postp/ssao.py:37-96), given the depth and normal buffers, the
      sample set and the rotation tile

Images are [H][W][3] float32 with y = 0 at the bottom (the reference's [x, y]
is img[y, x]).  Reads past the image clamp to the edge (the reference reads
outside its fields there: undefined).  Pinned (tests/test_postp_oracle.py) by
closed forms: a constant image passes FXAA unchanged; an image below the bloom
threshold is unchanged; a constant bright image gains exactly the filtered
value; a bright 2x2 block blooms into the separable Gaussian (float64 product
of the weights); a vertical step edge resamples the pixels beside it along x
with the smoothstep blend worked out by hand.
"""
import numpy as np

f32 = np.float32


def _bilerp(img, px, py):
    """bilerp (tina/common.py:140-149) with clamped texels; px, py float32 arrays."""
    h, w = img.shape[:2]
    ix, iy = np.floor(px).astype(np.int64), np.floor(py).astype(np.int64)
    x0, x1 = (px - ix.astype(f32)).astype(f32), (py - iy.astype(f32)).astype(f32)
    y0, y1 = (f32(1) - x0).astype(f32), (f32(1) - x1).astype(f32)
    xa, xb = np.clip(ix, 0, w - 1), np.clip(ix + 1, 0, w - 1)
    ya, yb = np.clip(iy, 0, h - 1), np.clip(iy + 1, 0, h - 1)
    f00, f01, f10, f11 = img[ya, xa], img[yb, xa], img[ya, xb], img[yb, xb]
    x0, x1, y0, y1 = x0[..., None], x1[..., None], y0[..., None], y1[..., None]
    r = x1 * (x0 * f11)
    r = r + y1 * (x0 * f10)
    r = r + y1 * (y0 * f00)
    r = r + x1 * (y0 * f01)
    return r.astype(f32)


def gauss_weights(radius, sigma=1.0):
    """Blooming.init_gwei in float32, summed in index order: w_i = exp(-(sigma i / radius)^2),
    divided by w_0 + 2 sum_{i>0} w_i."""
    w = np.array([np.exp(-(f32(f32(sigma) * f32(i)) / f32(radius)) ** 2) for i in range(radius + 1)], dtype=f32)
    total = f32(-1.0)
    for y in w:
        total = f32(total + f32(y * f32(2)))
    return (w / total).astype(f32)


def bloom(img, thresh=1.0, factor=1.0, scale=0.25, gwei=None, radius=None):
    img = np.asarray(img, dtype=f32)
    H, W = img.shape[:2]
    if radius is None:
        radius = min(W, H) // 16
    if gwei is None:
        gwei = gauss_weights(radius)
    hw, hh = W // 2, H // 2

    def filt(x):
        t = np.maximum(f32(0), (x - f32(thresh)).astype(f32))
        t = (f32(1) - f32(1) / (f32(1) + f32(scale) * t)).astype(f32)
        return (f32(factor) * t).astype(f32)

    half = np.zeros((hh, hw, 3), dtype=f32)
    for jx in (0, 1):
        for jy in (0, 1):
            half = (half + filt(img[jy:2 * hh:2, jx:2 * hw:2])).astype(f32)
    half = (half / f32(4)).astype(f32)
    xs, ys = np.arange(hw), np.arange(hh)
    tmp = (half * gwei[0]).astype(f32)
    for k in range(1, radius + 1):
        val = (half[:, np.maximum(0, xs - k)] + half[:, np.minimum(hw - 1, xs + k)]).astype(f32)
        tmp = (tmp + val * gwei[k]).astype(f32)
    out = (tmp * gwei[0]).astype(f32)
    for k in range(1, radius + 1):
        val = (tmp[np.maximum(0, ys - k)] + tmp[np.minimum(hh - 1, ys + k)]).astype(f32)
        out = (out + val * gwei[k]).astype(f32)
    Y, X = np.meshgrid(np.arange(H), np.arange(W), indexing='ij')
    up = _bilerp(out, (X.astype(f32) / f32(2)).astype(f32), (Y.astype(f32) / f32(2)).astype(f32))
    return (img + up).astype(f32)


def fxaa(img, abs_thresh=0.0625, rel_thresh=0.063, factor=1.0):
    img = np.asarray(img, dtype=f32)
    H, W = img.shape[:2]
    l = ((f32(0.2989) * img[..., 0] + f32(0.587) * img[..., 1]).astype(f32) + f32(0.114) * img[..., 2]).astype(f32)
    lumi = np.clip(l, f32(0), f32(1)).astype(f32)
    Y, X = np.meshgrid(np.arange(H), np.arange(W), indexing='ij')

    def L(dx, dy):
        return lumi[np.clip(Y + dy, 0, H - 1), np.clip(X + dx, 0, W - 1)]

    m, n, e, s, w = L(0, 0), L(0, 1), L(1, 0), L(0, -1), L(-1, 0)
    ne, nw, se, sw = L(1, 1), L(-1, 1), L(1, -1), L(-1, -1)
    hi = np.maximum.reduce([m, n, e, s, w])
    lo = np.minimum.reduce([m, n, e, s, w])
    c = (hi - lo).astype(f32)
    skip = (c < f32(abs_thresh)) | (c < (f32(rel_thresh) * hi).astype(f32))
    with np.errstate(divide='ignore', invalid='ignore'):
        filt = (f32(2) * (((n + e).astype(f32) + s).astype(f32) + w).astype(f32)).astype(f32)
        filt = (filt + (((ne + nw).astype(f32) + se).astype(f32) + sw).astype(f32)).astype(f32)
        filt = np.abs((filt / f32(12)).astype(f32) - m).astype(f32)
        filt = np.clip((filt / c).astype(f32), f32(0), f32(1)).astype(f32)
    t = filt
    sm = ((t * t).astype(f32) * (f32(3) - (f32(2) * t).astype(f32))).astype(f32)
    blend = ((sm * sm).astype(f32) * f32(factor)).astype(f32)
    blend = np.where(skip, f32(0), blend).astype(f32)  # skipped pixels keep their value below

    def ab(a, b, mid):
        return np.abs(((a + b).astype(f32) - (f32(2) * mid).astype(f32)).astype(f32)).astype(f32)

    hori = (ab(n, s, m) * f32(2)).astype(f32)
    hori = (hori + ab(ne, se, e)).astype(f32)
    hori = (hori + ab(nw, sw, w)).astype(f32)
    vert = (ab(e, w, m) * f32(2)).astype(f32)
    vert = (vert + ab(ne, nw, n)).astype(f32)
    vert = (vert + ab(se, sw, s)).astype(f32)
    is_hori = hori >= vert
    plumi, nlumi = np.where(is_hori, n, e), np.where(is_hori, s, w)
    blend = np.where(np.abs((plumi - m).astype(f32)) < np.abs((nlumi - m).astype(f32)), -blend, blend).astype(f32)
    px = (X.astype(f32) + (blend * np.where(is_hori, f32(0), f32(1))).astype(f32)).astype(f32)
    py = (Y.astype(f32) + (blend * np.where(is_hori, f32(1), f32(0))).astype(f32)).astype(f32)
    res = _bilerp(img, px, py)
    return np.where(skip[..., None], img, res).astype(f32)


def _mapply_pos(M, x, y, z):
    """mapply_pos (tina/common.py:169-183) on binary32 arrays, the reference's grouping."""
    M = np.asarray(M, dtype=f32)
    r = []
    for i in range(4):
        a = f32(M[i, 3] * f32(1))
        a = (a + M[i, 0] * x).astype(f32)
        a = (a + M[i, 1] * y).astype(f32)
        a = (a + M[i, 2] * z).astype(f32)
        r.append(a)
    return (r[0] / r[3]).astype(f32), (r[1] / r[3]).astype(f32), (r[2] / r[3]).astype(f32)


def _normalized(x, y, z):
    inv = (f32(1) / np.sqrt(((x * x).astype(f32) + (y * y).astype(f32)).astype(f32) + (z * z).astype(f32)).astype(f32))
    inv = inv.astype(f32)
    return (inv * x).astype(f32), (inv * y).astype(f32), (inv * z).astype(f32)


def _cross(a, b):
    return ((a[1] * b[2]).astype(f32) - (a[2] * b[1]).astype(f32)).astype(f32), \
        ((a[2] * b[0]).astype(f32) - (a[0] * b[2]).astype(f32)).astype(f32), \
        ((a[0] * b[1]).astype(f32) - (a[1] * b[0]).astype(f32)).astype(f32)


def ssao(W2V, V2W, bias, depth, norm, samples, rot, radius=0.2, thresh=0.0, factor=1.0):
    """SSAO.render_at (ssao.py:67-96) for every pixel.  depth [H, W] int32 (the
    engine's depth buffer, 2**30 where empty), norm [H, W, 3], samples [n, 3],
    rot [nr, nr, 2] indexed [x % nr, y % nr].  Returns ao [H, W] float32."""
    depth = np.asarray(depth)
    H, W = depth.shape
    norm = np.asarray(norm, dtype=f32)
    samples = np.asarray(samples, dtype=f32)
    rot = np.asarray(rot, dtype=f32)
    nr = rot.shape[0]
    radius, thresh, factor = f32(radius), f32(thresh), f32(factor)
    Y, X = np.meshgrid(np.arange(H), np.arange(W), indexing='ij')
    resx, resy = f32(W), f32(H)
    px = (X.astype(f32) + f32(bias[0])).astype(f32)
    py = (Y.astype(f32) + f32(bias[1])).astype(f32)
    vx = ((px / resx).astype(f32) * f32(2) - f32(1)).astype(f32)
    vy = ((py / resy).astype(f32) * f32(2) - f32(1)).astype(f32)
    vz = (depth.astype(f32) / f32(2 ** 30)).astype(f32)
    pos = _mapply_pos(V2W, vx, vy, vz)
    ro = _mapply_pos(V2W, vx, vy, np.full_like(vx, f32(-1)))
    ro1 = _mapply_pos(V2W, vx, vy, np.full_like(vx, f32(1)))
    rd = _normalized(*[(b - a).astype(f32) for a, b in zip(ro, ro1)])
    viewdir = [(-c).astype(f32) for c in rd]                           # calc_viewdir (shader.py:91-93)
    back = [(p - (radius * v).astype(f32)).astype(f32) for p, v in zip(pos, viewdir)]
    vradius = (_mapply_pos(W2V, *back)[2] - vz).astype(f32)
    n = [norm[..., k] for k in range(3)]
    up = [np.full_like(vx, f32(c)) for c in (233, 666, 512)]
    bitan = _normalized(*_cross(n, up))                                # tangentspace (advans.py:97-100)
    tan = _cross(bitan, n)
    r = rot[X % nr, Y % nr]
    rx, ry = r[..., 0], r[..., 1]
    occ = np.zeros_like(vx)
    for s in samples:
        sx = ((rx * s[0]).astype(f32) + (ry * s[1]).astype(f32)).astype(f32)
        sy = ((-rx * s[0]).astype(f32) + (ry * s[1]).astype(f32)).astype(f32)
        sz = s[2]
        smp = [((((tan[k] * sx).astype(f32) + (bitan[k] * sy).astype(f32)).astype(f32) + (n[k] * sz).astype(f32))
                .astype(f32)) for k in range(3)]
        smp = [(pos[k] + (smp[k] * radius).astype(f32)).astype(f32) for k in range(3)]
        smp = _mapply_pos(W2V, *smp)
        Dx = ((smp[0] * f32(0.5) + f32(0.5)).astype(f32) * resx).astype(f32)
        Dy = ((smp[1] * f32(0.5) + f32(0.5)).astype(f32) * resy).astype(f32)
        with np.errstate(invalid='ignore'):
            inside = (f32(0) <= Dx) & (Dx < resx) & (f32(0) <= Dy) & (Dy < resy)
        qx = np.where(inside, Dx, 0).astype(np.int64)
        qy = np.where(inside, Dy, 0).astype(np.int64)
        d = (depth[qy, qx].astype(f32) / f32(2 ** 30)).astype(f32)
        with np.errstate(invalid='ignore', divide='ignore'):
            hit = inside & (d < smp[2])
            rc = (vradius / (vz - d).astype(f32)).astype(f32)
            t = np.minimum(f32(1), np.maximum(f32(0), np.abs(rc))).astype(f32)
            rc = ((t * t).astype(f32) * (f32(3) - (f32(2) * t).astype(f32))).astype(f32)
        occ = np.where(hit, (occ + rc).astype(f32), occ).astype(f32)
    ao = (occ / f32(len(samples))).astype(f32)
    ao = (factor * (ao - thresh).astype(f32)).astype(f32)
    return np.minimum(f32(1), np.maximum(f32(0), ao)).astype(f32)


def ssao_apply(img, ao, nrot=4, box=True):
    """SSAO.apply (ssao.py:37-48): img * (1 - the nrot x nrot box sum of ao / nrot**2)
    with reads clamped to the screen, or img * (1 - ao) without the box."""
    img = np.asarray(img, dtype=f32)
    ao = np.asarray(ao, dtype=f32)
    H, W = ao.shape
    if not box:
        return (img * (f32(1) - ao)[..., None]).astype(f32)
    offs = nrot // 2
    Y, X = np.meshgrid(np.arange(H), np.arange(W), indexing='ij')
    r = np.zeros_like(ao)
    for k in range(nrot):
        for l in range(nrot):
            r = (r + ao[np.clip(Y + l - offs, 0, H - 1), np.clip(X + k - offs, 0, W - 1)]).astype(f32)
    m = (f32(1) - (r / f32(nrot * nrot)).astype(f32)).astype(f32)
    return (img * m[..., None]).astype(f32)
