"""Asset readers without a GPU: the PNG decoder against images encoded here with
each filter type of the PNG specification and the Paeth predictor's tie rules;
readgltf on the reference's glTF assets against what their geometry fixes (the
sphere's vertices on the unit sphere with outward unit normals, the index
ranges, the node transforms as rotations)."""
import os
import struct
import zlib

import numpy as np
import pytest

from taichi_three_amd.pngread import paeth, read_png
from taichi_three_amd.gltf import readgltf

ASSETS = os.path.join(os.path.dirname(__file__), 'golden', 'assets')


def _chunk(kind, body):
    return struct.pack('>I', len(body)) + kind + body + struct.pack('>I', zlib.crc32(kind + body) & 0xFFFFFFFF)


def _encode(img, ctype, depth=8, filters=(0, 1, 2, 3, 4), palette=None, trns=None):
    """A PNG of img ([h, w, c] or [h, w]) whose rows use the filter types in turn.
    The forward filters: Sub x - a, Up x - b, Average x - floor((a + b) / 2),
    Paeth x - PaethPredictor(a, b, c), bytes mod 256."""
    h = img.shape[0]
    if depth == 16:
        rows = img.astype('>u2').reshape(h, -1).view(np.uint8).reshape(h, -1)
        bpp = 2 * (img.shape[2] if img.ndim == 3 else 1)
    elif depth < 8:
        per = 8 // depth
        flat = img.reshape(h, -1)
        padded = np.zeros((h, -(-flat.shape[1] // per) * per), np.uint8)
        padded[:, :flat.shape[1]] = flat
        g = padded.reshape(h, -1, per)
        rows = np.zeros(g.shape[:2], np.uint8)
        for k in range(per):
            rows |= (g[:, :, k] << (8 - depth * (k + 1))).astype(np.uint8)
        bpp = 1
    else:
        rows = img.reshape(h, -1).astype(np.uint8)
        bpp = img.shape[2] if img.ndim == 3 else 1
    out = bytearray()
    prev = np.zeros(rows.shape[1], np.int32)
    for y in range(h):
        x = rows[y].astype(np.int32)
        a = np.concatenate([np.zeros(bpp, np.int32), x[:-bpp]])
        c = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]])
        f = filters[y % len(filters)]
        pred = [0, a, prev, (a + prev) // 2, paeth(a, prev, c)][f]
        out.append(f)
        out += ((x - pred) & 0xFF).astype(np.uint8).tobytes()
        prev = x
    w = img.shape[1]
    data = b'\x89PNG\r\n\x1a\n' + _chunk(b'IHDR', struct.pack('>IIBBBBB', w, h, depth, ctype, 0, 0, 0))
    if palette is not None:
        data += _chunk(b'PLTE', palette.astype(np.uint8).tobytes())
    if trns is not None:
        data += _chunk(b'tRNS', trns.astype(np.uint8).tobytes())
    raw = zlib.compress(bytes(out))
    data += _chunk(b'IDAT', raw[:len(raw) // 2]) + _chunk(b'IDAT', raw[len(raw) // 2:]) + _chunk(b'IEND', b'')
    return data


def test_paeth_tie_rules():
    # p = a + b - c; the nearest of a, b, c to p, ties to a, then b
    assert paeth(10, 20, 15) == 15      # pa = pb = 5 > pc = 0
    assert paeth(1, 2, 0) == 2          # pa = 2, pb = 1, pc = 3
    assert paeth(5, 5, 5) == 5
    assert paeth(7, 3, 5) == 5          # p = 5: pc = 0
    assert paeth(4, 10, 4) == 10        # p = 10: pb = 0
    assert paeth(10, 4, 4) == 10        # p = 10: pa = 0
    assert paeth(3, 9, 6) == 6          # p = 6: pa = pb = 3, pc = 0
    assert paeth(2, 8, 5) == 5


@pytest.mark.parametrize('ctype,ch', [(0, 1), (2, 3), (4, 2), (6, 4)])
def test_png_filters_round_trip(ctype, ch):
    rng = np.random.default_rng(ctype)
    img = rng.integers(0, 256, (13, 11, ch), dtype=np.uint8)
    img[5] = 255                        # rows where the filters wrap mod 256
    got = read_png(_encode(img if ch > 1 else img[..., 0], ctype))
    assert np.array_equal(got, img if ch > 1 else img[..., 0])


def test_png_16bit_palette_and_low_depth():
    rng = np.random.default_rng(9)
    img16 = rng.integers(0, 65536, (7, 5, 3), dtype=np.uint16)
    assert np.array_equal(read_png(_encode(img16, 2, depth=16)), img16)
    pal = rng.integers(0, 256, (6, 3))
    idx = rng.integers(0, 6, (9, 10), dtype=np.uint8)
    trns = np.array([0, 128, 255])
    got = read_png(_encode(idx, 3, palette=pal, trns=trns))
    alpha = np.array([0, 128, 255, 255, 255, 255])
    assert np.array_equal(got[..., :3], pal[idx]) and np.array_equal(got[..., 3], alpha[idx])
    grey4 = rng.integers(0, 16, (6, 7), dtype=np.uint8)
    assert np.array_equal(read_png(_encode(grey4, 0, depth=4)), grey4 * 17)   # scaled to 8 bits


def test_png_rejects_a_bad_crc():
    data = bytearray(_encode(np.zeros((2, 2, 3), np.uint8), 2))
    data[20] ^= 1                       # inside IHDR
    with pytest.raises(ValueError):
        read_png(bytes(data))


def test_reference_png_asset_and_its_gltf_copy():
    img = read_png(os.path.join(ASSETS, 'smallptwall.png'))
    assert img.shape == (512, 512, 3) and img.dtype == np.uint8
    scene = readgltf(os.path.join(ASSETS, 'cornell.gltf'))
    mats = [p.material for n in scene.nodes for p in (n.primitives or [])]
    textured = [m for m in mats if 'baseColorTexture' in (m.pbr or {})]
    assert len(textured) == 1
    emb = textured[0]._images(0)
    assert np.array_equal(emb, np.swapaxes(img, 0, 1))   # [x, y] with row 0 of the file at y = 0


def test_gltf_sphere_geometry():
    scene = readgltf(os.path.join(ASSETS, 'sphere.gltf'))
    nodes = {n.name: n for n in scene.nodes}
    assert set(nodes) == {'Light', 'Camera', 'Sphere', 'Plane', 'Point'}
    (sph,) = nodes['Sphere'].primitives
    o = sph.to_obj_mesh()
    r = np.linalg.norm(o['v'], axis=1)
    assert np.allclose(r, 1, atol=1e-5)                            # a unit UV sphere
    assert np.allclose(np.linalg.norm(o['vn'], axis=1), 1, atol=1e-4)
    assert np.all(np.sum(o['vn'] * o['v'], axis=1) > 0.9)          # outward smooth normals
    assert o['f'].min() == 0 and o['f'].max() == len(o['v']) - 1 and o['f'].shape == (960, 3)
    assert np.all((o['vt'] >= -1e-6) & (o['vt'] <= 1 + 1e-6))                 # Blender's UVs, to rounding
    assert sph.material.pbr['baseColorFactor'][:3] == [0, 1, 0]
    # Plane: translate @ rotate @ scale, the rotation part orthonormal with determinant 1
    M = nodes['Plane'].matrix()
    t, q, s = nodes['Plane'].trans
    L = M[:3, :3] / np.asarray(s, float)[None, :]
    assert np.allclose(L @ L.T, np.eye(3), atol=1e-6) and np.isclose(np.linalg.det(L), 1, atol=1e-6)
    assert np.allclose(M[:3, 3], t) and np.allclose(M[3], [0, 0, 0, 1])
    # the quaternion (x, y, z, w) rotates by 2 acos(w) about (x, y, z)
    axis = np.asarray(q[:3]) / np.linalg.norm(q[:3])
    assert np.allclose(L @ axis, axis, atol=1e-6)
    v = np.cross(axis, [1, 0, 0])
    v /= np.linalg.norm(v)
    assert np.isclose(np.dot(L @ v, v), np.cos(2 * np.arccos(q[3])), atol=1e-6)


def test_gltf_other_assets_parse():
    caus = readgltf(os.path.join(ASSETS, 'caustics.gltf'))
    cube = [n for n in caus.nodes if n.name == 'Cube'][0]
    o = cube.primitives[0].to_obj_mesh()
    assert o['v'].shape == (24, 3) and o['f'].shape == (12, 3) and cube.primitives[0].material is None
    assert np.allclose(np.abs(o['v']), 1)                           # the unit cube's corners
    corn = readgltf(os.path.join(ASSETS, 'cornell.gltf'))
    assert sum(len(n.primitives or []) for n in corn.nodes) == 3
