"""Writes the committed fixtures of tests/golden/ (run from the repository root):

  plain_512.npz    id and depth buffers [H, W] i32 and the tone-mapped colour (centre
                   128x128 crop, [192:320, 192:320]) of
                   the 512x512 sphere + soup frame (seed 0), computed by the CPU
                   oracle (the reference itself needs Taichi, absent here).
  lut_corner.json  texels [0..2] x [0..2] (x = VoH, y = roughness) of the reference's
                   IBL BRDF table assets/lut.jpg in ti.imread orientation ([x, y],
                   y up), the only texels Tina's bilerp reads (tina/matr/material.py:275).
                   Needs the reference checkout; the JSON is committed.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def plain_512():
    from oracle import orc
    from taichi_three_amd import workloads as wl, _abi
    from taichi_three_amd.matr import lower_material
    w = wl.plain_512()
    W, H = w.res
    cam = orc.camera(w.W2V, w.V2W)
    depth, ids = orc.raster(W, H, cam, w.verts)
    lo = lower_material(wl.make_material(w.material))
    flags = _abi.SMOOTH | _abi.VCOLOR | _abi.TONEMAP | _abi.BG
    img = orc.shade(W, H, cam, orc.HostMesh(w.verts, w.norms, None, w.colors), lo.struct,
                    wl.make_lighting(w).to_struct(), ids, flags)
    # colour: the 128x128 centre crop keeps the fixture small
    np.savez_compressed(os.path.join(HERE, "plain_512.npz"), ids=ids, depth=depth, img_crop=img[192:320, 192:320])


def lut_corner(ref='/root/reference/assets/lut.jpg'):
    from PIL import Image
    a = np.asarray(Image.open(ref).convert('RGB'))          # [row, col, c], row 0 at the top
    xy = np.transpose(a, (1, 0, 2))[:, ::-1]                 # [x, y, c] with y up (ti.imread)
    corner = (xy[:3, :3, :2].astype(np.float64) / 255).tolist()
    with open(os.path.join(HERE, 'lut_corner.json'), 'w') as f:
        json.dump({'source': 'assets/lut.jpg (720x718 RGB), texels [x][y][R,G] / 255', 'shape': list(a.shape),
                   'corner': corner}, f, indent=1)


if __name__ == '__main__':
    plain_512()
    if os.path.exists('/root/reference/assets/lut.jpg'):
        lut_corner()
