"""Differentiable shading (extension: the reference computes no gradients).

`DiffShading` renders an indexed mesh with Tina's PBR material
(tina/matr/material.py:322-362, 387-396, 701-706) under point lights
(tina/core/lighting.py:84-98), where albedo, smooth normal, roughness and
metallic are per-vertex tensors and every light has a scalar intensity.  The
rasterisation (tina_raster) fixes visibility; `render(...)` is a
torch.autograd.Function whose forward is k_dshade_fwd and whose backward is the
hand-written adjoint k_dshade_bwd, which recomputes each pixel's face weights
from the saved depth|id buffer rather than storing them.
"""
import ctypes

import numpy as np

from . import _abi

__all__ = ['DiffShading']


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('taichi_three_amd needs a ROCm GPU: torch.cuda.is_available() is False')
    return torch


class DiffShading:
    """Differentiable PBR shading of `faces` over `verts` seen by `engine`.

    lights: [(position (3), colour (3))] point lights; light l shines with
    intensity[l] * colour.  ambient: the ambient colour."""

    def __init__(self, engine, verts, faces, lights, ambient=(0.0, 0.0, 0.0), specular=0.5, culling=True):
        torch = _torch()
        self.engine = engine
        dev = engine.device
        verts = np.asarray(verts, np.float64)
        faces = np.asarray(faces, np.int64)
        if faces.ndim != 2 or faces.shape[1] != 3 or faces.min() < 0 or faces.max() >= len(verts):
            raise ValueError('faces must be [F][3] indices into verts')
        if len(lights) > _abi.MAX_LIGHTS:
            raise ValueError(f'at most {_abi.MAX_LIGHTS} lights')
        self.nverts, self.nfaces = len(verts), len(faces)
        self.soup = torch.from_numpy(verts[faces].astype(np.float32)).to(dev).contiguous()
        self.faces = torch.from_numpy(faces.astype(np.int32)).to(dev).contiguous()
        self.lights = _abi.Lights()
        self.lights.n = len(lights)
        for i, (p, c) in enumerate(lights):
            for k in range(3):
                self.lights.dirs[i][k] = float(p[k])
                self.lights.colors[i][k] = float(c[k])
            self.lights.dirs[i][3] = 1.0
        self.ambient = [float(x) for x in ambient]
        self.specular = float(specular)
        self.flags = (1 if culling else 0) | 2  # TINA_CULL | TINA_CLIP
        engine.reserve(self.nfaces)

    def rasterize(self):
        """Clears the engine's depth|id buffer and rasterises the mesh (ids 0..F-1)."""
        eng = self.engine
        eng.clear_depth()
        _abi.check(eng._lib.tina_raster(eng._ctx, _abi.ptr(self.soup), self.nfaces, 0, self.flags,
                                        _abi.stream_handle()), eng._ctx)

    def _mesh(self, colors, normals, rough, metal, intensity):
        m = _abi.DMesh()
        m.verts, m.faces = _abi.ptr(self.soup), _abi.ptr(self.faces)
        m.normals, m.colors = _abi.ptr(normals), _abi.ptr(colors)
        m.rough, m.metal, m.intensity = _abi.ptr(rough), _abi.ptr(metal), _abi.ptr(intensity)
        m.nfaces, m.nverts, m.id_base = self.nfaces, self.nverts, 0
        for k in range(3):
            m.ambient[k] = self.ambient[k]
        m.specular = self.specular
        return m

    def _check(self, colors, normals, rough, metal, intensity):
        torch = _torch()
        V, L = self.nverts, self.lights.n
        for name, t, shape in (('colors', colors, (V, 3)), ('normals', normals, (V, 3)), ('rough', rough, (V,)),
                               ('metal', metal, (V,)), ('intensity', intensity, (L,))):
            if tuple(t.shape) != shape or t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
                raise ValueError(f'{name} must be a contiguous float32 device tensor of shape {shape}')

    def render(self, colors, normals, rough, metal, intensity):
        """Image [H][W][3] of the last rasterisation, differentiable in every argument."""
        self._check(colors, normals, rough, metal, intensity)
        return _DiffShadeFn.apply(colors, normals, rough, metal, intensity, self)

    def forward_raw(self, colors, normals, rough, metal, intensity, out=None):
        torch = _torch()
        W, H = int(self.engine.res[0]), int(self.engine.res[1])
        img = out if out is not None else torch.zeros((H, W, 3), dtype=torch.float32, device=self.engine.device)
        m = self._mesh(colors, normals, rough, metal, intensity)
        eng = self.engine
        _abi.check(eng._lib.tina_dshade_forward(eng._ctx, ctypes.byref(m), ctypes.byref(self.lights), _abi.ptr(img),
                                                _abi.stream_handle()), eng._ctx)
        return img

    def backward_raw(self, gimg, colors, normals, rough, metal, intensity):
        torch = _torch()
        gs = [torch.zeros_like(t) for t in (colors, normals, rough, metal, intensity)]
        g = _abi.DGrads()
        g.colors, g.normals, g.rough, g.metal, g.intensity = (_abi.ptr(t) for t in gs)
        m = self._mesh(colors, normals, rough, metal, intensity)
        eng = self.engine
        _abi.check(eng._lib.tina_dshade_backward(eng._ctx, ctypes.byref(m), ctypes.byref(self.lights),
                                                 _abi.ptr(gimg), ctypes.byref(g), _abi.stream_handle()), eng._ctx)
        return gs


def _autograd():
    torch = _torch()

    class DiffShadeFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, colors, normals, rough, metal, intensity, ds):
            ctx.ds = ds
            ctx.save_for_backward(colors, normals, rough, metal, intensity)
            return ds.forward_raw(colors, normals, rough, metal, intensity)

        @staticmethod
        def backward(ctx, gimg):
            colors, normals, rough, metal, intensity = ctx.saved_tensors
            gs = ctx.ds.backward_raw(gimg.contiguous(), colors, normals, rough, metal, intensity)
            return (*gs, None)

    return DiffShadeFn


class _Lazy:
    fn = None

    @classmethod
    def apply(cls, *args):
        if cls.fn is None:
            cls.fn = _autograd()
        return cls.fn.apply(*args)


_DiffShadeFn = _Lazy
