"""Shader protocol (tina/core/shader.py:4-183).

A Taichi shader is a ti.func inlined into render_color.  Here a shader is a
description the raster lowers into one tina_shade launch: Shader(img,
lighting, material) gives the image, the light table and the material program;
ShaderGroup collects it with post shaders such as ProbeShader, which read the
resolved id buffer after the launch.
"""

import numpy as np

__all__ = ['IShader', 'Shader', 'ShaderGroup', 'BackgroundShader', 'NormalShader', 'ViewNormalShader',
           'TexcoordShader', 'ColorShader', 'PositionShader', 'DepthShader', 'ConstShader', 'ViewdirShader',
           'SimpleShader', 'ChessboardShader']


class IShader:
    def __init__(self, img):
        self.img = img

    def clear_buffer(self):
        self.img.fill(0)

    def post_shade(self, raster, id_base, nfaces):
        """Called after the object's pixels are shaded (ProbeShader hooks in here)."""


class Shader(IShader):
    """Lighting.shade_color(material, pos, normal, viewdir) into img (shader.py:112-135)."""

    def __init__(self, img, lighting, material):
        super().__init__(img)
        self.lighting = lighting
        self.material = material


class ShaderGroup(IShader):
    """Runs several shaders per pixel (shader.py:138-148)."""

    def __init__(self, shaders=()):
        self.shaders = list(shaders)

    def clear_buffer(self):
        for s in self.shaders:
            s.clear_buffer()

    def main(self):
        """The group's Shader(img, lighting, material), or None for a group of auxiliary shaders only."""
        mains = [s for s in self.shaders if isinstance(s, Shader)]
        if len(mains) > 1:
            raise NotImplementedError('a ShaderGroup holds at most one Shader(img, lighting, material)')
        return mains[0] if mains else None

    def post_shade(self, raster, id_base, nfaces):
        for s in self.shaders:
            if not isinstance(s, Shader):
                s.post_shade(raster, id_base, nfaces)


class BackgroundShader(IShader):
    """Skybox background where nothing was drawn (shader.py:174-183)."""

    def __init__(self, img, lighting):
        super().__init__(img)
        self.lighting = lighting


class _BufferShader(IShader):
    """An auxiliary shader (tina/core/shader.py:21-109): after an object's colour
    pass (or on its own, passed to render_color) it writes one value per pixel
    the object won into img (tina_gbuffer).  img: a Field [W, H, c] or [W, H]."""
    mode = 0

    def _value(self):
        return None

    def post_shade(self, raster, id_base, nfaces):
        gbuffer(raster, self.mode, self.img.storage, self._value(), getattr(self, 'size', 1))


def gbuffer(raster, mode, out, value=None, size=1):
    """tina_gbuffer: writes mode's per-pixel value (channels = out's last dimension,
    or 1 for [H, W]) into the device tensor `out` at every pixel the raster's
    object won.  Wires have no fragment (IShader.blend_color only)."""
    import ctypes
    from . import _abi
    if raster.nfaces == 0 or not hasattr(raster, 'mesh_struct'):
        return
    channels = int(out.shape[2]) if out.dim() == 3 else 1
    flags = 0
    if type(raster).__name__ == 'ParticleRaster':
        flags = _abi.PARTICLES | (_abi.VCOLOR if raster.coloring and raster._fb.get('colors') is not None else 0)
    else:
        flags |= _abi.SMOOTH if getattr(raster, 'smoothing', False) else 0
        flags |= _abi.TEX if getattr(raster, 'texturing', False) else 0
        flags |= _abi.VCOLOR if raster._fb is not None and raster._fb.get('colors') is not None else 0
    val = (ctypes.c_float * 3)(*([float(x) for x in np.broadcast_to(np.asarray(value, np.float32), 3)]
                                if value is not None else [0.0, 0.0, 0.0]))
    eng = raster.engine
    mesh = raster.mesh_struct()
    _abi.check(eng._lib.tina_gbuffer(eng._ctx, ctypes.byref(mesh), flags, mode, channels, ctypes.byref(val),
                                     ctypes.c_float(size), _abi.ptr(out), _abi.stream_handle()), eng._ctx)


class NormalShader(_BufferShader):
    """img[P] = normal (shader.py:43-46)."""
    mode = 0


class ViewNormalShader(_BufferShader):
    """img[P] = normalize(mapply_dir(W2V, normal)), its xy for a 2-channel img (shader.py:49-56)."""
    mode = 1


class TexcoordShader(_BufferShader):
    """img[P] = texcoord (shader.py:59-62)."""
    mode = 2


class ColorShader(_BufferShader):
    """img[P] = the Input('color') value (shader.py:65-68)."""
    mode = 3


class PositionShader(_BufferShader):
    """img[P] = pos (shader.py:30-33)."""
    mode = 4


class DepthShader(_BufferShader):
    """img[P] = engine.depth[P] (shader.py:36-40)."""
    mode = 5


class ConstShader(_BufferShader):
    """img[P] = value (shader.py:21-27)."""
    mode = 6

    def __init__(self, img, value):
        super().__init__(img)
        self.value = value

    def _value(self):
        return self.value


class ViewdirShader(_BufferShader):
    """img[P] = viewdir * 0.5 + 0.5 (shader.py:96-101)."""
    mode = 7


class SimpleShader(_BufferShader):
    """img[P] = |normal . viewdir| (shader.py:104-109)."""
    mode = 8


class ChessboardShader(_BufferShader):
    """img[P] = lerp((p // size).sum() % 2, 0.4, 0.9), p = P + bias (shader.py:71-78)."""
    mode = 9

    def __init__(self, img, size=8):
        super().__init__(img)
        self.size = size
