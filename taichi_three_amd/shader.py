"""Shader protocol (tina/core/shader.py:4-183).

A Taichi shader is a ti.func inlined into render_color.  Here a shader is a
description the raster lowers into one tina_shade launch: Shader(img,
lighting, material) gives the image, the light table and the material program;
ShaderGroup collects it with post shaders such as ProbeShader, which read the
resolved id buffer after the launch.
"""

__all__ = ['IShader', 'Shader', 'ShaderGroup', 'BackgroundShader', 'NormalShader']


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
        mains = [s for s in self.shaders if isinstance(s, Shader)]
        if len(mains) != 1:
            raise NotImplementedError('a ShaderGroup needs exactly one Shader(img, lighting, material)')
        return mains[0]

    def post_shade(self, raster, id_base, nfaces):
        for s in self.shaders:
            if not isinstance(s, Shader):
                s.post_shade(raster, id_base, nfaces)


class BackgroundShader(IShader):
    """Skybox background where nothing was drawn (shader.py:174-183)."""

    def __init__(self, img, lighting):
        super().__init__(img)
        self.lighting = lighting


class NormalShader(IShader):
    """img[P] = the shading normal (shader.py:43-46), for the pixels each object
    won; runs after that object's colour pass (tina_normals).  Wire pixels are
    left as they are (wires have no normal)."""

    def post_shade(self, raster, id_base, nfaces):
        from . import _abi
        if nfaces == 0 or not hasattr(raster, 'mesh_struct'):
            return
        flags = 0
        if type(raster).__name__ == 'ParticleRaster':
            flags = _abi.PARTICLES
        elif getattr(raster, 'smoothing', False):
            flags = _abi.SMOOTH
        eng = raster.engine
        mesh = raster.mesh_struct()
        import ctypes
        _abi.check(eng._lib.tina_normals(eng._ctx, ctypes.byref(mesh), flags, _abi.ptr(self.img.storage),
                                         _abi.stream_handle()), eng._ctx)
