"""TriangleRaster (tina/core/triangle.py:4-153) over the C ABI.

set_object binds the mesh's resident face tensors (no copy) and gives its faces
the next block of global ids; render_occup launches the coverage kernels
(tina_raster); render_color launches the deferred shade (tina_shade) over the
pixels this object won.  Faces of earlier objects hold lower ids, so the
packed-key minimum keeps the reference's rule that an earlier object wins a
depth tie (tina/scene/raster.py:183-187).
"""
import ctypes

import numpy as np

from . import _abi
from .engine import MAX
from .field import Field
from .shader import Shader, ShaderGroup
from .matr import lower_material

__all__ = ['TriangleRaster']


class TriangleRaster:
    def __init__(self, engine, maxfaces=MAX, smoothing=False, texturing=False,
                 culling=True, clipping=True, **extra_options):
        self.engine = engine
        self.res = engine.res
        self.maxfaces = int(maxfaces)
        self.smoothing = smoothing
        self.texturing = texturing
        self.culling = culling
        self.clipping = clipping
        self.nfaces = 0
        self.id_base = 0
        self._fb = None
        self._mesh = None
        self._lowered = {}
        engine.reserve(self.maxfaces)

    @property
    def raster_flags(self):
        return (_abi.CULL if self.culling else 0) | (_abi.CLIP if self.clipping else 0)

    def set_object(self, mesh):
        """Binds the mesh's faces (triangle.py:71-86) and assigns their ids."""
        mesh.pre_compute() if hasattr(mesh, 'pre_compute') else None
        fb = mesh.face_buffers()
        n = int(fb['nfaces'])
        if n > self.maxfaces:
            raise ValueError(f'mesh has {n} faces but the raster holds maxfaces={self.maxfaces} '
                             '(pass maxfaces= to the Scene or TriangleRaster)')
        if n and fb['verts'] is None:
            raise ValueError('mesh has no vertices (set_face_verts)')
        if self.smoothing and n and fb['norms'] is None:
            raise ValueError('smoothing=True needs per-vertex normals on the mesh')
        if self.texturing and n and fb['coors'] is None:
            raise ValueError('texturing=True needs texcoords on the mesh')
        self._fb = fb
        self._mesh = mesh
        self.nfaces = n
        self.id_base = self.engine._id_cursor
        self.engine._id_cursor += n

    def render_occup(self):
        """render_occup (triangle.py:88-131) -> tina_raster."""
        eng = self.engine
        if self.nfaces == 0:
            return
        _abi.check(eng._lib.tina_raster(eng._ctx, _abi.ptr(self._fb['verts']), int(self.nfaces),
                                        int(self.id_base), int(self.raster_flags), _abi.stream_handle()),
                   eng._ctx)

    def mesh_struct(self):
        fb = self._fb
        m = _abi.Mesh()
        m.verts = _abi.ptr(fb['verts'])
        m.norms = _abi.ptr(fb['norms']) if self.smoothing else 0
        m.coors = _abi.ptr(fb['coors']) if self.texturing else 0
        m.colors = _abi.ptr(fb['colors'])
        m.nfaces = int(self.nfaces)
        m.id_base = int(self.id_base)
        return m

    def lowered(self, material):
        key = (id(material), id(self._mesh), getattr(self._mesh, 'version', 0))
        lo = self._lowered.get(key)
        if lo is None:
            lo = lower_material(material, self._fb)
            self._lowered = {k: v for k, v in self._lowered.items() if k[1] != id(self._mesh)}
            self._lowered[key] = lo
        return lo

    def shade_flags(self, lowered):
        f = 0
        if self.smoothing:
            f |= _abi.SMOOTH
        if self.texturing:
            f |= _abi.TEX
        if self._fb is not None and self._fb['colors'] is not None:
            f |= _abi.VCOLOR
        return f

    def render_color(self, shader, extra_flags=0, bgcolor=(0.0, 0.0, 0.0), neural=None):
        """render_color (triangle.py:133-153) -> tina_shade for this object's pixels.

        With `neural` (a trained NeuralShading for this shader's lighting) a PBR
        material is shaded through the network instead (tina_shade_nn)."""
        if self.nfaces == 0 and not (extra_flags & _abi.BG):
            return
        main = shader.main() if isinstance(shader, ShaderGroup) else shader
        if main is None or hasattr(main, 'mode'):
            # auxiliary shaders only (NormalShader, DepthShader, ...): no colour pass
            if self.nfaces:
                shader.post_shade(self, self.id_base, self.nfaces)
            return
        if not isinstance(main, Shader):
            raise NotImplementedError(f'{type(main).__name__} is not a lowerable shader')
        eng = self.engine
        lo = self.lowered(main.material)
        flags = self.shade_flags(lo) | extra_flags
        lighting = main.lighting
        ibl = None
        from .lighting import SkyboxLighting
        if isinstance(lighting, SkyboxLighting):
            if not lo.ibl_ok:
                raise NotImplementedError('this material has no sample_ibl (only Lambert, CookTorrance, '
                                          'Mix and Scale do, tina/matr/material.py)')
            flags |= _abi.IBL
            ibl = lighting.ibltab.struct
            if lighting.nlights == 0:
                flags |= _abi.NOLIGHT
        lights = lighting.to_struct()
        img = main.img.storage
        mesh = self.mesh_struct() if self._fb is not None else _abi.Mesh()
        if self._fb is None or self.nfaces == 0:
            mesh.nfaces = 0
            mesh.verts = mesh.verts or 1  # never dereferenced: no face id can fall in an empty range
        bg = (ctypes.c_float * 3)(*[float(x) for x in np.broadcast_to(np.asarray(bgcolor, dtype=np.float32), 3)])
        if neural is not None and self._fb is not None and lo.is_pbr and neural.lighting is lighting \
                and (flags & _abi.TONEMAP) and not (flags & (_abi.TEX | _abi.NOLIGHT)):
            _abi.check(eng._lib.tina_shade_nn(eng._ctx, ctypes.byref(mesh), ctypes.byref(lo.struct),
                                              ctypes.byref(lights), ctypes.byref(ibl), _abi.ptr(neural.packed),
                                              _abi.ptr(img), ctypes.byref(bg), int(flags),
                                              ctypes.c_float(neural.threshold), _abi.stream_handle()), eng._ctx)
            shader.post_shade(self, self.id_base, self.nfaces)
            return
        _abi.check(eng._lib.tina_shade(eng._ctx, ctypes.byref(mesh), ctypes.byref(lo.struct),
                                       ctypes.byref(lights), ctypes.byref(ibl) if ibl is not None else None,
                                       _abi.ptr(img), ctypes.byref(bg), int(flags), _abi.stream_handle()),
                   eng._ctx)
        shader.post_shade(self, self.id_base, self.nfaces)

    @property
    def occup(self):
        """TriangleRaster.occup (triangle.py:16): face index of the last bound
        object per pixel, -1 elsewhere; shape (W, H)."""
        torch = self.engine.torch
        ids, _ = self.engine.resolve(want_depth=False)
        local = ids.long() - self.id_base
        inside = (ids >= 0) & (local >= 0) & (local < self.nfaces)
        return Field(torch.where(inside, local, torch.full_like(local, -1)).to(torch.int32))
