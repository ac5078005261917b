"""WireframeRaster (tina/core/wireframe.py:4-95) over the C ABI.

The reference draws wires during render_color: each draw_line step that beats
the depth buffer at that moment blends the line colour in.  Here the steps enter
the shared depth|id buffer in render_occup (tina_raster_wires) and render_color
paints the pixels the wires won (tina_paint_ids); since every later object's
colour pass writes only the pixels it won, the final image is the same: a wire
pixel keeps the line colour unless a nearer (or equally near and earlier)
object covers it.  The material of the object is not used (wireframe.py:94).
"""
import ctypes

import numpy as np

from . import _abi
from .engine import MAX
from .field import Field
from .mesh import SimpleMesh

__all__ = ['WireframeRaster']


class WireframeRaster:
    def __init__(self, engine, maxwires=MAX, linewidth=1.5, linecolor=(.9, .6, 0), clipping=False, **extra_options):
        self.engine = engine
        self.res = engine.res
        self.maxfaces = int(maxwires)
        self.linewidth = linewidth          # kept for the API; the reference draws one-pixel lines
        self.linecolor = np.asarray(linecolor, dtype=np.float32).reshape(3)
        self.clipping = clipping
        self.nfaces = 0
        self.id_base = 0
        self._fb = None
        self._mesh = None
        self._own = None

    @property
    def nwires(self):
        return self.nfaces

    @property
    def raster_flags(self):
        return _abi.CLIP if self.clipping else 0

    def set_object(self, mesh):
        """Binds the wire end points [n, 2, 3] of a mesh with npolygon 2 (wireframe.py:32-39)."""
        if hasattr(mesh, 'pre_compute'):
            mesh.pre_compute()
        if hasattr(mesh, 'get_npolygon') and mesh.get_npolygon() != 2:
            raise ValueError('WireframeRaster draws meshes of 2-corner faces (npolygon=2; see MeshToWire)')
        fb = mesh.face_buffers()
        n = int(fb['nfaces'])
        if n > self.maxfaces:
            raise ValueError(f'{n} wires but the raster holds maxwires={self.maxfaces}')
        v = fb['verts']
        if n and (v is None or tuple(v.shape[1:]) != (2, 3)):
            raise ValueError('wire vertices must be [n, 2, 3]')
        if n and not v.is_contiguous():
            fb = dict(fb)
            fb['verts'] = v.contiguous()
        self._fb = fb
        self._mesh = mesh
        self.nfaces = n
        self.id_base = self.engine._id_cursor
        self.engine._id_cursor += n

    def set_wire_verts(self, verts):
        """wireframe.py:41-47: end points [n, 2, 3] into the raster's own buffer, bound."""
        if self._own is None:
            self._own = SimpleMesh(self.maxfaces, npolygon=2)
        self._own.set_face_verts(verts)
        if self._mesh is self._own and self._fb is not None:
            self._fb = self._own.face_buffers()
            self.nfaces = int(self._fb['nfaces'])
            self.engine._id_cursor = max(self.engine._id_cursor, self.id_base + self.nfaces)
        else:
            self.set_object(self._own)

    def render_occup(self):
        """The depth test of wireframe.py:84-93 -> tina_raster_wires."""
        eng = self.engine
        if self.nfaces == 0:
            return
        _abi.check(eng._lib.tina_raster_wires(eng._ctx, _abi.ptr(self._fb['verts']), int(self.nfaces),
                                              int(self.id_base), int(self.raster_flags), _abi.stream_handle()),
                   eng._ctx)

    def render_color(self, shader, extra_flags=0, bgcolor=(0.0, 0.0, 0.0), neural=None):
        """blend_color(linecolor) on the pixels the wires won (wireframe.py:94-95) -> tina_paint_ids.

        extra_flags: TINA_BG paints the background where nothing was drawn (the
        caller handles a skybox background), TINA_TONEMAP maps through ACES."""
        from .shader import ShaderGroup
        eng = self.engine
        main = shader.main() if isinstance(shader, ShaderGroup) else shader
        if main is None or hasattr(main, 'mode'):
            return  # auxiliary shaders leave wire pixels alone (IShader.blend_color)
        flags = int(extra_flags) & (_abi.BG | _abi.TONEMAP)
        if self.nfaces == 0 and not (flags & _abi.BG):
            return
        col = (ctypes.c_float * 3)(*[float(x) for x in self.linecolor])
        bg = (ctypes.c_float * 3)(*[float(x) for x in np.broadcast_to(np.asarray(bgcolor, dtype=np.float32), 3)])
        _abi.check(eng._lib.tina_paint_ids(eng._ctx, int(self.id_base), int(self.nfaces), ctypes.byref(col),
                                           ctypes.byref(bg), flags, _abi.ptr(main.img.storage),
                                           _abi.stream_handle()), eng._ctx)
        shader.post_shade(self, self.id_base, self.nfaces)

    @property
    def occup(self):
        """Wire index of the bound set per pixel, -1 elsewhere; shape (W, H)."""
        torch = self.engine.torch
        ids, _ = self.engine.resolve(want_depth=False)
        local = ids.long() - self.id_base
        inside = (ids >= 0) & (local >= 0) & (local < self.nfaces)
        return Field(torch.where(inside, local, torch.full_like(local, -1)).to(torch.int32))
