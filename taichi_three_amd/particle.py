"""ParticleRaster (tina/core/particle.py:4-148) over the C ABI.

Each particle is drawn as the disc its sphere covers: render_occup launches
tina_raster_particles (depth = view-space z of the centre, ids from the
engine's shared id space, so particles and triangles of one scene resolve in
the same depth|id buffer); render_color launches tina_shade with
TINA_PARTICLES, which shades the front-hemisphere point behind each pixel
with the particle's colour, through the same material programs as triangles.
"""
import ctypes

import numpy as np

from . import _abi
from .engine import MAX
from .field import Field
from .matr import lower_material
from .pars import SimpleParticles
from .shader import Shader, ShaderGroup

__all__ = ['ParticleRaster']


class ParticleRaster:
    def __init__(self, engine, maxpars=MAX, coloring=True, clipping=True, **extra_options):
        self.engine = engine
        self.res = engine.res
        self.maxpars = int(maxpars)
        self.coloring = coloring
        self.clipping = clipping
        self.nfaces = 0          # particle count of the bound set (the scene's id bookkeeping)
        self.id_base = 0
        self._fb = None
        self._mesh = None
        self._own = None         # buffers for set_particles & co. used without set_object
        self._lowered = {}

    @property
    def npars(self):
        return self.nfaces

    @property
    def raster_flags(self):
        return _abi.CLIP if self.clipping else 0

    def _own_set(self):
        if self._own is None:
            self._own = SimpleParticles(self.maxpars, radius=0.1, device=self.engine.device)
        return self._own

    def _bind_own(self):
        # the raster's own set keeps its id block while it stays bound
        if self._mesh is self._own and self._fb is not None:
            self._fb = self._own.particle_buffers()
            self.nfaces = int(self._fb['npars'])
            self.engine._id_cursor = max(self.engine._id_cursor, self.id_base + self.nfaces)
        else:
            self.set_object(self._own)

    def set_particles(self, verts):
        """particle.py:45-50: up to maxpars centres into the raster's own buffers, bound for render_occup."""
        self._own_set().set_particles(verts)
        self._bind_own()

    def set_particle_radii(self, sizes):
        self._own_set().set_particle_radii(sizes)
        self._bind_own()

    def set_particle_colors(self, colors):
        if not self.coloring:
            raise ValueError('set_particle_colors needs coloring=True (particle.py:58-59)')
        self._own_set().set_particle_colors(colors)
        self._bind_own()

    def set_object(self, pars):
        """Binds the set's resident buffers (particle.py:64-75) and assigns their ids."""
        if hasattr(pars, 'pre_compute'):
            pars.pre_compute()
        fb = pars.particle_buffers()
        n = int(fb['npars'])
        if n > self.maxpars:
            raise ValueError(f'{n} particles but the raster holds maxpars={self.maxpars} '
                             '(pass maxpars= to the Scene or ParticleRaster)')
        for key in ('pos', 'radius', 'colors'):
            t = fb.get(key)
            if t is not None and n and not (t.is_contiguous() and t.dtype == self.engine.torch.float32):
                fb = dict(fb)
                fb[key] = t.contiguous().float()
        self._fb = fb
        self._mesh = pars
        self.nfaces = n
        self.id_base = self.engine._id_cursor
        self.engine._id_cursor += n

    def render_occup(self):
        """render_occup (particle.py:77-116) -> tina_raster_particles."""
        eng = self.engine
        if self.nfaces == 0:
            return
        _abi.check(eng._lib.tina_raster_particles(eng._ctx, _abi.ptr(self._fb['pos']), _abi.ptr(self._fb['radius']),
                                                  int(self.nfaces), int(self.id_base), int(self.raster_flags),
                                                  _abi.stream_handle()), eng._ctx)

    def mesh_struct(self):
        fb = self._fb
        m = _abi.Mesh()
        m.verts = _abi.ptr(fb['pos'])
        m.norms = _abi.ptr(fb['radius'])
        m.coors = 0
        colors = fb.get('colors') if self.coloring else None
        m.colors = _abi.ptr(colors) if colors is not None else 0
        m.nfaces = int(self.nfaces)
        m.id_base = int(self.id_base)
        return m

    def lowered(self, material):
        key = (id(material), id(self._mesh), getattr(self._mesh, 'version', 0))
        lo = self._lowered.get(key)
        if lo is None:
            lo = lower_material(material, {'attrs': {}})
            self._lowered = {k: v for k, v in self._lowered.items() if k[1] != id(self._mesh)}
            self._lowered[key] = lo
        return lo

    def render_color(self, shader, extra_flags=0, bgcolor=(0.0, 0.0, 0.0), neural=None):
        """render_color (particle.py:118-148) -> tina_shade with TINA_PARTICLES.

        `neural` is accepted for the scene's call signature; the learned shading
        covers triangle meshes only, so particles are always shaded analytically."""
        if self._fb is None or (self.nfaces == 0 and not (extra_flags & _abi.BG)):
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
        flags = _abi.PARTICLES | extra_flags
        if self.coloring and self._fb.get('colors') is not None and self.nfaces > 0:
            flags |= _abi.VCOLOR
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
        mesh = self.mesh_struct()
        if self.nfaces == 0:
            mesh.verts = mesh.verts or 1  # never dereferenced: no id can fall in an empty range
        bg = (ctypes.c_float * 3)(*[float(x) for x in np.broadcast_to(np.asarray(bgcolor, dtype=np.float32), 3)])
        _abi.check(eng._lib.tina_shade(eng._ctx, ctypes.byref(mesh), ctypes.byref(lo.struct), ctypes.byref(lights),
                                       ctypes.byref(ibl) if ibl is not None else None, _abi.ptr(main.img.storage),
                                       ctypes.byref(bg), int(flags), _abi.stream_handle()), eng._ctx)
        shader.post_shade(self, self.id_base, self.nfaces)

    @property
    def occup(self):
        """ParticleRaster.occup (particle.py:14): particle index of the bound set
        per pixel, -1 elsewhere; shape (W, H)."""
        torch = self.engine.torch
        ids, _ = self.engine.resolve(want_depth=False)
        local = ids.long() - self.id_base
        inside = (ids >= 0) & (local >= 0) & (local < self.nfaces)
        return Field(torch.where(inside, local, torch.full_like(local, -1)).to(torch.int32))
