"""Scene (tina/scene/raster.py:4-258): the orchestrator scene.render() drops into.

render() keeps the reference's order -- clear, then per object set_object /
render_occup / render_color, then tone mapping -- but launches it the MI355X
way: all coverage passes first into the shared depth|id buffer, then one
shade launch per object over the pixels that object won, with the background
fill and ACES fused into those launches.  The image equals the reference's:
every pixel ends up shaded by its overall depth winner, ties to earlier objects.
"""
import numpy as np

from . import _abi
from .engine import Engine, MAX
from .field import Field
from .lighting import Lighting, SkyboxLighting
from .matr import Diffuse
from .postp import ToneMapping, Blooming, FXAA, Accumator, SSAO, SSR
from .raster import TriangleRaster
from .particle import ParticleRaster
from .wireframe import WireframeRaster
from .shader import Shader, ShaderGroup, NormalShader, TexcoordShader

__all__ = ['Scene']


class _Obj:
    def __init__(self, material, raster):
        self.material = material
        self.raster = raster
        self.id_base = 0
        self.nfaces = 0
        self.fb = None


class Scene:
    def __init__(self, res_x=512, res_y=None, **options):
        self.options = options
        self.engine = Engine(res_x, res_y, maxfaces=options.get('maxfaces', MAX))
        self.res = self.engine.res
        self.ibl = options.get('ibl', False)
        # a trained NeuralShading for this scene's lighting shades its PBR objects (extension)
        self.neural = options.get('neural')
        self.tonemap = options.get('tonemap', True)
        self.bgcolor = options.get('bgcolor', 0)
        if not self.ibl:
            self.lighting = Lighting()
            # default lights (raster.py:89-93)
            self.lighting.add_light(dir=[1, 2, 3], color=[0.9, 0.9, 0.9])
            self.lighting.set_ambient_light([0.1, 0.1, 0.1])
        else:
            from .skybox import Skybox
            sky = options.get('skybox')
            if sky is None:
                sky = Skybox.procedural()
            self.lighting = SkyboxLighting(sky, precision=options.get('ibl_precision', 32))
        import torch
        W, H = int(self.res[0]), int(self.res[1])
        self.image = Field(torch.zeros((H, W, 3), dtype=torch.float32, device=self.engine.device))
        self.default_material = Diffuse()
        self.post_shaders = []
        self.pre_shaders = []
        self.materials = []
        self.shaders = {}
        self.objects = {}
        if self.tonemap:
            self.tonemap = ToneMapping(self.res)
        # post passes in the reference's order (raster.py:200-205): blooming on the
        # linear image, then tone mapping, then FXAA
        self.blooming = Blooming(self.res) if options.get('blooming', False) else None
        self.fxaa = FXAA(self.res) if options.get('fxaa', False) else None
        # temporal anti-aliasing: a jittered pixel centre per frame, averaged (raster.py:83-84, 173-174, 206-214)
        self.taa = bool(options.get('taa', False))
        self.accum = Accumator(self.res) if self.taa else None
        # SSAO and SSR read a normal buffer that a pre-shader fills per object, SSR
        # also a texcoord buffer when texturing (raster.py:42-69)
        self.ssao = None
        self.ssr = None
        if options.get('ssao', False) or options.get('ssr', False):
            self.norm_buffer = Field(torch.zeros((H, W, 3), dtype=torch.float32, device=self.engine.device))
            self.norm_shader = NormalShader(self.norm_buffer)
            self.pre_shaders.append(self.norm_shader)
        if options.get('ssr', False):
            self.coor_buffer = None
            if options.get('texturing', False):
                self.coor_buffer = Field(torch.zeros((H, W, 2), dtype=torch.float32, device=self.engine.device))
                self.pre_shaders.append(TexcoordShader(self.coor_buffer))
        if options.get('ssao', False):
            self.ssao = SSAO(self.res, self.norm_buffer, taa=self.taa)
        if options.get('ssr', False):
            self.ssr = SSR(self.res, self.norm_buffer, self.coor_buffer, taa=self.taa)
        if (self.blooming is not None or self.ssao is not None or self.ssr is not None) and self.neural is not None:
            raise ValueError('the learned shading writes tone-mapped colour: it cannot run before blooming, SSAO or SSR')

    def _ensure_material_shader(self, material):
        if material in self.materials:
            return
        shader = Shader(self.image, self.lighting, material)
        self.materials.append(material)
        # the pre-shaders (normal and texcoord buffers) run right after each
        # object's coverage pass instead (render), as the reference's per-object
        # loop leaves them
        self.shaders[material] = ShaderGroup([shader] + self.post_shaders)

    def add_object(self, object, material=None, raster=None):
        """Adds a mesh, a wire mesh or a particle set; the raster is picked by duck typing (raster.py:112-146)."""
        assert object not in self.objects
        if material is None:
            material = self.default_material
        if raster is None:
            if hasattr(object, 'get_nfaces'):
                if hasattr(object, 'get_npolygon') and object.get_npolygon() == 2:
                    if not hasattr(self, 'wireframe_raster'):
                        self.wireframe_raster = WireframeRaster(self.engine, **self.options)
                    raster = self.wireframe_raster
                else:
                    if not hasattr(self, 'triangle_raster'):
                        self.triangle_raster = TriangleRaster(self.engine, **self.options)
                    raster = self.triangle_raster
            elif hasattr(object, 'get_npars'):
                if not hasattr(self, 'particle_raster'):
                    self.particle_raster = ParticleRaster(self.engine, **self.options)
                raster = self.particle_raster
            elif hasattr(object, 'sample_volume'):
                raise NotImplementedError(f'{type(object).__name__}: the volume raster is not provided')
            else:
                raise ValueError(f'cannot determine raster type of object: {object}')
        self._ensure_material_shader(material)
        self.objects[object] = _Obj(material, raster)

    def set_camera(self, view, proj):
        self.engine.set_camera(view, proj)

    def input(self, gui):
        raise NotImplementedError('GUI camera control is not provided by taichi_three_amd; use scene.set_camera(view, proj)')

    def render(self, on_phase=None):
        """clear -> coverage of every object -> shade per object (+bg, +ACES).

        on_phase(name), if given, is called before the 'clear', 'raster' and
        'shade' stages and once more with 'end' (used to place timing events)."""
        eng = self.engine
        mark = on_phase or (lambda name: None)
        if self.taa:
            eng.randomize_bias(self.accum.count == 0)
        for s in self.pre_shaders + self.post_shaders:
            s.clear_buffer()
        mark('clear')
        eng.clear_depth()
        mark('raster')
        order = []
        for obj, info in self.objects.items():
            r = info.raster
            r.set_object(obj)
            info.id_base, info.nfaces, info.fb = r.id_base, r.nfaces, r._fb
            r.render_occup()
            # the reference runs the pre-shaders inside each object's render_color
            # (raster.py:183-187, 95-110): they write the pixels the object wins at
            # that point, and a wire drawn over them later (which only blends its
            # colour) leaves what they wrote
            for s in self.pre_shaders:
                s.post_shade(r, r.id_base, r.nfaces)
            order.append((obj, info))
        mark('shade')
        self._shade_all(order)
        if self.taa:
            self.accum.update(self.image)
        mark('end')

    def _shade_all(self, order):
        # tone mapping is fused into the shade kernels unless SSAO, SSR or blooming
        # has to see the linear image
        linear = self.blooming is not None or self.ssao is not None or self.ssr is not None
        tm = _abi.TONEMAP if self.tonemap and not linear else 0
        if linear and self.neural is not None:
            raise ValueError('the learned shading writes tone-mapped colour: it cannot run before blooming, SSAO or SSR')
        self._shade_objects(order, tm)
        if self.ssao is not None:               # raster.py:189-191
            self.ssao.render(self.engine)
            self.ssao.apply(self.image)
        if self.ssr is not None:                # raster.py:193-195
            self.ssr.begin()
            for obj, info in order:
                r = info.raster
                r._fb, r._mesh, r.nfaces, r.id_base = info.fb, obj, info.nfaces, info.id_base
                self.ssr.render_object(self.engine, self.image, r, info.material)
            self.ssr.apply(self.image)
        if (self.ssao is not None or self.ssr is not None) and isinstance(self.lighting, SkyboxLighting):
            self._fill_background(0)            # the sky is drawn after them, over empty pixels (:197-198)
        if self.blooming is not None:
            self.blooming.apply(self.image)
        if self.tonemap and linear:
            self.tonemap.apply(self.image)
        if self.fxaa is not None:
            self.fxaa.apply(self.image)

    def _shade_objects(self, order, tm):
        if not order:
            self._fill_background(tm)
            return
        first = True
        if isinstance(order[0][1].raster, WireframeRaster) and isinstance(self.lighting, SkyboxLighting):
            self._fill_background(tm)   # the paint pass has no skybox: the sky goes first
            first = False
        for obj, info in order:
            r = info.raster
            r._fb, r._mesh, r.nfaces, r.id_base = info.fb, obj, info.nfaces, info.id_base
            bg = _abi.BG if first else 0
            r.render_color(self.shaders[info.material], extra_flags=bg | tm, bgcolor=self.bgcolor,
                           neural=self.neural)
            first = False

    def _fill_background(self, tm):
        import ctypes
        if isinstance(self.lighting, SkyboxLighting):
            # no faces: a shade launch over an empty id range paints the sky
            r = TriangleRaster(self.engine, maxfaces=0)
            r._fb = None
            r.render_color(self.shaders.get(self.default_material) or
                           ShaderGroup([Shader(self.image, self.lighting, self.default_material)]),
                           extra_flags=_abi.BG | tm, bgcolor=self.bgcolor)
            return
        bg = (ctypes.c_float * 3)(*[float(x) for x in np.broadcast_to(np.asarray(self.bgcolor, dtype=np.float32), 3)])
        _abi.check(self.engine._lib.tina_fill(self.engine._ctx, _abi.ptr(self.image.storage), ctypes.byref(bg),
                                              tm, _abi.stream_handle()), self.engine._ctx)

    def load_gltf(self, path):
        """Adds the objects of a glTF file to the scene (raster.py:234-241); returns the scene."""
        from .gltf import readgltf
        return readgltf(path).extract(self)

    @property
    def img(self):
        """The final (tone-mapped) image, a Field of shape (W, H, 3); with taa the
        running mean of the frames since the last clear()."""
        return self.accum.img if self.taa else self.image

    def clear(self):
        """Restarts the TAA average (raster.py:230-232)."""
        if self.taa:
            self.accum.clear()

    def _fast_export_image(self, out):
        """Flat RGB export, out[(y * W + x) * 3 + c] (raster.py:243-250)."""
        out[:] = self.image.storage.detach().cpu().numpy().reshape(-1)
