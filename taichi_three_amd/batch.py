"""Batched multi-view rendering (extension: one tina.Scene renders one view,
tina/scene/raster.py:168-214).

`BatchScenes` uploads many independent scenes -- each with its own screen,
camera, faces, per-face/per-vertex material inputs and lights, all using one
material program.  `BatchRenderer` renders all of them with five persistent
kernels (clear, bin, giant faces, tiles, shade) that walk the work items of
every scene; `capture()` records those five launches once into a HIP graph and
`replay()` then renders the whole batch with a single graph launch.  The bin
sizes and launch shapes come from a `Tuning` (see taichi_three_amd.autotune).
"""
import ctypes

import numpy as np

from . import _abi
from .matr import lower_material

__all__ = ['BatchScenes', 'BatchRenderer', 'Tuning', 'DEFAULT_TUNING']


class Tuning(tuple):
    """(bin_faces, max_tiles, grid_bin, grid_tile, grid_shade): faces per binning
    workgroup, bbox tiles above which a face is rasterised as a giant, and the
    persistent workgroup counts of the bin, tile and shade passes (0: one
    workgroup per work item)."""
    __slots__ = ()
    FIELDS = ('bin_faces', 'max_tiles', 'grid_bin', 'grid_tile', 'grid_shade')

    def __new__(cls, bin_faces=4096, max_tiles=64, grid_bin=0, grid_tile=0, grid_shade=0):
        return super().__new__(cls, (int(bin_faces), int(max_tiles), int(grid_bin), int(grid_tile), int(grid_shade)))

    def struct(self):
        t = _abi.BatchTuning()
        t.bin_faces, t.max_tiles, t.grid_bin, t.grid_tile, t.grid_shade = self
        return t

    def as_dict(self):
        return dict(zip(self.FIELDS, self))


DEFAULT_TUNING = Tuning()


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('taichi_three_amd needs a ROCm GPU: torch.cuda.is_available() is False')
    return torch


class BatchScenes:
    """Device-resident inputs of a batch: workloads as made by
    taichi_three_amd.workloads (verts, colors, attrs, W2V/V2W, res, lights,
    ambient, material), one output image [H][W][3] per scene."""

    def __init__(self, workloads, tonemap=True, culling=True, clipping=True):
        from .lighting import Lighting
        from .workloads import make_material
        torch = _torch()
        self.dev = torch.device('cuda', torch.cuda.current_device())
        self.n = len(workloads)
        if self.n < 1 or self.n > 256:
            raise ValueError('a batch holds 1 to 256 scenes')
        self._keep = []
        self.descs = (_abi.SceneDesc * self.n)()
        self.images, self.res, self.nfaces = [], [], []
        for i, w in enumerate(workloads):
            W, H = int(w['res'][0]), int(w['res'][1])
            T = len(w['verts'])
            if T == 0:  # one dummy face keeps every pointer valid; nfaces stays 0
                w = dict(w, verts=np.zeros((1, 3, 3), np.float32), colors=np.zeros((1, 3, 3), np.float32),
                         norms=None, attrs={k: np.zeros((1,) + np.shape(v)[1:], np.float32)
                                            for k, v in w.get('attrs', {}).items()})
            verts = torch.from_numpy(np.ascontiguousarray(w['verts'], np.float32)).to(self.dev)
            colors = torch.from_numpy(np.ascontiguousarray(w['colors'], np.float32)).to(self.dev)
            norms = None
            if w.get('norms') is not None and w.get('smoothing'):
                norms = torch.from_numpy(np.ascontiguousarray(w['norms'], np.float32)).to(self.dev)
            attrs = {k: torch.from_numpy(np.ascontiguousarray(v, np.float32)).to(self.dev)
                     for k, v in w.get('attrs', {}).items()}
            img = torch.zeros((H, W, 3), dtype=torch.float32, device=self.dev)
            lo = lower_material(make_material(w['material']), dict(attrs=attrs, colors=colors))
            lt = Lighting()
            for l in w['lights']:
                lt.add_light(**l)
            lt.set_ambient_light(w['ambient'])
            d = self.descs[i]
            d.width, d.height, d.nfaces = W, H, T
            d.raster_flags = (1 if culling else 0) | (2 if clipping else 0)
            d.shade_flags = _abi.VCOLOR | _abi.BG | (_abi.TONEMAP if tonemap else 0) | \
                (_abi.SMOOTH if norms is not None else 0)
            d.img = _abi.ptr(img)
            d.cam.W2V[:] = [float(x) for x in np.asarray(w['W2V'], np.float32).reshape(-1)]
            d.cam.V2W[:] = [float(x) for x in np.asarray(w['V2W'], np.float32).reshape(-1)]
            d.cam.bias[:] = [0.5, 0.5]
            d.mesh.verts, d.mesh.colors = _abi.ptr(verts), _abi.ptr(colors)
            d.mesh.norms = _abi.ptr(norms) if norms is not None else 0
            d.mesh.nfaces, d.mesh.id_base = T, 0
            d.mat = lo.struct
            d.lights = lt.to_struct()
            self._keep += [verts, colors, norms, attrs, lo]
            self.images.append(img)
            self.res.append((W, H))
            self.nfaces.append(T)

    @property
    def total_faces(self):
        return int(sum(self.nfaces))

    @property
    def total_pixels(self):
        return int(sum(W * H for W, H in self.res))


class BatchRenderer:
    """A render plan for `scenes` (a BatchScenes, or workloads to upload) under
    one Tuning."""

    def __init__(self, scenes, tuning=DEFAULT_TUNING):
        self.scenes = scenes if isinstance(scenes, BatchScenes) else BatchScenes(scenes)
        self.tuning = Tuning(*tuning)
        self._lib = _abi.load()
        handle = ctypes.c_void_p()
        _abi.check(self._lib.tina_batch_create(self.scenes.n, self.scenes.descs,
                                               ctypes.byref(self.tuning.struct()), ctypes.byref(handle)))
        self._h = handle
        self.captured = False

    def close(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            _torch().cuda.synchronize()
            self._lib.tina_batch_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self):
        """Renders every scene (five kernel launches on the current stream)."""
        _abi.check(self._lib.tina_batch_render(self._h, _abi.stream_handle()))

    def capture(self):
        """Records render() into a HIP graph (once), on a side stream."""
        torch = _torch()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            _abi.check(self._lib.tina_batch_capture(self._h, _abi.stream_handle()))
        torch.cuda.current_stream().wait_stream(s)
        self.captured = True

    def replay(self):
        """One graph launch renders the whole batch."""
        if not self.captured:
            self.capture()
        _abi.check(self._lib.tina_batch_launch(self._h, _abi.stream_handle()))

    def image(self, i):
        return self.scenes.images[i]

    def resolve(self, i):
        """(ids, depth) of scene i as [H][W] int32 device tensors."""
        torch = _torch()
        W, H = self.scenes.res[i]
        p = ctypes.c_void_p()
        _abi.check(self._lib.tina_batch_keys(self._h, int(i), ctypes.byref(p)))
        ids = torch.empty((H, W), dtype=torch.int32, device=self.scenes.dev)
        dep = torch.empty((H, W), dtype=torch.int32, device=self.scenes.dev)
        _abi.check(self._lib.tina_resolve_keys(p, W * H, _abi.ptr(ids), _abi.ptr(dep), _abi.stream_handle()))
        return ids, dep
