"""Particle sets: the input side of ParticleRaster.set_object (tina/pars/*.py).

A particle set keeps its centres [n, 3], radii [n] and colours [n, 3] resident
in HBM as f32 tensors; `particle_buffers()` hands the raster those tensors
directly (the reference copies them into the raster's own fields each frame,
tina/core/particle.py:64-75).

  SimpleParticles   tina/pars/simple.py:4-55
  ParsEditBase      tina/pars/base.py:4-10
  ParsTransform     tina/pars/trans.py:5-30 (centres through a 4x4 matrix, radii scaled)
  export_simple_pars  tina/pars/export.py:30-47
"""
import numpy as np

from .field import Field

__all__ = ['SimpleParticles', 'ParsEditBase', 'ParsTransform', 'export_simple_pars']


def _torch():
    import torch
    return torch


class SimpleParticles:
    """Up to `maxpars` particles of default radius `radius` and colour 1 (simple.py:5-17)."""

    def __init__(self, maxpars=65536, radius=0.02, device=None):
        torch = _torch()
        dev = device or torch.device('cuda')
        self.maxpars = int(maxpars)
        self.verts = Field(torch.zeros((self.maxpars, 3), dtype=torch.float32, device=dev), screen=False)
        self.sizes = Field(torch.full((self.maxpars,), float(radius), dtype=torch.float32, device=dev), screen=False)
        self.colors = Field(torch.ones((self.maxpars, 3), dtype=torch.float32, device=dev), screen=False)
        self.npars = 0
        self.version = 0

    def pre_compute(self):
        pass

    def get_npars(self):
        return min(self.npars, self.maxpars)

    def _rows(self, arr, width):
        torch = _torch()
        dev = self.verts.storage.device
        if isinstance(arr, torch.Tensor):
            t = arr.to(device=dev, dtype=torch.float32)
        else:
            t = torch.as_tensor(np.ascontiguousarray(np.asarray(arr, dtype=np.float32)), device=dev)
        if width and (t.dim() != 2 or t.shape[1] != width):
            raise ValueError(f'expected an array of shape [n, {width}], got {tuple(t.shape)}')
        if not width and t.dim() != 1:
            raise ValueError(f'expected an array of shape [n], got {tuple(t.shape)}')
        return t

    def set_particles(self, verts):
        """Centres [n, 3]; the particle count becomes min(n, maxpars) (simple.py:39-44)."""
        t = self._rows(verts, 3)
        n = min(int(t.shape[0]), self.maxpars)
        self.verts.storage[:n].copy_(t[:n])
        self.npars = n
        self.version += 1

    def set_particle_radii(self, sizes):
        """Radii of the first npars particles (simple.py:46-49)."""
        t = self._rows(sizes, 0)
        n = self.get_npars()
        if t.shape[0] < n:
            raise ValueError(f'{t.shape[0]} radii for {n} particles')
        self.sizes.storage[:n].copy_(t[:n])
        self.version += 1

    def set_particle_colors(self, colors):
        """Colours of the first npars particles (simple.py:51-55)."""
        t = self._rows(colors, 3)
        n = self.get_npars()
        if t.shape[0] < n:
            raise ValueError(f'{t.shape[0]} colours for {n} particles')
        self.colors.storage[:n].copy_(t[:n])
        self.version += 1

    def get_particle_position(self, n):
        return self.verts.storage[n].cpu().numpy()

    def get_particle_radius(self, n):
        return float(self.sizes.storage[n])

    def get_particle_color(self, n):
        return self.colors.storage[n].cpu().numpy()

    def particle_buffers(self):
        n = self.get_npars()
        return dict(pos=self.verts.storage[:n], radius=self.sizes.storage[:n], colors=self.colors.storage[:n],
                    npars=n)


class ParsEditBase:
    """Wraps a particle set and forwards everything it does not override (base.py:4-10)."""

    def __init__(self, pars):
        self.pars = pars

    def __getattr__(self, attr):
        if attr == 'pars':
            raise AttributeError(attr)
        return getattr(self.pars, attr)

    def particle_buffers(self):
        return self.pars.particle_buffers()


class ParsTransform(ParsEditBase):
    """Centres mapped by mapply_pos(trans, p), radii multiplied by scale (trans.py:5-30).

    The product is evaluated element by element in binary32 in the order
    mapply writes it (tina/common.py:169-183): r_i = M[i][3] * 1 + M[i][0] p_0
    + M[i][1] p_1 + M[i][2] p_2, each step rounded, then r / w."""

    def __init__(self, pars):
        super().__init__(pars)
        self.trans = np.eye(4, dtype=np.float32)
        self.scale = np.float32(1)

    def set_transform(self, trans, scale):
        self.trans = np.asarray(trans, dtype=np.float32).reshape(4, 4)
        self.scale = np.float32(scale)

    def particle_buffers(self):
        fb = dict(self.pars.particle_buffers())
        p = fb['pos']
        M = [[float(x) for x in row] for row in self.trans]
        rows = []
        for i in range(4):
            acc = _torch().full_like(p[:, 0], M[i][3])
            for j in range(3):
                acc = acc + M[i][j] * p[:, j]
            rows.append(acc)
        fb['pos'] = _torch().stack([rows[0] / rows[3], rows[1] / rows[3], rows[2] / rows[3]], dim=1).contiguous()
        fb['radius'] = (float(self.scale) * fb['radius']).contiguous()
        return fb

    def get_particle_position(self, n):
        return self.particle_buffers()['pos'][n].cpu().numpy()

    def get_particle_radius(self, n):
        return float(self.particle_buffers()['radius'][n])


def export_simple_pars(pars):
    """{'v': centres [n, 3], 'vr': radii [n], 'vc': colours [n, 3]} as float32 arrays (export.py:30-47)."""
    if hasattr(pars, 'pre_compute'):
        pars.pre_compute()
    fb = pars.particle_buffers()
    out = {'v': fb['pos'].detach().cpu().numpy().astype(np.float32)}
    if fb.get('radius') is not None:
        out['vr'] = fb['radius'].detach().cpu().numpy().astype(np.float32)
    if fb.get('colors') is not None:
        out['vc'] = fb['colors'].detach().cpu().numpy().astype(np.float32)
    return out
