"""Lighting (tina/core/lighting.py:24-98) as a host-side table passed by value to
the shade kernel, and SkyboxLighting (lighting.py:4-21) backed by device IBL maps."""
import numpy as np

from . import _abi

__all__ = ['Lighting', 'SkyboxLighting']


class Lighting:
    """Up to 16 directional (w = 0) or point (w = 1) lights plus an ambient colour."""

    def __init__(self, maxlights=16):
        if maxlights > _abi.MAX_LIGHTS:
            raise ValueError(f'at most {_abi.MAX_LIGHTS} lights')
        self.maxlights = maxlights
        self.light_dirs = np.zeros((maxlights, 4), dtype=np.float32)
        self.light_dirs[:] = [0, 0, 1, 0]
        self.light_colors = np.ones((maxlights, 3), dtype=np.float32)
        self.ambient_color = np.zeros(3, dtype=np.float32)
        self.nlights = 0

    def set_lights(self, light_dirs):
        self.nlights = len(light_dirs)
        for i, (dir, color) in enumerate(light_dirs):
            self.light_dirs[i] = dir
            self.light_colors[i] = color

    def clear_lights(self):
        self.nlights = 0

    def add_light(self, dir=(0, 0, 1), pos=None, color=(1, 1, 1)):
        i = self.nlights
        if i >= self.maxlights:
            raise ValueError(f'at most {self.maxlights} lights')
        self.nlights = i + 1
        self.set_light(i, dir, pos, color)
        return i

    def set_light(self, i, dir=(0, 0, 1), pos=None, color=(1, 1, 1)):
        """Directions are normalised in FP64 before the f32 store (lighting.py:55-65)."""
        if pos is not None:
            d = np.array(pos, dtype=np.float64)
            dirw = 1
        else:
            d = np.array(dir, dtype=np.float64)
            d = d / np.linalg.norm(d)
            dirw = 0
        self.light_dirs[i] = np.array(d.tolist() + [dirw], dtype=np.float32)
        self.light_colors[i] = np.array(color, dtype=np.float32)

    def set_ambient_light(self, color):
        self.ambient_color[:] = np.array(color, dtype=np.float32)

    def to_struct(self):
        s = _abi.Lights()
        s.n = int(self.nlights)
        for i in range(self.nlights):
            for k in range(4):
                s.dirs[i][k] = float(self.light_dirs[i, k])
            for k in range(3):
                s.colors[i][k] = float(self.light_colors[i, k])
        for k in range(3):
            s.ambient[k] = float(self.ambient_color[k])
        return s


class SkyboxLighting(Lighting):
    """material.sample_ibl(ibltab, viewdir, normal) + emission (lighting.py:20-21).

    Direct lights added to it (add_light) are summed on top; the reference has
    no such combination, so with no lights this is exactly SkyboxLighting."""

    def __init__(self, skybox, ibltab=None, precision=32):
        super().__init__()
        from .skybox import IBLTables
        self.skybox = skybox
        self.ibltab = ibltab if ibltab is not None else IBLTables(skybox, precision=precision)

    def background(self):
        return self.skybox
