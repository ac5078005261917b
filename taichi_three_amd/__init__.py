"""taichi_three_amd -- Tina's rasterise-and-shade path on MI355X.

Drop-in for the tina.Engine / tina.Scene / tina.TriangleRaster / tina.matr API
of taichi_three (`import taichi_three_amd as tina`).  Every render call goes
through hand-written HIP kernels for gfx950 in libtina_hip.so (C ABI in
include/tina_hip.h); PyTorch-ROCm only holds the device buffers.
"""
__version__ = (0, 1, 1)

from .engine import Engine, MAX
from .field import Field
from .util.matrix import *
from .mesh import *
from .assimp import *
from .gltf import readgltf
from .matr import *
from .lighting import *
from .shader import *
from .raster import TriangleRaster
from .pars import *
from .particle import ParticleRaster
from .wireframe import WireframeRaster
from .postp import ToneMapping, Blooming, FXAA, Accumator, SSAO
from .probe import ProbeShader
from .scene import Scene
from . import _abi


def __getattr__(name):
    # IBL tables and the extension modules import lazily (they pull in more device code)
    if name in ('Skybox', 'IBLTables'):
        from . import skybox
        return getattr(skybox, name)
    raise AttributeError(name)
