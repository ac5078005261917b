"""tina.matr: material parameter nodes and materials (tina/matr/nodes.py,
tina/matr/material.py), lowered to the `tina_material` program the shade
kernel interprets.

The reference inlines the material tree into each Taichi kernel at JIT time.
Here `lower_material` walks the same tree once per (material, mesh) and emits
  - parameter slots: Const / Param / Input / Texture / FaceAttr / FresnelFactor
  - a post-order program: LAMBERT, PHONG, BLINNPHONG, COOKTORRANCE, EMISSION
    leaves push a value; SCALE, MIX, ADD combine
which the kernel evaluates in four modes (brdf, ambient, emission, sample_ibl)
with the reference's own arithmetic.
"""
import numpy as np

from . import _abi

__all__ = ['Node', 'Const', 'Param', 'Input', 'Texture', 'FaceAttr', 'FresnelFactor',
           'IMaterial', 'Lambert', 'Phong', 'BlinnPhong', 'CookTorrance', 'Emission',
           'MixMaterial', 'ScaleMaterial', 'AddMaterial',
           'Classic', 'Diffuse', 'Lamp', 'PBR', 'Blinn', 'lower_material']


class Node:
    """Parameter coercion of tina/matr/nodes.py:4-39: numbers and vectors become
    Const, image paths Texture, other strings Input."""
    arguments = []
    defaults = []

    def __init__(self, **kwargs):
        self.params = {}
        for dfl, key in zip(self.defaults, self.arguments):
            if key not in kwargs:
                if dfl is None:
                    raise ValueError(f'`{key}` must specified for `{type(self)}`')
                value = dfl
            else:
                value = kwargs.pop(key)
            self.params[key] = _coerce(value)
        for key in kwargs.keys():
            raise TypeError(f"{type(self).__name__}() got an unexpected keyword argument '{key}', "
                            f"supported keywords are: {self.arguments}")

    def param(self, key):
        return self.params[key]


def _coerce(value):
    if isinstance(value, Node):
        return value
    if isinstance(value, (bool, int, float, np.floating, np.integer)):
        return Const(float(value))
    if isinstance(value, (list, tuple, np.ndarray)):
        return Const(tuple(float(x) for x in np.asarray(value).reshape(-1)))
    if isinstance(value, str):
        if any(value.endswith(x) for x in ['.png', '.jpg', '.bmp']):
            return Texture(value)
        return Input(value)
    return value


class Const(Node):
    """Constant scalar or vector (nodes.py:42-49)."""

    def __init__(self, value):
        self.value = value

    @property
    def dim(self):
        return 1 if np.ndim(self.value) == 0 else len(self.value)


class _ScalarStore:
    """`param.value[None]` access to a device scalar or vector."""

    def __init__(self, tensor):
        self.tensor = tensor

    def __getitem__(self, idx):
        v = self.tensor.detach().cpu().numpy()
        return float(v[0]) if v.shape[0] == 1 else v

    def __setitem__(self, idx, value):
        import torch
        self.tensor.copy_(torch.as_tensor(np.asarray(value, dtype=np.float32).reshape(-1)))


class Param(Node):
    """Run-time parameter in device memory (nodes.py:52-76): changing it needs no
    re-lowering; the kernel reads it every frame."""

    def __init__(self, dtype=float, dim=None, initial=0):
        import torch
        self.dim = 1 if dim is None else int(dim)
        init = np.broadcast_to(np.asarray(initial, dtype=np.float32), (self.dim,)).copy()
        self.tensor = torch.as_tensor(init, device='cuda')
        self.value = _ScalarStore(self.tensor)
        self.initial = initial

    def make_slider(self, gui, title, min=0, max=1, step=0.01):
        raise NotImplementedError('GUI sliders are not provided by taichi_three_amd')


class Input(Node):
    """Per-pixel shading input: 'pos', 'color', 'normal' or 'texcoord' (nodes.py:79-96)."""
    _kinds = {'pos': _abi.P_POS, 'color': _abi.P_COLOR, 'normal': _abi.P_NORMAL, 'texcoord': _abi.P_TEXCOORD}

    def __init__(self, name):
        if name not in self._kinds:
            raise NotImplementedError(f"Input('{name}'): only {sorted(self._kinds)} are shading inputs")
        self.name = name


def _load_image(path):
    """ti.imread + texture_as_field (tina/advans.py:8-28): [x, y, c] with y up, u8 -> /255.
    PNG files are read without PIL when it is not installed."""
    try:
        from PIL import Image
        img = np.asarray(Image.open(path))
    except ImportError:
        from .pngread import read_png
        img = read_png(path)
    img = np.ascontiguousarray(np.transpose(img, (1, 0, 2) if img.ndim == 3 else (1, 0))[:, ::-1])
    if img.dtype == np.uint8:
        img = np.float32(img / 255)
    return img


class Texture(Node):
    """bilerp(texture, texcoord * (shape - 1)) (nodes.py:99-111).  `path` may be an
    image file or an array [w, h] / [w, h, c] indexed [x, y]."""
    arguments = ['texcoord']
    defaults = ['texcoord']

    def __init__(self, path, **kwargs):
        import torch
        if isinstance(path, str):
            arr = _load_image(path)
        else:
            arr = np.asarray(path)
            arr = np.float32(arr / 255) if arr.dtype == np.uint8 else arr.astype(np.float32)  # advans.py:13-14
        if arr.ndim == 2:
            arr = arr[:, :, None]
        self.shape = arr.shape
        self.tensor = torch.as_tensor(np.ascontiguousarray(arr, dtype=np.float32), device='cuda')
        super().__init__(**kwargs)


class FaceAttr(Node):
    """Per-face parameter array set with mesh.set_face_attr(name, values) (extension:
    the reference binds material parameters per object only)."""

    def __init__(self, name):
        self.name = name


class FresnelFactor(Node):
    """f0 = metallic * albedo + (1 - metallic) * 0.16 * specular**2 (material.py:62-83)."""
    arguments = ['metallic', 'albedo', 'specular']
    defaults = [0.0, 1.0, 0.5]


class IMaterial(Node):
    """Material protocol (material.py:5-60); combinators build Mix/Scale/Add trees."""

    def __add__(self, other):
        return AddMaterial(self, other)

    def mix(self, other, factor):
        return MixMaterial(self, other, factor)

    def __mul__(self, factor):
        return ScaleMaterial(self, factor)

    def __rmul__(self, factor):
        return ScaleMaterial(self, factor)


class Lambert(IMaterial):
    """brdf = 1/pi, ambient = 1 (material.py:387-442)."""
    arguments = []
    defaults = []


class Phong(IMaterial):
    """max(0, V.reflect(-L, N))**m * (m + 2) / 2 (material.py:445-472)."""
    arguments = ['shineness']
    defaults = [32.0]


class BlinnPhong(IMaterial):
    """max(0, N.H)**m * (m + 8) / 8 with H = normalize(L + V) (extension; the
    reference has only Phong).  The (m + 8) / 8 factor drops the 1/pi the same way
    Tina's Phong does."""
    arguments = ['shineness']
    defaults = [32.0]


class CookTorrance(IMaterial):
    """GGX with Tina's own normalisation (material.py:244-384)."""
    arguments = ['roughness', 'fresnel']
    defaults = [0.4, 1.0]
    rough_levels = [0.03, 0.08, 0.18, 0.35, 0.65, 1.0]
    rough_levels = [(a, a - b) for a, b in zip(rough_levels, [0] + rough_levels)]


class Emission(IMaterial):
    """Emission 1, brdf 0, ambient 0 (material.py:636-657)."""
    arguments = []
    defaults = []


class MixMaterial(IMaterial):
    """(1 - factor) * mat1 + factor * mat2 (material.py:86-145)."""
    arguments = ['factor']
    defaults = [0.5]

    def __init__(self, mat1, mat2, factor):
        super().__init__(factor=factor)
        self.mat1, self.mat2 = mat1, mat2


class ScaleMaterial(IMaterial):
    """factor * mat (material.py:148-191)."""
    arguments = ['factor']
    defaults = [1.0]

    def __init__(self, mat, factor):
        super().__init__(factor=factor)
        self.mat = mat


class AddMaterial(IMaterial):
    """mat1 + mat2 (material.py:194-238)."""
    arguments = []
    defaults = []

    def __init__(self, mat1, mat2):
        super().__init__()
        self.mat1, self.mat2 = mat1, mat2


def Classic(color='color', shineness=32, specular=0.4):
    """Mix(Lambert * color, Phong(shineness), specular) (material.py:684-688)."""
    return MixMaterial(Lambert() * color, Phong(shineness=shineness), specular)


def Blinn(color='color', shineness=32, specular=0.4):
    """Classic with the Blinn-Phong lobe (extension)."""
    return MixMaterial(Lambert() * color, BlinnPhong(shineness=shineness), specular)


def Diffuse(color='color'):
    """Lambert * color (material.py:691-693)."""
    return Lambert() * color


def Lamp(color='color'):
    return Emission() * color


def PBR(basecolor='color', metallic=0.0, roughness=0.4, specular=0.5):
    """Mix(Lambert * basecolor, CookTorrance(roughness, f0), f0) with
    f0 = FresnelFactor(metallic, basecolor, specular) (material.py:701-706)."""
    mat_diff = Lambert() * basecolor
    f0 = FresnelFactor(metallic=metallic, albedo=basecolor, specular=specular)
    mat_spec = CookTorrance(roughness=roughness, fresnel=f0)
    return MixMaterial(mat_diff, mat_spec, f0)


# ---------------------------------------------------------------- lowering

class Lowered:
    """A lowered material: the ctypes struct plus the tensors it points into."""

    def __init__(self, struct, keepalive, uses_color, uses_tex, uses_ibl_only_ok):
        self.struct = struct
        self.keepalive = keepalive
        self.uses_color = uses_color
        self.uses_tex = uses_tex
        self.ibl_ok = uses_ibl_only_ok

    @property
    def is_pbr(self):
        """True for the PBR preset's program [LAMBERT, SCALE, COOKTORRANCE, MIX]
        whose Cook-Torrance f0 is its FresnelFactor mix factor -- the material the
        learned surrogate replaces."""
        s = self.struct
        ops = [s.ops[i].op for i in range(s.nops)]
        if ops != [0, 5, 3, 6]:  # TINA_OP_LAMBERT, SCALE, COOKTORRANCE, MIX
            return False
        f0 = s.ops[2].b
        return f0 == s.ops[3].a and s.params[f0].kind == 8  # TINA_P_FRESNEL


def lower_material(material, mesh_buffers=None):
    """Walks a material tree into a tina_material program.

    mesh_buffers: the dict of the object's face_buffers() (for FaceAttr).
    Raises NotImplementedError for nodes with no device form."""
    slots = []          # list of _abi.Param
    slot_of = {}        # id(node) -> index
    ops = []
    keep = []
    flags = {'color': False, 'tex': False, 'ibl_ok': True}

    def dim_of(i):
        return slots[i].dim

    def slot(node):
        key = id(node)
        if key in slot_of:
            return slot_of[key]
        p = _abi.Param()
        if isinstance(node, Const):
            v = node.value
            if np.ndim(v) == 0:
                p.kind, p.dim = _abi.P_CONST, 1
                p.value[0] = float(v)
            else:
                vv = list(v) + [0.0] * (3 - len(v))
                if len(v) not in (2, 3):
                    raise NotImplementedError('Const vectors must have 2 or 3 components')
                p.kind, p.dim = _abi.P_CONST, 3
                p.value[0], p.value[1], p.value[2] = (float(x) for x in vv[:3])
        elif isinstance(node, Param):
            p.kind, p.dim, p.ptr = _abi.P_PARAM, (1 if node.dim == 1 else 3), _abi.ptr(node.tensor)
            if node.dim not in (1, 3):
                raise NotImplementedError('Param dim must be 1 or 3')
            keep.append(node.tensor)
        elif isinstance(node, Input):
            p.kind = Input._kinds[node.name]
            p.dim = 3
            if node.name == 'color':
                flags['color'] = True
            if node.name == 'texcoord':
                flags['tex'] = True
        elif isinstance(node, Texture):
            a = slot(node.params['texcoord'])
            p = _abi.Param()
            p.kind, p.a = _abi.P_TEXTURE, a
            p.tex_w, p.tex_h, p.tex_c = node.shape[0], node.shape[1], node.shape[2]
            p.dim = 1 if node.shape[2] == 1 else 3
            p.ptr = _abi.ptr(node.tensor)
            keep.append(node.tensor)
        elif isinstance(node, FaceAttr):
            arr = None if mesh_buffers is None else mesh_buffers.get('attrs', {}).get(node.name)
            if arr is None:
                raise ValueError(f"FaceAttr('{node.name}'): the mesh has no such face attribute "
                                 "(mesh.set_face_attr)")
            p.kind, p.ptr = _abi.P_FACE, _abi.ptr(arr)
            p.dim = 1 if arr.ndim == 1 or arr.shape[1] == 1 else 3
            keep.append(arr)
        elif isinstance(node, FresnelFactor):
            a = slot(node.params['metallic'])
            b = slot(node.params['albedo'])
            c = slot(node.params['specular'])
            p = _abi.Param()
            p.kind, p.a, p.b, p.c = _abi.P_FRESNEL, a, b, c
            p.dim = 1 if (dim_of(a) == 1 and dim_of(b) == 1 and dim_of(c) == 1) else 3
        else:
            raise NotImplementedError(f'{type(node).__name__} has no device form')
        if len(slots) >= _abi.MAX_PARAMS:
            raise NotImplementedError(f'material uses more than {_abi.MAX_PARAMS} parameter nodes')
        slots.append(p)
        slot_of[key] = len(slots) - 1
        return slot_of[key]

    def scalar_slot(node, what):
        i = slot(node)
        if dim_of(i) != 1:
            raise NotImplementedError(f'{what} must be a scalar parameter here')
        return i

    def emit(m):
        if isinstance(m, Lambert):
            ops.append((_abi.OP_LAMBERT, 0, 0))
        elif isinstance(m, Phong):
            ops.append((_abi.OP_PHONG, scalar_slot(m.params['shineness'], 'Phong shineness'), 0))
            flags['ibl_ok'] = False
        elif isinstance(m, BlinnPhong):
            ops.append((_abi.OP_BLINNPHONG, scalar_slot(m.params['shineness'], 'BlinnPhong shineness'), 0))
            flags['ibl_ok'] = False
        elif isinstance(m, CookTorrance):
            r = scalar_slot(m.params['roughness'], 'CookTorrance roughness')
            f = slot(m.params['fresnel'])
            ops.append((_abi.OP_COOKTORRANCE, r, f))
        elif isinstance(m, Emission):
            ops.append((_abi.OP_EMISSION, 0, 0))
            flags['ibl_ok'] = False
        elif isinstance(m, ScaleMaterial):
            emit(m.mat)
            ops.append((_abi.OP_SCALE, slot(m.params['factor']), 0))
        elif isinstance(m, MixMaterial):
            emit(m.mat1)
            emit(m.mat2)
            ops.append((_abi.OP_MIX, slot(m.params['factor']), 0))
        elif isinstance(m, AddMaterial):
            emit(m.mat1)
            emit(m.mat2)
            ops.append((_abi.OP_ADD, 0, 0))
            flags['ibl_ok'] = False
        else:
            raise NotImplementedError(f'material {type(m).__name__} has no device form')

    emit(material)
    if len(ops) > _abi.MAX_OPS:
        raise NotImplementedError(f'material program longer than {_abi.MAX_OPS} ops')
    depth = 0
    for op, _, _ in ops:
        depth += 1 if op <= _abi.OP_EMISSION else (-1 if op in (_abi.OP_MIX, _abi.OP_ADD) else 0)
        if depth > _abi.MAX_STACK:
            raise NotImplementedError(f'material tree deeper than {_abi.MAX_STACK} pending values')
    s = _abi.Material()
    s.nparams = len(slots)
    s.nops = len(ops)
    for i, p in enumerate(slots):
        s.params[i] = p
    for i, (op, a, b) in enumerate(ops):
        s.ops[i].op, s.ops[i].a, s.ops[i].b = op, a, b
    return Lowered(s, keep, flags['color'], flags['tex'], flags['ibl_ok'])
