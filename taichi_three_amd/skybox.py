"""Environment maps and image-based-lighting tables on the device.

`Skybox` (tina/skybox.py:39-145) holds an equirect HDR map as a [w][h][3] f32
device tensor, x-major like the reference's `Skybox.img[x, y]`; texel (x, y)
looks along `unspheremap((x, y) / (shape - 1))` (tina/skybox.py:31-35, 131-136),
so y runs from straight down (y = 0) to straight up (y = h - 1).

`IBLTables` bakes the prefiltered maps of `Lambert.cook_for_ibl`
(tina/matr/material.py:407-438) and `CookTorrance.cook_for_ibl`
(material.py:279-320) with the `tina_bake_ibl` kernel, at the reference's
sizes and sample counts, and fills the `tina_ibl` descriptor the shade kernel
reads.  The bakes use Hammersley points where the reference draws `ti.random`,
and skip the reference's KNN denoise of the Lambert map (Hammersley sums carry
no per-texel noise to remove).
"""
import ctypes
import math

import numpy as np

from . import _abi

__all__ = ['Skybox', 'IBLTables', 'procedural_sky', 'make_ibl_struct', 'ROUGH_LEVELS', 'LUT_CORNER']

# CookTorrance.rough_levels (material.py:251-253)
ROUGH_LEVELS = [0.03, 0.08, 0.18, 0.35, 0.65, 1.0]

# Texels [0..2] x [0..2] of the split-sum LUT assets/lut.jpg (720 x 718 RGB8),
# indexed [x = VoH][y = roughness][A, B] with the bottom-left origin of
# ti.imread, as bytes.  CookTorrance.sample_ibl looks the LUT up with
# bilerp(lut, (VoH, roughness)) at texel coordinates in [0, 1]^2
# (material.py:275, tina/common.py:140-149), so these are the only texels the
# reference ever reads; A = R / 255, B = G / 255.
LUT_CORNER_BYTES = (
    ((6, 249), (6, 248), (7, 246)),
    ((6, 249), (6, 249), (7, 247)),
    ((6, 249), (6, 249), (7, 247)),
)
LUT_CORNER = np.array(LUT_CORNER_BYTES, dtype=np.float64) / 255.0


def rough_levels():
    """(rough, rough - previous) per level, in f64 as material.py:253 computes them."""
    return [(a, a - b) for a, b in zip(ROUGH_LEVELS, [0] + ROUGH_LEVELS)]


def unmapcoor(w, h):
    """Unit direction of every texel of a w x h equirect map, shape (w, h, 3), f64
    (Skybox.unmapcoor + unspheremap + spherical, tina/skybox.py:31-35, 131-136;
    tina/advans.py:104-108)."""
    v = np.arange(w, dtype=np.float64)[:, None] / (w - 1)
    u = np.arange(h, dtype=np.float64)[None, :] / (h - 1)
    hh = u * 2.0 - 1.0
    r = np.sqrt(np.maximum(0.0, 1.0 - hh * hh))
    tau = 2.0 * math.pi
    dx, dy = r * np.cos(v * tau), r * np.sin(v * tau)
    return np.stack(np.broadcast_arrays(dx, hh, -dy), axis=-1)


# The three suns of the procedural sky: direction, peak radiance, angular sigma (rad)
SUNS = (
    ((0.45, 0.55, -0.70), (40.0, 36.0, 30.0), 0.05),
    ((-0.75, 0.30, 0.45), (8.0, 9.0, 12.0), 0.12),
    ((0.10, 0.90, 0.60), (3.0, 2.5, 2.0), 0.25),
)


def procedural_sky(w=512, h=256):
    """Procedural HDR sky, shape (w, h, 3) f32: a zenith-to-horizon gradient over
    a darker ground, plus three Gaussian suns, L += I exp(-theta^2 / (2 sigma^2))
    with theta the angle to the sun.  A deterministic stand-in for the
    reference's `Atmosphere` scattering sky (tina/skybox.py:149-302)."""
    d = unmapcoor(w, h)
    up = d[..., 1:2]
    zenith = np.array([0.15, 0.35, 0.85])
    horizon = np.array([0.85, 0.90, 1.00])
    ground = np.array([0.30, 0.27, 0.24])
    t = np.sqrt(np.clip(up, 0.0, 1.0))
    sky = horizon + (zenith - horizon) * t
    gnd = ground * (1.0 - 0.5 * np.clip(-up, 0.0, 1.0))
    img = np.where(up >= 0.0, sky, gnd)
    for sdir, power, sigma in SUNS:
        s = np.asarray(sdir, dtype=np.float64)
        s = s / np.linalg.norm(s)
        theta = np.arccos(np.clip(d @ s, -1.0, 1.0))
        img = img + np.asarray(power) * np.exp(-(theta * theta) / (2.0 * sigma * sigma))[..., None]
    return img.astype(np.float32)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('taichi_three_amd needs a ROCm GPU: torch.cuda.is_available() is False')
    return torch


class Skybox:
    """Skybox(path, scale=None, cubic=False) (tina/skybox.py:39-123).

    `path` may be an int resolution (a black (2r, r) map), a (w, h) shape, an
    HDR array or tensor of shape (w, h, 3), or an image file: LDR images are
    read as ti.imread does (x-major, bottom-left origin) and brought to HDR
    with ce_untonemap (tina/advans.py:64-69).  Cube maps are not provided."""

    def __init__(self, path, scale=None, cubic=False):
        if cubic:
            raise NotImplementedError('cube-map skyboxes (tina/skybox.py:6-20) are not provided; '
                                      'use an equirect map')
        torch = _torch()
        dev = torch.device('cuda', torch.cuda.current_device())
        self.cubic = False
        data = None
        if isinstance(path, (int, np.integer)):
            shape = (2 * int(path), int(path))
        elif isinstance(path, (tuple, list)):
            shape = (int(path[0]), int(path[1]))
        elif isinstance(path, str):
            data = self._read_image(path)
            shape = data.shape[:2]
        else:
            data = path
            shape = tuple(int(s) for s in data.shape[:2])
            if len(data.shape) != 3 or data.shape[2] != 3:
                raise ValueError(f'a skybox map has shape (w, h, 3), got {tuple(data.shape)}')
        if shape[0] < 2 or shape[1] < 2:
            raise ValueError(f'a skybox map needs at least 2 x 2 texels, got {shape}')
        if data is None:
            self.img = torch.zeros((shape[0], shape[1], 3), dtype=torch.float32, device=dev)
        else:
            self.img = torch.as_tensor(np.asarray(data, dtype=np.float32) if not hasattr(data, 'data_ptr')
                                       else data, dtype=torch.float32).to(dev).contiguous()
        if scale is not None:
            self.img.mul_(float(scale))
        self.shape = tuple(shape)
        self.resolution = self.shape[1]

    @staticmethod
    def _read_image(path):
        from PIL import Image
        arr = np.asarray(Image.open(path).convert('RGB'))
        img = np.float32(arr / 255)
        img = img.transpose(1, 0, 2)[:, ::-1]  # rows top-down -> [x][y], y up
        return -np.log(1 - np.clip(img * 0.996, 0, 1))

    @classmethod
    def procedural(cls, w=512, h=256):
        """The 512 x 256 HDR sky of `procedural_sky` as a Skybox."""
        return cls(procedural_sky(w, h))

    def to_numpy(self):
        return self.img.detach().cpu().numpy()


def bake_level(env, ew, eh, out, ow, oh, roughness, nsamples):
    """Enqueues one tina_bake_ibl on the current stream (roughness < 0: Lambert)."""
    lib = _abi.load()
    _abi.check(lib.tina_bake_ibl(_abi.ptr(env), int(ew), int(eh), _abi.ptr(out), int(ow), int(oh),
                                 ctypes.c_float(roughness), int(nsamples), _abi.stream_handle()))


def make_ibl_struct(diff, spec):
    """tina_ibl descriptor over the Lambert map `diff` and the levels `spec`
    (env first), each (w, h, 3) f32 -- device tensors for the shade kernel, or
    NumPy arrays for host-side consumers -- with the LUT corner and the
    roughness brackets of CookTorrance.sample_ibl (material.py:264-275)."""
    s = _abi.IBLTables()
    s.diff = _abi.ptr(diff)
    s.diff_w, s.diff_h = int(diff.shape[0]), int(diff.shape[1])
    if len(spec) != _abi.IBL_LEVELS:
        raise ValueError(f'expected {_abi.IBL_LEVELS} specular maps (env + 6 levels), got {len(spec)}')
    for i, m in enumerate(spec):
        s.spec[i] = _abi.ptr(m)
        s.spec_w[i], s.spec_h[i] = int(m.shape[0]), int(m.shape[1])
    for x in range(3):
        for y in range(3):
            for k in range(2):
                s.lut[x][y][k] = float(LUT_CORNER[x, y, k])
    for i, (rough, step) in enumerate(rough_levels()):
        s.rough_lo[i] = float(rough - step)
        s.rough_hi[i] = float(rough)
        s.rough_step[i] = float(step)
    return s


class IBLTables:
    """The `ibltab` of SkyboxLighting (tina/core/lighting.py:6-13): 'env', the
    Lambert map 'diff' (env.resolution // 6, 128 * precision samples) and the six
    CookTorrance levels 'spec' (resolution shrinking by 2**(4 step) and samples
    growing by 3**(4 step) from 8 * precision, material.py:301-316).  `struct`
    is the tina_ibl descriptor; the tensors stay owned here."""

    def __init__(self, skybox, precision=32):
        torch = _torch()
        self.skybox = skybox
        self.precision = int(precision)
        dev = skybox.img.device
        self.levels = rough_levels()

        res = skybox.resolution // 6
        self.diff = torch.empty((2 * res, res, 3), dtype=torch.float32, device=dev)
        self.diff_samples = 128 * self.precision
        bake_level(skybox.img, *skybox.shape, self.diff, 2 * res, res, -1.0, self.diff_samples)

        self.spec = [skybox.img]
        self.spec_samples = [0]
        resolution = skybox.resolution
        nsamples = 8 * self.precision
        for rough, step in self.levels:
            resolution = int(resolution / 2 ** (step * 4))
            if resolution < 2:
                raise ValueError(f'skybox resolution {skybox.resolution} is too small for the IBL levels')
            lvl = torch.empty((2 * resolution, resolution, 3), dtype=torch.float32, device=dev)
            bake_level(skybox.img, *skybox.shape, lvl, 2 * resolution, resolution, rough, nsamples)
            self.spec.append(lvl)
            self.spec_samples.append(nsamples)
            nsamples = int(nsamples * 3 ** (step * 4))
        self.struct = make_ibl_struct(self.diff, self.spec)

    def save(self, path):
        """Writes the baked maps (the Lambert map and levels 1-6; level 0 is the
        skybox itself) and the precision to an .npz file.  The reference
        pickles its Taichi fields for the same purpose (tina/shield.py:7-57);
        an .npz holds only arrays, so loading it executes nothing."""
        arrays = {'precision': np.array(self.precision), 'env_shape': np.array(self.skybox.img.shape),
                  'diff': self.diff.detach().cpu().numpy()}
        for i, m in enumerate(self.spec[1:], 1):
            arrays[f'spec{i}'] = m.detach().cpu().numpy()
        np.savez(path, **arrays)

    @classmethod
    def load(cls, path, skybox):
        """The tables of `save`, on `skybox`'s device, without baking.  The file
        must have been written for a skybox of the same shape."""
        torch = _torch()
        with np.load(path, allow_pickle=False) as z:
            if tuple(int(v) for v in z['env_shape']) != tuple(skybox.img.shape):
                raise ValueError(f'tables baked for a {tuple(z["env_shape"])} skybox, not {tuple(skybox.img.shape)}')
            self = cls.__new__(cls)
            self.skybox = skybox
            self.precision = int(z['precision'])
            self.levels = rough_levels()
            dev = skybox.img.device
            res = skybox.resolution // 6
            self.diff = torch.from_numpy(np.ascontiguousarray(z['diff'], np.float32)).to(dev)
            self.diff_samples = 128 * self.precision
            if tuple(self.diff.shape) != (2 * res, res, 3):
                raise ValueError(f'Lambert map of shape {tuple(self.diff.shape)}, expected {(2 * res, res, 3)}')
            self.spec, self.spec_samples = [skybox.img], [0]
            resolution, nsamples = skybox.resolution, 8 * self.precision
            for i, (rough, step) in enumerate(self.levels, 1):
                resolution = int(resolution / 2 ** (step * 4))
                lvl = torch.from_numpy(np.ascontiguousarray(z[f'spec{i}'], np.float32)).to(dev)
                if tuple(lvl.shape) != (2 * resolution, resolution, 3):
                    raise ValueError(f'level {i} of shape {tuple(lvl.shape)}, expected {(2 * resolution, resolution, 3)}')
                self.spec.append(lvl)
                self.spec_samples.append(nsamples)
                nsamples = int(nsamples * 3 ** (step * 4))
        self.struct = make_ibl_struct(self.diff, self.spec)
        return self

    def __getitem__(self, key):
        """ibltab['env' | 'diff' | 'spec' | 'lut'] as the reference's dict has them."""
        return {'env': self.skybox, 'diff': self.diff, 'spec': tuple(self.spec), 'lut': LUT_CORNER}[key]

    def host_copy(self):
        """(struct, arrays): the same tables copied to host memory, with a tina_ibl
        whose pointers address them -- for host-side consumers of the bakes."""
        diff = np.ascontiguousarray(self.diff.cpu().numpy())
        spec = [np.ascontiguousarray(m.cpu().numpy()) for m in self.spec]
        return make_ibl_struct(diff, spec), (diff, spec)
