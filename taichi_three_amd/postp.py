"""Image post-processing on the [H, W, 3] f32 image: ToneMapping
(tina/postp/tonemap.py:4-12, in-place ACES), Blooming (tina/postp/blooming.py),
FXAA (tina/postp/fxaa.py), SSAO (tina/postp/ssao.py) and the TAA Accumator
(tina/util/accumator.py), each one or two calls into libtina_hip.so."""
import ctypes

from . import _abi

__all__ = ['ToneMapping', 'Blooming', 'FXAA', 'Accumator', 'SSAO', 'SSR']


class ToneMapping:
    def __init__(self, res):
        self.res = res

    def apply(self, image):
        t = image.storage
        if not t.is_contiguous() or t.dtype.itemsize != 4:
            raise ValueError('ToneMapping.apply needs a contiguous f32 RGB image')
        npix = t.numel() // 3
        _abi.check(_abi.load().tina_tonemap_aces(_abi.ptr(t), int(npix), _abi.stream_handle()))


def _dims(image):
    t = image.storage
    if not t.is_contiguous() or t.dtype.itemsize != 4 or t.dim() != 3 or t.shape[2] != 3:
        raise ValueError('post-processing needs a contiguous f32 [H, W, 3] image')
    return t, int(t.shape[1]), int(t.shape[0])


class Blooming:
    """Blooming (tina/postp/blooming.py:4-68): the image plus the blurred bright
    part of it.  Parameters as the reference: thresh 1, factor 1, scale 0.25,
    sigma 1, radius min(res) // 16; call init_gwei() after changing radius or
    sigma.  Runs before tone mapping, on the linear image."""

    def __init__(self, res):
        import torch
        self.res = (int(res[0]), int(res[1]))
        W, H = self.res
        dev = torch.device('cuda', torch.cuda.current_device())
        self.img = torch.zeros((H // 2, W // 2, 3), dtype=torch.float32, device=dev)
        self.tmp = torch.zeros_like(self.img)
        self.thresh, self.factor, self.scale, self.sigma = 1.0, 1.0, 0.25, 1.0
        self.radius = min(W, H) // 16
        self.gwei = None
        self.init_gwei()

    def init_gwei(self):
        """w_i = exp(-(sigma i / radius)^2) for i = 0..radius, divided by
        w_0 + 2 sum_{i>0} w_i (blooming.py:28-38), in float32 in index order."""
        import numpy as np
        import torch
        r = max(int(self.radius), 1)
        f32 = np.float32
        w = np.array([np.exp(-(f32(f32(self.sigma) * f32(i)) / f32(r)) ** 2) for i in range(r + 1)], dtype=f32)
        total = f32(-1.0)
        for y in w:
            total = f32(total + f32(y * f32(2)))
        self.radius = r
        self.gwei = torch.as_tensor((w / total).astype(f32), device=self.img.device)

    def apply(self, image):
        t, W, H = _dims(image)
        if (W, H) != self.res:
            raise ValueError(f'Blooming made for {self.res}, image is {(W, H)}')
        _abi.check(_abi.load().tina_bloom(_abi.ptr(t), W, H, ctypes.c_float(self.thresh), ctypes.c_float(self.factor),
                                          ctypes.c_float(self.scale), _abi.ptr(self.gwei), int(self.radius),
                                          _abi.ptr(self.img), _abi.ptr(self.tmp), _abi.stream_handle()))


class FXAA:
    """FXAA (tina/postp/fxaa.py:9-68) with the reference's thresholds
    (abs_thresh 0.0625, rel_thresh 0.063, factor 1).  Runs after tone mapping."""

    def __init__(self, res):
        import torch
        self.res = (int(res[0]), int(res[1]))
        W, H = self.res
        dev = torch.device('cuda', torch.cuda.current_device())
        self.lumi = torch.zeros((H, W), dtype=torch.float32, device=dev)
        self.img = torch.zeros((H, W, 3), dtype=torch.float32, device=dev)
        self.abs_thresh, self.rel_thresh, self.factor = 0.0625, 0.063, 1.0

    def apply(self, image):
        t, W, H = _dims(image)
        if (W, H) != self.res:
            raise ValueError(f'FXAA made for {self.res}, image is {(W, H)}')
        _abi.check(_abi.load().tina_fxaa(_abi.ptr(t), W, H, ctypes.c_float(self.abs_thresh),
                                         ctypes.c_float(self.rel_thresh), ctypes.c_float(self.factor),
                                         _abi.ptr(self.lumi), _abi.ptr(self.img), _abi.stream_handle()))


class Accumator:
    """Running mean of a sequence of frames (tina/util/accumator.py:4-23): the
    temporal anti-aliasing buffer of Scene(taa=True)."""

    def __init__(self, res):
        import torch
        from .field import Field
        W, H = int(res[0]), int(res[1])
        self._t = torch.zeros((H, W, 3), dtype=torch.float32,
                              device=torch.device('cuda', torch.cuda.current_device()))
        self.img = Field(self._t)
        self.count = 0

    def clear(self):
        self.count = 0
        self._t.zero_()

    def update(self, src):
        t, W, H = _dims(src)
        if tuple(t.shape) != tuple(self._t.shape):
            raise ValueError('Accumator.update: image size differs')
        self.count += 1
        _abi.check(_abi.load().tina_accumulate(_abi.ptr(self._t), _abi.ptr(t), int(t.numel()), int(self.count),
                                               _abi.stream_handle()))


class SSAO:
    """Screen-space ambient occlusion (tina/postp/ssao.py:5-96) from the depth
    buffer and a normal buffer (a Field [W, H, 3] filled by NormalShader).

    The nsamples hemisphere samples spherical(lerp(u, .01, 1), v) *
    lerp(w ** 1.5, .01, 1) and the noise_size x noise_size tile of rotations
    (cos t, sin t) are drawn on the host (NumPy, `seed`), where the reference
    draws them with ti.random; seed_samples() draws a new set.  With taa the
    reference takes nsamples // 4 fresh samples per pixel and skips the box
    filter; here each frame draws a fresh set of nsamples // 4 samples and
    rotations and skips the box filter."""

    def __init__(self, res, norm, nsamples=64, thresh=0.0, radius=0.2, factor=1.0, noise_size=4, taa=False,
                 seed=0):
        import numpy as np
        import torch
        self.res = (int(res[0]), int(res[1]))
        W, H = self.res
        dev = torch.device('cuda', torch.cuda.current_device())
        self.norm = norm
        self.radius, self.thresh, self.factor = float(radius), float(thresh), float(factor)
        self.taa = bool(taa)
        self.nsamples = int(nsamples) if not self.taa else int(nsamples) // 4
        self.noise_size = int(noise_size)
        self.img = torch.zeros((H, W), dtype=torch.float32, device=dev)
        self._rng = np.random.default_rng(seed)
        self.seed_samples()

    @staticmethod
    def make_samples(rng, n):
        """make_sample (ssao.py:50-55) for n samples, in binary32."""
        import numpy as np
        f32 = np.float32
        u, v, w = (rng.random(n).astype(f32) for _ in range(3))
        r = (f32(0.01) * (f32(1) - w ** f32(1.5)) + f32(1.0) * w ** f32(1.5)).astype(f32)
        u = (f32(0.01) * (f32(1) - u) + f32(1.0) * u).astype(f32)
        t = (v * f32(2 * np.pi)).astype(f32)
        s = np.sqrt(np.maximum(f32(0), f32(1) - u * u)).astype(f32)
        return (np.stack([s * np.cos(t), s * np.sin(t), u], axis=1).astype(f32) * r[:, None]).astype(f32)

    def seed_samples(self):
        """seed_samples (ssao.py:28-35): the sample set and the rotation tile."""
        import numpy as np
        import torch
        dev = self.img.device
        self.samples = torch.as_tensor(self.make_samples(self._rng, self.nsamples), device=dev).contiguous()
        t = (self._rng.random((self.noise_size, self.noise_size)) * 2 * np.pi).astype(np.float32)
        rot = np.stack([np.cos(t), np.sin(t)], axis=2).astype(np.float32)
        self.rotations = torch.as_tensor(rot, device=dev).contiguous()

    def render(self, engine):
        """render (ssao.py:61-65): the occlusion of every pixel into self.img."""
        if self.taa:
            self.seed_samples()
        n = self.norm.storage
        _abi.check(engine._lib.tina_ssao(engine._ctx, _abi.ptr(n), _abi.ptr(self.samples), int(self.nsamples),
                                         _abi.ptr(self.rotations), int(self.noise_size), ctypes.c_float(self.radius),
                                         ctypes.c_float(self.thresh), ctypes.c_float(self.factor),
                                         _abi.ptr(self.img), _abi.stream_handle()), engine._ctx)

    def apply(self, image):
        """apply (ssao.py:37-48): image *= 1 - (box-filtered) occlusion."""
        t, W, H = _dims(image)
        _abi.check(_abi.load().tina_ssao_apply(_abi.ptr(t), W, H, _abi.ptr(self.img), int(self.noise_size),
                                               0 if self.taa else 1, _abi.stream_handle()))


class SSR:
    """Screen-space reflections (tina/postp/ssr.py:6-103) from the depth buffer,
    a normal buffer (NormalShader) and, with texturing, a texcoord buffer
    (TexcoordShader).  render() runs one pass per object over the pixels it won,
    drawing reflection directions from that object's material (the reference
    looks the material up in a per-pixel material-id buffer; the object that won
    a pixel is the one whose material that buffer holds).  The non-TAA mode
    draws the reference's own WangHashRNG(P % blurring) stream; with taa a
    stream hashed from the pixel and the frame, where the reference draws
    ti.random."""

    def __init__(self, res, norm, coor=None, taa=False):
        import torch
        self.res = (int(res[0]), int(res[1]))
        W, H = self.res
        self.norm = norm
        self.coor = coor
        self.taa = bool(taa)
        self.nsamples = 32 if not self.taa else 12
        self.nsteps = 32 if not self.taa else 64
        self.stepsize = 2.0
        self.tolerance = 15.0
        self.blurring = 4
        self.frame = 0
        self.img = torch.zeros((H, W, 4), dtype=torch.float32,
                               device=torch.device('cuda', torch.cuda.current_device()))

    def begin(self):
        """Empties the reflection buffer for a new frame (render's pixels of no object)."""
        self.img.zero_()
        self.frame += 1

    def render_object(self, engine, image, raster, material):
        """render (ssr.py:51-56) over the pixels the object bound to `raster`
        (set_object) won, drawing from `material`."""
        if raster.nfaces == 0 or not hasattr(raster, 'mesh_struct'):
            return                           # wires write no material id (IShader.blend_color)
        t, _, _ = _dims(image)
        lo = raster.lowered(material)
        mesh = raster.mesh_struct()
        _abi.check(engine._lib.tina_ssr(
            engine._ctx, ctypes.byref(mesh), ctypes.byref(lo.struct), _abi.ptr(self.norm.storage),
            _abi.ptr(self.coor.storage) if self.coor is not None else None, _abi.ptr(t), int(self.nsamples),
            int(self.nsteps), ctypes.c_float(self.stepsize), ctypes.c_float(self.tolerance), int(self.blurring),
            1 if self.taa else 0, self.frame & 0xffffffff, _abi.ptr(self.img), _abi.stream_handle()), engine._ctx)

    def apply(self, image):
        """apply (ssr.py:31-45): image * (1 - a) + rgb of the box-filtered buffer."""
        t, W, H = _dims(image)
        _abi.check(_abi.load().tina_ssr_apply(_abi.ptr(t), W, H, _abi.ptr(self.img), int(self.blurring),
                                              1 if self.taa else 0, _abi.stream_handle()))
