"""Image post-processing on the [H, W, 3] f32 image: ToneMapping
(tina/postp/tonemap.py:4-12, in-place ACES), Blooming (tina/postp/blooming.py),
FXAA (tina/postp/fxaa.py) and the TAA Accumator (tina/util/accumator.py), each
one call into libtina_hip.so."""
import ctypes

from . import _abi

__all__ = ['ToneMapping', 'Blooming', 'FXAA', 'Accumator']


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
