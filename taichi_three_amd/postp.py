"""ToneMapping (tina/postp/tonemap.py:4-12): in-place ACES on an image Field."""
from . import _abi

__all__ = ['ToneMapping']


class ToneMapping:
    def __init__(self, res):
        self.res = res

    def apply(self, image):
        t = image.storage
        if not t.is_contiguous() or t.dtype.itemsize != 4:
            raise ValueError('ToneMapping.apply needs a contiguous f32 RGB image')
        npix = t.numel() // 3
        _abi.check(_abi.load().tina_tonemap_aces(_abi.ptr(t), int(npix), _abi.stream_handle()))
