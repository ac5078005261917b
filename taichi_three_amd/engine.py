"""Engine: the per-screen state of the rasteriser (tina/core/engine.py:4-76).

The reference keeps an i32 depth field and, per raster, an i32 `occup` field
written under a race (tina/core/triangle.py:123-125).  Here both live in one
packed u64 key per pixel, ((depth ^ 0x80000000) << 32) | face_id, owned by the
C-ABI context and reduced with a 64-bit atomic minimum; depth and ids are
unpacked from it on demand.
"""
import ctypes

import numpy as np

from . import _abi
from .field import Field

MAX = 2 ** 20  # tina/common.py:7


def _require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('taichi_three_amd needs a ROCm GPU: torch.cuda.is_available() is False')
    return torch


class Engine:
    """Screen, camera and the packed depth|id buffer (tina/core/engine.py:6-28)."""

    def __init__(self, res_x=512, res_y=None, maxfaces=MAX):
        if res_y is None:
            res_y = res_x
        if isinstance(res_x, (tuple, list, np.ndarray)):
            res_x, res_y = int(res_x[0]), int(res_x[1])
        self.res = np.array([int(res_x), int(res_y)])
        self.maxdepth = _abi.MAXDEPTH
        self.torch = _require_gpu()
        self.device = self.torch.device('cuda', self.torch.cuda.current_device())
        lib = _abi.load()
        handle = ctypes.c_void_p()
        _abi.check(lib.tina_ctx_create(int(res_x), int(res_y), int(maxfaces), ctypes.byref(handle)))
        self._ctx = handle
        self._cap = int(maxfaces)
        self._lib = lib
        # default camera: identity with z flipped, bias (0.5, 0.5)  (engine.py:20-26)
        self.W2V = np.eye(4, dtype=np.float32)
        self.W2V[2, 2] = -1
        self.V2W = self.W2V.copy()
        self.bias = np.array([0.5, 0.5], dtype=np.float32)
        self._push_camera()
        self._id_cursor = 0

    def __del__(self):
        ctx = getattr(self, '_ctx', None)
        if ctx is not None and ctx.value:
            try:
                self._lib.tina_ctx_destroy(ctx)
            except Exception:
                pass
            self._ctx = None

    # ---- camera ------------------------------------------------------------
    def _camera_struct(self):
        cam = _abi.Camera()
        cam.W2V[:] = [float(x) for x in np.asarray(self.W2V, dtype=np.float32).reshape(-1)]
        cam.V2W[:] = [float(x) for x in np.asarray(self.V2W, dtype=np.float32).reshape(-1)]
        cam.bias[:] = [float(self.bias[0]), float(self.bias[1])]
        return cam

    def _push_camera(self):
        cam = self._camera_struct()
        _abi.check(self._lib.tina_set_camera(self._ctx, ctypes.byref(cam)), self._ctx)

    def set_camera(self, view, proj):
        """W2V = proj @ view and V2W = inv(W2V) in FP64, then cast to f32 (engine.py:72-76)."""
        from .util.matrix import camera_matrices
        self.W2V, self.V2W = camera_matrices(view, proj)
        self._push_camera()

    def randomize_bias(self, center):
        """Sub-pixel jitter for TAA (engine.py:30-39)."""
        if center:
            self.bias = np.array([0.5, 0.5], dtype=np.float32)
        else:
            self.bias = np.random.random_sample(2).astype(np.float32)
        self._push_camera()

    # ---- host-side helpers mirroring the ti.func ones (engine.py:51-65) ----
    def to_viewspace(self, p):
        p = np.asarray(p, dtype=np.float32)
        r = self.W2V[:3, :3] @ p + self.W2V[:3, 3]
        w = self.W2V[3, :3] @ p + self.W2V[3, 3]
        return r / w

    def from_viewspace(self, p):
        p = np.asarray(p, dtype=np.float32)
        r = self.V2W[:3, :3] @ p + self.V2W[:3, 3]
        w = self.V2W[3, :3] @ p + self.V2W[3, 3]
        return r / w

    def to_viewport(self, p):
        return (np.asarray(p)[:2] * 0.5 + 0.5) * self.res

    def from_viewport(self, p):
        return np.asarray(p) / self.res * 2 - 1

    # ---- buffers -----------------------------------------------------------
    def reserve(self, nfaces):
        if nfaces > self._cap:
            _abi.check(self._lib.tina_ctx_reserve(self._ctx, int(nfaces)), self._ctx)
            self._cap = int(nfaces)

    def clear_depth(self):
        """Engine.clear_depth (engine.py:67-70): every key -> (2**30, no face)."""
        _abi.check(self._lib.tina_clear(self._ctx, _abi.stream_handle()), self._ctx)
        self._id_cursor = 0

    def resolve(self, want_id=True, want_depth=True):
        """Unpacks the keys into (id, depth) device tensors [H, W] i32."""
        torch = self.torch
        H, W = int(self.res[1]), int(self.res[0])
        ids = torch.empty((H, W), dtype=torch.int32, device=self.device) if want_id else None
        dep = torch.empty((H, W), dtype=torch.int32, device=self.device) if want_depth else None
        _abi.check(self._lib.tina_resolve(self._ctx, _abi.ptr(ids) or None, _abi.ptr(dep) or None,
                                          _abi.stream_handle()), self._ctx)
        return ids, dep

    @property
    def depth(self):
        """Engine.depth as a Field of shape (W, H): i32, 2**30 where nothing was drawn."""
        return Field(self.resolve(want_id=False)[1])

    @property
    def ids(self):
        """Per-pixel winning global face id (W, H) i32, -1 where nothing was drawn."""
        return Field(self.resolve(want_depth=False)[0])

    def keys(self):
        """The packed u64 key buffer as an int64 tensor view [H, W] (no copy).
        A clear is applied by the next library call that touches the buffer
        (this one included), so take the view after the clear you want to see."""
        p = ctypes.c_void_p()
        _abi.check(self._lib.tina_keys(self._ctx, ctypes.byref(p)), self._ctx)
        H, W = int(self.res[1]), int(self.res[0])
        return _wrap_device_ptr(self.torch, p.value, (H, W), self.torch.int64, self.device)


def _wrap_device_ptr(torch, addr, shape, dtype, device):
    """Zero-copy tensor over library-owned device memory (lifetime: the Engine)."""
    n = int(np.prod(shape))

    class _Holder:
        __cuda_array_interface__ = {
            'shape': (n,), 'typestr': '<i8' if dtype == torch.int64 else '<i4',
            'data': (int(addr), False), 'version': 3, 'strides': None}

    return torch.as_tensor(_Holder(), device=device).view(*shape)
