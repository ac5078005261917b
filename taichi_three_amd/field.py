"""A small stand-in for the Taichi fields Tina's API hands out.

Tina exposes its buffers as Taichi fields indexed [x, y] with y = 0 at the
bottom row (tina/core/engine.py:9, tina/scene/raster.py:34); user code reads
them with `.to_numpy()` (shape (W, H, ...)) or `field[None]`.  Here every
buffer is a device tensor stored [y][x] (x fastest, the layout the kernels
stream through), and Field presents the [x, y] view on top of it.
"""
import numpy as np


class Field:
    """Device buffer with Taichi-field-like accessors.

    `storage` is the torch tensor in kernel layout: [H, W, *channels] for
    screen buffers (`screen=True`), or any shape for plain buffers.
    """

    def __init__(self, storage, screen=True):
        self.storage = storage
        self.screen = screen

    @property
    def shape(self):
        s = tuple(self.storage.shape)
        if self.screen:
            return (s[1], s[0]) + s[2:]
        return s

    @property
    def n(self):
        return self.storage.shape[-1] if self.storage.dim() > 2 else 1

    def to_torch(self):
        """[x, y, ...] view of the device tensor (no copy)."""
        if self.screen:
            return self.storage.transpose(0, 1)
        return self.storage

    def to_numpy(self):
        return self.to_torch().detach().cpu().numpy()

    def __array__(self, dtype=None):
        a = self.to_numpy()
        return a if dtype is None else a.astype(dtype)

    def from_numpy(self, arr):
        import torch
        t = torch.as_tensor(np.asarray(arr), dtype=self.storage.dtype, device=self.storage.device)
        if self.screen:
            t = t.transpose(0, 1)
        self.storage.copy_(t)

    def fill(self, value):
        if np.ndim(value) == 0:
            self.storage.fill_(value)
        else:
            import torch
            self.storage.copy_(torch.as_tensor(value, dtype=self.storage.dtype,
                                               device=self.storage.device).expand_as(self.storage))

    def __getitem__(self, idx):
        if idx is None:
            return self.storage.detach().cpu().numpy()
        if self.screen and isinstance(idx, tuple) and len(idx) >= 2:
            idx = (idx[1], idx[0]) + tuple(idx[2:])
        v = self.storage[idx]
        return v.item() if v.numel() == 1 and v.dim() == 0 else v.detach().cpu().numpy()

    def __setitem__(self, idx, value):
        import torch
        if idx is None:
            self.storage.copy_(torch.as_tensor(np.asarray(value), dtype=self.storage.dtype,
                                               device=self.storage.device).reshape(self.storage.shape))
            return
        if self.screen and isinstance(idx, tuple) and len(idx) >= 2:
            idx = (idx[1], idx[0]) + tuple(idx[2:])
        self.storage[idx] = torch.as_tensor(np.asarray(value), dtype=self.storage.dtype)

    def __repr__(self):
        return f'Field(shape={self.shape}, dtype={self.storage.dtype}, device={self.storage.device})'
