"""Learned shading surrogate for the PBR material under image-based lighting.

No counterpart in the reference.  A 12-64-64-64-4 ReLU network stands in for
the analytic evaluation of `PBR` (tina/matr/material.py:322-362, 701-706) under
`SkyboxLighting` plus one directional light (tina/core/lighting.py:4-21,
84-98), the `Lighting.shade_color` of every covered pixel.  Its inputs are the
normal n, the view direction as its mirror image about n (R = 2 (n.V) n - V),
the cosine R.L to light 0, log(roughness), metallic and albedo; its outputs are
the ACES-mapped colour and an estimate of that colour's largest channel error.
Pixels whose estimate exceeds `threshold` are shaded by the analytic kernel in
the same call, so the analytic shader is both the fallback and the teacher.

Everything runs in libtina_hip.so: training pairs come from the analytic shader
(`tina_nn_samples`), the network is trained by the library's float32 forward /
backward / Adam kernels (`tina_mlp_train_step`), packed to bf16 MFMA fragments
(`tina_mlp_pack`) and evaluated inside the deferred shading pass
(`tina_shade_nn`) on v_mfma_f32_16x16x32_bf16.
"""
import ctypes
import math

import numpy as np

from . import _abi

__all__ = ['NeuralShading']

NPARAMS = 9412
PACK_BYTES = 22 * 64 * 16 + 196 * 4
BLOCK = 64
SIZES = (12, 64, 64, 64, 4)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('taichi_three_amd needs a ROCm GPU: torch.cuda.is_available() is False')
    return torch


def init_params(seed):
    """He-uniform weights (limit sqrt(6 / fan_in)), zero biases, in the flat
    layout W1 b1 W2 b2 W3 b3 W4 b4."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(4):
        o, k = SIZES[i + 1], SIZES[i]
        lim = math.sqrt(6.0 / k)
        out.append(rng.uniform(-lim, lim, o * k))
        out.append(np.zeros(o))
    p = np.concatenate(out).astype(np.float32)
    assert p.size == NPARAMS
    return p


class NeuralShading:
    """The surrogate for one lighting environment: `lighting` is a
    SkyboxLighting whose light 0 is directional (the sun the network sees).

        ns = NeuralShading(scene.lighting)
        ns.train()                  # 16M pairs from the analytic shader
        scene.neural = ns           # scene.render() now shades PBR objects with it
    """

    def __init__(self, lighting, specular=0.5, seed=0, threshold=1e-2):
        from .lighting import SkyboxLighting
        if not isinstance(lighting, SkyboxLighting):
            raise TypeError('NeuralShading learns SkyboxLighting (image-based lighting)')
        if lighting.nlights < 1 or float(lighting.light_dirs[0, 3]) != 0.0:
            raise ValueError('NeuralShading needs a directional light 0 (add_light(dir=...))')
        torch = _torch()
        self.lighting = lighting
        self.specular = float(specular)
        self.threshold = float(threshold)
        self.device = torch.device('cuda', torch.cuda.current_device())
        self.params = torch.from_numpy(init_params(seed)).to(self.device)
        self.m1 = torch.zeros(NPARAMS, dtype=torch.float32, device=self.device)
        self.m2 = torch.zeros_like(self.m1)
        self.step = 0
        self.packed = torch.zeros(PACK_BYTES, dtype=torch.uint8, device=self.device)
        self._partial = None
        self.pack()

    # ---- data and training ----------------------------------------------------
    def samples(self, n, seed=1, roughness=(0.05, 1.0), p_metal=0.3):
        """(X [n][12], Y [n][4]) device tensors from the analytic shader."""
        torch = _torch()
        X = torch.empty((n, 12), dtype=torch.float32, device=self.device)
        Y = torch.empty((n, 4), dtype=torch.float32, device=self.device)
        lt = self.lighting.to_struct()
        _abi.check(_abi.load().tina_nn_samples(
            ctypes.byref(lt), ctypes.byref(self.lighting.ibltab.struct), int(seed) & 0xffffffff, int(n),
            ctypes.c_float(roughness[0]), ctypes.c_float(roughness[1]), ctypes.c_float(p_metal),
            ctypes.c_float(self.specular), _abi.ptr(X), _abi.ptr(Y), _abi.stream_handle()))
        return X, Y

    def _workspace(self, batch):
        torch = _torch()
        rows = (batch + BLOCK - 1) // BLOCK
        if self._partial is None or self._partial.numel() < rows * NPARAMS:
            self._partial = torch.empty(rows * NPARAMS, dtype=torch.float32, device=self.device)
        return self._partial

    def train_step(self, X, Y, first, batch, lr):
        """One Adam step on rows [first, first + batch) of (X, Y)."""
        self.step += 1
        part = self._workspace(batch)
        _abi.check(_abi.load().tina_mlp_train_step(
            _abi.ptr(X), _abi.ptr(Y), int(first), int(X.shape[0]), int(batch), _abi.ptr(self.params),
            _abi.ptr(self.m1), _abi.ptr(self.m2), _abi.ptr(part), ctypes.c_float(lr), int(self.step), None,
            _abi.stream_handle()))

    def gradient(self, X, Y, first=0, batch=None):
        """The loss gradient over rows [first, first + batch), parameters unchanged."""
        torch = _torch()
        batch = int(X.shape[0] - first if batch is None else batch)
        g = torch.empty(NPARAMS, dtype=torch.float32, device=self.device)
        part = self._workspace(batch)
        _abi.check(_abi.load().tina_mlp_train_step(
            _abi.ptr(X), _abi.ptr(Y), int(first), int(X.shape[0]), batch, _abi.ptr(self.params),
            _abi.ptr(self.m1), _abi.ptr(self.m2), _abi.ptr(part), ctypes.c_float(0.0), 1, _abi.ptr(g),
            _abi.stream_handle()))
        return g

    def train(self, nsamples=1 << 24, steps=20000, batch=16384, lr=3e-3, lr_end=3e-5, seed=1, p_metal=0.3,
              data=None):
        """Trains on `nsamples` analytic pairs (or `data` = (X, Y)) for `steps`
        Adam steps of `batch` rows, cosine learning-rate decay lr -> lr_end, then
        packs the weights for inference."""
        X, Y = data if data is not None else self.samples(nsamples, seed=seed, p_metal=p_metal)
        n = int(X.shape[0])
        if n < batch:
            raise ValueError(f'{n} samples cannot fill a batch of {batch}')
        nb = n // batch
        for s in range(steps):
            f = lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * s / max(1, steps)))
            self.train_step(X, Y, (s % nb) * batch, batch, f)
        self.pack()
        return self

    # ---- checkpoint ---------------------------------------------------------------
    def save(self, path):
        """Writes the float32 master weights, both Adam moments, the step count
        and the settings to an .npz file (arrays only: loading it executes
        nothing, where the reference pickles its fields, tina/shield.py:7-57)."""
        np.savez(path, params=self.params.detach().cpu().numpy(), m1=self.m1.detach().cpu().numpy(),
                 m2=self.m2.detach().cpu().numpy(), step=np.array(self.step),
                 specular=np.array(self.specular), threshold=np.array(self.threshold))

    @classmethod
    def load(cls, path, lighting):
        """A surrogate for `lighting` resumed from `save`: training continues
        from the saved moments and step, and the packed weights are ready."""
        torch = _torch()
        with np.load(path, allow_pickle=False) as z:
            arrays = {k: z[k] for k in ('params', 'm1', 'm2')}
            for k, v in arrays.items():
                if v.shape != (NPARAMS,) or v.dtype != np.float32:
                    raise ValueError(f'{k}: expected {NPARAMS} float32 values, got {v.shape} {v.dtype}')
            self = cls(lighting, specular=float(z['specular']), threshold=float(z['threshold']))
            self.step = int(z['step'])
        for k, v in arrays.items():
            getattr(self, k).copy_(torch.from_numpy(v))
        self.pack()
        return self

    # ---- inference --------------------------------------------------------------
    def pack(self):
        _abi.check(_abi.load().tina_mlp_pack(_abi.ptr(self.params), _abi.ptr(self.packed), _abi.stream_handle()))

    def infer(self, X):
        """The packed network on features X [n][12]: [n][4] (colour, error estimate)."""
        torch = _torch()
        O = torch.empty((X.shape[0], 4), dtype=torch.float32, device=self.device)
        _abi.check(_abi.load().tina_mlp_inference(_abi.ptr(self.packed), _abi.ptr(X), int(X.shape[0]),
                                                  _abi.ptr(O), _abi.stream_handle()))
        return O

    def analytic(self, X):
        """The analytic shader on the same feature rows: [n][4] (ACES colour, 0)."""
        torch = _torch()
        O = torch.empty((X.shape[0], 4), dtype=torch.float32, device=self.device)
        lt = self.lighting.to_struct()
        _abi.check(_abi.load().tina_nn_analytic_rows(
            ctypes.byref(lt), ctypes.byref(self.lighting.ibltab.struct), _abi.ptr(X), int(X.shape[0]),
            ctypes.c_float(self.specular), _abi.ptr(O), _abi.stream_handle()))
        return O

    def fallback_count(self, engine):
        """Pixels the last neural shade of `engine` handed to the analytic kernel."""
        torch = _torch()
        buf = torch.zeros(1, dtype=torch.int32, device=self.device)
        _abi.check(engine._lib.tina_nn_fallback_count(engine._ctx, _abi.ptr(buf), _abi.stream_handle()),
                   engine._ctx)
        return int(buf.item()) & 0xffffffff
