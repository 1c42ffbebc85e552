"""Learned subgrid form: a periodic 3-D CNN on a coarse grid (libld.so, bf16
MFMA with FP32 accumulation, trained on the GPU by its own forward and
backward kernels).  No counterpart in the reference.

Fields are rows ``x[node, c]`` (node = (i N + j) N + k, the layout of a
``(N, N, N, c)`` array), float32 CUDA tensors.  The parameters live on the
device in the library's padded layout (channel counts rounded up to 16 or 32,
padded entries zero); ``pack``/``unpack`` convert from/to per-layer
``(W (27, co, ci), b (co,))`` arrays with the real channel counts, tap =
(dx+1) 9 + (dy+1) 3 + (dz+1).
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import _lib


def _pad(c: int) -> int:
    return 16 if c <= 16 else 32


def _ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


class SubgridCNN:
    def __init__(self, N: int, channels: Sequence[int] = (3, 32, 32, 32, 3), device: str = "cuda"):
        if not torch.cuda.is_available():
            raise RuntimeError("SubgridCNN needs a GPU (libld.so has no CPU path)")
        self.N = int(N)
        self.channels = tuple(int(c) for c in channels)
        self.L = len(self.channels) - 1
        self.device = torch.device(device)
        self.lib = _lib.load()
        p = _lib.LdCnnParams()
        p.N = self.N
        p.nlayers = self.L
        for i, c in enumerate(self.channels):
            p.channels[i] = c
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.ld_cnn_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        n = ctypes.c_int64()
        _lib.check(self.lib.ld_cnn_nparams(self.h, ctypes.byref(n)))
        self.nparams = n.value
        self.padded = [_pad(c) for c in self.channels]

    def close(self):
        if getattr(self, "h", None):
            self.lib.ld_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        self.lib.ld_cnn_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))

    # ------------------------------------------------------------ parameters
    def pack(self, layers: List[Tuple[np.ndarray, np.ndarray]]) -> np.ndarray:
        """Per-layer (W (27, co, ci), b (co,)) with real channel counts -> flat padded FP32."""
        out = []
        for l, (W, b) in enumerate(layers):
            ci, co = self.channels[l], self.channels[l + 1]
            CI, CO = self.padded[l], self.padded[l + 1]
            if W.shape != (27, co, ci) or b.shape != (co,):
                raise ValueError(f"layer {l}: expected W (27, {co}, {ci}) and b ({co},)")
            Wp = np.zeros((27, CO, CI), dtype=np.float32)
            Wp[:, :co, :ci] = W
            bp = np.zeros(CO, dtype=np.float32)
            bp[:co] = b
            out += [Wp.ravel(), bp]
        flat = np.concatenate(out)
        assert flat.size == self.nparams
        return flat

    def unpack(self, flat: np.ndarray) -> List[Tuple[np.ndarray, np.ndarray]]:
        layers, o = [], 0
        for l in range(self.L):
            ci, co = self.channels[l], self.channels[l + 1]
            CI, CO = self.padded[l], self.padded[l + 1]
            W = flat[o:o + 27 * CO * CI].reshape(27, CO, CI)[:, :co, :ci]
            o += 27 * CO * CI
            b = flat[o:o + CO][:co]
            o += CO
            layers.append((W.copy(), b.copy()))
        return layers

    def init_params(self, seed: int = 0, gain: float = 1.0) -> List[Tuple[np.ndarray, np.ndarray]]:
        """He-normal weights (fan-in 27 ci), zero biases; the last layer scaled by `gain`."""
        rng = np.random.default_rng(seed)
        layers = []
        for l in range(self.L):
            ci, co = self.channels[l], self.channels[l + 1]
            W = rng.normal(0.0, np.sqrt(2.0 / (27 * ci)), (27, co, ci)).astype(np.float32)
            if l == self.L - 1:
                W *= gain
            layers.append((W, np.zeros(co, dtype=np.float32)))
        self.set_params(layers)
        return layers

    def set_params(self, layers):
        flat = self.pack(layers) if isinstance(layers, list) else np.asarray(layers, dtype=np.float32)
        t = torch.from_numpy(flat).to(self.device)
        self._stream()
        _lib.check(self.lib.ld_cnn_set_params(self.h, _ptr(t)))
        torch.cuda.synchronize(self.device)

    def get_params(self) -> np.ndarray:
        t = torch.empty(self.nparams, dtype=torch.float32, device=self.device)
        self._stream()
        _lib.check(self.lib.ld_cnn_get_params(self.h, _ptr(t)))
        return t.cpu().numpy()

    # ------------------------------------------------------------ compute
    def _rows(self, x: torch.Tensor, c: int, name: str) -> torch.Tensor:
        n3 = self.N ** 3
        if x.device.type != "cuda" or x.dtype != torch.float32:
            raise TypeError(f"{name} must be a float32 CUDA tensor")
        x = x.reshape(n3, c)
        if not x.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        return x

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        x = self._rows(x, self.channels[0], "x")
        y = torch.empty((x.shape[0], self.channels[-1]), dtype=torch.float32, device=self.device)
        self._stream()
        _lib.check(self.lib.ld_cnn_forward(self.h, _ptr(x), _ptr(y)))
        return y

    def grad(self, x: torch.Tensor, target: torch.Tensor):
        """(loss, flat padded parameter gradient) of the mean squared error."""
        x = self._rows(x, self.channels[0], "x")
        t = self._rows(target, self.channels[-1], "target")
        g = torch.empty(self.nparams, dtype=torch.float32, device=self.device)
        loss = ctypes.c_double()
        self._stream()
        _lib.check(self.lib.ld_cnn_grad(self.h, _ptr(x), _ptr(t), _ptr(g), ctypes.byref(loss)))
        return loss.value, g

    def train_step(self, x, target, lr=1e-3, betas=(0.9, 0.999), eps=1e-8) -> float:
        x = self._rows(x, self.channels[0], "x")
        t = self._rows(target, self.channels[-1], "target")
        loss = ctypes.c_double()
        self._stream()
        _lib.check(self.lib.ld_cnn_train_step(self.h, _ptr(x), _ptr(t), float(lr), float(betas[0]), float(betas[1]),
                                              float(eps), ctypes.byref(loss)))
        return loss.value
