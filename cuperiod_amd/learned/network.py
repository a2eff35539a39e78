"""The proposer network on the device: parameters, forward pass, training step.

Four 7-tap convolutions (2 -> 64 -> 64 -> 64 -> 64 channels, GELU) and a per-position
linear head over L = 4096 positions; softmax over the positions gives the probability
that each coarse frequency cell holds the full-grid BLS best period. The 64 -> 64
layers and their gradients run on the BF16 MFMA kernels of
``include/cuperiod_proposer.h``; the small first layer, the head and Adam run on the
vector ALUs. Master parameters are FP32; the MFMA layers read BF16 copies refreshed by
every Adam step. torch only allocates the buffers.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from cuperiod_amd import _lib
from cuperiod_amd.core.errors import BackendUnavailableError

C = 64
TAPS = 7
L = 4096
HIST_BINS = 512
#: Parameter names and shapes, in the order they are saved.
SHAPES = {
    "w1": (C, TAPS, 2), "b1": (C,),
    "w2": (C, TAPS, C), "b2": (C,),
    "w3": (C, TAPS, C), "b3": (C,),
    "w4": (C, TAPS, C), "b4": (C,),
    "wh": (C,), "bh": (1,),
}
_MFMA = ("w2", "w3", "w4")


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(None)


def _stream(device) -> ctypes.c_void_p:
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _check(status: int, what: str) -> None:
    if status != _lib.CPB_OK:
        raise _lib.status_error(status, f"proposer {what}: status {status}")


def init_parameters(seed: int) -> dict[str, np.ndarray]:
    """He-normal weights (fan-in = taps x input channels), zero biases, small head."""
    rng = np.random.default_rng(seed)
    out: dict[str, np.ndarray] = {}
    for name, shape in SHAPES.items():
        if name.startswith("b"):
            out[name] = np.zeros(shape, dtype=np.float32)
        elif name == "wh":
            out[name] = rng.normal(0.0, 0.1 / np.sqrt(C), shape).astype(np.float32)
        else:
            fan_in = shape[1] * shape[2]
            out[name] = rng.normal(0.0, np.sqrt(2.0 / fan_in), shape).astype(np.float32)
    return out


def save_parameters(params: dict[str, np.ndarray], path: str | Path) -> None:
    np.savez(path, **{k: np.asarray(v, dtype=np.float32) for k, v in params.items()})


def load_parameters(path: str | Path) -> dict[str, np.ndarray]:
    with np.load(path, allow_pickle=False) as data:
        params = {k: np.asarray(data[k], dtype=np.float32) for k in SHAPES}
    for k, shape in SHAPES.items():
        if params[k].shape != shape:
            raise ValueError(f"proposer weights {path}: {k} has shape {params[k].shape}, "
                             f"expected {shape}")
    return params


@dataclass
class Activations:
    """Per-batch buffers: pre-activations z and activations h of layers 1-4 (BF16)."""

    z: list
    h: list


class ProposerNet:
    """Device-resident parameters with forward, window and training entry points."""

    def __init__(self, params: dict[str, np.ndarray], device: int = 0) -> None:
        import torch

        self.lib = _lib.load_library()
        if not torch.cuda.is_available():
            raise BackendUnavailableError("proposer: torch sees no HIP device")
        self.device = torch.device("cuda", device)
        self.p = {k: torch.from_numpy(np.ascontiguousarray(v, dtype=np.float32)).to(self.device)
                  for k, v in params.items()}
        self.g = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.pb = {k: self.p[k].to(torch.bfloat16) for k in _MFMA}
        self.pt = {k: torch.empty_like(self.pb[k]) for k in _MFMA}
        self.step_count = 0
        self._flip()
        self._buffers: dict[tuple[int, bool], Activations] = {}
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.device)

    def parameters(self) -> dict[str, np.ndarray]:
        return {k: v.detach().cpu().numpy().copy() for k, v in self.p.items()}

    def _flip(self) -> None:
        s = _stream(self.device)
        for k in _MFMA:
            _check(self.lib.cpp_flip_weights(s, _ptr(self.pb[k]), _ptr(self.pt[k])), "flip")

    def _acts(self, n: int, train: bool) -> Activations:
        import torch

        key = (n, train)
        acts = self._buffers.get(key)
        if acts is None:
            def buf():
                return torch.empty((n, L, C), dtype=torch.bfloat16, device=self.device)

            acts = Activations(z=[buf() for _ in range(4)] if train else [None] * 4,
                               h=[buf() for _ in range(4)])
            self._buffers[key] = acts
        return acts

    def forward(self, x, *, train: bool = False) -> Activations:
        """Layers 1-4 on ``x`` ([n, 4096, 2] FP32, contiguous); returns the buffers."""
        n = int(x.shape[0])
        if tuple(x.shape) != (n, L, 2) or not x.is_contiguous():
            raise ValueError(f"proposer input must be a contiguous [n, {L}, 2] tensor")
        a = self._acts(n, train)
        s = _stream(self.device)
        _check(self.lib.cpp_layer1_forward(s, _ptr(x), _ptr(self.p["w1"]), _ptr(self.p["b1"]), n,
                                           _ptr(a.z[0]), _ptr(a.h[0])), "layer 1")
        for i, k in enumerate(_MFMA, start=1):
            _check(self.lib.cpp_conv_forward(s, _ptr(a.h[i - 1]), _ptr(self.pb[k]),
                                             _ptr(self.p["b" + k[1]]), _ptr(None), _ptr(a.z[i]),
                                             _ptr(a.h[i]), n, 0), f"layer {i + 1}")
        return a

    def windows(self, x, n_windows: int, radius: int, f_lo, dlog, half_width):
        """``(centre [n, W] int32, mass [n, W] FP32, span [n, W, 2] int32)`` device tensors
        for inputs ``x``; ``f_lo``, ``dlog``, ``half_width`` are per-curve FP64 device
        tensors (cell x has frequency ``f_lo * exp(x * dlog)``; a window claims the cells
        within ``half_width`` of its centre, at least ``radius`` each side, and ``span``
        holds the first and last of them)."""
        import torch

        n = int(x.shape[0])
        a = self.forward(x)
        centre = torch.empty((n, n_windows), dtype=torch.int32, device=self.device)
        mass = torch.empty((n, n_windows), dtype=torch.float32, device=self.device)
        span = torch.empty((n, n_windows, 2), dtype=torch.int32, device=self.device)
        _check(self.lib.cpp_head_windows(_stream(self.device), _ptr(a.h[3]), _ptr(self.p["wh"]),
                                         _ptr(self.p["bh"]), _ptr(f_lo), _ptr(dlog),
                                         _ptr(half_width), n, int(n_windows), int(radius),
                                         _ptr(centre), _ptr(mass), _ptr(span)), "head")
        return centre, mass, span

    def backward(self, x, a: Activations, labels, scale: float) -> None:
        """Accumulate the gradients of ``scale * sum(cross-entropy)`` into ``self.g``;
        ``labels`` is an [n, 2] int32 device tensor of target cell ranges."""
        import torch

        n = int(x.shape[0])
        s = _stream(self.device)
        dz = torch.empty((n, L, C), dtype=torch.bfloat16, device=self.device)
        _check(self.lib.cpp_head_train(s, _ptr(a.h[3]), _ptr(a.z[3]), _ptr(self.p["wh"]),
                                       _ptr(self.p["bh"]), _ptr(labels), n, float(scale),
                                       _ptr(dz), _ptr(self.g["wh"]), _ptr(self.g["bh"]),
                                       _ptr(self.loss)), "head backward")
        dprev = torch.empty_like(dz)
        for i in (3, 2, 1):  # layer i+1 = w{i+1}: input h[i-1], pre-activation of input z[i-1]
            k = _MFMA[i - 1]
            _check(self.lib.cpp_conv_wgrad(s, _ptr(dz), _ptr(a.h[i - 1]), n, _ptr(self.g[k]),
                                           _ptr(self.g["b" + k[1]])), f"wgrad {i + 1}")
            _check(self.lib.cpp_conv_forward(s, _ptr(dz), _ptr(self.pt[k]), _ptr(None),
                                             _ptr(a.z[i - 1]), _ptr(None), _ptr(dprev), n, 1),
                   f"dgrad {i + 1}")
            dz, dprev = dprev, dz
        _check(self.lib.cpp_layer1_wgrad(s, _ptr(dz), _ptr(x), n, _ptr(self.g["w1"]),
                                         _ptr(self.g["b1"])), "wgrad 1")

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999,
             eps: float = 1e-8) -> None:
        """One Adam step on every parameter (gradients are consumed and zeroed)."""
        self.step_count += 1
        s = _stream(self.device)
        for k, p in self.p.items():
            pb = self.pb.get(k)
            _check(self.lib.cpp_adam(s, _ptr(p), _ptr(self.g[k]), _ptr(self.m[k]),
                                     _ptr(self.v[k]), _ptr(pb), p.numel(), float(lr),
                                     float(beta1), float(beta2), float(eps), self.step_count),
                   "adam")
        self._flip()

    def train_step(self, x, labels, lr: float) -> None:
        """Forward, backward (mean cross-entropy over the batch) and Adam."""
        a = self.forward(x, train=True)
        self.backward(x, a, labels, 1.0 / float(x.shape[0]))
        self.adam(lr)


__all__ = ["C", "HIST_BINS", "L", "SHAPES", "TAPS", "ProposerNet", "init_parameters",
           "load_parameters", "save_parameters"]
