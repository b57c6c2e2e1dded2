"""Learned warm start for the TGV-MRI solvers (no counterpart in the
reference, which starts every solve from u = A^*f, v = 0:
tgvmrioptimizer.cu:63-79,112-118).

A conv3x3 network (5 layers 2 -> 32 -> 32 -> 32 -> 32 -> 6, ReLU between
layers) maps the zero-filled reconstruction A^*f, divided by its RMS s, to a
residual for u and a prediction of v (the TGV vector field):

    x = A^*f / s,  (r, v') = net(x),  u0 = A^*f + s r,  v0 = s v'.

Convolutions run on the bf16 MFMA of the GPU (fp32 accumulation); weights,
gradients and Adam state are fp32.  ``WarmStartCNN.warm_start(solver)`` sets
(u0, v0) on a configured solver in one call; ``training_pair`` builds the
(x, t) pair of a solved slice.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from . import _capi as C

CIN, COUT = 2, 6
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_WEIGHTS = os.path.join(DATA_DIR, "warmstart_tgv_mri_320.npz")


def layer_channels(depth: int = 5, width: int = 32) -> list[int]:
    return [CIN] + [width] * (depth - 1) + [COUT]


def num_params(ch) -> int:
    return sum(9 * a * b + b for a, b in zip(ch[:-1], ch[1:]))


def he_init(ch, seed: int = 0, last_scale: float = 0.1) -> np.ndarray:
    """He-normal weights (std sqrt(2 / (9 cin))), zero biases; the last
    layer scaled by last_scale so that an untrained net starts near u0 = A^*f."""
    rng = np.random.default_rng(seed)
    out = []
    for l, (a, b) in enumerate(zip(ch[:-1], ch[1:])):
        w = rng.standard_normal(9 * a * b) * np.sqrt(2.0 / (9 * a))
        if l == len(ch) - 2:
            w *= last_scale
        out += [w, np.zeros(b)]
    return np.concatenate(out).astype(np.float32)


class WarmStartCNN:
    """The network on the GPU (handle of pdt_cnn)."""

    def __init__(self, depth: int = 5, width: int = 32, seed: int = 0):
        self.depth, self.width = depth, width
        self.channels = layer_channels(depth, width)
        h = ct.c_void_p()
        C.check(C.lib().pdt_cnn_create(CIN, width, depth, COUT, ct.byref(h)))
        self._h = h
        n = ct.c_size_t()
        C.check(C.lib().pdt_cnn_num_params(self._h, ct.byref(n)))
        self.n_params = n.value
        assert self.n_params == num_params(self.channels)
        self.set_params(he_init(self.channels, seed))

    # -- parameters -----------------------------------------------------
    def set_params(self, p) -> None:
        """Uploads the parameters and resets the Adam state."""
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        C.check(C.lib().pdt_cnn_set_params(self._h, p.ctypes.data_as(ct.POINTER(ct.c_float)), p.size))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        C.check(C.lib().pdt_cnn_get_params(self._h, p.ctypes.data_as(ct.POINTER(ct.c_float)), p.size))
        return p

    def save(self, path: str, **meta) -> None:
        np.savez(path, params=self.get_params(), depth=self.depth, width=self.width,
                 **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS) -> "WarmStartCNN":
        with np.load(path, allow_pickle=False) as z:
            net = cls(int(z["depth"]), int(z["width"]))
            net.set_params(z["params"])
        return net

    # -- compute ----------------------------------------------------------
    @staticmethod
    def _dev(a, ch):
        if isinstance(a, C.DeviceArray):
            return a
        a = np.ascontiguousarray(a, np.float32)
        assert a.ndim == 4 and a.shape[-1] == ch, a.shape
        return C.DeviceArray.from_numpy(a.reshape(-1))

    def forward(self, x) -> np.ndarray:
        """x: (n, h, w, 2) -> (n, h, w, 6)."""
        n, h, w, _ = x.shape
        dx = self._dev(x, CIN)
        y = C.DeviceArray.empty(C.PDT_F32, [n * h * w * COUT])
        C.check(C.lib().pdt_cnn_forward(self._h, dx.h, y.h, n, h, w))
        return y.numpy().reshape(n, h, w, COUT)

    def loss_grad(self, x, t):
        """(loss, gradient) of 1/(2m) sum (net(x) - t)^2."""
        n, h, w, _ = x.shape
        g = np.empty(self.n_params, np.float32)
        loss = ct.c_double()
        dx, dt = self._dev(x, CIN), self._dev(t, COUT)  # keep the device arrays alive across the call
        C.check(C.lib().pdt_cnn_loss_grad(self._h, dx.h, dt.h, n, h, w,
                                          ct.byref(loss), g.ctypes.data_as(ct.POINTER(ct.c_float)), g.size))
        return loss.value, g

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        opt = C.PdtAdam(lr, beta1, beta2, eps)
        C.check(C.lib().pdt_cnn_adam_step(self._h, ct.byref(opt)))

    def train_step(self, x, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, shape=None) -> float:
        """One Adam step on a batch; x, t numpy (n, h, w, c) or flat DeviceArrays
        with shape = (n, h, w)."""
        n, h, w = shape if shape is not None else x.shape[:3]
        opt = C.PdtAdam(lr, beta1, beta2, eps)
        loss = ct.c_double()
        dx, dt = self._dev(x, CIN), self._dev(t, COUT)
        C.check(C.lib().pdt_cnn_train_step(self._h, dx.h, dt.h, n, h, w, ct.byref(opt), ct.byref(loss)))
        return loss.value

    def warm_start(self, solver) -> None:
        """Sets (u0, v0) of a configured complex64 TGV-MRI / TGV-with-operator
        solver from its data (replaces the A^*f cold start of solve())."""
        C.check(C.lib().pdt_solver_warm_start(solver._h, self._h))

    def __del__(self):
        if getattr(self, "_h", None) and C._lib is not None:
            C._lib.pdt_cnn_destroy(self._h)


def training_pair(atf: np.ndarray, u: np.ndarray, v: np.ndarray):
    """(x, t) of one solved slice: atf = A^*f and u (H, W) complex, v (2, H, W)
    complex; returns x (H, W, 2) and t (H, W, 6) float32 on the scale
    s = RMS(A^*f)."""
    s = float(np.sqrt(np.mean(np.abs(atf.astype(np.complex128)) ** 2))) or 1.0
    x = np.stack([atf.real, atf.imag], -1) / s
    r = (u - atf) / s
    t = np.stack([r.real, r.imag, v[0].real / s, v[0].imag / s, v[1].real / s, v[1].imag / s], -1)
    return x.astype(np.float32), t.astype(np.float32)
