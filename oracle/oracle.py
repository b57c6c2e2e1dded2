"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is an FP64 C++ restatement of the reference's TV/TGV primal-dual
solvers and Cartesian MRI operators (``pdt_oracle.hpp``), plus the warm-start
network and Adam (``pdt_cnn_oracle.cpp``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
module, always as the checker or the reported CPU baseline, never as part of
the product path.

Arrays follow the reference layout: numpy ``(H, W)`` images and
``(C, Hk, W)`` k-space in C order (iu size ``{W, H[, C]}``).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libpdt_oracle.so")

TV, TGV, TGV_MRI, TV_OP, TGV_OP = 0, 1, 2, 3, 4
OP_CARTESIAN, OP_CARTESIAN_ROOS, OP_SAMPLING = 0, 1, 2


class _OpDesc(ct.Structure):
    _fields_ = [("kind", ct.c_int), ("W", ct.c_int), ("Hk", ct.c_int), ("C", ct.c_int),
                ("S", ct.c_void_p), ("M", ct.c_void_p)]


class _Params(ct.Structure):
    _fields_ = [("lambda_", ct.c_double), ("alpha0", ct.c_double), ("alpha1", ct.c_double),
                ("reduction", ct.c_double), ("max_iter", ct.c_uint), ("check", ct.c_uint),
                ("gap_tol", ct.c_double), ("gap_every", ct.c_uint),
                ("radius_u", ct.c_double), ("radius_v", ct.c_double)]


class _Stats(ct.Structure):
    _fields_ = [("iters", ct.c_uint), ("n_gap", ct.c_uint), ("primal", ct.c_double),
                ("dual", ct.c_double), ("gap", ct.c_double), ("rel_gap", ct.c_double),
                ("seconds", ct.c_double)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ct.CDLL(LIB_PATH)
        L.pdto_solve.restype = ct.c_int
        L.pdto_solve.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.POINTER(_OpDesc), ct.c_int, ct.c_int,
                                 ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.POINTER(_Params),
                                 ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_uint, ct.POINTER(_Stats)]
        L.pdto_op_forward.argtypes = [ct.c_int, ct.POINTER(_OpDesc), ct.c_void_p, ct.c_void_p]
        L.pdto_op_adjoint.argtypes = [ct.c_int, ct.POINTER(_OpDesc), ct.c_void_p, ct.c_void_p]
        L.pdto_fft1.argtypes = [ct.c_int, ct.c_int, ct.c_void_p]
        L.pdto_fft2c.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p]
        L.pdto_stencil.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p]
        L.pdto_set_threads.argtypes = [ct.c_int]
        L.pdto_max_threads.restype = ct.c_int
        ip = ct.POINTER(ct.c_int)
        L.pdto_cnn_num_params.argtypes = [ct.c_int, ip]
        L.pdto_cnn_forward.argtypes = [ct.c_int, ip, ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                       ct.c_void_p, ct.c_void_p]
        L.pdto_cnn_loss_grad.argtypes = [ct.c_int, ip, ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                         ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_double), ct.c_void_p]
        L.pdto_adam.argtypes = [ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_double,
                                ct.c_double, ct.c_double, ct.c_double, ct.c_int]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().pdto_set_threads(int(n))


def max_threads() -> int:
    return int(lib().pdto_max_threads())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ct.c_void_p)


def _dt(prec: int, cplx: bool):
    if prec == 64:
        return np.complex128 if cplx else np.float64
    return np.complex64 if cplx else np.float32


class Op:
    """Cartesian operator description: sens (C, Hk, W) complex, mask (Hk, W)."""

    def __init__(self, kind: int, sens, mask, prec: int = 64):
        self.kind = kind
        self.prec = prec
        self.mask = np.ascontiguousarray(mask, dtype=np.float64 if prec == 64 else np.float32)
        Hk, W = self.mask.shape
        if kind == OP_SAMPLING:
            self.sens = None
            C = 1
        else:
            self.sens = np.ascontiguousarray(sens, dtype=_dt(prec, True))
            C = self.sens.shape[0]
            assert self.sens.shape == (C, Hk, W)
        self.W, self.Hk, self.C = W, Hk, C
        self.desc = _OpDesc(kind, W, Hk, C, _ptr(self.sens), _ptr(self.mask))

    @property
    def H(self) -> int:
        return self.Hk // 2 if self.kind == OP_CARTESIAN_ROOS else self.Hk

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=_dt(self.prec, True))
        assert x.shape == (self.H, self.W)
        y = np.empty((self.C, self.Hk, self.W), _dt(self.prec, True))
        lib().pdto_op_forward(self.prec, ct.byref(self.desc), _ptr(x), _ptr(y))
        return y if self.kind != OP_SAMPLING else y[0]

    def adjoint(self, y):
        y = np.ascontiguousarray(y, dtype=_dt(self.prec, True)).reshape(self.C, self.Hk, self.W)
        x = np.empty((self.H, self.W), _dt(self.prec, True))
        lib().pdto_op_adjoint(self.prec, ct.byref(self.desc), _ptr(y), _ptr(x))
        return x


def solve(kind: int, f, *, op: Op | None = None, lambda_: float = 1.0, alpha0: float = 1.0,
          alpha1: float = 1.0, reduction: float = 1.0, max_iter: int = 1, gap_tol: float = 0.0,
          gap_every: int = 0, radius_u: float = 0.0, radius_v: float = 0.0, u0=None, v0=None,
          prec: int = 64, want_v: bool = False):
    """Run one oracle solve; returns (u, v or None, stats dict, gap history)."""
    cplx = np.iscomplexobj(f) or op is not None
    dt = _dt(prec, cplx)
    f = np.ascontiguousarray(f, dtype=dt)
    if op is None:
        H, W = f.shape
    else:
        assert op.prec == prec
        H, W = op.H, op.W
    u0a = None if u0 is None else np.ascontiguousarray(u0, dtype=dt)
    v0a = None if v0 is None else np.ascontiguousarray(v0, dtype=dt)
    u = np.empty((H, W), dt)
    v = np.empty((2, H, W), dt) if want_v else None
    p = _Params(lambda_, alpha0, alpha1, reduction, max_iter, 100, gap_tol, gap_every,
                radius_u, radius_v)
    cap = (max_iter + 1) // gap_every + 1 if gap_every else 0
    hist = np.zeros((max(cap, 1), 6), np.float64)
    st = _Stats()
    rc = lib().pdto_solve(prec, kind, int(cplx), ct.byref(op.desc) if op else None, W, H,
                          _ptr(f), _ptr(u0a), _ptr(v0a), ct.byref(p), _ptr(u), _ptr(v),
                          _ptr(hist), cap, ct.byref(st))
    if rc != 0:
        raise RuntimeError("oracle solve failed")
    stats = {k: getattr(st, k) for k, _ in _Stats._fields_}
    return u, v, stats, hist[:min(st.n_gap, cap)]


def fft1(a, sign: int = -1):
    a = np.array(a, dtype=np.complex128, copy=True)
    lib().pdto_fft1(a.size, sign, _ptr(a))
    return a


def fft2c(z, sign: int = -1):
    z = np.ascontiguousarray(z, dtype=np.complex128)
    H, W = z.shape
    out = np.empty_like(z)
    lib().pdto_fft2c(W, H, sign, _ptr(z), _ptr(out))
    return out


def stencil(which: int, a):
    """0: grad (H,W)->(2,H,W); 1: div (2,H,W)->(H,W); 2: symgrad (2,H,W)->(3,H,W);
    3: symdiv (3,H,W)->(2,H,W).  FP64."""
    cplx = np.iscomplexobj(a)
    a = np.ascontiguousarray(a, dtype=np.complex128 if cplx else np.float64)
    H, W = a.shape[-2:]
    nout = {0: 2, 1: 1, 2: 3, 3: 2}[which]
    out = np.empty((nout, H, W), a.dtype)
    lib().pdto_stencil(which, int(cplx), W, H, _ptr(a), _ptr(out))
    return out[0] if nout == 1 else out


# --------------------------------------------------------------------------
# warm-start network (pdt_cnn_oracle.cpp)
# --------------------------------------------------------------------------
def _chans(ch):
    ch = [int(c) for c in ch]
    return len(ch) - 1, (ct.c_int * len(ch))(*ch)


def cnn_num_params(ch) -> int:
    nl, c = _chans(ch)
    return int(lib().pdto_cnn_num_params(nl, c))


def cnn_forward(ch, params, x, bf16: bool = False):
    """x: (n, h, w, ch[0]) -> (n, h, w, ch[-1]), FP64 (bf16 rounding emulated if asked)."""
    nl, c = _chans(ch)
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    n, h, w, _ = x.shape
    y = np.empty((n, h, w, ch[-1]), np.float64)
    lib().pdto_cnn_forward(nl, c, _ptr(p), n, h, w, int(bf16), _ptr(x), _ptr(y))
    return y


def cnn_loss_grad(ch, params, x, t, bf16: bool = False):
    """Loss 1/(2m) sum (net(x) - t)^2 and its gradient (parameter order of the network)."""
    nl, c = _chans(ch)
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    t = np.ascontiguousarray(t, np.float64)
    n, h, w, _ = x.shape
    g = np.empty_like(p)
    loss = ct.c_double()
    lib().pdto_cnn_loss_grad(nl, c, _ptr(p), n, h, w, int(bf16), _ptr(x), _ptr(t), ct.byref(loss), _ptr(g))
    return loss.value, g


def adam(w, g, m, v, lr, b1, b2, eps, step):
    """One Adam step in place on FP64 arrays (step counted from 1)."""
    for a in (w, g, m, v):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    lib().pdto_adam(w.size, _ptr(w), _ptr(g), _ptr(m), _ptr(v), lr, b1, b2, eps, int(step))
