"""An independent numpy model of the reference's TGV-family solve loops
(TEST INFRASTRUCTURE: the checker of the oracle's loops, never the product).

Everything here is built from definitions, not from the oracle's code:

* the finite differences as dense matrices: forward differences with a zero
  last row/column (tgv_kernels.cuh:40-50), ``div = -grad^T`` and
  ``div_sym = -E^T W`` with W = diag(1, 1, 2) (the xy plane counted twice,
  tgv_kernels.cuh:25-38,105-145), which equal the reference's backward
  differences for sizes >= 2;
* the Cartesian operator as a dense matrix whose columns are numpy's centred
  orthonormal FFT of S_c times the zero-padded unit image
  (mricartesianoperator.cu:196-265,320-414; the read-out window starts at
  FE row unsigned(Hk * 0.25f + 1), :341), and its adjoint as the conjugate
  transpose;
* the loops with the step sizes, order and alpha continuation read from
  tgvoptimizer.cu:102-135, tgvmrioptimizer.cu:127-198 and
  tgvoptimizerwithop.cu:123-176;
* the primal energy and the partial primal-dual gap from their definitions
  (Chambolle-Pock 2011, Sec. 2): Fenchel-Young gaps plus ball radius times
  the stationarity residuals.

Arrays are numpy images (H, W), x fastest; stacked planes (k, H, W).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


# ---------------------------------------------------------------------------
# stencils as matrices
# ---------------------------------------------------------------------------
def _fwd_1d(n: int):
    """(n, n) forward difference, last entry zero (sparse)."""
    D = sp.lil_matrix((n, n))
    for i in range(n - 1):
        D[i, i] = -1.0
        D[i, i + 1] = 1.0
    return D.tocsr()


def diff_matrices(H: int, W: int):
    """Dx, Dy acting on the flattened (H, W) image (index y * W + x)."""
    Dx = sp.kron(sp.identity(H), _fwd_1d(W), format="csr")
    Dy = sp.kron(_fwd_1d(H), sp.identity(W), format="csr")
    return Dx, Dy


class Stencils:
    def __init__(self, H: int, W: int):
        self.H, self.W, self.N = H, W, H * W
        Dx, Dy = diff_matrices(H, W)
        self.G = sp.vstack([Dx, Dy], format="csr")                                  # (2N, N)
        self.E = sp.bmat([[Dx, None], [None, Dy], [0.5 * Dy, 0.5 * Dx]], format="csr")  # (3N, 2N): xx, yy, xy
        self.Wq = np.concatenate([np.ones(2 * self.N), 2 * np.ones(self.N)])
        self.Div = (-self.G.T).tocsr()                                               # (N, 2N)
        self.DivSym = (-(self.E.T @ sp.diags(self.Wq))).tocsr()                      # (2N, 3N)

    def grad(self, u):
        return (self.G @ u.reshape(-1)).reshape(2, self.H, self.W)

    def div(self, p):
        return (self.Div @ p.reshape(-1)).reshape(self.H, self.W)

    def symgrad(self, v):
        return (self.E @ v.reshape(-1)).reshape(3, self.H, self.W)

    def symdiv(self, q):
        return (self.DivSym @ q.reshape(-1)).reshape(2, self.H, self.W)


def proj_p(p, alpha):
    """p / max(1, |p| / alpha), |p| = sqrt(|px|^2 + |py|^2) per pixel."""
    n = np.sqrt(np.abs(p[0]) ** 2 + np.abs(p[1]) ** 2)
    return p / np.maximum(1.0, n / alpha)[None]


def proj_q(q, alpha):
    """Frobenius ball with the xy entry counted twice (norm_sym)."""
    n = np.sqrt(np.abs(q[0]) ** 2 + np.abs(q[1]) ** 2 + 2 * np.abs(q[2]) ** 2)
    return q / np.maximum(1.0, n / alpha)[None]


def norm_p(p):
    return np.sqrt(np.abs(p[0]) ** 2 + np.abs(p[1]) ** 2)


def norm_q(q):
    return np.sqrt(np.abs(q[0]) ** 2 + np.abs(q[1]) ** 2 + 2 * np.abs(q[2]) ** 2)


def redot(a, b) -> float:
    return float(np.sum(np.real(np.conj(a) * b)))


# ---------------------------------------------------------------------------
# Cartesian operator as a dense matrix
# ---------------------------------------------------------------------------
def roos_offset(Hk: int) -> int:
    return int(np.float32(Hk) * np.float32(0.25) + np.float32(1))


def cartesian_ops(sens, mask, roos: bool):
    """(forward, adjoint) of the Cartesian operator as numpy functions on
    flattened vectors: forward x -> M * fftshift(fft2(ifftshift(S_c pad(x)),
    ortho)) per coil; adjoint its conjugate transpose, written out (mask,
    inverse centred transform, conj(S_c), coil sum, crop)."""
    sens = np.asarray(sens, complex)
    mask = np.asarray(mask, float)
    C, Hk, W = sens.shape
    H = Hk // 2 if roos else Hk
    y0 = roos_offset(Hk) if roos else 0

    def fwd(x):
        pad = np.zeros((Hk, W), complex)
        pad[y0:y0 + H] = x.reshape(H, W)
        out = mask[None] * np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(sens * pad[None], axes=(1, 2)),
                                                        norm="ortho"), axes=(1, 2))
        return out.reshape(-1)

    def adj(y):
        z = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(mask[None] * y.reshape(C, Hk, W), axes=(1, 2)),
                                         norm="ortho"), axes=(1, 2))
        return np.sum(np.conj(sens) * z, axis=0)[y0:y0 + H].reshape(-1)

    return fwd, adj


def cartesian_matrix(sens, mask, roos: bool) -> np.ndarray:
    """Dense A (image (H, W) -> k-space (C, Hk, W)): the forward of
    cartesian_ops applied to every unit image."""
    C, Hk, W = np.shape(sens)
    H = Hk // 2 if roos else Hk
    fwd, _ = cartesian_ops(sens, mask, roos)
    return np.array([fwd(np.eye(1, H * W, k).reshape(-1).astype(complex)) for k in range(H * W)]).T


# ---------------------------------------------------------------------------
# loops
# ---------------------------------------------------------------------------
def tgv_denoise(f, lam, alpha0, alpha1, iters, *, tau=1 / np.sqrt(12), sigma=1 / np.sqrt(12),
                primal_first=True):
    """TgvOptimizer::solve: u = ubar = f, v = p = q = 0; per iteration the u
    step (prox of lam/2 |u - f|^2), the v step, then p and q."""
    H, W = f.shape
    S = Stencils(H, W)
    u = f.copy()
    ub = f.copy()
    v = np.zeros((2, H, W), f.dtype)
    vb = v.copy()
    p = np.zeros((2, H, W), f.dtype)
    q = np.zeros((3, H, W), f.dtype)

    def primal():
        nonlocal u, ub, v, vb
        un = (u + tau * (S.div(p) + lam * f)) / (1 + tau * lam)
        ub = 2 * un - u
        u = un
        vn = v + tau * (p + S.symdiv(q))
        vb = 2 * vn - v
        v = vn

    def dual():
        nonlocal p, q
        p = proj_p(p + sigma * (S.grad(ub) - vb), alpha1)
        q = proj_q(q + sigma * S.symgrad(vb), alpha0)

    for _ in range(iters):
        if primal_first:
            primal()
            dual()
        else:
            dual()
            primal()
    return dict(u=u, v=v, ub=ub, vb=vb, p=p, q=q)


class OpSolve:
    """TgvMriOptimizer (kind "mri": tau 1/16, sigma 1/8, lambda 1, alpha
    continuation) and TgvOptimizerWithOp (kind "op": tau = sigma = 1/8, fixed
    alpha, r prox with lambda).  ``alpha_at(k, K)`` gives the continuation
    factor red^(k/K) by default."""

    def __init__(self, A, f, H, W, *, kind="mri", lam=1.0, alpha0=1.0, alpha1=1.0, reduction=1.0,
                 tau=None, sigma=None, alpha_exponent=None, lam_in_prox=True):
        """A: a dense matrix, or a (forward, adjoint) pair of functions."""
        if isinstance(A, tuple):
            self.fwd, self.adj = A
        else:
            AH = A.conj().T
            self.fwd, self.adj = (lambda x: A @ x), (lambda y: AH @ y)
        self.f = f.reshape(-1).astype(complex)
        self.H, self.W = H, W
        self.S = Stencils(H, W)
        self.kind = kind
        mri = kind == "mri"
        self.tau = tau if tau is not None else (1 / 16 if mri else 1 / 8)
        self.sigma = sigma if sigma is not None else 1 / 8
        self.lam = 1.0 if mri else lam
        self.lam_in_prox = lam_in_prox
        self.a0, self.a1, self.red = alpha0, alpha1, reduction
        self.alpha_exponent = alpha_exponent or (lambda k, K: k / K)
        u0 = self.adj(self.f).reshape(H, W)
        self.reset(u0)

    def reset(self, u0, v0=None):
        H, W = self.H, self.W
        self.u = u0.astype(complex).copy()
        self.ub = self.u.copy()
        self.v = np.zeros((2, H, W), complex) if v0 is None else v0.astype(complex).copy()
        self.vb = self.v.copy()
        self.p = np.zeros((2, H, W), complex)
        self.q = np.zeros((3, H, W), complex)
        self.r = np.zeros_like(self.f)
        self.k = 0

    def alphas(self, k, K):
        if self.kind != "mri":
            return self.a0, self.a1
        e = self.alpha_exponent(k, K)
        return self.a0 * self.red ** e, self.a1 * self.red ** e

    def step(self, K):
        """Iteration k = self.k of a solve with max_iter = K; returns the
        quantities at which the gap of this iteration is evaluated."""
        S, s, t = self.S, self.sigma, self.tau
        a0, a1 = self.alphas(self.k, K)
        ub, vb = self.ub, self.vb
        res = self.fwd(ub.reshape(-1)) - self.f
        lp = self.lam if self.lam_in_prox else 1.0
        self.r = (self.r + s * res) / (1 + s * lp)
        self.p = proj_p(self.p + s * (S.grad(ub) - vb), a1)
        self.q = proj_q(self.q + s * S.symgrad(vb), a0)
        w = self.adj(self.r).reshape(self.H, self.W)
        un = self.u + t * (S.div(self.p) - w)
        self.ub = 2 * un - self.u
        self.u = un
        vn = self.v + t * (self.p + S.symdiv(self.q))
        self.vb = 2 * vn - self.v
        self.v = vn
        self.k += 1
        return dict(ub=ub, vb=vb, r=self.r, p=self.p, q=self.q, w=w, a0=a0, a1=a1)

    # ---- energies ----
    def primal(self, u, v, a0, a1) -> float:
        """1/(2 lambda) |A u - f|^2 + a1 sum|grad u - v| + a0 sum|E v|_F."""
        S = self.S
        d = self.fwd(u.reshape(-1)) - self.f
        return (float(np.sum(np.abs(d) ** 2)) / (2 * self.lam) + a1 * float(np.sum(norm_p(S.grad(u) - v)))
                + a0 * float(np.sum(norm_q(S.symgrad(v)))))

    def partial_gap(self, st, rho_u, rho_v) -> tuple[float, float]:
        """(P, G) at (ubar, vbar; r, p, q) of one iteration: G = Fenchel-Young
        gaps of the data, first- and second-order terms plus rho_u |A^* r -
        div p| + rho_v |p + div_sym q|."""
        S, lam = self.S, self.lam
        ub, vb, r, p, q = st["ub"], st["vb"], st["r"], st["p"], st["q"]
        a0, a1 = st["a0"], st["a1"]
        d = self.fwd(ub.reshape(-1)) - self.f
        fy_data = float(np.sum(np.abs(d - lam * r) ** 2)) / (2 * lam)
        g = S.grad(ub) - vb
        e = S.symgrad(vb)
        fy_1 = a1 * float(np.sum(norm_p(g))) - redot(g, p)
        fy_0 = a0 * float(np.sum(norm_q(e))) - (redot(e[0], q[0]) + redot(e[1], q[1]) + 2 * redot(e[2], q[2]))
        res_u = np.linalg.norm(st["w"] - S.div(p))
        res_v = np.linalg.norm(p + S.symdiv(q))
        P = self.primal(ub, vb, a0, a1)
        return P, fy_data + fy_1 + fy_0 + rho_u * res_u + rho_v * res_v


def tgv_denoise_gap(f, u, v, p, q, lam, alpha0, alpha1) -> tuple[float, float]:
    """TGV-L2 denoising: primal energy at (u, v) and the dual value of the
    feasible pair p~ = -s div_sym q, s q with s in [0, min(1, alpha1 /
    max|div_sym q|)] chosen to maximise D(s) = s Re<div d, f> - s^2 |div d|^2
    / (2 lam), d = div_sym q."""
    S = Stencils(*f.shape)
    P = (alpha1 * float(np.sum(norm_p(S.grad(u) - v))) + alpha0 * float(np.sum(norm_q(S.symgrad(v))))
         + 0.5 * lam * float(np.sum(np.abs(u - f) ** 2)))
    d = S.symdiv(q)
    m = float(np.max(norm_p(d)))
    g = S.div(d)
    a, b = redot(g, f), float(np.sum(np.abs(g) ** 2))
    smax = min(1.0, alpha1 / m) if m > 0 else 1.0
    s = min(max(lam * a / b, 0.0), smax) if b > 0 else smax
    return P, s * a - s * s * b / (2 * lam)
