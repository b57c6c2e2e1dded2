"""Writes the golden fixtures of the test suite (run once; outputs committed).

* fft.npz        -- random complex vectors and their numpy FFT / inverse FFT
                    (lengths 1..736, including the radix-23 lengths 368, 736),
                    and centered orthonormal 2D transforms of small planes.
* operator.npz   -- the three Cartesian operators at the reference adjointness
                    size {10, 10, 5} (and an odd-size Cartesian case): inputs
                    and numpy-computed forward/adjoint outputs.  numpy, not
                    the oracle, computes the expected values.
* masks.npz      -- the reference's own sampling-mask fixtures
                    (data/mask_cartesian_random4.npy, regular4.npy), stored as
                    their shape and sampled PE-line lists (the masks are
                    constant along FE).  Needs the reference tree mounted at
                    /root/reference; the outputs are data only.

Usage: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DATA = "/root/reference/data"


def centered(z, inverse=False):
    ax = (-2, -1)
    f = np.fft.ifft2 if inverse else np.fft.fft2
    return np.fft.fftshift(f(np.fft.ifftshift(z, axes=ax), axes=ax, norm="ortho"), axes=ax)


def roos_offset(hk):
    return int(np.float32(hk) * np.float32(0.25) + np.float32(1.0))


def fwd(x, S, M, kind):
    C, Hk, W = S.shape
    if kind == "samp":
        return M * centered(x)
    pad = np.zeros((Hk, W), np.complex128)
    if kind == "roos":
        y0 = roos_offset(Hk)
        pad[y0:y0 + Hk // 2] = x
    else:
        pad[:] = x
    return M[None] * centered(S * pad[None])


def adj(y, S, M, kind):
    if kind == "samp":
        return centered(M * y, inverse=True)
    C, Hk, W = S.shape
    img = np.sum(np.conj(S) * centered(M[None] * y, inverse=True), axis=0)
    if kind == "roos":
        y0 = roos_offset(Hk)
        return img[y0:y0 + Hk // 2]
    return img


def main():
    rng = np.random.default_rng(20180101)
    out = {}
    for n in [1, 2, 3, 4, 5, 7, 8, 10, 12, 16, 23, 46, 64, 320, 368, 384, 640, 736, 768]:
        a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"x{n}"] = a
        out[f"F{n}"] = np.fft.fft(a)
        out[f"I{n}"] = np.fft.ifft(a) * n
    for (h, w) in [(10, 7), (12, 10), (5, 6), (9, 9)]:
        z = rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))
        out[f"z{h}x{w}"] = z
        out[f"Z{h}x{w}"] = centered(z)
        out[f"IZ{h}x{w}"] = centered(z, inverse=True)
    np.savez_compressed(os.path.join(HERE, "fft.npz"), **out)

    ops = {}
    # the reference adjointness test sizes (tests/main_adjointness_test.cu:31-32):
    # S in U[0,1)^2, binary mask
    for kind, (W, Hk, C) in [("cart", (10, 10, 5)), ("roos", (10, 10, 5)), ("samp", (10, 10, 1)),
                             ("cart", (7, 9, 3)), ("roos", (6, 14, 2))]:
        key = f"{kind}_{W}_{Hk}_{C}"
        S = rng.random((C, Hk, W)) + 1j * rng.random((C, Hk, W))
        M = rng.integers(0, 2, (Hk, W)).astype(np.float64)
        H = Hk // 2 if kind == "roos" else Hk
        x = rng.random((H, W)) + 1j * rng.random((H, W))
        y = rng.random((C, Hk, W)) + 1j * rng.random((C, Hk, W))
        if kind == "samp":
            y = y[0]
        ops[key + "_S"] = S
        ops[key + "_M"] = M
        ops[key + "_x"] = x
        ops[key + "_y"] = y
        ops[key + "_Ax"] = fwd(x, S, M, kind)
        ops[key + "_Aty"] = adj(y, S, M, kind)
    np.savez_compressed(os.path.join(HERE, "operator.npz"), **ops)

    masks = {}
    for name in ["random4", "regular4"]:
        path = os.path.join(REF_DATA, f"mask_cartesian_{name}.npy")
        m = np.load(path, allow_pickle=False)
        assert np.all(m == m[0:1]), "mask is expected constant along FE"
        masks[f"{name}_shape"] = np.array(m.shape)
        masks[f"{name}_lines"] = np.flatnonzero(m[0]).astype(np.int32)
        masks[f"{name}_values"] = np.unique(m)
    np.savez_compressed(os.path.join(HERE, "masks.npz"), **masks)
    print("wrote fft.npz, operator.npz, masks.npz")


if __name__ == "__main__":
    main()
