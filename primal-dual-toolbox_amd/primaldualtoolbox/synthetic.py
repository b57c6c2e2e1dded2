"""Seeded synthetic inputs for the TV/TGV solvers (phantoms, coil maps, masks).

The reference ships no raw k-space for its demos (the knee raw data and
sensitivities are missing blobs), so every workload here is generated from a
seed.  Arrays use numpy's C order with the reference's axis convention:
image ``(FE, PE)``, k-space ``(coils, FE, PE)``, mask ``(FE, PE)``, i.e. the
iu size ``{PE, FE[, coils]}`` with PE (dim 0) contiguous
(``bin/mri_cartesian_test.py:29-53``).
"""
from __future__ import annotations

import numpy as np


def ellipse_phantom(rng: np.random.Generator, n: int, m: int | None = None,
                    count: int = 20) -> np.ndarray:
    """Sum of ``count`` random ellipses on an ``(m, n)`` grid, clipped to [0, 1].

    Centres U[0, N), semi-axes U[0.05, 0.4] * N, angle U[0, pi), intensity
    U[0, 1]; overlapping intensities add before the clip.
    """
    m = n if m is None else m
    yy, xx = np.mgrid[0:m, 0:n].astype(np.float64)
    img = np.zeros((m, n), np.float64)
    size = float(max(n, m))
    for _ in range(count):
        cx, cy = rng.uniform(0, n), rng.uniform(0, m)
        a, b = rng.uniform(0.05, 0.4, size=2) * size
        th = rng.uniform(0, np.pi)
        c, s = np.cos(th), np.sin(th)
        dx, dy = xx - cx, yy - cy
        u = (c * dx + s * dy) / a
        v = (-s * dx + c * dy) / b
        img[(u * u + v * v) <= 1.0] += rng.uniform(0, 1)
    return np.clip(img, 0.0, 1.0)


def tv_denoising_case(seed: int = 0, n: int = 256, sigma: float = 0.1):
    """Real float32 ROF case: ellipse phantom plus N(0, sigma^2) noise."""
    rng = np.random.default_rng(seed)
    clean = ellipse_phantom(rng, n)
    noisy = clean + sigma * rng.standard_normal(clean.shape)
    return clean.astype(np.float32), noisy.astype(np.float32)


def complex_phantom(rng: np.random.Generator, n: int, m: int | None = None) -> np.ndarray:
    """Complex phantom: ellipses plus a quadratic ramp U[0, 0.3] in magnitude,
    a random 2nd-order polynomial phase of amplitude pi."""
    m = n if m is None else m
    mag = ellipse_phantom(rng, n, m)
    yy, xx = np.mgrid[0:m, 0:n].astype(np.float64)
    xs = 2.0 * xx / max(n - 1, 1) - 1.0
    ys = 2.0 * yy / max(m - 1, 1) - 1.0
    ramp = (xs - rng.uniform(-1, 1)) ** 2 + (ys - rng.uniform(-1, 1)) ** 2
    ramp = ramp / max(ramp.max(), 1e-12) * rng.uniform(0.0, 0.3)
    coef = rng.uniform(-1, 1, size=6)
    phi = (coef[0] + coef[1] * xs + coef[2] * ys + coef[3] * xs * xs
           + coef[4] * xs * ys + coef[5] * ys * ys)
    phi = phi / max(np.abs(phi).max(), 1e-12) * np.pi
    return (mag + ramp) * np.exp(1j * phi)


def tgv_denoising_case(seed: int = 0, n: int = 512, sigma: float = 0.05):
    """Complex64 TGV case: complex phantom plus complex N(0, sigma^2) noise."""
    rng = np.random.default_rng(seed)
    clean = complex_phantom(rng, n)
    noise = sigma * (rng.standard_normal(clean.shape) + 1j * rng.standard_normal(clean.shape))
    return clean.astype(np.complex64), (clean + noise).astype(np.complex64)


def coil_sensitivities(rng: np.random.Generator, coils: int, n_fe: int, n_pe: int,
                       size: int | None = None) -> np.ndarray:
    """Gaussian magnitude bumps (width 0.4 N) centred on a circle of radius
    0.6 N, times a smooth random first-order phase, normalised so that
    sum_c |S_c|^2 = 1 at every pixel.  Shape ``(coils, n_fe, n_pe)``."""
    size = float(size if size is not None else min(n_fe, n_pe))
    ys = np.arange(n_fe, dtype=np.float64)
    xs = np.arange(n_pe, dtype=np.float64)
    cy, cx = (n_fe - 1) / 2.0, (n_pe - 1) / 2.0
    out = np.empty((coils, n_fe, n_pe), np.complex128)
    w = 0.4 * size
    for c in range(coils):
        ang = 2.0 * np.pi * c / coils + rng.uniform(-0.2, 0.2)
        px = cx + 0.6 * size * np.cos(ang)
        py = cy + 0.6 * size * np.sin(ang)
        k = rng.uniform(-1, 1, size=3)
        # bump and phase are separable in (y, x): an outer product of 1D factors
        fy = np.exp(-((ys - py) ** 2) / (2.0 * w * w) + 1j * (k[0] * np.pi + k[2] * np.pi * (ys - cy) / n_fe))
        fx = np.exp(-((xs - px) ** 2) / (2.0 * w * w) + 1j * (k[1] * np.pi * (xs - cx) / n_pe))
        out[c] = np.outer(fy, fx)
    norm2 = np.sum(out.real ** 2 + out.imag ** 2, axis=0)
    out *= (1.0 / np.sqrt(norm2))[None]
    return out


def line_mask(rng: np.random.Generator, n_fe: int, n_pe: int, accel: float = 4.0,
              center: int = 24, regular: bool = False) -> np.ndarray:
    """Cartesian PE-line mask ``(n_fe, n_pe)`` float32, constant along FE.

    ``round(n_pe / accel)`` lines in total, ``center`` of them the fully
    sampled centre; the rest random (or every ``accel``-th line if regular).
    """
    total = int(round(n_pe / accel))
    c0 = n_pe // 2 - center // 2
    lines = set(range(c0, c0 + center))
    if regular:
        lines |= set(range(0, n_pe, int(accel)))
    else:
        rest = [i for i in range(n_pe) if i not in lines]
        extra = max(total - len(lines), 0)
        lines |= set(rng.choice(rest, size=extra, replace=False).tolist())
    m = np.zeros((n_fe, n_pe), np.float32)
    m[:, sorted(lines)] = 1.0
    return m


def roos_offset(n_fe_k: int) -> int:
    """FE offset of the image inside the read-out oversampled grid:
    ``unsigned(n_fe_k * 0.25f + 1)`` (mricartesianoperator.cu:341)."""
    return int(np.float32(n_fe_k) * np.float32(0.25) + np.float32(1.0))


def centered_fft2(z: np.ndarray, inverse: bool = False) -> np.ndarray:
    """fftshift(FFT2(ifftshift(z))) orthonormal over the last two axes."""
    ax = (-2, -1)
    try:
        import scipy.fft as sfft
        f = sfft.ifft2 if inverse else sfft.fft2
        return sfft.fftshift(f(sfft.ifftshift(z, axes=ax), axes=ax, norm="ortho", workers=-1), axes=ax)
    except ImportError:
        f = np.fft.ifft2 if inverse else np.fft.fft2
        return np.fft.fftshift(f(np.fft.ifftshift(z, axes=ax), axes=ax, norm="ortho"), axes=ax)


def cartesian_forward_np(x: np.ndarray, sens: np.ndarray, mask: np.ndarray,
                         roos: bool) -> np.ndarray:
    """numpy restatement of the Cartesian forward operator (data generation)."""
    coils, n_fe_k, n_pe = sens.shape
    pad = np.zeros((n_fe_k, n_pe), np.complex128)
    if roos:
        y0 = roos_offset(n_fe_k)
        pad[y0:y0 + x.shape[0]] = x
    else:
        pad[:] = x
    return mask[None] * centered_fft2(sens * pad[None])


def mri_case(seed: int = 0, n: int = 320, coils: int = 15, n_pe: int | None = None,
             accel: float = 4.0, center: int = 24, regular: bool = False,
             snr_db: float = 30.0, roos: bool = True, lines=None):
    """Cartesian multi-coil case on a read-out oversampled grid.

    Returns a dict with ``image`` (n, n_pe) complex, ``sens`` (coils, 2n, n_pe)
    complex64, ``mask`` (2n, n_pe) float32 and ``kspace`` (coils, 2n, n_pe)
    complex64 = mask * (A x + noise) at the requested SNR
    (20 log10 ||A x|| / ||noise|| over sampled entries).  ``lines`` (PE-line
    indices) replaces the random mask, e.g. by one of the reference's masks.
    """
    n_pe = n if n_pe is None else n_pe
    rng = np.random.default_rng(seed)
    img = complex_phantom(rng, n_pe, n)
    n_fe_k = 2 * n if roos else n
    sens = coil_sensitivities(rng, coils, n_fe_k, n_pe, size=min(n, n_pe))
    if lines is None:
        mask = line_mask(rng, n_fe_k, n_pe, accel=accel, center=center, regular=regular)
    else:
        mask = np.zeros((n_fe_k, n_pe), np.float32)
        mask[:, np.asarray(lines, dtype=np.int64)] = 1.0
    y = cartesian_forward_np(img, sens, mask, roos)
    sampled = np.broadcast_to(mask[None] != 0, y.shape)
    noise = rng.standard_normal(y.shape) + 1j * rng.standard_normal(y.shape)
    noise *= sampled
    scale = np.linalg.norm(y) / (np.linalg.norm(noise) * 10.0 ** (snr_db / 20.0))
    f = y + scale * noise
    return {
        "image": img.astype(np.complex64),
        "sens": sens.astype(np.complex64),
        "mask": mask,
        "kspace": f.astype(np.complex64),
        "roos": roos,
    }


# Matrix sizes (FE x PE of the image) and ranges of the batched workload.
BATCH_SIZES = ((256, 256), (320, 320), (368, 368), (384, 320))


def batch_case(seed: int):
    """One slice of the batched workload: size from BATCH_SIZES, coils
    U{8..20}, random or regular 4x mask with 16..32 centre lines."""
    rng = np.random.default_rng(seed)
    n, n_pe = BATCH_SIZES[int(rng.integers(0, len(BATCH_SIZES)))]
    coils = int(rng.integers(8, 21))
    center = int(rng.integers(16, 33))
    regular = bool(rng.integers(0, 2))
    case = mri_case(seed=seed, n=n, coils=coils, n_pe=n_pe, center=center, regular=regular)
    case["meta"] = {"seed": seed, "n": n, "n_pe": n_pe, "coils": coils, "center": center, "regular": regular}
    return case


def process_pool(workers: int):
    """A process pool whose workers start fresh interpreters (spawn), so that
    no worker inherits a parent that has already opened the GPU."""
    import concurrent.futures as cf
    import multiprocessing as mp
    return cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn"))


def batch_cases(seeds, workers: int = 16):
    """Generates batch_case(seed) for each seed in a process pool."""
    seeds = list(seeds)
    if workers <= 1 or len(seeds) <= 1:
        return [batch_case(s) for s in seeds]
    with process_pool(min(workers, len(seeds))) as ex:
        return list(ex.map(batch_case, seeds))


def mri320_case(seed: int):
    """A slice like the 320 x 320, 15-coil benchmark workload with its own seed
    (4x random lines, 24 centre lines, 30 dB)."""
    return mri_case(seed=seed, n=320, coils=15, accel=4.0, center=24, snr_db=30.0, roos=True)
