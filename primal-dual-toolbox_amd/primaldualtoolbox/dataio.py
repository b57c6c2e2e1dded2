"""Reading and writing data in the reference's on-disk conventions.

The reference's demos load Cartesian k-space data from numpy ``.npy`` files
(bin/mri_cartesian_test.py:29-36) and from MATLAB ``.mat`` files
(bin/mri_cartesian_test.m:25-33):

* numpy files hold arrays in the layout the Python classes take: rawdata and
  coil sensitivities of shape (C, FE, PE) complex, the sampling mask and the
  reference image of shape (FE, PE);
* MATLAB files index the same arrays as (FE, PE, C) and (FE, PE)
  (``nFE = size(rawdata, 1)``, bin/mri_cartesian_test.m:41).  scipy reads a
  MATLAB array with MATLAB's index order, so a 2D array needs no change and a
  3D array moves its coil axis to the front.  The reference's two mask files
  (data/mask_cartesian_{random4,regular4}.{npy,mat}) hold identical masks
  under these rules.

Masks are Cartesian line masks: constant along FE, one value per PE line;
``mask_lines`` and ``line_mask`` convert between the mask and its sorted list
of sampled PE lines.

Nothing here unpickles: ``.npy``/``.npz`` are read with numpy's default
``allow_pickle=False`` and ``.mat`` through ``scipy.io.loadmat`` (MATLAB 5/7
files; 7.3 files are HDF5 and are refused with a message).
"""
from __future__ import annotations

import os

import numpy as np

# Variable names the reference's MATLAB demo expects in its .mat files
# (bin/mri_cartesian_test.m:25-33 loads them by name).
MAT_KEYS = {"rawdata": "rawdata", "sensitivities": "sensitivities", "mask": "mask", "reference": "reference"}


def _mat_vars(path: str) -> dict:
    import scipy.io as sio
    try:
        d = sio.loadmat(path)
    except NotImplementedError as e:  # scipy's message for HDF5-based 7.3 files
        raise ValueError(f"{path}: MATLAB 7.3 (HDF5) files are not supported; save with -v7 ({e})") from None
    return {k: v for k, v in d.items() if not k.startswith("__")}


def load(path: str, key: str | None = None) -> np.ndarray:
    """Loads one array from ``.npy``, ``.npz`` or ``.mat``.

    For ``.npz`` and ``.mat`` files ``key`` names the variable; without it
    the file must hold exactly one.  ``.mat`` arrays are returned as MATLAB
    indexes them (see ``from_matlab`` for the conversion to the numpy layout).
    """
    ext = os.path.splitext(path)[1].lower()
    if ext == ".npy":
        if key is not None:
            raise ValueError(f"{path}: a .npy file holds one array; no key expected")
        return np.load(path)
    if ext == ".npz":
        with np.load(path) as z:
            names = list(z.keys())
            return _pick(path, {k: z[k] for k in names}, key)
    if ext == ".mat":
        return _pick(path, _mat_vars(path), key)
    raise ValueError(f"{path}: unknown file type {ext!r} (expected .npy, .npz or .mat)")


def _pick(path: str, vars_: dict, key: str | None) -> np.ndarray:
    if key is None:
        if len(vars_) != 1:
            raise ValueError(f"{path}: holds {sorted(vars_)}; name one with key=")
        return next(iter(vars_.values()))
    if key not in vars_:
        raise KeyError(f"{path}: no variable {key!r} (has {sorted(vars_)})")
    return vars_[key]


def from_matlab(a: np.ndarray) -> np.ndarray:
    """MATLAB indexing (FE, PE[, C]) -> numpy layout ((C,) FE, PE), C-contiguous."""
    a = np.asarray(a)
    if a.ndim == 2:
        return np.ascontiguousarray(a)
    if a.ndim == 3:
        return np.ascontiguousarray(np.moveaxis(a, 2, 0))
    raise ValueError(f"expected a 2D or 3D array, got shape {a.shape}")


def to_matlab(a: np.ndarray) -> np.ndarray:
    """numpy layout ((C,) FE, PE) -> MATLAB indexing (FE, PE[, C])."""
    a = np.asarray(a)
    if a.ndim == 2:
        return a
    if a.ndim == 3:
        return np.moveaxis(a, 0, 2)
    raise ValueError(f"expected a 2D or 3D array, got shape {a.shape}")


def save_mat(path: str, **arrays) -> None:
    """Writes numpy-layout arrays to a MATLAB 5 file in MATLAB indexing
    (3D arrays get their coil axis last), readable by ``load``/``from_matlab``
    and by MATLAB's ``load``."""
    import scipy.io as sio
    sio.savemat(path, {k: to_matlab(v) for k, v in arrays.items()}, do_compression=True)


def mask_lines(mask: np.ndarray) -> np.ndarray:
    """Sorted PE-line indices sampled by a Cartesian mask of shape (FE, PE).

    Raises ValueError unless every PE column is constant along FE and the
    values are 0 and 1 (the reference's masks, data/mask_cartesian_*.npy)."""
    m = np.asarray(mask)
    if m.ndim != 2:
        raise ValueError(f"a line mask is 2D (FE, PE), got shape {m.shape}")
    if np.iscomplexobj(m):
        raise ValueError("a line mask is real")
    if not np.all((m == 0) | (m == 1)):
        raise ValueError("a line mask holds only 0 and 1")
    if not np.all(m == m[:1]):
        raise ValueError("not a Cartesian line mask: values vary along FE")
    return np.nonzero(m[0])[0].astype(np.int64)


def line_mask(lines, n_fe: int, n_pe: int, dtype=np.float32) -> np.ndarray:
    """The (n_fe, n_pe) mask sampling the PE lines ``lines``."""
    idx = np.asarray(lines, dtype=np.int64)
    if idx.size and (idx.min() < 0 or idx.max() >= n_pe):
        raise ValueError(f"line index out of range [0, {n_pe})")
    m = np.zeros((n_fe, n_pe), dtype)
    m[:, idx] = 1
    return m


def roos_window(n_fe: int) -> slice:
    """FE rows of the image inside a read-out oversampled grid of n_fe rows:
    ``[int(n_fe * 0.25 + 1), int(n_fe * 0.75 + 1))``, the crop the reference
    applies to its reference image (bin/mri_cartesian_test.py:45) and the
    window of ``MriCartesianRemoveROOSOperator`` (FE_idx,
    mricartesianoperator.cu:341)."""
    return slice(int(np.float32(n_fe) * np.float32(0.25) + np.float32(1.0)),
                 int(np.float32(n_fe) * np.float32(0.75) + np.float32(1.0)))


def crop_roos(image: np.ndarray) -> np.ndarray:
    """The FE window of an image on the oversampled grid (axis -2 is FE)."""
    return np.asarray(image)[..., roos_window(np.shape(image)[-2]), :]


def load_cartesian(rawdata, sensitivities, mask, reference=None) -> dict:
    """Loads a Cartesian case in the numpy layout.

    Each argument is an array (numpy layout) or a path; ``.mat`` files are
    read by the variable names of the reference's MATLAB demo and converted
    from MATLAB indexing.  Returns ``{"kspace", "sens", "mask", "lines"[,
    "reference"]}`` with kspace/sens complex64 (C, FE, PE) and mask float32
    (FE, PE); raises ValueError when the shapes disagree."""
    def get(x, name):
        if isinstance(x, (str, os.PathLike)):
            p = os.fspath(x)
            if p.lower().endswith(".mat"):
                return from_matlab(load(p, MAT_KEYS[name]))
            return load(p)
        return np.asarray(x)

    f = get(rawdata, "rawdata").astype(np.complex64)
    s = get(sensitivities, "sensitivities").astype(np.complex64)
    m = get(mask, "mask").astype(np.float32)
    if f.ndim != 3 or s.shape != f.shape:
        raise ValueError(f"rawdata {f.shape} and sensitivities {s.shape} must both be (C, FE, PE)")
    if m.shape != f.shape[1:]:
        raise ValueError(f"mask {m.shape} must be (FE, PE) = {f.shape[1:]}")
    out = {"kspace": f, "sens": s, "mask": m, "lines": mask_lines(m)}
    if reference is not None:
        out["reference"] = get(reference, "reference")
    return out
