"""The reference's MATLAB entry points (matlab/CMakeLists.txt:14-30) as
functions with the same names, argument order, checks and messages.

Arrays are taken and returned in MATLAB indexing, as the mex functions see
them: sensitivities and rawdata (FE, PE, C), images and masks (FE, PE)
(bin/mri_cartesian_test.m:25-45); complex arrays are complex128 and the mask
float64, as the mex type checks require.  The operators compute in single
precision and the reconstructions in double precision, as the mex sources
do (``typedef float real_type`` in mexMriCartesianFwdOp.cpp:43,
``typedef double real_type`` in mexMriCartesianReconstructTgv.cpp:44);
results are returned as complex128.  A failed check raises ``MexError``
carrying the mex error identifier and message.
"""
from __future__ import annotations

import numpy as np

from . import _core as core
from .dataio import from_matlab, to_matlab


class MexError(RuntimeError):
    """An argument error of a MATLAB entry point (mexErrMsgIdAndTxt)."""

    def __init__(self, identifier: str, message: str):
        super().__init__(f"{identifier}: {message}" if identifier else message)
        self.identifier = identifier
        self.message = message


def _check_args(name: str, args, n: int, what: str) -> None:
    # mexMriCartesianFwdOp.cpp:54-76 (and the same block in the other five)
    if len(args) != n:
        raise MexError(f"MATLAB:{name}:invalidNumInputs", f"{what} inputs required ({_ARGS[n]})")
    for i, label in ((0, "sensitivities"), (1, "rawdata")):
        a = args[i]
        if not (isinstance(a, np.ndarray) and a.dtype == np.complex128):
            raise MexError(f"MATLAB:{name}:wrongDataType", f"Input {i + 1} ({label}) must be double complex!")
    m = args[2]
    if not (isinstance(m, np.ndarray) and m.dtype == np.float64):
        raise MexError(f"MATLAB:{name}:wrongDataType", "Input 3 (mask) must be double!")


_ARGS = {3: "sensitivities, rawdata, mask", 4: "sensitivities, rawdata, mask, tgv_params"}


def _operator(cls, sens, mask, dtype):
    if sens.ndim != 3 or mask.ndim != 2:
        raise MexError("", "sensitivities must be (FE, PE, C) and the mask (FE, PE)")
    op = cls(dtype=dtype)
    op.setCoilSens(from_matlab(sens).astype(dtype))  # constant [0]
    op.setMask(mask)                                 # constant [1]
    return op


def _op_call(name, cls, adjoint, args):
    _check_args(name, args, 3, "Three")
    sens, x, mask = args
    op = _operator(cls, sens, mask, np.complex64)
    if adjoint:
        return op.adjoint(from_matlab(x).astype(np.complex64)).astype(np.complex128)
    return to_matlab(op.forward(np.ascontiguousarray(x, np.complex64))).astype(np.complex128)


def gpuMriCartesianFwd(*args):
    """gpuMriCartesianFwd(sensitivities, image, mask) -> rawdata (FE, PE, C)
    (mexMriCartesianFwdOp.cpp)."""
    return _op_call("gpuMriCartesianFwd", core.MriCartesianOperator, False, args)


def gpuMriCartesianAdj(*args):
    """gpuMriCartesianAdj(sensitivities, rawdata, mask) -> image (FE, PE)
    (mexMriCartesianAdjOp.cpp)."""
    return _op_call("gpuMriCartesianAdj", core.MriCartesianOperator, True, args)


def gpuMriCartesianRemoveROOSFwd(*args):
    """gpuMriCartesianRemoveROOSFwd(sensitivities, image, mask) -> rawdata
    (FE, PE, C); the image has FE/2 rows (mexMriCartesianRemoveROOSFwdOp.cpp)."""
    return _op_call("gpuMriCartesianRemoveROOSFwd", core.MriCartesianRemoveROOSOperator, False, args)


def gpuMriCartesianRemoveROOSAdj(*args):
    """gpuMriCartesianRemoveROOSAdj(sensitivities, rawdata, mask) -> image
    (FE/2, PE) (mexMriCartesianRemoveROOSAdjOp.cpp)."""
    return _op_call("gpuMriCartesianRemoveROOSAdj", core.MriCartesianRemoveROOSOperator, True, args)


def _field(params, key):
    if isinstance(params, dict):
        return params[key]
    return getattr(params, key)


def _tgv_call(name, cls, args):
    _check_args(name, args, 4, "Four")
    sens, rawdata, mask, prm = args
    # mexMriCartesianReconstructTgv.cpp:79-89: a struct with the four fields
    if not (isinstance(prm, dict) or all(hasattr(prm, k) for k in ("alpha0", "alpha1", "max_iter", "reduction"))):
        raise MexError("", "Expect struct with tgv params!")
    try:
        alpha0, alpha1 = float(_field(prm, "alpha0")), float(_field(prm, "alpha1"))
        max_iter, reduction = int(_field(prm, "max_iter")), float(_field(prm, "reduction"))
    except (KeyError, AttributeError) as e:
        raise MexError("", f"Expect struct with tgv params! (missing {e})") from None
    if sens.shape != rawdata.shape:  # IU_SIZE_CHECK(&h_sensitivities, &h_rawdata)
        raise MexError("", f"size mismatch: sensitivities {sens.shape} vs rawdata {rawdata.shape}")
    op = _operator(cls, sens, mask, np.complex128)
    opt = core.TgvMriOptimizer_2c3c(dtype=np.complex128)
    p = opt.getParameters()
    p.reduction, p.max_iter, p.alpha0, p.alpha1 = reduction, max_iter, alpha0, alpha1
    opt.setOperator(op)
    opt.setNoisyData(from_matlab(rawdata))
    opt.solve()
    return np.array(opt.getResult(), dtype=np.complex128)


def gpuTgvMriOptimizer_cartesian(*args):
    """gpuTgvMriOptimizer_cartesian(sensitivities, rawdata, mask, tgv_params)
    -> image (FE, PE); tgv_params has alpha0, alpha1, max_iter, reduction
    (mexMriCartesianReconstructTgv.cpp)."""
    return _tgv_call("gpuTgvMriOptimizer_cartesian", core.MriCartesianOperator, args)


def gpuTgvMriOptimizer_cartesianRemoveROOS(*args):
    """gpuTgvMriOptimizer_cartesianRemoveROOS(sensitivities, rawdata, mask,
    tgv_params) -> image (FE/2, PE) (mexMriCartesianRemoveROOSReconstructTgv.cpp)."""
    return _tgv_call("gpuTgvMriOptimizer_cartesianRemoveROOS", core.MriCartesianRemoveROOSOperator, args)
