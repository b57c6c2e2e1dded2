"""ctypes binding of libpdt_hip.so (the C ABI declared in include/pdt_capi.h).

The library is loaded from this package directory.  There is no fallback: if
the HIP library is missing or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

LIB_NAME = "libpdt_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

PDT_OK, PDT_ERR_ARG, PDT_ERR_SIZE, PDT_ERR_STATE, PDT_ERR_CONST, PDT_ERR_HIP, PDT_ERR_UNSUPPORTED = range(7)
PDT_F32, PDT_C64, PDT_F64, PDT_C128 = range(4)
PDT_OP_CARTESIAN, PDT_OP_CARTESIAN_REMOVE_ROOS, PDT_OP_SAMPLING = range(3)
PDT_SOLVER_TV, PDT_SOLVER_TGV, PDT_SOLVER_TGV_MRI, PDT_SOLVER_TV_OP, PDT_SOLVER_TGV_OP = range(5)

# Every symbol the header declares (checked by the CPU test suite).
EXPORTS = [
    "pdt_last_error", "pdt_version", "pdt_set_device",
    "pdt_array_create", "pdt_array_wrap", "pdt_array_upload", "pdt_array_download",
    "pdt_array_info", "pdt_array_device_ptr", "pdt_array_destroy",
    "pdt_op_create", "pdt_op_add_constant", "pdt_op_num_constants", "pdt_op_forward",
    "pdt_op_adjoint", "pdt_op_input_size", "pdt_op_output_size", "pdt_op_adjointness",
    "pdt_op_name", "pdt_op_destroy", "pdt_op_apply_noalloc",
    "pdt_solver_create", "pdt_params_default", "pdt_solver_get_params", "pdt_solver_set_params",
    "pdt_solver_set_operator", "pdt_solver_set_noisy_data", "pdt_solver_set_input0",
    "pdt_solver_solve", "pdt_solver_get_result", "pdt_solver_get_v", "pdt_solver_get_stats",
    "pdt_solver_destroy", "pdt_solver_profile", "pdt_batch_profile", "pdt_autotune", "pdt_set_tiles", "pdt_batch_solve", "pdt_solve_concurrent",
    "pdt_cnn_create", "pdt_cnn_destroy", "pdt_cnn_num_params", "pdt_cnn_set_params", "pdt_cnn_get_params",
    "pdt_cnn_forward", "pdt_cnn_loss_grad", "pdt_cnn_adam_step", "pdt_cnn_train_step", "pdt_solver_warm_start",
]


class PdtParams(ct.Structure):
    _fields_ = [("lambda_", ct.c_double), ("alpha0", ct.c_double), ("alpha1", ct.c_double),
                ("reduction", ct.c_double), ("max_iter", ct.c_uint32), ("check", ct.c_uint32),
                ("gap_tol", ct.c_double), ("gap_every", ct.c_uint32), ("dual_fp16", ct.c_int32),
                ("use_graph", ct.c_int32), ("radius_u", ct.c_double), ("radius_v", ct.c_double)]


class PdtStats(ct.Structure):
    _fields_ = [("iters", ct.c_uint32), ("n_gap", ct.c_uint32), ("primal", ct.c_double),
                ("dual", ct.c_double), ("gap", ct.c_double), ("rel_gap", ct.c_double),
                ("ms", ct.c_double), ("converged", ct.c_int32), ("ms_to_stop", ct.c_double)]


class PdtTiles(ct.Structure):
    _fields_ = [("cb", ct.c_int32), ("cg", ct.c_int32), ("rpi", ct.c_int32), ("cb_a", ct.c_int32),
                ("wpc", ct.c_int32 * 4), ("tb_x", ct.c_int32), ("tb_y", ct.c_int32), ("tb_k", ct.c_int32),
                ("dual_px", ct.c_int32), ("unpaired", ct.c_int32)]


class PdtKernelProf(ct.Structure):
    _fields_ = [("name", ct.c_char * 32), ("ms", ct.c_double), ("bytes", ct.c_double)]


class PdtAdam(ct.Structure):
    _fields_ = [("lr", ct.c_double), ("beta1", ct.c_double), ("beta2", ct.c_double), ("eps", ct.c_double)]


class PdtError(RuntimeError):
    """Error raised by the HIP library (the IuException of the reference)."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_NAME} is not built (run __graft_entry__.build())")
        L = ct.CDLL(LIB_PATH)
        vp, i64p = ct.c_void_p, ct.POINTER(ct.c_int64)
        L.pdt_last_error.restype = ct.c_char_p
        L.pdt_array_create.argtypes = [ct.c_int, ct.c_int, i64p, ct.POINTER(vp)]
        L.pdt_array_wrap.argtypes = [ct.c_int, ct.c_int, i64p, vp, ct.POINTER(vp)]
        L.pdt_array_upload.argtypes = [vp, vp, ct.c_size_t]
        L.pdt_array_download.argtypes = [vp, vp, ct.c_size_t]
        L.pdt_array_info.argtypes = [vp, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), i64p]
        L.pdt_array_device_ptr.argtypes = [vp]
        L.pdt_array_device_ptr.restype = vp
        L.pdt_array_destroy.argtypes = [vp]
        L.pdt_op_create.argtypes = [ct.c_int, ct.c_int, ct.POINTER(vp)]
        L.pdt_op_add_constant.argtypes = [vp, vp, ct.c_int, ct.c_int, i64p]
        L.pdt_op_num_constants.argtypes = [vp, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]
        L.pdt_op_forward.argtypes = [vp, vp, vp]
        L.pdt_op_adjoint.argtypes = [vp, vp, vp]
        L.pdt_op_apply_noalloc.argtypes = [vp, ct.c_int, vp, vp, vp]
        L.pdt_op_input_size.argtypes = [vp, ct.c_int, i64p, ct.POINTER(ct.c_int), i64p]
        L.pdt_op_output_size.argtypes = [vp, ct.c_int, i64p, ct.POINTER(ct.c_int), i64p]
        dp = ct.POINTER(ct.c_double)
        L.pdt_op_adjointness.argtypes = [vp, dp, dp, dp, dp]
        L.pdt_op_name.argtypes = [vp, ct.c_char_p, ct.c_size_t]
        L.pdt_op_destroy.argtypes = [vp]
        L.pdt_solver_create.argtypes = [ct.c_int, ct.c_int, ct.POINTER(vp)]
        L.pdt_params_default.argtypes = [ct.c_int, ct.POINTER(PdtParams)]
        L.pdt_solver_get_params.argtypes = [vp, ct.POINTER(PdtParams)]
        L.pdt_solver_set_params.argtypes = [vp, ct.POINTER(PdtParams)]
        L.pdt_solver_set_operator.argtypes = [vp, vp]
        L.pdt_solver_set_noisy_data.argtypes = [vp, vp]
        L.pdt_solver_set_input0.argtypes = [vp, vp, vp]
        L.pdt_solver_solve.argtypes = [vp, ct.c_int]
        L.pdt_solver_get_result.argtypes = [vp, ct.POINTER(vp)]
        L.pdt_solver_get_v.argtypes = [vp, ct.POINTER(vp)]
        L.pdt_solver_get_stats.argtypes = [vp, ct.POINTER(PdtStats)]
        L.pdt_solver_destroy.argtypes = [vp]
        L.pdt_solver_profile.argtypes = [vp, ct.c_int, ct.POINTER(PdtKernelProf), ct.c_int,
                                         ct.POINTER(ct.c_int)]
        L.pdt_batch_profile.argtypes = [ct.POINTER(vp), ct.c_int, ct.c_int, ct.POINTER(PdtKernelProf), ct.c_int,
                                         ct.POINTER(ct.c_int)]
        ip = ct.POINTER(ct.c_int)
        L.pdt_autotune.argtypes = [ct.POINTER(vp), ct.c_int, ct.c_int, ct.POINTER(PdtTiles), dp]
        L.pdt_set_tiles.argtypes = [ct.POINTER(vp), ct.c_int, ct.POINTER(PdtTiles)]
        L.pdt_batch_solve.argtypes = [ct.POINTER(vp), ct.c_int, ct.c_int]
        L.pdt_solve_concurrent.argtypes = [ct.POINTER(vp), ct.c_int, ct.c_int, ct.c_int]
        fp, szp = ct.POINTER(ct.c_float), ct.POINTER(ct.c_size_t)
        L.pdt_cnn_create.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.POINTER(vp)]
        L.pdt_cnn_destroy.argtypes = [vp]
        L.pdt_cnn_num_params.argtypes = [vp, szp]
        L.pdt_cnn_set_params.argtypes = [vp, fp, ct.c_size_t]
        L.pdt_cnn_get_params.argtypes = [vp, fp, ct.c_size_t]
        L.pdt_cnn_forward.argtypes = [vp, vp, vp, ct.c_int, ct.c_int, ct.c_int]
        L.pdt_cnn_loss_grad.argtypes = [vp, vp, vp, ct.c_int, ct.c_int, ct.c_int, dp, fp, ct.c_size_t]
        L.pdt_cnn_adam_step.argtypes = [vp, ct.POINTER(PdtAdam)]
        L.pdt_cnn_train_step.argtypes = [vp, vp, vp, ct.c_int, ct.c_int, ct.c_int, ct.POINTER(PdtAdam), dp]
        L.pdt_solver_warm_start.argtypes = [vp, vp]
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != PDT_OK:
        msg = lib().pdt_last_error().decode(errors="replace")
        raise PdtError(rc, msg)


def np_dtype_code(dt) -> int:
    dt = np.dtype(dt)
    return {np.dtype(np.float32): PDT_F32, np.dtype(np.complex64): PDT_C64,
            np.dtype(np.float64): PDT_F64, np.dtype(np.complex128): PDT_C128}[dt]


CODE_NP = {PDT_F32: np.float32, PDT_C64: np.complex64, PDT_F64: np.float64, PDT_C128: np.complex128}


def iu_size(shape) -> list[int]:
    """numpy shape (C order) -> iu size (dim 0 fastest)."""
    return [int(s) for s in reversed(shape)]


class DeviceArray:
    """Owning (or wrapping) handle of an iu-layout device array."""

    def __init__(self, handle, owner=None):
        self.h = handle
        self._owner = owner  # keeps wrapped memory alive

    @classmethod
    def empty(cls, dtype_code: int, iu: list[int]):
        h = ct.c_void_p()
        arr = (ct.c_int64 * len(iu))(*iu)
        check(lib().pdt_array_create(dtype_code, len(iu), arr, ct.byref(h)))
        return cls(h)

    @classmethod
    def from_numpy(cls, a: np.ndarray, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype if dtype is not None else a.dtype)
        d = cls.empty(np_dtype_code(a.dtype), iu_size(a.shape))
        check(lib().pdt_array_upload(d.h, a.ctypes.data_as(ct.c_void_p), a.nbytes))
        return d

    @classmethod
    def wrap(cls, dtype_code: int, iu: list[int], device_ptr: int, owner=None):
        h = ct.c_void_p()
        arr = (ct.c_int64 * len(iu))(*iu)
        check(lib().pdt_array_wrap(dtype_code, len(iu), arr, ct.c_void_p(device_ptr), ct.byref(h)))
        return cls(h, owner)

    def info(self):
        dt, nd = ct.c_int(), ct.c_int()
        size = (ct.c_int64 * 3)()
        check(lib().pdt_array_info(self.h, ct.byref(dt), ct.byref(nd), size))
        return dt.value, [size[i] for i in range(nd.value)]

    def numpy(self) -> np.ndarray:
        dt, iu = self.info()
        out = np.empty(tuple(reversed(iu)), CODE_NP[dt])
        check(lib().pdt_array_download(self.h, out.ctypes.data_as(ct.c_void_p), out.nbytes))
        return out

    def device_ptr(self) -> int:
        return lib().pdt_array_device_ptr(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.pdt_array_destroy(self.h)
            self.h = None


def borrowed_numpy(handle) -> np.ndarray:
    """Downloads a borrowed array handle (not destroyed here)."""
    dt, nd = ct.c_int(), ct.c_int()
    size = (ct.c_int64 * 3)()
    check(lib().pdt_array_info(handle, ct.byref(dt), ct.byref(nd), size))
    out = np.empty(tuple(reversed([size[i] for i in range(nd.value)])), CODE_NP[dt.value])
    check(lib().pdt_array_download(handle, out.ctypes.data_as(ct.c_void_p), out.nbytes))
    return out
