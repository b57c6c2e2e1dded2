"""Python classes mirroring the reference's boost-python modules.

Names, methods and parameter fields follow python/pyDenoising.cpp:39-94 and
python/pyMriReconstruction.cpp:112-268 (``Lambda`` stands for ``lambda``,
which is a Python keyword).  numpy arrays cross the boundary in the
reference's convention: a numpy array of shape (C, FE, PE) is the iu array of
size {PE, FE, C} (python/moduleutils.h:29-72).
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _capi as C


# ---------------------------------------------------------------------------
# parameters
# ---------------------------------------------------------------------------
class _Params:
    """Mutable parameter object returned by ``getParameters()``.

    Fields of the reference (tvoptimizer.h:32-61, tgvoptimizer.h:34-67,
    tgvmrioptimizer.h:33-66) plus ``gap_tol``, ``gap_every``, ``dual_fp16``,
    ``use_graph``, ``radius_u`` and ``radius_v``, which have no reference
    counterpart.
    """

    _ref_fields: tuple = ()
    _extra = ("gap_tol", "gap_every", "dual_fp16", "use_graph", "radius_u", "radius_v")

    def __init__(self, kind: int):
        object.__setattr__(self, "_p", C.PdtParams())
        C.check(C.lib().pdt_params_default(kind, ct.byref(self._p)))

    def __getattr__(self, name):
        if name == "Lambda":
            return self._p.lambda_
        if name in self._ref_fields or name in self._extra:
            return getattr(self._p, name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name == "Lambda" and "Lambda" in self._ref_fields:
            self._p.lambda_ = float(value)
        elif name in self._ref_fields or name in self._extra:
            if name in ("max_iter", "check", "gap_every"):
                value = int(value)
                if value < 0:
                    raise OverflowError("can't convert negative value to unsigned int")
            elif name in ("dual_fp16", "use_graph"):
                value = int(bool(value))
            else:
                value = float(value)
            setattr(self._p, name, value)
        else:
            raise AttributeError(f"'{type(self).__name__}' object has no attribute '{name}'")

    def __str__(self):
        lines = [type(self).__name__ + ":"]
        for f in self._ref_fields:
            lines.append(f" {f.lower() if f == 'Lambda' else f}={getattr(self, f)}")
        return "\n".join(lines) + "\n"


class TvParameters(_Params):
    _ref_fields = ("check", "Lambda", "max_iter")

    def __init__(self):
        super().__init__(C.PDT_SOLVER_TV)


class TgvParameters(_Params):
    _ref_fields = ("check", "Lambda", "max_iter", "alpha0", "alpha1")

    def __init__(self):
        super().__init__(C.PDT_SOLVER_TGV)


class TgvMriParameters(_Params):
    _ref_fields = ("alpha0", "alpha1", "check", "reduction", "max_iter")

    def __init__(self):
        super().__init__(C.PDT_SOLVER_TGV_MRI)


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------
class _Operator:
    _kind = C.PDT_OP_CARTESIAN

    def __init__(self, dtype=np.complex64):
        self._dtype = np.dtype(dtype)
        self._real = np.float64 if self._dtype == np.complex128 else np.float32
        h = ct.c_void_p()
        C.check(C.lib().pdt_op_create(self._kind, C.np_dtype_code(self._dtype), ct.byref(h)))
        self._h = h

    def _add(self, arr, dtype):
        a = np.ascontiguousarray(arr, dtype=dtype)
        iu = C.iu_size(a.shape)
        size = (ct.c_int64 * len(iu))(*iu)
        C.check(C.lib().pdt_op_add_constant(self._h, a.ctypes.data_as(ct.c_void_p), C.np_dtype_code(a.dtype),
                                            len(iu), size))

    # pyMriReconstruction.cpp:47-65: setMask adds a real 2D constant,
    # setCoilSens a complex 3D one; the order of calls is the constant order.
    def setMask(self, mask):
        self._add(mask, self._real)

    def setCoilSens(self, sens):
        self._add(sens, self._dtype)

    def _out_size(self, in_iu):
        nd = ct.c_int()
        out = (ct.c_int64 * 3)()
        src = (ct.c_int64 * len(in_iu))(*in_iu)
        C.check(C.lib().pdt_op_output_size(self._h, len(in_iu), src, ct.byref(nd), out))
        return [out[i] for i in range(nd.value)]

    def _in_size(self, out_iu):
        nd = ct.c_int()
        res = (ct.c_int64 * 3)()
        src = (ct.c_int64 * len(out_iu))(*out_iu)
        C.check(C.lib().pdt_op_input_size(self._h, len(out_iu), src, ct.byref(nd), res))
        return [res[i] for i in range(nd.value)]

    # pyMriReconstruction.cpp:67-93
    def forward(self, x):
        src = C.DeviceArray.from_numpy(x, self._dtype)
        dst = C.DeviceArray.empty(C.np_dtype_code(self._dtype), self._out_size(C.iu_size(np.shape(x))))
        C.check(C.lib().pdt_op_forward(self._h, src.h, dst.h))
        return dst.numpy()

    def adjoint(self, y):
        src = C.DeviceArray.from_numpy(y, self._dtype)
        dst = C.DeviceArray.empty(C.np_dtype_code(self._dtype), self._in_size(C.iu_size(np.shape(y))))
        C.check(C.lib().pdt_op_adjoint(self._h, src.h, dst.h))
        return dst.numpy()

    def adjointnessCheck(self):
        """Returns (<Au, p>, <u, A^*p>) for seeded random u, p
        (mricartesianoperator.cu:267-282)."""
        v = [ct.c_double() for _ in range(4)]
        C.check(C.lib().pdt_op_adjointness(self._h, *[ct.byref(x) for x in v]))
        return complex(v[0].value, v[1].value), complex(v[2].value, v[3].value)

    def __str__(self):
        buf = ct.create_string_buffer(256)
        C.check(C.lib().pdt_op_name(self._h, buf, 256))
        return buf.value.decode()

    def __del__(self):
        if getattr(self, "_h", None) and C._lib is not None:
            C._lib.pdt_op_destroy(self._h)
            self._h = None


class MriCartesianOperator(_Operator):
    _kind = C.PDT_OP_CARTESIAN


class MriCartesianRemoveROOSOperator(_Operator):
    _kind = C.PDT_OP_CARTESIAN_REMOVE_ROOS


class MriSamplingOperator(_Operator):
    _kind = C.PDT_OP_SAMPLING

    def setCoilSens(self, sens):  # not bound for the sampling operator in the reference
        raise AttributeError("MriSamplingOperator has no coil sensitivities")


# ---------------------------------------------------------------------------
# optimizers
# ---------------------------------------------------------------------------
class _Optimizer:
    _kind = C.PDT_SOLVER_TV
    _params_cls = TvParameters
    _img_dtype = np.float32
    _has_operator = False

    def __init__(self, dtype=None):
        self._dtype = np.dtype(dtype if dtype is not None else self._img_dtype)
        h = ct.c_void_p()
        C.check(C.lib().pdt_solver_create(self._kind, C.np_dtype_code(self._dtype), ct.byref(h)))
        self._h = h
        self._params = self._params_cls()
        self._op = None

    def getParameters(self):
        return self._params

    def setNoisyData(self, f):
        d = C.DeviceArray.from_numpy(f, self._dtype)
        C.check(C.lib().pdt_solver_set_noisy_data(self._h, d.h))

    def setInput0(self, u0, v0=None):
        """u0 as in the reference; v0 (shape (2, H, W)) warm-starts v."""
        du = C.DeviceArray.from_numpy(u0, self._dtype)
        dv = C.DeviceArray.from_numpy(v0, self._dtype) if v0 is not None else None
        C.check(C.lib().pdt_solver_set_input0(self._h, du.h, dv.h if dv is not None else None))

    def setOperator(self, op):
        if not self._has_operator:
            raise AttributeError("this optimizer takes no operator")
        C.check(C.lib().pdt_solver_set_operator(self._h, op._h))
        self._op = op

    def _push_params(self):
        C.check(C.lib().pdt_solver_set_params(self._h, ct.byref(self._params._p)))

    def solve(self, verbose: bool = False):
        self._push_params()
        C.check(C.lib().pdt_solver_solve(self._h, int(bool(verbose))))

    def getResult(self):
        h = ct.c_void_p()
        C.check(C.lib().pdt_solver_get_result(self._h, ct.byref(h)))
        return C.borrowed_numpy(h)

    def getV(self):
        h = ct.c_void_p()
        C.check(C.lib().pdt_solver_get_v(self._h, ct.byref(h)))
        return C.borrowed_numpy(h)

    def getStats(self) -> dict:
        st = C.PdtStats()
        C.check(C.lib().pdt_solver_get_stats(self._h, ct.byref(st)))
        return {k: getattr(st, k) for k, _ in C.PdtStats._fields_}

    def profileKernels(self, reps: int = 20) -> list:
        """Per fused kernel: name, ms per launch inside the graph-replayed
        iteration loop (the loop time per iteration, from hipEvents around a
        graph of reps iterations, divided in proportion to event pairs around
        each kernel's direct launches) and algorithmic bytes per launch.
        The iterates are put back after the timing runs."""
        self._push_params()
        out = (C.PdtKernelProf * 8)()
        n = ct.c_int()
        C.check(C.lib().pdt_solver_profile(self._h, int(reps), out, 8, ct.byref(n)))
        return [{"name": out[i].name.decode(), "ms": out[i].ms, "bytes": out[i].bytes} for i in range(n.value)]

    def __del__(self):
        if getattr(self, "_h", None) and C._lib is not None:
            C._lib.pdt_solver_destroy(self._h)
            self._h = None


def _tiles_dict(t) -> dict:
    return {"cb": t.cb, "cg": t.cg, "rpi": t.rpi, "cb_a": t.cb_a, "wg_per_cu": list(t.wpc),
            "tv_block": [t.tb_x, t.tb_y, t.tb_k], "dual_px": t.dual_px, "unpaired": t.unpaired}


def autotune(optimizers, reps: int = 20) -> dict:
    """Times the candidate tile shapes (for one ROF or TGV denoising slice
    also temporally blocked launches), then occupancy caps (workgroups per
    CU) per pass, on these configured optimizers (one batch), caches the
    fastest and returns it.  The optimizers' iterates are put back after
    the timing runs, so a solve after tuning starts where it would have."""
    for o in optimizers:
        o._push_params()
    arr = (ct.c_void_p * len(optimizers))(*[o._h.value for o in optimizers])
    t, ms = C.PdtTiles(), ct.c_double()
    C.check(C.lib().pdt_autotune(arr, len(optimizers), int(reps), ct.byref(t), ct.byref(ms)))
    return dict(_tiles_dict(t), ms_per_iteration=ms.value)


def set_tiles(optimizers, cb: int = 0, cg: int = 0, rpi: int = 0, cb_a: int = 0, wg_per_cu=None,
              tv_block=None, dual_px: int = 0, unpaired: int = 0) -> None:
    """Stores a manual tile choice for the shape of these optimizers (one
    batch) in the tile cache the autotuner uses: coils per forward and
    adjoint row item (cb, cb_a), columns per column item (cg), rows per
    denoising item (rpi), 0 = the default; wg_per_cu: 4 occupancy caps
    (forward row, column, adjoint row, denoising passes; 0 = none);
    tv_block: (tile x, tile y, iterations per launch) of the temporally
    blocked ROF / TGV denoising launches, None or iterations 0 = off
    (the one-iteration kernels); dual_px: pixels per
    thread of the TGV denoising dual step (1, 2; 0 = the default);
    unpaired: 1 = an MRI iteration always as the three passes (forward rows,
    columns, adjoint rows), 0 = two launches where the adjoint row pass
    holds every coil of a row (the default)."""
    for o in optimizers:
        o._push_params()
    arr = (ct.c_void_p * len(optimizers))(*[o._h.value for o in optimizers])
    t = C.PdtTiles(int(cb), int(cg), int(rpi), int(cb_a), (ct.c_int32 * 4)(*(wg_per_cu or (0, 0, 0, 0))),
                   *[int(v) for v in (tv_block or (0, 0, 0))], int(dual_px), int(unpaired))
    C.check(C.lib().pdt_set_tiles(arr, len(optimizers), ct.byref(t)))


def batch_profile(optimizers, reps: int = 20) -> list:
    """profileKernels over the launches of a batch of configured optimizers:
    per fused kernel, ms per launch over every slice and the algorithmic bytes
    of that launch summed over the slices.  The iterates are put back after
    the timing runs."""
    for o in optimizers:
        o._push_params()
    arr = (ct.c_void_p * len(optimizers))(*[o._h.value for o in optimizers])
    out = (C.PdtKernelProf * 8)()
    n = ct.c_int()
    C.check(C.lib().pdt_batch_profile(arr, len(optimizers), int(reps), out, 8, ct.byref(n)))
    return [{"name": out[i].name.decode(), "ms": out[i].ms, "bytes": out[i].bytes} for i in range(n.value)]


def batch_solve(optimizers, verbose: bool = False) -> None:
    """Solves configured optimizers of one class together, one persistent
    work-list launch per half-iteration for the whole batch."""
    if not optimizers:
        return
    for o in optimizers:
        o._push_params()
    arr = (ct.c_void_p * len(optimizers))(*[o._h.value for o in optimizers])
    C.check(C.lib().pdt_batch_solve(arr, len(optimizers), int(bool(verbose))))


def solve_concurrent(optimizers, streams: int = 4, verbose: bool = False) -> None:
    """Solves configured optimizers each alone (as ``solve()`` would),
    ``streams`` of them at a time: native worker threads take the next
    optimizer in order and run its solve on its own HIP stream, so the
    launches of different slices overlap on the GPU.  The optimizers may
    differ in class, size and parameters; each one's result, stop and
    statistics are those of its own ``solve()``."""
    if not optimizers:
        return
    for o in optimizers:
        o._push_params()
    arr = (ct.c_void_p * len(optimizers))(*[o._h.value for o in optimizers])
    C.check(C.lib().pdt_solve_concurrent(arr, len(optimizers), int(streams), int(bool(verbose))))


# pyDenoising.cpp:63-93
class TvOptimizer_2f(_Optimizer):
    _kind, _params_cls, _img_dtype = C.PDT_SOLVER_TV, TvParameters, np.float32


class TvOptimizer_2c(_Optimizer):
    _kind, _params_cls, _img_dtype = C.PDT_SOLVER_TV, TvParameters, np.complex64


class TgvOptimizer_2f(_Optimizer):
    _kind, _params_cls, _img_dtype = C.PDT_SOLVER_TGV, TgvParameters, np.float32


class TgvOptimizer_2c(_Optimizer):
    _kind, _params_cls, _img_dtype = C.PDT_SOLVER_TGV, TgvParameters, np.complex64


# pyMriReconstruction.cpp:197-267
class TgvMriOptimizer_2c3c(_Optimizer):
    _kind, _params_cls, _img_dtype, _has_operator = C.PDT_SOLVER_TGV_MRI, TgvMriParameters, np.complex64, True


class TgvOptimizer_2c3c(_Optimizer):
    _kind, _params_cls, _img_dtype, _has_operator = C.PDT_SOLVER_TGV_OP, TgvParameters, np.complex64, True


class TvOptimizerWithOp_2c3c(_Optimizer):
    _kind, _params_cls, _img_dtype, _has_operator = C.PDT_SOLVER_TV_OP, TvParameters, np.complex64, True


class TgvMriOptimizer_2c2c(_Optimizer):
    _kind, _params_cls, _img_dtype, _has_operator = C.PDT_SOLVER_TGV_MRI, TgvMriParameters, np.complex64, True


class TgvOptimizer_2c2c(_Optimizer):
    _kind, _params_cls, _img_dtype, _has_operator = C.PDT_SOLVER_TGV_OP, TgvParameters, np.complex64, True


class TvOptimizer_2c2c(_Optimizer):
    _kind, _params_cls, _img_dtype, _has_operator = C.PDT_SOLVER_TV_OP, TvParameters, np.complex64, True
