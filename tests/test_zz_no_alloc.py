"""The allocation-free operator application (pdt_op_apply_noalloc) against
the allocating pdt_op_forward / pdt_op_adjoint on the same operator and
inputs: the same kernels and tiles run on the per-GPU scratch array, so the
outputs agree bit for bit.  Two operators in a row, the second smaller, then
back to back on one stream without a wait between them; FP32 and FP64; the
state and size errors (raised on the host, nothing launched).
"""
import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _capi():
    from primaldualtoolbox import _capi as C
    return C


def _op(case, kind, dtype):
    import primaldualtoolbox as pdt
    cls = {"roos": pdt.mri.MriCartesianRemoveROOSOperator, "cart": pdt.mri.MriCartesianOperator,
           "samp": pdt.mri.MriSamplingOperator}[kind]
    op = cls(dtype=dtype)
    if kind != "samp":
        op.setCoilSens(case["sens"])
    op.setMask(case["mask"])
    return op


def _case(seed, n, n_pe, coils, roos, kind="roos"):
    from primaldualtoolbox import synthetic as syn
    c = syn.mri_case(seed=seed, n=n, n_pe=n_pe, coils=coils, center=4, roos=roos)
    if kind == "samp":
        c = dict(c, kspace=c["kspace"][0])
    return c


def _noalloc(op, forward, src, dst, stream=None):
    C = _capi()
    return C.lib().pdt_op_apply_noalloc(op._h, int(forward), src.h, dst.h, stream)


def _pair(op, case, dtype, seed):
    """(x, y, A x, A^* y) with the allocating calls, and device inputs."""
    C = _capi()
    rng = np.random.default_rng(seed)
    y = np.asarray(case["kspace"], dtype)
    shape = case["image"].shape
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
    Ax = op.forward(x)   # allocating: also prepares the operator
    Aty = op.adjoint(y)
    return (C.DeviceArray.from_numpy(x, dtype), C.DeviceArray.from_numpy(y, dtype), Ax, Aty)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_noalloc_equals_allocating_calls(dtype):
    C = _capi()
    cases = [("roos", _case(1, 96, 80, 6, True)),            # two coil groups: partial adjoints
             ("cart", _case(2, 40, 48, 3, False, "cart")),   # smaller, one coil group
             ("samp", _case(3, 24, 20, 1, False, "samp"))]   # smaller still, no sensitivities
    runs = []
    for i, (kind, case) in enumerate(cases):
        op = _op(case, kind, dtype)
        x, y, Ax, Aty = _pair(op, case, dtype, 10 + i)
        out_f = C.DeviceArray.empty(C.np_dtype_code(dtype), C.iu_size(Ax.shape))
        out_a = C.DeviceArray.empty(C.np_dtype_code(dtype), C.iu_size(Aty.shape))
        runs.append((kind, op, x, y, Ax, Aty, out_f, out_a))
    # one call after another, each checked before the next (larger first)
    for kind, op, x, y, Ax, Aty, out_f, out_a in runs:
        print(f"noalloc {kind} {np.dtype(dtype).name} {Ax.shape}", flush=True)
        C.check(_noalloc(op, True, x, out_f))
        np.testing.assert_array_equal(out_f.numpy(), Ax)
        C.check(_noalloc(op, False, y, out_a))
        np.testing.assert_array_equal(out_a.numpy(), Aty)
    # back to back on the null stream, no wait in between: stream order keeps
    # each call's scratch use apart
    for _, op, x, y, _, _, out_f, out_a in runs:
        C.check(_noalloc(op, True, x, out_f))
        C.check(_noalloc(op, False, y, out_a))
    for _, _, _, _, Ax, Aty, out_f, out_a in runs:
        np.testing.assert_array_equal(out_f.numpy(), Ax)
        np.testing.assert_array_equal(out_a.numpy(), Aty)


def test_noalloc_state_and_size_errors():
    C = _capi()
    case = _case(4, 32, 32, 2, True)
    op = _op(case, "roos", np.complex64)
    x = C.DeviceArray.from_numpy(np.zeros((32, 32), np.complex64))
    y = C.DeviceArray.empty(C.PDT_C64, [32, 64, 2])
    # not prepared yet: no allocating call has run on this operator
    assert _noalloc(op, True, x, y) == C.PDT_ERR_STATE
    assert b"not prepared" in C.lib().pdt_last_error()
    # wrong output size: the reference's size check
    bad = C.DeviceArray.empty(C.PDT_C64, [32, 64, 3])
    assert _noalloc(op, True, x, bad) == C.PDT_ERR_SIZE
    # intermediates above the scratch maximum: 16 coils, 512 x 512 image,
    # every PE line sampled -> 2 * 16 * 512 * 512 * 8 bytes of T1 and T2
    W, Hk, coils = 512, 1024, 16
    big = C.DeviceArray.empty(C.PDT_C64, [W, Hk // 2])
    bigk = C.DeviceArray.empty(C.PDT_C64, [W, Hk, coils])
    import primaldualtoolbox as pdt
    op2 = pdt.mri.MriCartesianRemoveROOSOperator()
    op2.setCoilSens(np.ones((coils, Hk, W), np.complex64))
    op2.setMask(np.ones((Hk, W), np.float32))
    C.check(C.lib().pdt_op_forward(op2._h, big.h, bigk.h))  # prepares (allocating call)
    assert _noalloc(op2, True, big, bigk) == C.PDT_ERR_SIZE
    assert b"scratch" in C.lib().pdt_last_error()
