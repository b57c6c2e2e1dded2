"""Concurrent single-slice solves (pdt_solve_concurrent / solve_concurrent).

Each solver's result, v, iteration count and gap equal those of its own
solve() bit for bit: for a mixed set of solver classes, sizes, precisions
and stops; with fresh solvers, whose engines are built and graphs captured
on several host threads at once; and with solvers that share one operator.
Argument checks and the report of a failing solve run on the host before
any launch of that solve.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pdt():
    import primaldualtoolbox as pdt
    return pdt


def _mixed_set(pdt):
    """Configured solvers, the same values on every call: TGV-MRI and TV with
    the operator on images of 64-160 pixels with 4-10 coils, FP32 and FP64,
    stops at the gap or at max_iter; ROF (float) and TGV (complex) denoising."""
    from primaldualtoolbox import synthetic as syn
    keep = []
    shapes = [(64, 64, 4, "tgvmri", np.complex64), (96, 80, 6, "tvop", np.complex64),
              (128, 128, 8, "tgvmri", np.complex64), (64, 96, 5, "tgvmri", np.complex128),
              (160, 160, 10, "tgvmri", np.complex64)]
    for k, (n, n_pe, coils, solver, dt) in enumerate(shapes):
        case = syn.mri_case(seed=100 + k, n=n, n_pe=n_pe, coils=coils)
        op = pdt.mri.MriCartesianRemoveROOSOperator(dtype=dt)
        op.setCoilSens(case["sens"])
        op.setMask(case["mask"])
        cls = {"tgvmri": pdt.mri.TgvMriOptimizer_2c3c, "tvop": pdt.mri.TvOptimizerWithOp_2c3c}[solver]
        o = cls(dtype=dt)
        o.setOperator(op)
        o.setNoisyData(case["kspace"])
        p = o.getParameters()
        if solver == "tgvmri":
            a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
            p.alpha1, p.alpha0 = a1, 2 * a1
        else:
            p.Lambda = 0.5
        p.max_iter, p.gap_every, p.gap_tol = 400, 50, 2e-3
        keep.append(op)
        keep.append(o)
    _, noisy = syn.tv_denoising_case(3, 96)
    o = pdt.denoising.TvOptimizer_2f()
    o.setNoisyData(noisy)
    p = o.getParameters()
    p.Lambda, p.max_iter, p.gap_every = 10, 299, 100
    keep.append(o)
    _, noisy = syn.tgv_denoising_case(4, 128)
    o = pdt.denoising.TgvOptimizer_2c()
    o.setNoisyData(noisy)
    p = o.getParameters()
    p.Lambda, p.alpha0, p.alpha1, p.max_iter = 20, 2, 1, 199
    keep.append(o)
    return keep


def _outcome(o):
    st = o.getStats()
    v = o.getV() if type(o).__name__.startswith("Tgv") else None  # TGV solvers hold v
    return o.getResult(), v, int(st["iters"]), float(st["rel_gap"])


def _assert_same(a, b):
    assert a[2] == b[2], (a[2], b[2])
    assert a[3] == b[3] or (np.isnan(a[3]) and np.isnan(b[3]))
    assert np.array_equal(a[0], b[0])
    if a[1] is not None:
        assert np.array_equal(a[1], b[1])


@pytest.mark.parametrize("streams", [1, 3, 16])
def test_concurrent_equals_single_solves(streams):
    """Fresh solvers solved concurrently against the same configuration
    solved one at a time with solve()."""
    pdt = _pdt()
    alone = [o for o in _mixed_set(pdt) if not hasattr(o, "forward")]
    for o in alone:
        o.solve()
    ref = [_outcome(o) for o in alone]
    together = [o for o in _mixed_set(pdt) if not hasattr(o, "forward")]
    pdt.mri.solve_concurrent(together, streams=streams)
    iters = [r[2] for r in ref]
    print(f"  streams {streams}: iterations {iters}", flush=True)
    assert len(set(iters)) > 2  # stops at different iterations
    for a, b in zip(ref, [_outcome(o) for o in together]):
        _assert_same(a, b)
    # a second concurrent solve of the same solvers (cached engines) continues
    # from their state like a second solve() of each
    for o in alone:
        o.solve()
    pdt.denoising.solve_concurrent(together, streams=streams)
    for a, b in zip([_outcome(o) for o in alone], [_outcome(o) for o in together]):
        _assert_same(a, b)


def test_concurrent_solvers_sharing_an_operator():
    """Four TGV-MRI solvers on one operator (different regularisation), each
    starting from A^* f computed with that operator's own intermediates."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=7, n=128, coils=8)
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))

    def make():
        op = pdt.mri.MriCartesianRemoveROOSOperator(dtype=np.complex64)
        op.setCoilSens(case["sens"])
        op.setMask(case["mask"])
        out = []
        for f in (0.5, 1.0, 2.0, 4.0):
            o = pdt.mri.TgvMriOptimizer_2c3c(dtype=np.complex64)
            o.setOperator(op)
            o.setNoisyData(case["kspace"])
            p = o.getParameters()
            p.alpha1, p.alpha0, p.max_iter, p.gap_every = f * a1, 2 * f * a1, 150, 50
            out.append(o)
        return op, out

    op1, alone = make()
    for o in alone:
        o.solve()
    op2, together = make()
    pdt.mri.solve_concurrent(together, streams=4)
    for a, b in zip([_outcome(o) for o in alone], [_outcome(o) for o in together]):
        _assert_same(a, b)
    assert not np.array_equal(alone[0].getResult(), alone[3].getResult())


def test_concurrent_arguments_and_errors():
    pdt = _pdt()
    import primaldualtoolbox._capi as C
    from primaldualtoolbox import synthetic as syn
    _, noisy = syn.tv_denoising_case(5, 64)
    good = []
    for _ in range(3):
        o = pdt.denoising.TvOptimizer_2f()
        o.setNoisyData(noisy)
        o.getParameters().max_iter = 20
        good.append(o)
    with pytest.raises(C.PdtError, match="twice") as e:
        pdt.denoising.solve_concurrent([good[0], good[1], good[0]])
    assert e.value.code == C.PDT_ERR_ARG
    for bad in (0, 17):
        with pytest.raises(C.PdtError, match="streams") as e:
            pdt.denoising.solve_concurrent(good, streams=bad)
        assert e.value.code == C.PDT_ERR_ARG
    # a solver without data fails in its own solve: the error of solve()
    nodata = pdt.denoising.TvOptimizer_2f()
    with pytest.raises(C.PdtError) as single:
        nodata.solve()
    with pytest.raises(C.PdtError) as e:
        pdt.denoising.solve_concurrent([good[0], nodata, good[1], good[2]], streams=2)
    assert e.value.code == single.value.code == C.PDT_ERR_STATE
    assert str(e.value) == str(single.value)
