"""GPU parity: the HIP solvers and operators against the FP64 CPU oracle.

Tolerances (fixed for every test here): relative L2 error of u <= 1e-3 after a
fixed iteration count, relative gap values within 1 %, operator outputs within
2e-5 relative L2 (FP32 FFT), adjointness within 1e-5 relative, and 1e-12 for
the FP64 variants.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_U = 1e-3
TOL_OP32 = 2e-5


def _pdt():
    import primaldualtoolbox as pdt
    return pdt


def _rng(seed):
    return np.random.default_rng(seed)


def _crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# --------------------------------------------------------------------------
# denoising
# --------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(16, 16), (13, 29), (64, 48)])
@pytest.mark.parametrize("cplx", [False, True])
def test_tv_denoise_matches_oracle(oracle, shape, cplx):
    pdt = _pdt()
    rng = _rng(1)
    f = rng.standard_normal(shape) if not cplx else _crandn(rng, shape)
    f = f.astype(np.complex64 if cplx else np.float32)
    print(f"TV denoise shape={shape} cplx={cplx}", flush=True)
    opt = (pdt.denoising.TvOptimizer_2c if cplx else pdt.denoising.TvOptimizer_2f)()
    opt.setNoisyData(f)
    opt.getParameters().Lambda = 10
    opt.getParameters().max_iter = 199
    opt.solve()
    u = opt.getResult()
    ref, _, st, _ = oracle.solve(oracle.TV, f.astype(np.complex128 if cplx else np.float64), lambda_=10, max_iter=199)
    assert u.shape == ref.shape and u.dtype == f.dtype
    assert rel_l2(u, ref) < TOL_U
    assert opt.getStats()["iters"] == 200


@pytest.mark.parametrize("W", [300, 700, 1100])
def test_denoise_row_widths_match_oracle(oracle, W):
    """Row widths for each denoising workgroup size: 300 (512 threads, some
    idle), 700 (1024 threads), 1100 (1024 threads, two pixels for some);
    TV and TGV (complex), 100 iterations vs the FP64 oracle."""
    pdt = _pdt()
    rng = _rng(9)
    f = _crandn(rng, (6, W)).astype(np.complex64)
    tv = pdt.denoising.TvOptimizer_2c()
    tv.setNoisyData(f)
    tv.getParameters().Lambda, tv.getParameters().max_iter = 10, 99
    tv.solve()
    ref, _, _, _ = oracle.solve(oracle.TV, f.astype(np.complex128), lambda_=10, max_iter=99)
    e_tv = rel_l2(tv.getResult(), ref)
    tgv = pdt.denoising.TgvOptimizer_2c()
    tgv.setNoisyData(f)
    prm = tgv.getParameters()
    prm.Lambda, prm.alpha0, prm.alpha1, prm.max_iter = 20, 2, 1, 99
    tgv.solve()
    ref, _, _, _ = oracle.solve(oracle.TGV, f.astype(np.complex128), lambda_=20, alpha0=2, alpha1=1, max_iter=99)
    e_tgv = rel_l2(tgv.getResult(), ref)
    print(f"  W={W}: TV rel L2 {e_tv:.3e}, TGV rel L2 {e_tgv:.3e}", flush=True)
    assert e_tv < TOL_U and e_tgv < TOL_U


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("fp16", [False, True])
def test_tgv_denoise_matches_oracle(oracle, cplx, fp16):
    pdt = _pdt()
    rng = _rng(2)
    shape = (40, 56)
    f = (rng.standard_normal(shape) if not cplx else _crandn(rng, shape)).astype(np.complex64 if cplx else np.float32)
    print(f"TGV denoise cplx={cplx} fp16={fp16}", flush=True)
    opt = (pdt.denoising.TgvOptimizer_2c if cplx else pdt.denoising.TgvOptimizer_2f)()
    opt.setNoisyData(f)
    prm = opt.getParameters()
    prm.Lambda, prm.alpha0, prm.alpha1, prm.max_iter = 20, 2, 1, 299
    prm.dual_fp16 = fp16
    opt.solve()
    u = opt.getResult()
    ref, _, _, _ = oracle.solve(oracle.TGV, f.astype(np.complex128 if cplx else np.float64), lambda_=20,
                                alpha0=2, alpha1=1, max_iter=299)
    err = rel_l2(u, ref)
    print(f"  rel L2 vs oracle = {err:.3e}", flush=True)
    # fp16 storage of p, q is reported, bounded loosely
    assert err < (2e-2 if fp16 else TOL_U)


def test_tv_gap_matches_oracle(oracle):
    pdt = _pdt()
    rng = _rng(3)
    f = rng.standard_normal((48, 40)).astype(np.float32)
    opt = pdt.denoising.TvOptimizer_2f()
    opt.setNoisyData(f)
    prm = opt.getParameters()
    prm.Lambda, prm.max_iter, prm.gap_every = 4, 299, 50
    opt.solve()
    st = opt.getStats()
    _, _, ost, hist = oracle.solve(oracle.TV, f.astype(np.float64), lambda_=4, max_iter=299, gap_every=50)
    assert st["n_gap"] == ost["n_gap"] == 6
    assert abs(st["rel_gap"] - ost["rel_gap"]) <= 0.01 * abs(ost["rel_gap"])
    assert abs(st["primal"] - ost["primal"]) <= 1e-4 * abs(ost["primal"])


def test_tgv_gap_matches_oracle(oracle):
    pdt = _pdt()
    rng = _rng(4)
    f = _crandn(rng, (32, 36)).astype(np.complex64)
    opt = pdt.denoising.TgvOptimizer_2c()
    opt.setNoisyData(f)
    prm = opt.getParameters()
    prm.Lambda, prm.alpha0, prm.alpha1, prm.max_iter, prm.gap_every = 20, 2, 1, 199, 100
    opt.solve()
    st = opt.getStats()
    _, _, ost, _ = oracle.solve(oracle.TGV, f.astype(np.complex128), lambda_=20, alpha0=2, alpha1=1, max_iter=199,
                                gap_every=100)
    assert abs(st["rel_gap"] - ost["rel_gap"]) <= 0.01 * abs(ost["rel_gap"])


# --------------------------------------------------------------------------
# operators
# --------------------------------------------------------------------------
def _op_case(rng, kind, W, Hk, C, line_mask=False):
    S = _crandn(rng, (C, Hk, W))
    if line_mask:
        cols = rng.random(W) < 0.4
        M = np.repeat(cols[None, :], Hk, axis=0).astype(np.float64)
    else:
        M = rng.integers(0, 2, (Hk, W)).astype(np.float64)
    return S, M


@pytest.mark.parametrize("kind,W,Hk,C", [
    ("cart", 10, 10, 5), ("roos", 10, 10, 5), ("samp", 10, 10, 1),
    ("cart", 7, 9, 3), ("roos", 12, 22, 2), ("roos", 320, 640, 3), ("roos", 368, 736, 2),
    ("cart", 23, 46, 2), ("roos", 384, 640, 2),
])
def test_operator_matches_oracle(oracle, kind, W, Hk, C):
    pdt = _pdt()
    rng = _rng(W * 1000 + Hk)
    S, M = _op_case(rng, kind, W, Hk, C, line_mask=W > 100)
    print(f"operator {kind} W={W} Hk={Hk} C={C}", flush=True)
    okind = {"cart": oracle.OP_CARTESIAN, "roos": oracle.OP_CARTESIAN_ROOS, "samp": oracle.OP_SAMPLING}[kind]
    oop = oracle.Op(okind, S if kind != "samp" else None, M)
    cls = {"cart": pdt.mri.MriCartesianOperator, "roos": pdt.mri.MriCartesianRemoveROOSOperator,
           "samp": pdt.mri.MriSamplingOperator}[kind]
    op = cls()
    if kind != "samp":
        op.setCoilSens(S.astype(np.complex64))
    op.setMask(M.astype(np.float32))
    x = _crandn(rng, (oop.H, W))
    y = oop.forward(x)
    Ax = op.forward(x.astype(np.complex64))
    assert Ax.shape == y.shape
    assert rel_l2(Ax, y) < TOL_OP32
    yy = _crandn(rng, y.shape)
    Aty = op.adjoint(yy.astype(np.complex64))
    assert rel_l2(Aty, oop.adjoint(yy)) < TOL_OP32


@pytest.mark.parametrize("kind", ["cart", "roos", "samp"])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_operator_adjointness(kind, dtype):
    """The reference's adjointness test (tests/main_adjointness_test.cu:28-67)
    with sizes {10, 10, 5}, asserted: relative 1e-5 in FP32, 1e-12 in FP64."""
    pdt = _pdt()
    rng = _rng(7)
    cls = {"cart": pdt.mri.MriCartesianOperator, "roos": pdt.mri.MriCartesianRemoveROOSOperator,
           "samp": pdt.mri.MriSamplingOperator}[kind]
    op = cls(dtype=dtype)
    if kind != "samp":
        op.setCoilSens(rng.random((5, 10, 10)) + 1j * rng.random((5, 10, 10)))
    op.setMask(rng.integers(0, 2, (10, 10)).astype(float))
    lhs, rhs = op.adjointnessCheck()
    tol = 1e-12 if dtype == np.complex128 else 1e-5
    assert abs(lhs - rhs) <= tol * abs(lhs)


# --------------------------------------------------------------------------
# MRI solvers
# --------------------------------------------------------------------------
def _mri_case(n=32, coils=4, seed=0):
    from primaldualtoolbox import synthetic as syn
    return syn.mri_case(seed=seed, n=n, coils=coils, center=max(4, 24 * n // 320))


@pytest.mark.parametrize("solver", ["tgvmri", "tgvop", "tvop"])
def test_mri_solver_matches_oracle(oracle, solver):
    pdt = _pdt()
    case = _mri_case()
    f = case["kspace"]
    a1 = 1e-5 * float(np.linalg.norm(f))
    print(f"MRI solver {solver}", flush=True)
    op = pdt.mri.MriCartesianRemoveROOSOperator()
    op.setCoilSens(case["sens"])
    op.setMask(case["mask"])
    cls = {"tgvmri": pdt.mri.TgvMriOptimizer_2c3c, "tgvop": pdt.mri.TgvOptimizer_2c3c,
           "tvop": pdt.mri.TvOptimizerWithOp_2c3c}[solver]
    opt = cls()
    opt.setOperator(op)
    opt.setNoisyData(f)
    prm = opt.getParameters()
    prm.max_iter = 199
    if solver != "tvop":
        prm.alpha1, prm.alpha0 = a1, 2 * a1
    if solver != "tgvmri":
        prm.Lambda = 0.5
    prm.gap_every = 100
    opt.solve()
    u = opt.getResult()
    oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"].astype(np.complex128), case["mask"].astype(np.float64))
    kind = {"tgvmri": oracle.TGV_MRI, "tgvop": oracle.TGV_OP, "tvop": oracle.TV_OP}[solver]
    ref, _, ost, _ = oracle.solve(kind, f.astype(np.complex128), op=oop, lambda_=0.5 if solver != "tgvmri" else 1.0,
                                  alpha1=a1 if solver != "tvop" else 1.0, alpha0=2 * a1 if solver != "tvop" else 1.0,
                                  max_iter=199, gap_every=100)
    err = rel_l2(u, ref)
    st = opt.getStats()
    print(f"  rel L2 = {err:.3e}; gap gpu {st['rel_gap']:.6e} oracle {ost['rel_gap']:.6e}", flush=True)
    assert err < TOL_U
    assert abs(st["rel_gap"] - ost["rel_gap"]) <= 0.01 * abs(ost["rel_gap"])


def test_mri_stops_at_gap(oracle):
    pdt = _pdt()
    case = _mri_case(n=24, coils=3, seed=5)
    f = case["kspace"]
    a1 = 1e-2 * float(np.linalg.norm(f))
    op = pdt.mri.MriCartesianRemoveROOSOperator()
    op.setCoilSens(case["sens"])
    op.setMask(case["mask"])
    opt = pdt.mri.TgvMriOptimizer_2c3c()
    opt.setOperator(op)
    opt.setNoisyData(f)
    prm = opt.getParameters()
    prm.alpha1, prm.alpha0, prm.max_iter, prm.gap_every, prm.gap_tol = a1, 2 * a1, 20000, 50, 1e-3
    opt.solve()
    st = opt.getStats()
    oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"].astype(np.complex128), case["mask"].astype(np.float64))
    _, _, ost, _ = oracle.solve(oracle.TGV_MRI, f.astype(np.complex128), op=oop, alpha1=a1, alpha0=2 * a1,
                                max_iter=20000, gap_every=50, gap_tol=1e-3)
    print(f"  gpu iters {st['iters']} oracle iters {ost['iters']}", flush=True)
    assert st["converged"] == 1
    assert st["rel_gap"] <= 1e-3
    assert abs(int(st["iters"]) - int(ost["iters"])) <= 100


def test_batch_equals_single(oracle):
    pdt = _pdt()
    cases = [_mri_case(n=n, coils=c, seed=s) for n, c, s in [(24, 3, 1), (32, 5, 2), (20, 2, 3)]]
    singles, batch = [], []
    for case in cases:
        for store in (singles, batch):
            op = pdt.mri.MriCartesianRemoveROOSOperator()
            op.setCoilSens(case["sens"])
            op.setMask(case["mask"])
            opt = pdt.mri.TgvMriOptimizer_2c3c()
            opt.setOperator(op)
            opt.setNoisyData(case["kspace"])
            a1 = 1e-4 * float(np.linalg.norm(case["kspace"]))
            prm = opt.getParameters()
            prm.alpha1, prm.alpha0, prm.max_iter = a1, 2 * a1, 99
            store.append(opt)
    for o in singles:
        o.solve()
    pdt.mri.batch_solve(batch)
    for a, b in zip(singles, batch):
        ua, ub = a.getResult(), b.getResult()
        assert np.array_equal(ua, ub)
