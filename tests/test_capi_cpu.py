"""C ABI and host logic, CPU only (no kernel is launched).

The library must load here (it links libamdhip64 but needs no GPU for the
calls below), export every symbol include/pdt_capi.h declares, and raise the
reference's errors for the host-detectable failures.
"""
import ctypes as ct
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pdt_capi.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"PDT_API\s+[\w\s\*]+?\b(pdt_\w+)\s*\(", txt)))


def test_header_declares_and_library_exports_every_symbol():
    from primaldualtoolbox import _capi as C
    names = _declared()
    assert len(names) >= 30
    L = C.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []
    out = subprocess.run(["nm", "-D", "--defined-only", C.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (pdt_\w+)", out))
    assert set(names) == exported
    assert sorted(C.EXPORTS) == names


def test_version_and_errors_without_gpu():
    from primaldualtoolbox import _capi as C
    L = C.lib()
    assert L.pdt_version() >= 100
    h = ct.c_void_p()
    assert L.pdt_solver_create(99, C.PDT_F32, ct.byref(h)) == C.PDT_ERR_ARG
    assert b"unknown solver kind" in L.pdt_last_error()
    assert L.pdt_op_create(C.PDT_OP_CARTESIAN, C.PDT_F32, ct.byref(h)) == C.PDT_ERR_ARG
    assert L.pdt_solver_create(C.PDT_SOLVER_TGV_MRI, C.PDT_F32, ct.byref(h)) == C.PDT_ERR_UNSUPPORTED


def test_reference_state_errors():
    """tvoptimizer.cu:79-85, tgvmrioptimizer.cu:96-110, tgvoptimizerwithop.cu:41-47."""
    import primaldualtoolbox as pdt
    from primaldualtoolbox import _capi as C
    opt = pdt.denoising.TvOptimizer_2f()
    with pytest.raises(C.PdtError, match="No noisy data is set!") as e:
        opt.solve()
    assert e.value.code == C.PDT_ERR_STATE
    mri = pdt.mri.TgvMriOptimizer_2c3c()
    with pytest.raises(C.PdtError, match="No operator is set!"):
        mri.solve()
    with pytest.raises(C.PdtError, match="There is no result yet!"):
        pdt.mri.TgvOptimizer_2c3c().getResult()
    with pytest.raises(AttributeError):
        opt.setOperator(None)


def test_parameter_defaults_and_fields():
    """TvParameters/TgvParameters/TgvMriParameters constructors
    (tvoptimizer.cu:26-29, tgvoptimizer.cu:27-30, tgvmrioptimizer.cu:28-31)
    and the Python field names (pyDenoising.cpp:48-61, pyMriReconstruction.cpp:121-143)."""
    import primaldualtoolbox as pdt
    tv = pdt.denoising.TvOptimizer_2f().getParameters()
    assert (tv.Lambda, tv.max_iter, tv.check) == (1.0, 1, 100)
    tgv = pdt.denoising.TgvOptimizer_2c().getParameters()
    assert (tgv.Lambda, tgv.alpha0, tgv.alpha1, tgv.max_iter, tgv.check) == (1.0, 1.0, 1.0, 1, 100)
    m = pdt.mri.TgvMriOptimizer_2c3c().getParameters()
    assert (m.alpha0, m.alpha1, m.reduction, m.max_iter, m.check) == (1.0, 1.0, 1.0, 1, 100)
    with pytest.raises(AttributeError):
        m.Lambda = 3  # TgvMriParameters has no lambda
    with pytest.raises(AttributeError):
        tv.alpha0 = 2  # TvParameters has no alpha
    with pytest.raises(OverflowError):
        tv.max_iter = -1
    tv.max_iter = 499
    assert tv.max_iter == 499
    m.gap_tol, m.gap_every = 1e-4, 50
    assert (m.gap_tol, m.gap_every) == (1e-4, 50)
    assert "max_iter=1" in str(tgv) and "alpha0=1.0" in str(tgv)


def test_python_api_surface():
    """Class and method names of the reference modules."""
    import primaldualtoolbox as pdt
    for name in ["TvOptimizer_2f", "TvOptimizer_2c", "TgvOptimizer_2f", "TgvOptimizer_2c", "TvParameters",
                 "TgvParameters"]:
        assert hasattr(pdt.denoising, name)
    for name in ["MriCartesianOperator", "MriCartesianRemoveROOSOperator", "MriSamplingOperator",
                 "TgvMriOptimizer_2c3c", "TgvOptimizer_2c3c", "TvOptimizerWithOp_2c3c", "TgvMriOptimizer_2c2c",
                 "TgvOptimizer_2c2c", "TvOptimizer_2c2c", "TgvMriParameters"]:
        assert hasattr(pdt.mri, name)
    for m in ["setNoisyData", "solve", "getResult", "getParameters", "setInput0", "setOperator"]:
        assert callable(getattr(pdt.mri.TgvMriOptimizer_2c3c, m))
    for m in ["setMask", "setCoilSens", "forward", "adjoint"]:
        assert callable(getattr(pdt.mri.MriCartesianRemoveROOSOperator, m))


def test_numpy_iu_axis_convention():
    """numpy (C, FE, PE) <-> iu {PE, FE, C} (python/moduleutils.h:29-72)."""
    from primaldualtoolbox import _capi as C
    assert C.iu_size((15, 640, 320)) == [320, 640, 15]
    assert C.iu_size((320, 320)) == [320, 320]


def test_synthetic_generators_follow_spec():
    from primaldualtoolbox import synthetic as syn
    rng = np.random.default_rng(0)
    img = syn.ellipse_phantom(rng, 64)
    assert img.min() >= 0 and img.max() <= 1
    S = syn.coil_sensitivities(rng, 6, 40, 20)
    assert np.allclose(np.sum(np.abs(S) ** 2, axis=0), 1.0)
    m = syn.line_mask(rng, 40, 320, accel=4, center=24)
    lines = np.flatnonzero(m[0])
    assert len(lines) == 80 and np.all(m == m[0:1])
    assert set(range(160 - 12, 160 + 12)) <= set(lines.tolist())
    case = syn.mri_case(seed=0, n=32, coils=3, center=4)
    f, mask = case["kspace"], case["mask"]
    assert np.all(f[:, mask == 0] == 0)  # data only on sampled lines


def test_cpp_mirror_compiles():
    """The header-only C++ mirror of the reference classes compiles against
    the C ABI (host compiler, no GPU)."""
    src = os.path.join(ROOT, "tests", "cpp", "mirror_example.cpp")
    out = os.path.join("/tmp", "pdt_mirror_example")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", f"-I{os.path.join(ROOT, 'include')}", src, "-o", out,
           f"-L{os.path.dirname(_lib_path())}", "-lpdt_hip", f"-Wl,-rpath,{os.path.dirname(_lib_path())}",
           "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _lib_path():
    from primaldualtoolbox import _capi as C
    return C.LIB_PATH


def test_committed_warm_start_weights_match_the_network():
    from primaldualtoolbox import warmstart as ws
    with np.load(ws.DEFAULT_WEIGHTS, allow_pickle=False) as z:
        depth, width, p = int(z["depth"]), int(z["width"]), z["params"]
    assert (depth, width) == (5, 32)
    assert p.dtype == np.float32 and p.size == ws.num_params(ws.layer_channels(depth, width))
    assert np.isfinite(p).all()


def test_reference_module_names():
    """The reference package imports the binding modules pymrireconstruction
    and pydenoising and exposes them as mri and denoising
    (primaldualtoolbox/__init__.py:21-22)."""
    import importlib
    import primaldualtoolbox as pdt
    m = importlib.import_module("primaldualtoolbox.pymrireconstruction")
    d = importlib.import_module("primaldualtoolbox.pydenoising")
    assert pdt.mri is m and pdt.denoising is d
    assert m.TgvMriOptimizer_2c3c is pdt.mri.TgvMriOptimizer_2c3c
    assert d.TgvOptimizer_2c is pdt.denoising.TgvOptimizer_2c


def _solver(kind, dtype, fp16=0):
    from primaldualtoolbox import _capi as C
    L = C.lib()
    h = ct.c_void_p()
    C.check(L.pdt_solver_create(kind, dtype, ct.byref(h)))
    p = C.PdtParams()
    C.check(L.pdt_params_default(kind, ct.byref(p)))
    p.dual_fp16 = fp16
    C.check(L.pdt_solver_set_params(h, ct.byref(p)))
    return h


def test_batch_checks_fire_before_any_launch():
    """Solvers batched into one engine must share kind, dtype, gap_every and
    the storage size of p, q; a handle may appear once.  The checks run on
    the host before any state is touched (no GPU needed here), for the
    batched solve, the autotuner and the manual tile choice alike."""
    from primaldualtoolbox import _capi as C
    L = C.lib()
    a = _solver(C.PDT_SOLVER_TGV, C.PDT_C64, 0)
    b = _solver(C.PDT_SOLVER_TGV, C.PDT_C64, 1)  # fp16 p, q: half the bytes
    c = _solver(C.PDT_SOLVER_TV, C.PDT_C64)
    d = _solver(C.PDT_SOLVER_TGV, C.PDT_C64, 0)
    try:
        arr = lambda *hs: (ct.c_void_p * len(hs))(*hs)  # noqa: E731
        ms = ct.c_double()
        cases = [(arr(a, b), "share dual_fp16"), (arr(a, c), "share kind and dtype"),
                 (arr(a, d, a), "appears twice")]
        for hs, msg in cases:
            assert L.pdt_batch_solve(hs, len(hs), 0) == C.PDT_ERR_ARG
            assert msg.encode() in L.pdt_last_error()
            tiles = C.PdtTiles(1, 1, 1, 1)
            assert L.pdt_autotune(hs, len(hs), 1, ct.byref(tiles), ct.byref(ms)) == C.PDT_ERR_ARG
            assert msg.encode() in L.pdt_last_error()
            assert L.pdt_set_tiles(hs, len(hs), ct.byref(tiles)) == C.PDT_ERR_ARG
            assert msg.encode() in L.pdt_last_error()
        # a consistent batch gets past the checks to the reference's state error
        assert L.pdt_batch_solve(arr(a, d), 2, 0) == C.PDT_ERR_STATE
        assert b"No noisy data is set!" in L.pdt_last_error()
    finally:
        for h in (a, b, c, d):
            L.pdt_solver_destroy(h)


def test_concurrent_solve_checks_fire_before_any_launch():
    """pdt_solve_concurrent refuses null or repeated solvers and a stream
    count outside [1, PDT_MAX_CONCURRENT = 16] on the host, before any solve
    starts (no GPU needed here).  It takes solvers of different kinds and
    dtypes, which one batched engine would refuse."""
    from primaldualtoolbox import _capi as C
    L = C.lib()
    a = _solver(C.PDT_SOLVER_TGV, C.PDT_C64, 0)
    b = _solver(C.PDT_SOLVER_TV, C.PDT_F32)
    try:
        arr = lambda *hs: (ct.c_void_p * len(hs))(*hs)  # noqa: E731
        cases = [(arr(a, b, a), 4, "appears twice"), (arr(a, None), 2, "null solver"),
                 (arr(a, b), 0, "streams"), (arr(a, b), 17, "streams")]
        for hs, streams, msg in cases:
            assert L.pdt_solve_concurrent(hs, len(hs), streams, 0) == C.PDT_ERR_ARG
            assert msg.encode() in L.pdt_last_error()
        assert L.pdt_solve_concurrent(arr(a), 0, 4, 0) == C.PDT_ERR_ARG
        assert L.pdt_batch_solve(arr(a, b), 2, 0) == C.PDT_ERR_ARG  # one engine: same kind and dtype
        import primaldualtoolbox as pdt
        assert pdt.mri.solve_concurrent is pdt.denoising.solve_concurrent
    finally:
        for h in (a, b):
            L.pdt_solver_destroy(h)


def test_params_before_any_state_are_free_to_change():
    """Before the first setInput0/solve no state exists, so dual_fp16 may
    still change (the GPU tests check the refusal once p, q exist)."""
    from primaldualtoolbox import _capi as C
    L = C.lib()
    h = _solver(C.PDT_SOLVER_TGV, C.PDT_C64, 1)
    try:
        p = C.PdtParams()
        C.check(L.pdt_solver_get_params(h, ct.byref(p)))
        p.dual_fp16, p.max_iter = 0, 77
        assert L.pdt_solver_set_params(h, ct.byref(p)) == C.PDT_OK  # no state yet: allowed
        p.dual_fp16 = 1
        assert L.pdt_solver_set_params(h, ct.byref(p)) == C.PDT_OK
        q = C.PdtParams()
        C.check(L.pdt_solver_get_params(h, ct.byref(q)))
        assert (q.dual_fp16, q.max_iter) == (1, 77)
    finally:
        L.pdt_solver_destroy(h)
