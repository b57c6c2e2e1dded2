"""GPU tests of the engine around the kernels: the mixed-size batch of the
batched workload (every slice size, 8-20 coils, random and regular masks),
parameter changes refused once state exists, state buffers kept across
setInput0, and the autotuner with its cache file.
"""
import ctypes as ct
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# seeds of batch_case covering 256^2, 320^2, 368^2 and 384 x 320 slices,
# 8..20 coils, random and regular masks
BATCH_SEEDS = (1000, 1001, 1002, 1003, 1004, 1006, 1017, 1034)


def _solver(case, max_iter, **prm):
    import primaldualtoolbox as pdt
    op = pdt.mri.MriCartesianRemoveROOSOperator()
    op.setCoilSens(case["sens"])
    op.setMask(case["mask"])
    opt = pdt.mri.TgvMriOptimizer_2c3c()
    opt.setOperator(op)
    opt.setNoisyData(case["kspace"])
    p = opt.getParameters()
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    p.alpha1, p.alpha0, p.max_iter = a1, 2 * a1, max_iter
    for k, v in prm.items():
        setattr(p, k, v)
    return op, opt, a1


@pytest.fixture(scope="module")
def batch_cases():
    from primaldualtoolbox import synthetic as syn
    cases = [syn.batch_case(s) for s in BATCH_SEEDS]
    sizes = {(c["meta"]["n"], c["meta"]["n_pe"]) for c in cases}
    assert sizes == set(syn.BATCH_SIZES)
    assert {c["meta"]["regular"] for c in cases} == {False, True}
    assert min(c["meta"]["coils"] for c in cases) == 8 and max(c["meta"]["coils"] for c in cases) == 20
    return cases


def test_mixed_batch_equals_single_solves(batch_cases):
    """The eight slices batched (one launch per FFT-length group and pass)
    give each slice's single-slice iterates bit for bit (same tiles)."""
    import primaldualtoolbox as pdt
    built = [_solver(c, 29) for c in batch_cases]
    solvers = [o for _, o, _ in built]
    pdt.mri.set_tiles(solvers, cb=4, cg=2, cb_a=4)
    pdt.mri.batch_solve(solvers)
    got = [(o.getResult().copy(), o.getV().copy(), o.getStats()["iters"]) for o in solvers]
    for (u_b, v_b, it), c in zip(got, batch_cases):
        _, o, _ = _solver(c, 29)
        pdt.mri.set_tiles([o], cb=4, cg=2, cb_a=4)
        o.solve()
        print(f"batch slice {c['meta']}: {it} iterations", flush=True)
        assert it == 30
        np.testing.assert_array_equal(u_b, o.getResult())
        np.testing.assert_array_equal(v_b, o.getV())


def test_mixed_batch_slices_match_oracle(oracle, batch_cases):
    """One 368^2 and one 384 x 320 slice of the batch against the FP64
    oracle after 10 iterations (rel. L2 of u <= 1e-3)."""
    import primaldualtoolbox as pdt
    built = [_solver(c, 9) for c in batch_cases]
    pdt.mri.batch_solve([o for _, o, _ in built])
    for (_, o, a1), c in zip(built, batch_cases):
        if (c["meta"]["n"], c["meta"]["n_pe"]) not in ((368, 368), (384, 320)) or c["meta"]["seed"] not in (1001, 1002):
            continue
        oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, c["sens"], c["mask"])
        ref, _, _, _ = oracle.solve(oracle.TGV_MRI, c["kspace"].astype(np.complex128), op=oop, alpha1=a1,
                                    alpha0=2 * a1, max_iter=9)
        err = rel_l2(o.getResult(), ref)
        print(f"slice {c['meta']}: rel L2 vs FP64 oracle {err:.3e}", flush=True)
        assert err < 1e-3


def test_refused_dual_fp16_change_keeps_parameters():
    """Once p, q exist, turning fp16 storage off is refused and changes
    nothing: the parameters still say fp16 (p, q keep their size)."""
    import primaldualtoolbox as pdt
    from primaldualtoolbox import _capi as C
    rng = np.random.default_rng(0)
    f = (rng.standard_normal((24, 20)) + 1j * rng.standard_normal((24, 20))).astype(np.complex64)
    opt = pdt.denoising.TgvOptimizer_2c()
    opt.setNoisyData(f)
    p = opt.getParameters()
    p.max_iter, p.dual_fp16 = 4, 1
    opt.solve()
    q = C.PdtParams()
    C.check(C.lib().pdt_solver_get_params(opt._h, ct.byref(q)))
    q.dual_fp16, q.max_iter = 0, 9
    assert C.lib().pdt_solver_set_params(opt._h, ct.byref(q)) == C.PDT_ERR_STATE
    r = C.PdtParams()
    C.check(C.lib().pdt_solver_get_params(opt._h, ct.byref(r)))
    assert (r.dual_fp16, r.max_iter) == (1, 4)


def test_state_buffers_kept_across_set_input0():
    """setInput0 of the same shape overwrites the state in place: the result
    array keeps its device address (captured graphs hold it), and a solve
    from the same input repeats its iterates bit for bit."""
    import primaldualtoolbox as pdt
    from primaldualtoolbox import _capi as C
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=7, n=64, coils=4, center=8)
    op, opt, _ = _solver(case, 120)
    u0 = op.adjoint(case["kspace"])
    opt.setInput0(u0)
    opt.solve()
    h = ct.c_void_p()
    C.check(C.lib().pdt_solver_get_result(opt._h, ct.byref(h)))
    ptr0 = C.lib().pdt_array_device_ptr(h)
    u_a = opt.getResult().copy()
    for _ in range(2):
        opt.setNoisyData(case["kspace"])
        opt.setInput0(u0)
        opt.solve()
        C.check(C.lib().pdt_solver_get_result(opt._h, ct.byref(h)))
        assert C.lib().pdt_array_device_ptr(h) == ptr0
        np.testing.assert_array_equal(opt.getResult(), u_a)


_TUNE_CHILD = r"""
import json, os, sys
sys.path.insert(0, os.path.join(sys.argv[1], "primal-dual-toolbox_amd"))
import numpy as np
import primaldualtoolbox as pdt
from primaldualtoolbox import synthetic as syn
case = syn.mri_case(seed=11, n=96, coils=6, center=8)
op = pdt.mri.MriCartesianRemoveROOSOperator()
op.setCoilSens(case["sens"]); op.setMask(case["mask"])
opt = pdt.mri.TgvMriOptimizer_2c3c()
opt.setOperator(op); opt.setNoisyData(case["kspace"])
a1 = 1e-3 * float(np.linalg.norm(case["kspace"]))
p = opt.getParameters(); p.alpha1, p.alpha0, p.max_iter = a1, 2 * a1, 39
tiles = pdt.mri.autotune([opt], reps=3) if sys.argv[2] == "tune" else None
opt.setNoisyData(case["kspace"]); opt.setInput0(op.adjoint(case["kspace"]))
opt.solve()
np.save(sys.argv[3], opt.getResult())
print(json.dumps({"tiles": tiles, "a1": a1}))
"""


def test_autotune_cache_file_is_deterministic_and_tuned_solve_matches_oracle(tmp_path, oracle):
    """Process 1 autotunes (tile shapes and workgroups per CU) into a
    PDT_TUNE_CACHE file and solves; process 2 reads the file and solves
    without tuning: the file is unchanged, the iterates are bit for bit
    those of process 1 (the same tiles), and both match the FP64 oracle."""
    import json
    from primaldualtoolbox import synthetic as syn
    cache = tmp_path / "tune.cache"
    env = dict(os.environ, PDT_TUNE_CACHE=str(cache))
    out1, out2 = tmp_path / "u1.npy", tmp_path / "u2.npy"
    r1 = subprocess.run([sys.executable, "-c", _TUNE_CHILD, ROOT, "tune", str(out1)], env=env, capture_output=True,
                        text=True, timeout=300)
    assert r1.returncode == 0, r1.stderr[-2000:]
    info = json.loads(r1.stdout.strip().splitlines()[-1])
    print(f"autotuned tiles {info['tiles']}", flush=True)
    text1 = cache.read_text()
    assert text1.strip()
    r2 = subprocess.run([sys.executable, "-c", _TUNE_CHILD, ROOT, "cached", str(out2)], env=env,
                        capture_output=True, text=True, timeout=300)
    assert r2.returncode == 0, r2.stderr[-2000:]
    assert cache.read_text() == text1
    u1, u2 = np.load(out1), np.load(out2)
    np.testing.assert_array_equal(u1, u2)
    case = syn.mri_case(seed=11, n=96, coils=6, center=8)
    oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"], case["mask"])
    a1 = info["a1"]
    ref, _, _, _ = oracle.solve(oracle.TGV_MRI, case["kspace"].astype(np.complex128), op=oop, alpha1=a1,
                                alpha0=2 * a1, max_iter=39)
    assert rel_l2(u1, ref) < 1e-3


@pytest.mark.parametrize("kind", ["mri", "tgv"])
def test_occupancy_caps_change_only_the_schedule(kind):
    """A cap on the workgroups per CU (extra LDS per workgroup) leaves the
    kernels and tiles as they were: the iterates are bit for bit those
    without a cap."""
    import primaldualtoolbox as pdt
    from primaldualtoolbox import synthetic as syn
    res = []
    for caps in (None, [1, 2, 1, 3], [3, 1, 2, 4]):
        if kind == "mri":
            case = syn.mri_case(seed=12, n=96, coils=9, center=8)
            _, opt, _ = _solver(case, 30)
        else:
            _, noisy = syn.tgv_denoising_case(seed=3, n=200)
            opt = pdt.denoising.TgvOptimizer_2c()
            opt.setNoisyData(noisy)
            p = opt.getParameters()
            p.Lambda, p.alpha0, p.alpha1, p.max_iter = 20, 2, 1, 30
        pdt.mri.set_tiles([opt], cb=4, cg=2, cb_a=4, wg_per_cu=caps)
        opt.solve()
        res.append(opt.getResult().copy())
    np.testing.assert_array_equal(res[0], res[1])
    np.testing.assert_array_equal(res[0], res[2])


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("fp16", [False, True])
def test_tgv_dual_two_pixels_per_thread_equals_one(cplx, fp16):
    """The TGV denoising dual step with two pixels per thread (vector
    accesses) gives the one-pixel kernel's iterates bit for bit, for even
    widths including the last column and row; the FP64 gap sums, taken in
    another order, agree to 1e-12."""
    import primaldualtoolbox as pdt
    rng = np.random.default_rng(5)
    for shape in [(64, 48), (30, 512), (17, 10), (9, 2)]:
        f = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
        f = f.astype(np.complex64 if cplx else np.float32)
        res = []
        for dpx in (1, 2):
            opt = (pdt.denoising.TgvOptimizer_2c if cplx else pdt.denoising.TgvOptimizer_2f)()
            opt.setNoisyData(f)
            p = opt.getParameters()
            p.Lambda, p.alpha0, p.alpha1, p.max_iter, p.gap_every, p.dual_fp16 = 8, 0.4, 0.2, 59, 30, int(fp16)
            pdt.denoising.set_tiles([opt], rpi=1, dual_px=dpx)
            opt.solve()
            res.append((opt.getResult().copy(), opt.getV().copy(), opt.getStats()["rel_gap"]))
        np.testing.assert_array_equal(res[0][0], res[1][0])
        np.testing.assert_array_equal(res[0][1], res[1][1])
        assert abs(res[0][2] - res[1][2]) <= 1e-12 * abs(res[0][2])


def test_autotune_and_profile_leave_the_solver_state():
    """The autotuner and the kernel profile run iterations on the solvers'
    own buffers to time them and put the iterates back: a solve after them
    equals one without them (TGV denoising, bit-equal for every tile shape;
    MRI, with the tuned tiles set on the other solver)."""
    import primaldualtoolbox as pdt
    from primaldualtoolbox import synthetic as syn
    rng = np.random.default_rng(21)
    f = (rng.standard_normal((70, 52)) + 1j * rng.standard_normal((70, 52))).astype(np.complex64)
    res = []
    for tune in (False, True):
        opt = pdt.denoising.TgvOptimizer_2c()
        opt.setNoisyData(f)
        p = opt.getParameters()
        p.Lambda, p.alpha0, p.alpha1, p.max_iter = 20.0, 2.0, 1.0, 39
        if tune:
            pdt.denoising.autotune([opt], reps=2)
            opt.profileKernels(3)
        opt.solve()
        res.append((opt.getResult().copy(), opt.getV().copy()))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    np.testing.assert_array_equal(res[0][1], res[1][1])
    case = syn.mri_case(seed=12, n=64, coils=4, center=8)
    _, o1, _ = _solver(case, 20)
    t = pdt.mri.autotune([o1], reps=2)
    o1.profileKernels(3)
    o1.solve()
    _, o2, _ = _solver(case, 20)
    pdt.mri.set_tiles([o2], cb=t["cb"], cg=t["cg"], cb_a=t["cb_a"], wg_per_cu=t["wg_per_cu"])
    o2.solve()
    print(f"tuned tiles {t}", flush=True)
    np.testing.assert_array_equal(o1.getResult(), o2.getResult())
    np.testing.assert_array_equal(o1.getV(), o2.getV())
