"""More GPU parity: full-size workloads, the remaining solver/operator
combinations, precision variants, control flow and error paths.

Tolerances: rel. L2 of u <= 1e-3 against the FP64 oracle after a fixed
iteration count (FP32), <= 1e-9 for the FP64 GPU build; relative gap values
within 1 %; FP32 vs FP64 GPU at full size within 1e-3.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _pdt():
    import primaldualtoolbox as pdt
    return pdt


def _mri_solver(pdt, case, cls=None, op_cls=None, dtype=np.complex64, a1=None, max_iter=99, **prm):
    op_cls = op_cls or pdt.mri.MriCartesianRemoveROOSOperator
    op = op_cls(dtype=dtype)
    if op_cls is not pdt.mri.MriSamplingOperator:
        op.setCoilSens(case["sens"])
    op.setMask(case["mask"])
    opt = (cls or pdt.mri.TgvMriOptimizer_2c3c)(dtype=dtype)
    opt.setOperator(op)
    opt.setNoisyData(case["kspace"])
    p = opt.getParameters()
    if a1 is not None:
        p.alpha1, p.alpha0 = a1, 2 * a1
    p.max_iter = max_iter
    for k, v in prm.items():
        setattr(p, k, v)
    return op, opt


def test_tv_256_500_iterations(oracle):
    """TV-L2 256x256 float32, lambda 10, 500 iterations, vs FP64 oracle; gap within 1 %."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    _, noisy = syn.tv_denoising_case(0, 256)
    opt = pdt.denoising.TvOptimizer_2f()
    opt.setNoisyData(noisy)
    p = opt.getParameters()
    p.Lambda, p.max_iter, p.gap_every = 10, 499, 500
    opt.solve()
    ref, _, ost, _ = oracle.solve(oracle.TV, noisy.astype(np.float64), lambda_=10, max_iter=499, gap_every=500)
    err = rel_l2(opt.getResult(), ref)
    st = opt.getStats()
    print(f"tv 256: rel L2 {err:.3e} gap gpu {st['rel_gap']:.5e} oracle {ost['rel_gap']:.5e}", flush=True)
    assert err < 1e-3
    assert abs(st["rel_gap"] - ost["rel_gap"]) <= 0.01 * ost["rel_gap"]


@pytest.mark.parametrize("fp16", [False, True])
def test_tgv_complex_512_1000_iterations(oracle, fp16):
    """TGV complex64 512x512, alpha1 1, alpha0 2, lambda 20, 1000 iterations.
    FP32 duals: rel L2 <= 1e-3; FP16-storage duals: error reported, <= 5e-2."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    _, noisy = syn.tgv_denoising_case(0, 512)
    opt = pdt.denoising.TgvOptimizer_2c()
    opt.setNoisyData(noisy)
    p = opt.getParameters()
    p.Lambda, p.alpha0, p.alpha1, p.max_iter, p.dual_fp16 = 20, 2, 1, 999, fp16
    opt.solve()
    ref = _tgv512_ref(oracle, noisy)
    err = rel_l2(opt.getResult(), ref)
    print(f"tgv 512 fp16={fp16}: rel L2 vs FP64 oracle {err:.3e}", flush=True)
    assert err < (5e-2 if fp16 else 1e-3)


_CFG2 = {}


def _tgv512_ref(oracle, noisy):
    if "u" not in _CFG2:
        _CFG2["u"] = oracle.solve(oracle.TGV, noisy.astype(np.complex128), lambda_=20, alpha0=2, alpha1=1,
                                  max_iter=999)[0]
    return _CFG2["u"]


@pytest.mark.parametrize("tiles", [None, (5, 1, 15), (4, 2, 5)])
def test_mri_320_full_size_parity_fixed_iterations(oracle, tiles):
    """320x320, 15 coils, ROOS 640: 20 iterations vs the FP64 oracle, with
    the default tiles and with forced ones: (5, 1, 15) is the autotuned shape
    of this workload (all 15 coils of a row in one 1024-thread adjoint block,
    u and v steps fused, coil products summed by three thread groups),
    (4, 2, 5) the path with per-group partial adjoints and k_mri_primal."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=0, n=320, coils=15)
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    _, opt = _mri_solver(pdt, case, a1=a1, max_iter=19)
    if tiles:
        pdt.mri.set_tiles([opt], cb=tiles[0], cg=tiles[1], cb_a=tiles[2])
    try:
        opt.solve()
    finally:
        if tiles:
            pdt.mri.set_tiles([opt])
    oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"].astype(np.complex128), case["mask"].astype(np.float64))
    ref, _, _, _ = oracle.solve(oracle.TGV_MRI, case["kspace"].astype(np.complex128), op=oop, alpha1=a1,
                                alpha0=2 * a1, max_iter=19)
    err = rel_l2(opt.getResult(), ref)
    print(f"mri 320 20 iterations, tiles {tiles}: rel L2 {err:.3e}", flush=True)
    assert err < 1e-3


@pytest.mark.parametrize("n,n_pe", [(256, 256), (368, 368), (384, 320), (160, 384)])
def test_fft_lengths_solver_parity(oracle, n, n_pe):
    """TGV-MRI through the per-length FFT kernels: PE 256/368/320/384 and FE
    512/736/768/320 (320 = 2n for n = 160), 4 coils, 10 iterations vs the
    FP64 oracle."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=3, n=n, n_pe=n_pe, coils=4)
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    print(f"fft lengths PE {n_pe} FE {2 * n}", flush=True)
    _, opt = _mri_solver(pdt, case, a1=a1, max_iter=9)
    opt.solve()
    oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"].astype(np.complex128), case["mask"].astype(np.float64))
    ref, _, _, _ = oracle.solve(oracle.TGV_MRI, case["kspace"].astype(np.complex128), op=oop, alpha1=a1,
                                alpha0=2 * a1, max_iter=9)
    err = rel_l2(opt.getResult(), ref)
    print(f"  rel L2 {err:.3e}", flush=True)
    assert err < 1e-3


@pytest.mark.parametrize("cb,cg,cb_a", [(1, 1, 1), (5, 4, 10), (10, 8, 10), (10, 2, 5)])
def test_odd_radix_split_over_waves(oracle, cb, cg, cb_a):
    """The radix-23 stage of the 368 / 736 transforms splits each butterfly's
    output pairs over 4, 2 or 1 waves depending on how many butterflies a
    block holds; these tiles reach every split in the forward rows (16 cb
    butterflies on 256 threads), the columns (32 cg) and the adjoint rows
    (1024 threads when cb_a = 10 coils).  10 iterations vs the FP64 oracle."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=6, n=368, coils=10)
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    _, opt = _mri_solver(pdt, case, a1=a1, max_iter=9)
    pdt.mri.set_tiles([opt], cb=cb, cg=cg, cb_a=cb_a)
    try:
        opt.solve()
    finally:
        pdt.mri.set_tiles([opt])  # back to the default tiles for this shape
    oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"].astype(np.complex128), case["mask"].astype(np.float64))
    ref, _, _, _ = oracle.solve(oracle.TGV_MRI, case["kspace"].astype(np.complex128), op=oop, alpha1=a1,
                                alpha0=2 * a1, max_iter=9)
    err = rel_l2(opt.getResult(), ref)
    print(f"tiles cb={cb} cg={cg} cb_a={cb_a}: rel L2 {err:.3e}", flush=True)
    assert err < 1e-3


@pytest.mark.parametrize("n_pe,coils", [(256, 8), (384, 6), (512, 4), (640, 4)])
def test_wide_adjoint_permuted_slots(oracle, n_pe, coils):
    """Every coil of a row in one 1024-thread adjoint workgroup (cb_a = C,
    cb_a * W >= 2048) for each row length whose wide FFT buffers have
    XOR-permuted slots: 256, 384 and 512 with the register coil sum and the
    fused primal step, 640 (2 W > 1024 threads) without them.  64 FE rows on
    the 128 ROOS grid, 10 iterations vs the FP64 oracle."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=12, n=64, n_pe=n_pe, coils=coils)
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    _, opt = _mri_solver(pdt, case, a1=a1, max_iter=9)
    pdt.mri.set_tiles([opt], cb=4, cg=1, cb_a=coils)
    try:
        opt.solve()
    finally:
        pdt.mri.set_tiles([opt])
    oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"].astype(np.complex128), case["mask"].astype(np.float64))
    ref, _, _, _ = oracle.solve(oracle.TGV_MRI, case["kspace"].astype(np.complex128), op=oop, alpha1=a1,
                                alpha0=2 * a1, max_iter=9)
    err = rel_l2(opt.getResult(), ref)
    print(f"wide adjoint PE {n_pe}, {coils} coils: rel L2 {err:.3e}", flush=True)
    assert err < 1e-3


def _batch_pair(pdt, specs):
    """Two identical lists of configured solvers (for single vs batched)."""
    from primaldualtoolbox import synthetic as syn
    out = ([], [])
    for n, n_pe, coils, seed, max_iter, tol in specs:
        case = syn.mri_case(seed=seed, n=n, n_pe=n_pe, coils=coils)
        a1 = 1e-4 * float(np.linalg.norm(case["kspace"]))
        for store in out:
            store.append(_mri_solver(pdt, case, a1=a1, max_iter=max_iter, gap_every=50, gap_tol=tol))
    return out


@pytest.mark.parametrize("solver,coils,n", [("tgvmri", 8, 320), ("tgvmri", 15, 320), ("tvop", 15, 320),
                                            ("tgvmri", 20, 256)])
def test_single_slice_step_workgroups_equal_batch(solver, coils, n):
    """A single-slice solve runs the dual step and, with wide adjoint blocks,
    the v step in workgroups of their own, and its adjoint row blocks take
    the straight-line path (sensitivity loads in flight during the FFT); a
    batch runs both steps inside the row workgroups and the general adjoint
    path.  320 x 320 with 8 or 15 coils (three thread groups of up to five
    coil products per adjoint block) and 256 x 256 with 20 coils (four
    groups of five), TGV-MRI and TV with the operator, the same tiles (all
    coils of a row in one adjoint block) for both: the iterates agree bit for
    bit after 100 iterations, the gap to 1e-12."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=31, n=n, coils=coils)
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    cls = {"tgvmri": pdt.mri.TgvMriOptimizer_2c3c, "tvop": pdt.mri.TvOptimizerWithOp_2c3c}[solver]
    prm = dict(gap_every=50) if solver == "tgvmri" else dict(gap_every=50, Lambda=0.5)
    single = _mri_solver(pdt, case, cls=cls, a1=a1 if solver == "tgvmri" else None, max_iter=99, **prm)[1]
    batch = [_mri_solver(pdt, case, cls=cls, a1=a1 if solver == "tgvmri" else None, max_iter=99, **prm)[1]
             for _ in range(2)]
    pdt.mri.set_tiles([single], cb=4, cg=1, cb_a=coils)
    pdt.mri.set_tiles(batch, cb=4, cg=1, cb_a=coils)
    try:
        single.solve()
        pdt.mri.batch_solve(batch)
    finally:
        pdt.mri.set_tiles([single])
        pdt.mri.set_tiles(batch)
    s0 = single.getStats()
    for b in batch:
        sb = b.getStats()
        print(f"  single {s0['iters']} it. gap {s0['rel_gap']:.12e}; batch {sb['iters']} it. gap {sb['rel_gap']:.12e}",
              flush=True)
        assert sb["iters"] == s0["iters"] == 100
        assert abs(sb["rel_gap"] - s0["rel_gap"]) <= 1e-12 * abs(s0["rel_gap"])
        assert np.array_equal(b.getResult(), single.getResult())
        if solver == "tgvmri":
            assert np.array_equal(b.getV(), single.getV())


@pytest.mark.parametrize("solver,coils,n,n_pe,gap_every,max_iter,graph", [
    ("tgvmri", 15, 320, 320, 50, 120, 1),  # runs of 49 pairs, a gap iteration, a short last run
    ("tgvmri", 15, 320, 320, 3, 20, 1),    # runs of two iterations: a first and a last one only
    ("tgvmri", 15, 320, 320, 50, 60, 0),   # direct launches
    ("tvop", 15, 320, 320, 50, 60, 1),
    ("tgvmri", 20, 256, 256, 0, 74, 1),    # no gap evaluation: chunks of 50 and 25; two coils per wave
    ("tgvmri", 8, 320, 320, 7, 30, 1),
    ("tgvmri", 6, 384, 384, 50, 30, 1),    # rows of 384 (radix-6 last stage)
    ("tgvmri", 4, 64, 512, 50, 30, 1),     # rows of 512
    ("tgvmri", 10, 368, 368, 50, 30, 1),   # rows of 368: two unpadded buffers, workgroup-wide stages
])
def test_paired_iterations_equal_three_passes(solver, coils, n, n_pe, gap_every, max_iter, graph):
    """A single slice whose adjoint row blocks hold every coil runs an
    iteration without the gap evaluation as two launches (the column pass
    with the dual step's workgroups, then the adjoint row pass that also
    computes the next iteration's forward row FFTs from its new ubar row);
    the first iteration of each run of them starts with the forward row pass
    and the last ends without the next FFTs.  Against the same solve forced
    to three passes per iteration (unpaired): the same iterates bit for bit,
    the same stop and gap, over run lengths 1, 2, 6, 49 and 50 iterations and
    every row length of the all-coil adjoint path (256, 320, 384 and 512 with
    per-wave transforms on padded slots, 368 with workgroup-wide stages)."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=37, n=n, n_pe=n_pe, coils=coils)
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    cls = {"tgvmri": pdt.mri.TgvMriOptimizer_2c3c, "tvop": pdt.mri.TvOptimizerWithOp_2c3c}[solver]
    prm = dict(gap_every=gap_every, use_graph=graph, gap_tol=0.0)
    if solver == "tvop":
        prm["Lambda"] = 0.5
    a = None if solver == "tvop" else a1
    _, paired = _mri_solver(pdt, case, cls=cls, a1=a, max_iter=max_iter, **prm)
    _, three = _mri_solver(pdt, case, cls=cls, a1=a, max_iter=max_iter, **prm)
    try:
        pdt.mri.set_tiles([paired], cb=5, cg=1, cb_a=coils)
        names = [k["name"] for k in paired.profileKernels(3)]
        paired.solve()
        pdt.mri.set_tiles([three], cb=5, cg=1, cb_a=coils, unpaired=1)
        names3 = [k["name"] for k in three.profileKernels(3)]
        three.solve()
    finally:
        pdt.mri.set_tiles([paired])
    print(f"  launches: {names} / {names3}", flush=True)
    assert names == ["k_cols_dual", "k_rows_adj_next"]
    assert names3 == ["k_rows_fwd", "k_cols", "k_rows_adj"]
    sp, s3 = paired.getStats(), three.getStats()
    print(f"  paired {sp['iters']} it. gap {sp['rel_gap']:.12e}; three passes {s3['iters']} it. "
          f"gap {s3['rel_gap']:.12e}", flush=True)
    assert sp["iters"] == s3["iters"] == max_iter + 1
    assert sp["rel_gap"] == s3["rel_gap"]
    assert np.array_equal(paired.getResult(), three.getResult())
    if solver == "tgvmri":
        assert np.array_equal(paired.getV(), three.getV())


def test_batch_ragged_stops_equal_single_solves():
    """Slices of one batch that stop at different iterations (iteration
    limits and gap tolerances) leave the work lists as they finish; every
    slice's result and iteration count equal those of its own solve, and its
    gap to 1e-12 (a single solve sums the dual step's FP64 gap terms over
    other workgroups than a batch does)."""
    pdt = _pdt()
    specs = [(32, 32, 3, 1, 49, 0.0), (24, 24, 2, 2, 299, 0.0), (32, 32, 4, 3, 999, 1e-2),
             (20, 28, 2, 4, 149, 0.0), (32, 32, 3, 5, 999, 3e-3), (24, 24, 2, 6, 99, 0.0)]
    singles, batch = _batch_pair(pdt, specs)
    for _, o in singles:
        o.solve()
    pdt.mri.batch_solve([o for _, o in batch])
    for (_, a), (_, b) in zip(singles, batch):
        sa, sb = a.getStats(), b.getStats()
        print(f"  iters single {sa['iters']} batch {sb['iters']}", flush=True)
        assert sa["iters"] == sb["iters"]
        assert abs(sa["rel_gap"] - sb["rel_gap"]) <= 1e-12 * abs(sb["rel_gap"])
        assert np.array_equal(a.getResult(), b.getResult())


def test_batch_mixed_compiled_and_runtime_lengths():
    """One batch holding slices with compile-time FFT plans (256 / 512) and
    runtime plans (PE 40, FE 80) equals the single solves bit for bit."""
    pdt = _pdt()
    specs = [(256, 256, 2, 7, 29, 0.0), (40, 40, 3, 8, 29, 0.0), (256, 256, 3, 9, 29, 0.0)]
    singles, batch = _batch_pair(pdt, specs)
    for _, o in singles:
        o.solve()
    pdt.mri.batch_solve([o for _, o in batch])
    for (_, a), (_, b) in zip(singles, batch):
        assert np.array_equal(a.getResult(), b.getResult())


def test_mri_320_time_to_gap_fp32_matches_fp64_gpu():
    """Size-independent check at full size: the FP32 and FP64 GPU solvers stop
    within one gap interval of each other at rel. gap 1e-4 and agree on u."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=0, n=320, coils=15)
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    res = {}
    for dt in (np.complex64, np.complex128):
        _, opt = _mri_solver(pdt, case, dtype=dt, a1=a1, max_iter=100000, gap_tol=1e-4, gap_every=50)
        opt.solve()
        res[dt] = (opt.getResult(), opt.getStats())
    (u32, s32), (u64, s64) = res[np.complex64], res[np.complex128]
    print(f"fp32 iters {s32['iters']} gap {s32['rel_gap']:.3e}; fp64 iters {s64['iters']} gap {s64['rel_gap']:.3e}",
          flush=True)
    assert s32["converged"] and s64["converged"]
    assert abs(int(s32["iters"]) - int(s64["iters"])) <= 50
    assert rel_l2(u32, u64) < 1e-3


def test_alpha_continuation_matches_oracle(oracle):
    """TGV-MRI with reduction != 1 (tgvmrioptimizer.cu:127-155)."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=3, n=32, coils=4, center=4)
    a1 = 1e-2 * float(np.linalg.norm(case["kspace"]))
    _, opt = _mri_solver(pdt, case, a1=a1, max_iter=149, reduction=1e-3)
    opt.solve()
    oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"].astype(np.complex128), case["mask"].astype(np.float64))
    ref, _, _, _ = oracle.solve(oracle.TGV_MRI, case["kspace"].astype(np.complex128), op=oop, alpha1=a1,
                                alpha0=2 * a1, reduction=1e-3, max_iter=149)
    assert rel_l2(opt.getResult(), ref) < 1e-3


def test_cartesian_without_roos_and_sampling_solvers(oracle):
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=4, n=32, coils=3, center=4, roos=False)
    a1 = 1e-3 * float(np.linalg.norm(case["kspace"]))
    _, opt = _mri_solver(pdt, case, op_cls=pdt.mri.MriCartesianOperator, a1=a1, max_iter=99)
    opt.solve()
    oop = oracle.Op(oracle.OP_CARTESIAN, case["sens"].astype(np.complex128), case["mask"].astype(np.float64))
    ref, _, _, _ = oracle.solve(oracle.TGV_MRI, case["kspace"].astype(np.complex128), op=oop, alpha1=a1,
                                alpha0=2 * a1, max_iter=99)
    assert rel_l2(opt.getResult(), ref) < 1e-3
    # single coil, no sensitivities: MriSamplingOperator + TgvMriOptimizer_2c2c
    rng = np.random.default_rng(0)
    x = syn.complex_phantom(rng, 24, 20)
    mask = syn.line_mask(rng, 20, 24, accel=2, center=4)
    k = (mask * syn.centered_fft2(x)).astype(np.complex64)
    sc = {"kspace": k, "mask": mask, "sens": None}
    _, opt = _mri_solver(pdt, sc, cls=pdt.mri.TgvMriOptimizer_2c2c, op_cls=pdt.mri.MriSamplingOperator,
                         a1=1e-2, max_iter=99)
    opt.solve()
    oop = oracle.Op(oracle.OP_SAMPLING, None, mask.astype(np.float64))
    ref, _, _, _ = oracle.solve(oracle.TGV_MRI, k.astype(np.complex128), op=oop, alpha1=1e-2, alpha0=2e-2,
                                max_iter=99)
    assert rel_l2(opt.getResult(), ref) < 1e-3


@pytest.mark.parametrize("kind", ["tv", "tgv", "tgvmri"])
def test_fp64_gpu_matches_oracle_tightly(oracle, kind):
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    if kind == "tgvmri":
        case = syn.mri_case(seed=1, n=24, coils=3, center=4)
        a1 = 1e-3 * float(np.linalg.norm(case["kspace"]))
        _, opt = _mri_solver(pdt, case, dtype=np.complex128, a1=a1, max_iter=199)
        opt.solve()
        oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"], case["mask"])
        ref = oracle.solve(oracle.TGV_MRI, case["kspace"].astype(np.complex128), op=oop, alpha1=a1, alpha0=2 * a1,
                           max_iter=199)[0]
    else:
        _, noisy = syn.tgv_denoising_case(1, 40)
        cls = pdt.denoising.TvOptimizer_2c if kind == "tv" else pdt.denoising.TgvOptimizer_2c
        opt = cls(dtype=np.complex128)
        opt.setNoisyData(noisy)
        p = opt.getParameters()
        p.Lambda, p.max_iter = 20, 199
        if kind == "tgv":
            p.alpha0, p.alpha1 = 2, 1
        opt.solve()
        ref = oracle.solve(oracle.TV if kind == "tv" else oracle.TGV, noisy.astype(np.complex128), lambda_=20,
                           alpha0=2, alpha1=1, max_iter=199)[0]
    err = rel_l2(opt.getResult(), ref)
    print(f"fp64 {kind}: rel L2 {err:.3e}", flush=True)
    assert err < 1e-9


def test_graph_replay_equals_direct_launches():
    """hipGraph-replayed chunks give bit-identical results to direct launches,
    replay after replay (with the solver's own synchronisations between)."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=2, n=32, coils=4, center=4)
    outs = []
    for use_graph in (0, 1):
        _, opt = _mri_solver(pdt, case, a1=1e-3, max_iter=499, gap_every=50, use_graph=use_graph)
        opt.solve()
        outs.append((opt.getResult(), opt.getStats()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1]["rel_gap"] == outs[1][1]["rel_gap"]


def test_resolve_continues_and_input0_warm_start(oracle):
    """A second solve() continues from the state (tvoptimizer.cu:87); setInput0
    resets the duals; v0 = 0 reproduces the reference cold start."""
    pdt = _pdt()
    rng = np.random.default_rng(5)
    f = (rng.standard_normal((20, 24)) + 1j * rng.standard_normal((20, 24))).astype(np.complex64)
    a = pdt.denoising.TgvOptimizer_2c()
    a.setNoisyData(f)
    pa = a.getParameters()
    pa.Lambda, pa.alpha0, pa.alpha1, pa.max_iter = 10, 2, 1, 99
    a.solve()
    a.solve()
    b = pdt.denoising.TgvOptimizer_2c()
    b.setNoisyData(f)
    pb = b.getParameters()
    pb.Lambda, pb.alpha0, pb.alpha1, pb.max_iter = 10, 2, 1, 199
    b.solve()
    assert rel_l2(a.getResult(), b.getResult()) < 1e-6
    c = pdt.denoising.TgvOptimizer_2c()
    c.setNoisyData(f)
    c.setInput0(f, np.zeros((2, 20, 24), np.complex64))
    pc = c.getParameters()
    pc.Lambda, pc.alpha0, pc.alpha1, pc.max_iter = 10, 2, 1, 199
    c.solve()
    assert np.array_equal(c.getResult(), b.getResult())
    v0 = (0.1 * rng.standard_normal((2, 20, 24))).astype(np.complex64)
    c.setInput0(f, v0)
    c.solve()
    ref, _, _, _ = oracle.solve(oracle.TGV, f.astype(np.complex128), lambda_=10, alpha0=2, alpha1=1, max_iter=199,
                                v0=v0.astype(np.complex128))
    assert rel_l2(c.getResult(), ref) < 1e-3


def test_new_data_same_shape_equals_fresh_solver():
    """setNoisyData with new data of the same shape overwrites the held copy
    and the cached engine (same buffers) must pick the new data up: the
    result equals that of a fresh solver, for denoising (data read by the
    kernels directly) and for MRI (data compacted per solve)."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    rng = np.random.default_rng(12)
    f1, f2 = (rng.standard_normal((2, 24, 20)) + 1j * rng.standard_normal((2, 24, 20))).astype(np.complex64)

    def tgv():
        o = pdt.denoising.TgvOptimizer_2c()
        p = o.getParameters()
        p.Lambda, p.alpha0, p.alpha1, p.max_iter = 10, 2, 1, 49
        return o

    a = tgv()
    a.setNoisyData(f1)
    a.solve()
    a.setNoisyData(f2)
    a.setInput0(f2)
    a.solve()
    b = tgv()
    b.setNoisyData(f2)
    b.solve()
    assert np.array_equal(a.getResult(), b.getResult())
    c1 = syn.mri_case(seed=21, n=32, coils=3, center=4)
    c2 = syn.mri_case(seed=22, n=32, coils=3, center=4)
    c2["mask"], c2["sens"] = c1["mask"], c1["sens"]  # same operator, new data
    a1 = 1e-3 * float(np.linalg.norm(c1["kspace"]))
    _, m = _mri_solver(pdt, c1, a1=a1, max_iter=29)
    m.solve()
    op = pdt.mri.MriCartesianRemoveROOSOperator()
    op.setCoilSens(c1["sens"])
    op.setMask(c1["mask"])
    m.setNoisyData(c2["kspace"])
    m.setInput0(op.adjoint(c2["kspace"]))
    m.solve()
    _, n = _mri_solver(pdt, {**c1, "kspace": c2["kspace"]}, a1=a1, max_iter=29)
    n.solve()
    assert np.array_equal(m.getResult(), n.getResult())


def test_operator_constant_and_size_errors():
    """operatorbase.h:242-256,276-287,493-502; mricartesianoperator.cu:310-318."""
    pdt = _pdt()
    from primaldualtoolbox import _capi as C
    op = pdt.mri.MriCartesianRemoveROOSOperator()
    x = np.zeros((5, 10), np.complex64)
    # the binding asks getOutputSize first, which reads constant 0
    # (pyMriReconstruction.cpp:76, operatorbase.h:267-274)
    with pytest.raises(C.PdtError, match=r"Requested index \(0\) is not valid. Constants available: 0"):
        op.forward(x)
    src = C.DeviceArray.from_numpy(x)
    dst = C.DeviceArray.empty(C.PDT_C64, [10, 10, 2])
    with pytest.raises(C.PdtError, match=r"Number of constants \(0\) does not match number of required constants \(2\)"):
        C.check(C.lib().pdt_op_forward(op._h, src.h, dst.h))
    op.setCoilSens(np.ones((2, 10, 10), np.complex64))
    op.setMask(np.ones((10, 10), np.float32))
    with pytest.raises(C.PdtError, match="Too many InputTypeConstants"):
        op.setMask(np.ones((10, 10), np.float32))
    with pytest.raises(C.PdtError) as e:
        op.forward(np.zeros((6, 10), np.complex64))
    assert e.value.code == C.PDT_ERR_SIZE
    bad = pdt.mri.MriCartesianOperator()
    bad.setMask(np.ones((10, 10), np.float32))  # wrong order: mask first
    bad.setCoilSens(np.ones((2, 10, 10), np.complex64))
    with pytest.raises(C.PdtError, match="cannot be cast"):
        bad.forward(np.zeros((10, 10), np.complex64))


def test_reference_mask_fixture_operator(oracle):
    """The reference's random4 mask (data/mask_cartesian_random4.npy, stored
    as its line list) on its 768 x 384 grid with synthetic coils."""
    pdt = _pdt()
    from primaldualtoolbox import synthetic as syn
    m = np.load(os.path.join(ROOT, "tests", "golden", "masks.npz"))
    Hk, W = m["random4_shape"]
    mask = np.zeros((Hk, W), np.float32)
    mask[:, m["random4_lines"]] = 1
    rng = np.random.default_rng(9)
    S = syn.coil_sensitivities(rng, 4, Hk, W, size=384).astype(np.complex64)
    op = pdt.mri.MriCartesianRemoveROOSOperator()
    op.setCoilSens(S)
    op.setMask(mask)
    x = syn.complex_phantom(rng, W, Hk // 2)
    oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, S.astype(np.complex128), mask.astype(np.float64))
    y = oop.forward(x)
    assert rel_l2(op.forward(x.astype(np.complex64)), y) < 2e-5
    assert rel_l2(op.adjoint(y.astype(np.complex64)), oop.adjoint(y)) < 2e-5


def test_cpp_mirror_runs():
    src = os.path.join(ROOT, "tests", "cpp", "mirror_example.cpp")
    libdir = os.path.join(ROOT, "primal-dual-toolbox_amd", "primaldualtoolbox")
    exe = os.path.join("/tmp", f"pdt_mirror_{os.getpid()}")
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{os.path.join(ROOT, 'include')}", src, "-o", exe, f"-L{libdir}",
                    "-lpdt_hip", f"-Wl,-rpath,{libdir}", "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
