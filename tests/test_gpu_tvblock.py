"""Temporally blocked ROF and TGV denoising launches (k iterations per launch
on tiles held in LDS with a k-pixel halo, ping-pong state) against the
one-iteration kernels: the same iterates bit for bit, for real and complex,
FP32 and FP64 images, FP16-stored TGV duals, sizes that no tile divides,
tiles larger than the image, and solves whose chunks end in a gap
evaluation; and against the FP64 oracle."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _solve(f, tiles, max_iter, gap_every=0, lam=10.0):
    import primaldualtoolbox as pdt
    opt = pdt.denoising.TvOptimizer_2c(dtype=f.dtype) if np.iscomplexobj(f) else \
        pdt.denoising.TvOptimizer_2f(dtype=f.dtype)
    opt.setNoisyData(f)
    p = opt.getParameters()
    p.Lambda, p.max_iter, p.gap_every = lam, max_iter, gap_every
    pdt.denoising.set_tiles([opt], rpi=1, tv_block=tiles)
    opt.solve()
    return opt.getResult().copy(), opt.getStats()


@pytest.mark.parametrize("dtype", [np.float32, np.complex64, np.float64, np.complex128])
@pytest.mark.parametrize("shape,tiles", [((256, 256), (32, 16, 8)), ((37, 100), (16, 16, 4)),
                                         ((29, 13), (7, 5, 3)), ((20, 24), (64, 32, 8))])
def test_blocked_launches_equal_one_iteration_kernels(dtype, shape, tiles):
    rng = np.random.default_rng(sum(shape))
    f = rng.standard_normal(shape)
    if np.dtype(dtype).kind == "c":
        f = f + 1j * rng.standard_normal(shape)
    f = f.astype(dtype)
    print(f"tv block {np.dtype(dtype).name} {shape} tiles {tiles}", flush=True)
    for max_iter, gap_every in ((49, 0), (99, 50), (36, 0)):
        u0, s0 = _solve(f, None, max_iter, gap_every)
        u1, s1 = _solve(f, tiles, max_iter, gap_every)
        assert s0["iters"] == s1["iters"] == max_iter + 1
        np.testing.assert_array_equal(u0, u1)
        if gap_every:
            assert s0["rel_gap"] == s1["rel_gap"]


def test_blocked_launches_match_oracle(oracle):
    """ROF 256^2, lambda 10, 500 iterations through the blocked launches,
    against the FP64 oracle (rel. L2 <= 1e-3, gap within 1 %)."""
    from primaldualtoolbox import synthetic as syn
    _, noisy = syn.tv_denoising_case(0, 256)
    u, st = _solve(noisy.astype(np.float32), (32, 16, 8), 499, gap_every=500)
    ref, _, ost, _ = oracle.solve(oracle.TV, noisy.astype(np.float64), lambda_=10, max_iter=499, gap_every=500)
    err = rel_l2(u, ref)
    print(f"tv 256 blocked: rel L2 {err:.3e}, gap {st['rel_gap']:.5e} vs {ost['rel_gap']:.5e}", flush=True)
    assert err < 1e-3
    assert abs(st["rel_gap"] - ost["rel_gap"]) <= 0.01 * ost["rel_gap"]


def _solve_tgv(f, tiles, max_iter, gap_every=0, fp16=False, prm=(8.0, 0.4, 0.2)):
    import primaldualtoolbox as pdt
    opt = pdt.denoising.TgvOptimizer_2c(dtype=f.dtype) if np.iscomplexobj(f) else \
        pdt.denoising.TgvOptimizer_2f(dtype=f.dtype)
    opt.setNoisyData(f)
    p = opt.getParameters()
    p.Lambda, p.alpha0, p.alpha1 = prm
    p.max_iter, p.gap_every, p.dual_fp16 = max_iter, gap_every, int(fp16)
    pdt.denoising.set_tiles([opt], rpi=1, tv_block=tiles, dual_px=1)
    opt.solve()
    return opt.getResult().copy(), opt.getV().copy(), opt.getStats()


@pytest.mark.parametrize("prm", [(8.0, 0.4, 0.2), (20.0, 2.0, 1.0)])
@pytest.mark.parametrize("dtype", [np.float32, np.complex64, np.float64, np.complex128])
@pytest.mark.parametrize("shape,tiles", [((128, 96), (16, 16, 4)), ((37, 100), (32, 8, 2)), ((29, 13), (7, 5, 3)),
                                         ((20, 24), (64, 32, 3)), ((130, 100), (32, 32, 4)),
                                         ((200, 160), (32, 32, 6))])
def test_tgv_blocked_launches_equal_one_iteration_kernels(dtype, shape, tiles, prm):
    """(lambda, alpha0, alpha1): a power-of-two lambda makes lambda f exact,
    lambda = 20 does not (a kernel that rounded lambda f on its own, instead
    of fusing it into div + lambda f, would differ there only)."""
    rng = np.random.default_rng(7 + sum(shape))
    f = rng.standard_normal(shape)
    if np.dtype(dtype).kind == "c":
        f = f + 1j * rng.standard_normal(shape)
    f = f.astype(dtype)
    print(f"tgv block {np.dtype(dtype).name} {shape} tiles {tiles} parameters {prm}", flush=True)
    for max_iter, gap_every in ((49, 0), (99, 50), (36, 0)):
        u0, v0, s0 = _solve_tgv(f, None, max_iter, gap_every, prm=prm)
        u1, v1, s1 = _solve_tgv(f, tiles, max_iter, gap_every, prm=prm)
        assert s0["iters"] == s1["iters"] == max_iter + 1
        bad = np.argwhere(u0 != u1)
        if len(bad):
            print(f"  {len(bad)} pixels differ after {max_iter + 1} iterations, first at {bad[0]}, "
                  f"max |du| {np.abs(u0 - u1).max():.3e}", flush=True)
        np.testing.assert_array_equal(u0, u1)
        np.testing.assert_array_equal(v0, v1)
        if gap_every:
            assert s0["rel_gap"] == s1["rel_gap"]


@pytest.mark.parametrize("cplx", [False, True])
def test_tgv_blocked_launches_fp16_duals_equal_one_iteration_kernels(cplx):
    """With p, q stored as half, the blocked launch rounds them to half after
    every dual step in LDS: the iterates equal the one-iteration kernels'."""
    rng = np.random.default_rng(11)
    f = rng.standard_normal((70, 52)) + (1j * rng.standard_normal((70, 52)) if cplx else 0)
    f = f.astype(np.complex64 if cplx else np.float32)
    for max_iter, gap_every, prm in ((59, 30, (8.0, 0.4, 0.2)), (40, 0, (20.0, 2.0, 1.0))):
        u0, v0, s0 = _solve_tgv(f, None, max_iter, gap_every, fp16=True, prm=prm)
        u1, v1, s1 = _solve_tgv(f, (16, 16, 4), max_iter, gap_every, fp16=True, prm=prm)
        assert s0["iters"] == s1["iters"] == max_iter + 1
        np.testing.assert_array_equal(u0, u1)
        np.testing.assert_array_equal(v0, v1)
        if gap_every:
            assert s0["rel_gap"] == s1["rel_gap"]


def test_tgv_blocked_launches_match_oracle(oracle):
    """TGV 512^2 complex, lambda 20, alpha0 2, alpha1 1, 1000 iterations
    through the blocked launches, against the FP64 oracle (rel. L2 <= 1e-3,
    gap within 1 %)."""
    import primaldualtoolbox as pdt
    from primaldualtoolbox import synthetic as syn
    _, noisy = syn.tgv_denoising_case(seed=0, n=512)
    opt = pdt.denoising.TgvOptimizer_2c()
    opt.setNoisyData(noisy.astype(np.complex64))
    p = opt.getParameters()
    p.Lambda, p.alpha0, p.alpha1, p.max_iter, p.gap_every = 20, 2, 1, 999, 1000
    pdt.denoising.set_tiles([opt], rpi=1, tv_block=(16, 16, 4))
    opt.solve()
    st = opt.getStats()
    ref, _, ost, _ = oracle.solve(oracle.TGV, noisy.astype(np.complex128), lambda_=20, alpha0=2, alpha1=1,
                                  max_iter=999, gap_every=1000)
    err = rel_l2(opt.getResult(), ref)
    print(f"tgv 512 blocked: rel L2 {err:.3e}, gap {st['rel_gap']:.5e} vs {ost['rel_gap']:.5e}", flush=True)
    assert err < 1e-3
    assert abs(st["rel_gap"] - ost["rel_gap"]) <= 0.01 * ost["rel_gap"]
