"""Temporally blocked ROF launches (k iterations per launch on tiles held in
LDS with a k-pixel halo, ping-pong state) against the one-iteration kernels:
the same iterates bit for bit, for real and complex, FP32 and FP64 images,
sizes that no tile divides, tiles larger than the image, and solves whose
chunks end in a gap evaluation; and against the FP64 oracle."""
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
