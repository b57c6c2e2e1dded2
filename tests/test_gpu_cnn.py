"""GPU parity of the warm-start network (bf16 MFMA convolutions, fp32
accumulation) against the FP64 oracle.

Tolerances: with the oracle rounding to bf16 at the same points as the GPU,
outputs within rel. L2 2e-3 and gradients within 2e-2 per layer (the
remaining differences are fp32-vs-fp64 sums that flip a bf16 rounding now
and then); against the unrounded FP64 network 5e-2.  Adam: the fp32 update
equals the FP64 closed form given the GPU gradient to 1e-6 absolute.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _ws():
    from primaldualtoolbox import warmstart as ws
    return ws


@pytest.mark.parametrize("n,h,w", [(2, 10, 12), (1, 7, 70), (3, 5, 129)])
def test_cnn_forward_matches_oracle(oracle, n, h, w):
    ws = _ws()
    net = ws.WarmStartCNN(seed=n * 100 + w)
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(w)
    x = rng.standard_normal((n, h, w, 2)).astype(np.float32)
    print(f"cnn forward n={n} h={h} w={w}", flush=True)
    y = net.forward(x)
    ch = net.channels
    assert rel_l2(y, oracle.cnn_forward(ch, p, x, bf16=True)) < 2e-3
    assert rel_l2(y, oracle.cnn_forward(ch, p, x)) < 5e-2


def test_cnn_gradients_match_oracle(oracle):
    ws = _ws()
    net = ws.WarmStartCNN(seed=7)
    ch = net.channels
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((2, 9, 70, 2)).astype(np.float32)
    t = (0.3 * rng.standard_normal((2, 9, 70, 6))).astype(np.float32)
    loss, g = net.loss_grad(x, t)
    ref_loss, ref = oracle.cnn_loss_grad(ch, p, x, t, bf16=True)
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    o = 0
    for l, (a, b) in enumerate(zip(ch[:-1], ch[1:])):
        nw = 9 * a * b
        ew, eb = rel_l2(g[o:o + nw], ref[o:o + nw]), rel_l2(g[o + nw:o + nw + b], ref[o + nw:o + nw + b])
        print(f"  layer {l} ({a}->{b}): dW rel {ew:.2e} db rel {eb:.2e}", flush=True)
        assert ew < 2e-2 and eb < 2e-2
        o += nw + b
    _, ref64 = oracle.cnn_loss_grad(ch, p, x, t)
    assert rel_l2(g, ref64) < 5e-2


def test_adam_steps_match_closed_form(oracle):
    ws = _ws()
    net = ws.WarmStartCNN(seed=3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((1, 6, 20, 2)).astype(np.float32)
    t = rng.standard_normal((1, 6, 20, 6)).astype(np.float32)
    w = net.get_params().astype(np.float64)
    m, v = np.zeros_like(w), np.zeros_like(w)
    for step in (1, 2):
        _, g = net.loss_grad(x, t)
        net.adam_step(lr=1e-3)
        oracle.adam(w, g.astype(np.float64), m, v, 1e-3, 0.9, 0.999, 1e-8, step)
        assert np.abs(net.get_params() - w).max() < 1e-6


def test_training_reduces_loss():
    ws = _ws()
    net = ws.WarmStartCNN(seed=5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((4, 16, 16, 2)).astype(np.float32)
    t = np.concatenate([0.5 * x, np.roll(x, 1, axis=2), -x], axis=-1).astype(np.float32)
    first = net.train_step(x, t, lr=3e-3)
    for _ in range(60):
        last = net.train_step(x, t, lr=3e-3)
    print(f"  loss {first:.4e} -> {last:.4e}", flush=True)
    assert last < 0.5 * first


def test_warm_start_sets_network_prediction():
    """pdt_solver_warm_start = u0 = A^*f + s r, v0 = s v' with (r, v') the
    network output on A^*f / s, s = RMS(A^*f)."""
    import primaldualtoolbox as pdt
    from primaldualtoolbox import synthetic as syn
    ws = _ws()
    case = syn.mri_case(seed=11, n=24, coils=3)
    op = pdt.mri.MriCartesianRemoveROOSOperator()
    op.setCoilSens(case["sens"])
    op.setMask(case["mask"])
    opt = pdt.mri.TgvMriOptimizer_2c3c()
    opt.setOperator(op)
    opt.setNoisyData(case["kspace"])
    net = ws.WarmStartCNN(seed=9)
    net.warm_start(opt)
    u, v = opt.getResult().copy(), opt.getV().copy()
    atf = op.adjoint(case["kspace"]).astype(np.complex128)
    s = np.sqrt(np.mean(np.abs(atf) ** 2))
    x = np.stack([atf.real, atf.imag], -1)[None] / s
    y = net.forward(x.astype(np.float32))[0]
    assert rel_l2(u, atf + s * (y[..., 0] + 1j * y[..., 1])) < 1e-5
    assert rel_l2(v[0], s * (y[..., 2] + 1j * y[..., 3])) < 1e-5
    assert rel_l2(v[1], s * (y[..., 4] + 1j * y[..., 5])) < 1e-5
    # and the solver runs from there
    opt.getParameters().max_iter = 49
    opt.solve()
    assert np.isfinite(opt.getResult()).all()


def test_trained_warm_start_reduces_iterations():
    """What the trained network gains: on a held-out slice (seed 3000, not in
    the training seeds 1000-2999) the solve from the network's (u, v) reaches
    the same relative gap 1e-4 in fewer iterations than the cold start, both
    within rel. L2 1e-3 of the FP64 oracle's solution."""
    import os
    import primaldualtoolbox as pdt
    from primaldualtoolbox import synthetic as syn
    ws = _ws()
    if not os.path.exists(ws.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights in the package")
    net = ws.WarmStartCNN.load()
    case = syn.mri320_case(3000)
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    res = {}
    for warm in (False, True):
        op = pdt.mri.MriCartesianRemoveROOSOperator()
        op.setCoilSens(case["sens"])
        op.setMask(case["mask"])
        opt = pdt.mri.TgvMriOptimizer_2c3c()
        opt.setOperator(op)
        opt.setNoisyData(case["kspace"])
        p = opt.getParameters()
        p.alpha1, p.alpha0, p.max_iter, p.gap_tol, p.gap_every = a1, 2 * a1, 100000, 1e-4, 50
        if warm:
            net.warm_start(opt)
        opt.solve()
        st = opt.getStats()
        assert st["converged"] == 1
        res[warm] = (int(st["iters"]), opt.getResult().copy())
    # both at the FP64 oracle's solution (its own stop at the same gap,
    # tests/golden/endstate_320x320x15.npz)
    ref = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                               "endstate_320x320x15.npz"))["u_3000"]
    err = {w: rel_l2(res[w][1], ref) for w in (False, True)}
    print(f"  iterations cold {res[False][0]} warm {res[True][0]}; rel L2 vs oracle cold {err[False]:.3e} "
          f"warm {err[True]:.3e}", flush=True)
    assert res[True][0] <= 0.9 * res[False][0]
    assert err[True] < 1e-3 and err[False] < 1e-3
