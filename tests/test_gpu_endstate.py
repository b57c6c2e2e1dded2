"""Full-size end state of the 320 x 320, 15-coil Cartesian TGV-MRI workload
against the FP64 oracle: the GPU solve to relative partial gap 1e-4
(checked every 50 iterations), from the cold start and from the trained
network's warm start, must land within rel. L2 1e-3 of the oracle's solution
at the oracle's own stopping point (tests/golden/endstate_320x320x15.npz,
written by tests/golden/make_endstate.py), and the cold solve must stop at
the oracle's iteration give or take one evaluation interval.
"""
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "endstate_320x320x15.npz")
GAP_TOL, GAP_EVERY = 1e-4, 50
# the seeds with an oracle end state: 0, the benchmark slice, and held-out
# slices of the warm-start network (3000 and up)
with np.load(GOLD) as _z:
    SEEDS = sorted(int(k[6:]) for k in _z.files if k.startswith("iters_"))
HELD_OUT = [s for s in SEEDS if s >= 3000]


@pytest.fixture(scope="module")
def endstate():
    return np.load(GOLD)


def _solve(seed, a1, warm_net=None):
    import primaldualtoolbox as pdt
    from primaldualtoolbox import synthetic as syn
    case = syn.mri320_case(seed)  # alpha1 from the fixture (the generator's data may differ by ulps across CPUs)
    op = pdt.mri.MriCartesianRemoveROOSOperator()
    op.setCoilSens(case["sens"])
    op.setMask(case["mask"])
    opt = pdt.mri.TgvMriOptimizer_2c3c()
    opt.setOperator(op)
    opt.setNoisyData(case["kspace"])
    p = opt.getParameters()
    p.alpha1, p.alpha0, p.reduction, p.max_iter = a1, 2 * a1, 1.0, 100000
    p.gap_tol, p.gap_every = GAP_TOL, GAP_EVERY
    if warm_net is not None:
        warm_net.warm_start(opt)
    opt.solve()
    st = opt.getStats()
    assert st["converged"] == 1 and st["rel_gap"] <= GAP_TOL
    return opt.getResult().copy(), int(st["iters"])


@pytest.mark.parametrize("seed", SEEDS)
def test_cold_solve_to_gap_matches_oracle_end_state(endstate, seed):
    u, it = _solve(seed, float(endstate[f"alpha1_{seed}"]))
    ref, it_ref = endstate[f"u_{seed}"], int(endstate[f"iters_{seed}"])
    err = rel_l2(u, ref)
    print(f"seed {seed}: GPU cold {it} iterations, FP64 oracle {it_ref}; rel L2 vs oracle {err:.3e}", flush=True)
    assert abs(it - it_ref) <= GAP_EVERY
    assert err < 1e-3


@pytest.mark.parametrize("seed", HELD_OUT)
def test_warm_start_solve_to_gap_matches_oracle_end_state(endstate, seed):
    """Held-out slices (the network trained on seeds 1000-2999): the solve
    from the network's (u, v) reaches the same gap, at the oracle's solution,
    in fewer iterations than the oracle's cold solve."""
    from primaldualtoolbox import warmstart as ws
    if not os.path.exists(ws.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights in the package")
    net = ws.WarmStartCNN.load()
    u, it = _solve(seed, float(endstate[f"alpha1_{seed}"]), warm_net=net)
    ref, it_ref = endstate[f"u_{seed}"], int(endstate[f"iters_{seed}"])
    err = rel_l2(u, ref)
    print(f"seed {seed}: GPU warm {it} iterations, oracle cold {it_ref}; rel L2 vs oracle {err:.3e}", flush=True)
    assert it < it_ref
    assert err < 1e-3
