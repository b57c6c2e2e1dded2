"""Pins the oracle's TGV-family solve loops, the alpha continuation and the
primal-dual gaps against an independent numpy model (tests/pdhg_model.py,
built from stencil/operator matrices and the reference's loop constants),
and checks the partial gap's certificate against a converged solution.

Every check has a planted-mistake twin: the model with one plausible
transcription error (dual step first, the wrong step size, the wrong alpha
exponent, lambda outside the r prox, a dropped gap term, a zero ball radius)
must disagree with the oracle by far more than the tolerance, so the check
has the power to catch that error.
"""
import os

import numpy as np
import pytest

import pdhg_model as M
from conftest import rel_l2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _mri_case(oracle, seed, n, coils, center):
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=seed, n=n, coils=coils, center=center)
    f = case["kspace"].astype(np.complex128)
    op = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"], case["mask"])
    return case, f, op


# ---------------------------------------------------------------------------
# the model's operator is the numpy definition
# ---------------------------------------------------------------------------
def test_model_operator_is_numpy_definition(oracle):
    """The model's dense A reproduces the numpy-computed operator goldens,
    and its adjoint function is exactly A^H (conjugate transpose)."""
    g = np.load(os.path.join(GOLD, "operator.npz"))
    for key, roos in (("roos_10_10_5", True), ("cart_10_10_5", False), ("roos_6_14_2", True)):
        S, Mk = g[key + "_S"], g[key + "_M"]
        A = M.cartesian_matrix(S, Mk, roos)
        assert rel_l2(A @ g[key + "_x"].reshape(-1), g[key + "_Ax"].reshape(-1)) < 1e-13
        _, adj = M.cartesian_ops(S, Mk, roos)
        y = g[key + "_y"].reshape(-1)
        assert rel_l2(adj(y), A.conj().T @ y) < 1e-13
        assert rel_l2(adj(y), g[key + "_Aty"].reshape(-1)) < 1e-13


# ---------------------------------------------------------------------------
# TGV denoising (tgvoptimizer.cu:102-135): tau = sigma = 1/sqrt(12), primal step first
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("cplx", [False, True])
def test_tgv_denoising_loop_matches_model(oracle, cplx):
    rng = np.random.default_rng(11 + cplx)
    f = rng.standard_normal((7, 9)) + (1j * rng.standard_normal((7, 9)) if cplx else 0)
    lam, a0, a1 = 3.0, 0.3, 0.2
    for K in (1, 2, 3):
        u, v, _, _ = oracle.solve(oracle.TGV, f, lambda_=lam, alpha0=a0, alpha1=a1, max_iter=K - 1, want_v=True)
        m = M.tgv_denoise(f, lam, a0, a1, K)
        assert rel_l2(u, m["u"]) < 1e-13
        if K > 1:
            assert rel_l2(v, m["v"]) < 1e-13
    # planted mistakes: the dual step first, the ROF step sizes 1/sqrt(8)
    assert rel_l2(u, M.tgv_denoise(f, lam, a0, a1, 3, primal_first=False)["u"]) > 1e-3
    assert rel_l2(u, M.tgv_denoise(f, lam, a0, a1, 3, tau=1 / np.sqrt(8), sigma=1 / np.sqrt(8))["u"]) > 1e-3


def test_tgv_denoising_gap_matches_model(oracle):
    """Primal energy at (ubar, vbar) after the primal step, and the restored
    dual value from q after the dual step, of the evaluating iteration."""
    rng = np.random.default_rng(5)
    f = rng.standard_normal((8, 10)) + 1j * rng.standard_normal((8, 10))
    lam, a0, a1, K = 3.0, 0.3, 0.2, 6
    _, _, _, hist = oracle.solve(oracle.TGV, f, lambda_=lam, alpha0=a0, alpha1=a1, max_iter=K - 1, gap_every=K)
    m = M.tgv_denoise(f, lam, a0, a1, K)
    P, D = M.tgv_denoise_gap(f, m["ub"], m["vb"], m["p"], m["q"], lam, a0, a1)
    assert abs(hist[0, 0] - P) < 1e-12 * abs(P)
    assert abs(hist[0, 1] - D) < 1e-12 * abs(P)
    # planted mistake: the restoration without the feasibility scale
    S = M.Stencils(*f.shape)
    g = S.div(S.symdiv(m["q"]))
    D1 = M.redot(g, f) - float(np.sum(np.abs(g) ** 2)) / (2 * lam)
    assert abs(hist[0, 1] - D1) > 1e-3 * abs(P)


# ---------------------------------------------------------------------------
# TGV-MRI (tgvmrioptimizer.cu:127-198): tau 1/16, sigma 1/8, lambda 1, alpha continuation
# ---------------------------------------------------------------------------
def test_tgv_mri_loop_and_alpha_schedule_match_model(oracle):
    """Five iterations (max_iter K = 4) with reduction 0.1, alpha small enough
    that the projections are active: |p| reaches alpha1(k) at k = 0, K/2 and
    K, |q| reaches alpha0(k) at K/2 and K (q is 0 at k = 0), so the result
    depends on alpha_i(k) = alpha_i red^(k/K) at those k."""
    case, f, op = _mri_case(oracle, seed=1, n=8, coils=2, center=2)
    A = M.cartesian_matrix(case["sens"], case["mask"], True)
    nf = np.linalg.norm(f)
    a1, a0, red, K = 3e-3 * nf, 1e-4 * nf, 0.1, 4
    u, v, _, _ = oracle.solve(oracle.TGV_MRI, f, op=op, alpha1=a1, alpha0=a0, reduction=red, max_iter=K,
                              want_v=True)
    m = M.OpSolve(A, f, op.H, op.W, kind="mri", alpha0=a0, alpha1=a1, reduction=red)
    for k in range(K + 1):
        st = m.step(K)
        assert np.isclose(st["a1"], a1 * red ** (k / K), rtol=1e-14)
        if k in (0, K // 2, K):
            assert abs(M.norm_p(st["p"]).max() / st["a1"] - 1) < 1e-12
        if k in (K // 2, K):
            assert abs(M.norm_q(st["q"]).max() / st["a0"] - 1) < 1e-12
    assert rel_l2(u, m.u) < 1e-13 and rel_l2(v, m.v) < 1e-13
    # planted mistakes
    twins = {"tau = 1/8": dict(tau=1 / 8), "exponent (k+1)/K": dict(alpha_exponent=lambda k, K: (k + 1) / K),
             "reversed schedule": dict(alpha_exponent=lambda k, K: 1 - k / K)}
    for name, kw in twins.items():
        t = M.OpSolve(A, f, op.H, op.W, kind="mri", alpha0=a0, alpha1=a1, reduction=red, **kw)
        for _ in range(K + 1):
            t.step(K)
        assert rel_l2(u, t.u) > 1e-5, name  # 8 orders above the agreement


# ---------------------------------------------------------------------------
# TGV with operator (tgvoptimizerwithop.cu:123-176): tau = sigma = 1/8,
# r <- (r + sigma (A ubar - f)) / (1 + sigma lambda), fixed alpha
# ---------------------------------------------------------------------------
def test_tgv_with_operator_loop_matches_model(oracle):
    case, f, op = _mri_case(oracle, seed=3, n=8, coils=2, center=2)
    A = M.cartesian_matrix(case["sens"], case["mask"], True)
    nf = np.linalg.norm(f)
    lam, a1, a0, K = 0.7, 2e-3 * nf, 4e-3 * nf, 3
    u, v, _, _ = oracle.solve(oracle.TGV_OP, f, op=op, lambda_=lam, alpha1=a1, alpha0=a0, max_iter=K,
                              want_v=True)
    m = M.OpSolve(A, f, op.H, op.W, kind="op", lam=lam, alpha0=a0, alpha1=a1)
    for _ in range(K + 1):
        m.step(K)
    assert rel_l2(u, m.u) < 1e-13 and rel_l2(v, m.v) < 1e-13
    # planted mistakes: lambda left out of the r prox; the TGV-MRI step sizes
    t = M.OpSolve(A, f, op.H, op.W, kind="op", lam=lam, alpha0=a0, alpha1=a1, lam_in_prox=False)
    t2 = M.OpSolve(A, f, op.H, op.W, kind="op", lam=lam, alpha0=a0, alpha1=a1, tau=1 / 16)
    for _ in range(K + 1):
        t.step(K)
        t2.step(K)
    assert rel_l2(u, t.u) > 1e-4 and rel_l2(u, t2.u) > 1e-3


# ---------------------------------------------------------------------------
# the partial gap and its certificate
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cert(oracle):
    """The 24 x 24 x 3 case, its converged solution (u*, v*) (FP64 oracle,
    200000 iterations, tests/golden/make_certificate.py) and the model run
    up to the oracle's first evaluation with relative gap <= 1e-4."""
    z = np.load(os.path.join(GOLD, "certificate_24x24x3.npz"))
    case, f, op = _mri_case(oracle, int(z["seed"]), int(z["n"]), int(z["coils"]), int(z["center"]))
    a1, a0 = float(z["alpha1"]), float(z["alpha0"])
    every = 10
    _, _, st, hist = oracle.solve(oracle.TGV_MRI, f, op=op, alpha1=a1, alpha0=a0, max_iter=100000,
                                  gap_every=every, gap_tol=1e-4)
    ops = M.cartesian_ops(case["sens"], case["mask"], True)
    m = M.OpSolve(ops, f, op.H, op.W, kind="mri", alpha0=a0, alpha1=a1)
    rho_u = 0.01 * np.linalg.norm(m.adj(m.f))  # the default ball radii
    rho_v = np.sqrt(8) * rho_u
    evals = []
    for k in range(st["iters"]):
        s = m.step(100000)
        if (k + 1) % every == 0:
            P, G = m.partial_gap(s, rho_u, rho_v)
            evals.append(dict(k=k, P=P, G=G, **s))
    return dict(z=z, m=m, st=st, hist=hist, evals=evals, rho_u=rho_u, rho_v=rho_v, a0=a0, a1=a1)


def test_mri_partial_gap_matches_model(cert):
    """P(ubar, vbar) and G = Fenchel-Young gaps + rho_u |A^*r - div p| +
    rho_v |p + div_sym q| of every evaluation, model vs oracle."""
    hist, ev = cert["hist"], cert["evals"]
    assert len(ev) == len(hist) > 100
    P = np.array([e["P"] for e in ev])
    G = np.array([e["G"] for e in ev])
    assert np.max(np.abs(hist[:, 0] - P) / P) < 1e-12
    assert np.max(np.abs((hist[:, 0] - hist[:, 1]) - G) / G) < 1e-9
    # planted mistake: the data Fenchel-Young gap without lambda r
    e = ev[-1]
    m = cert["m"]
    d = m.fwd(e["ub"].reshape(-1)) - m.f
    wrong = e["G"] - float(np.sum(np.abs(d - e["r"]) ** 2)) / 2 + float(np.sum(np.abs(d) ** 2)) / 2
    assert abs(wrong - (hist[-1, 0] - hist[-1, 1])) > 1.0 * e["G"]


def test_partial_gap_certificate(cert):
    """At the first evaluation with relative gap <= 1e-4 (where the solve
    stops), the minimiser lies in the balls B(ubar, rho_u) x B(vbar, rho_v)
    and G bounds P(ubar, vbar) - P*."""
    z, m, ev = cert["z"], cert["m"], cert["evals"]
    first = next(e for e in ev if e["G"] / e["P"] <= 1e-4)
    assert first["k"] + 1 == cert["st"]["iters"]  # the oracle stopped at the same evaluation
    u_star, v_star = z["u"], z["v"]
    P_star = m.primal(u_star, v_star, cert["a0"], cert["a1"])
    du = np.linalg.norm(first["ub"] - u_star)
    dv = np.linalg.norm(first["vb"] - v_star)
    assert du <= cert["rho_u"] and dv <= cert["rho_v"]
    assert 0 <= first["P"] - P_star <= first["G"]
    # planted mistake: a zero ball radius certifies nothing (the minimiser is
    # outside a ball of radius 0)
    assert not (du <= 0.0 and dv <= 0.0)


def test_with_operator_partial_gap_matches_model(oracle):
    """With lambda != 1 the data term is 1/(2 lambda) |A u - f|^2 and its
    Fenchel-Young gap 1/(2 lambda) |A ubar - f - lambda r|^2."""
    case, f, op = _mri_case(oracle, seed=4, n=12, coils=2, center=2)
    nf = np.linalg.norm(f)
    lam, a1, a0, K, every = 0.6, 5e-3 * nf, 1e-2 * nf, 59, 20
    _, _, _, hist = oracle.solve(oracle.TGV_OP, f, op=op, lambda_=lam, alpha1=a1, alpha0=a0, max_iter=K,
                                 gap_every=every)
    m = M.OpSolve(M.cartesian_ops(case["sens"], case["mask"], True), f, op.H, op.W, kind="op", lam=lam,
                  alpha0=a0, alpha1=a1)
    rho_u = 0.01 * np.linalg.norm(m.adj(m.f))
    rows = []
    for k in range(K + 1):
        s = m.step(K)
        if (k + 1) % every == 0:
            rows.append(m.partial_gap(s, rho_u, np.sqrt(8) * rho_u))
    P, G = np.array(rows).T
    assert len(rows) == len(hist) == 3
    assert np.max(np.abs(hist[:, 0] - P) / P) < 1e-12
    assert np.max(np.abs((hist[:, 0] - hist[:, 1]) - G) / G) < 1e-9
    # planted mistake: lambda = 1 in the energies
    m.lam = 1.0
    assert abs(m.partial_gap(s, rho_u, np.sqrt(8) * rho_u)[0] - hist[-1, 0]) > 1e-2 * hist[-1, 0]
