"""Pins the FP64 CPU oracle (CPU only, no GPU).

Each check is chosen so that a plausible transcription mistake fails it:
golden numpy vectors and brute-force DFT matrices for the FFT and the shift
conventions; the reference's own adjointness test (relative 1e-12 in FP64) and
numpy-computed operator outputs for the three Cartesian operators; exact
adjoint identities and brute-force matrices for the finite differences;
operator norms for the step-size conditions; closed-form ROF solutions and
fixed points for the solvers; non-negativity and decay of the gaps.  The last
tests plant mistakes in copies of the oracle's inputs/outputs to show that the
checks catch them.
"""
import os

import numpy as np
import pytest

from conftest import rel_l2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def fft_gold():
    return np.load(os.path.join(GOLD, "fft.npz"))


@pytest.fixture(scope="module")
def op_gold():
    return np.load(os.path.join(GOLD, "operator.npz"))


# ---------------------------------------------------------------------------
# FFT and shift conventions
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 10, 12, 16, 23, 46, 64, 320, 368, 384, 640, 736, 768])
def test_fft1_matches_numpy_golden(oracle, fft_gold, n):
    x = fft_gold[f"x{n}"]
    assert rel_l2(oracle.fft1(x, -1), fft_gold[f"F{n}"]) < 1e-13
    assert rel_l2(oracle.fft1(x, +1), fft_gold[f"I{n}"]) < 1e-13


@pytest.mark.parametrize("n", [2, 3, 5, 6, 7, 9, 11, 12])
def test_fft1_matches_dft_matrix(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    D = np.exp(-2j * np.pi * np.outer(j, j) / n)
    assert rel_l2(oracle.fft1(x, -1), D @ x) < 1e-13
    assert rel_l2(oracle.fft1(x, +1), D.conj() @ x) < 1e-13


@pytest.mark.parametrize("shape", [(10, 7), (12, 10), (5, 6), (9, 9)])
def test_fft2c_matches_numpy_golden(oracle, fft_gold, shape):
    h, w = shape
    z = fft_gold[f"z{h}x{w}"]
    assert rel_l2(oracle.fft2c(z, -1), fft_gold[f"Z{h}x{w}"]) < 1e-13
    assert rel_l2(oracle.fft2c(z, +1), fft_gold[f"IZ{h}x{w}"]) < 1e-13


@pytest.mark.parametrize("shape", [(7, 5), (6, 4), (5, 8), (640, 320)])
def test_centered_fft_of_centre_delta_is_flat(oracle, shape):
    """ifftshift moves index n//2 to 0, so the centered orthonormal transform
    of a delta at (H//2, W//2) is the constant 1/sqrt(HW), odd or even."""
    h, w = shape
    z = np.zeros(shape)
    z[h // 2, w // 2] = 1
    assert np.abs(oracle.fft2c(z) - 1 / np.sqrt(h * w)).max() < 1e-14


def test_fft2c_parseval_and_inverse(oracle):
    rng = np.random.default_rng(3)
    z = rng.standard_normal((22, 13)) + 1j * rng.standard_normal((22, 13))
    Z = oracle.fft2c(z)
    assert abs(np.linalg.norm(Z) - np.linalg.norm(z)) < 1e-12 * np.linalg.norm(z)
    assert rel_l2(oracle.fft2c(Z, +1), z) < 1e-13


# ---------------------------------------------------------------------------
# Cartesian operators
# ---------------------------------------------------------------------------
OP_CASES = ["cart_10_10_5", "roos_10_10_5", "samp_10_10_1", "cart_7_9_3", "roos_6_14_2"]


def _op(oracle, g, key):
    kind = {"cart": oracle.OP_CARTESIAN, "roos": oracle.OP_CARTESIAN_ROOS, "samp": oracle.OP_SAMPLING}[key[:4]]
    return oracle.Op(kind, None if key.startswith("samp") else g[key + "_S"], g[key + "_M"])


@pytest.mark.parametrize("key", OP_CASES)
def test_operator_matches_numpy_golden(oracle, op_gold, key):
    op = _op(oracle, op_gold, key)
    assert rel_l2(op.forward(op_gold[key + "_x"]), op_gold[key + "_Ax"]) < 1e-13
    assert rel_l2(op.adjoint(op_gold[key + "_y"]), op_gold[key + "_Aty"]) < 1e-13


@pytest.mark.parametrize("key", OP_CASES)
def test_operator_adjointness_fp64(oracle, op_gold, key):
    """tests/main_adjointness_test.cu:28-67 asserted (relative 1e-12)."""
    op = _op(oracle, op_gold, key)
    x, y = op_gold[key + "_x"], op_gold[key + "_y"]
    lhs = np.vdot(y.reshape(-1), op.forward(x).reshape(-1))
    rhs = np.vdot(op.adjoint(y).reshape(-1), x.reshape(-1))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_roos_window_offsets(oracle):
    """FE_idx = unsigned(2 Ny * 0.25f + 1) (mricartesianoperator.cu:341)."""
    from primaldualtoolbox import synthetic as syn
    expect = {10: 3, 14: 4, 640: 161, 736: 185, 768: 193, 512: 129}
    for hk, y0 in expect.items():
        assert syn.roos_offset(hk) == y0
        # the oracle places image row 0 at FE row y0: probe with a row delta
        op = oracle.Op(oracle.OP_CARTESIAN_ROOS, np.ones((1, hk, 4)), np.ones((hk, 4)))
        x = np.zeros((hk // 2, 4), complex)
        x[0, :] = 1.0
        y = oracle.fft2c(op.forward(x)[0], +1)  # back to the padded image
        rows = np.flatnonzero(np.abs(y).sum(axis=1) > 1e-9)
        assert rows.tolist() == [y0]


def test_operator_norm_at_most_one(oracle):
    """||A|| <= 1 when sum_c |S_c|^2 = 1 and the mask is binary."""
    from primaldualtoolbox import synthetic as syn
    rng = np.random.default_rng(0)
    S = syn.coil_sensitivities(rng, 4, 32, 16)
    M = syn.line_mask(rng, 32, 16, accel=2, center=4)
    op = oracle.Op(oracle.OP_CARTESIAN_ROOS, S, M)
    x = rng.standard_normal((16, 16)) + 0j
    for _ in range(60):
        x = op.adjoint(op.forward(x))
        nrm = np.linalg.norm(x)
        x /= nrm
    assert nrm <= 1 + 1e-12
    assert nrm > 0.5


# ---------------------------------------------------------------------------
# finite differences
# ---------------------------------------------------------------------------
def _mat(oracle, which, shape_in, cplx=False):
    n = int(np.prod(shape_in))
    cols = []
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1
        cols.append(np.asarray(oracle.stencil(which, e.reshape(shape_in))).reshape(-1))
    return np.array(cols).T


@pytest.mark.parametrize("H,W", [(4, 5), (6, 7), (2, 3)])
def test_grad_div_brute_force_adjoint(oracle, H, W):
    """div = -grad^T exactly (tgv_kernels.cuh:40-82) for W, H >= 2."""
    G = _mat(oracle, 0, (H, W))
    Dv = _mat(oracle, 1, (2, H, W))
    assert np.abs(Dv + G.T).max() == 0.0


@pytest.mark.parametrize("H,W", [(4, 5), (3, 6)])
def test_symgrad_symdiv_brute_force_adjoint(oracle, H, W):
    """<E v, q>_w = -<v, div_sym q> with weight 2 on the xy plane."""
    E = _mat(oracle, 2, (2, H, W))
    Dq = _mat(oracle, 3, (3, H, W))
    wts = np.concatenate([np.ones(2 * H * W), 2 * np.ones(H * W)])
    assert np.abs(Dq + (wts[:, None] * E).T).max() < 1e-15


def test_size_one_axis_quirk(oracle):
    """dm with a size-1 axis returns p(0) (tgv_kernels.cuh:72-81), which is
    not -grad^T: the reference behaviour is kept, not repaired."""
    G = _mat(oracle, 0, (1, 4))
    Dv = _mat(oracle, 1, (2, 1, 4))
    assert np.abs(Dv + G.T).max() > 0.5


def test_stencil_norms(oracle):
    """||grad||^2 < 8 and ||K_TGV||^2 < 12, justifying tau sigma ||K||^2 < 1 for
    tau = sigma = 1/sqrt(8) and 1/sqrt(12)."""
    rng = np.random.default_rng(1)
    H = W = 24
    u = rng.standard_normal((H, W))
    for _ in range(300):
        u = -oracle.stencil(1, oracle.stencil(0, u))
        lam = np.linalg.norm(u)
        u /= lam
    assert 7.5 < lam < 8.0
    # K(u, v) = (grad u - v, E v); K^T (p, q) = (-div p, -p - div_sym q)
    # (div_sym is the adjoint under the xy-weighted pairing, so no weights here)
    u = rng.standard_normal((H, W))
    v = rng.standard_normal((2, H, W))
    for _ in range(300):
        p = oracle.stencil(0, u) - v
        q = oracle.stencil(2, v)
        u2 = -oracle.stencil(1, p)
        v2 = -p - oracle.stencil(3, q)
        nrm = np.sqrt(np.sum(u2 ** 2) + np.sum(v2 ** 2))
        u, v = u2 / nrm, v2 / nrm
    assert 10.0 < nrm < 12.0


# ---------------------------------------------------------------------------
# solvers
# ---------------------------------------------------------------------------
def test_rof_step_closed_form(oracle):
    """N x 2 image, step of height h at x = m: the ROF minimiser is the
    plateau pair 1/(lambda m), h - 1/(lambda (N - m)) (tvoptimizer.cu loop)."""
    N, m, lam, h = 40, 15, 2.0, 1.0
    f = np.zeros((2, N))
    f[:, m:] = h
    u, _, st, _ = oracle.solve(oracle.TV, f, lambda_=lam, max_iter=19999)
    assert st["iters"] == 20000
    assert np.abs(u[:, :m] - 1 / (lam * m)).max() < 1e-12
    assert np.abs(u[:, m:] - (h - 1 / (lam * (N - m)))).max() < 1e-12


@pytest.mark.parametrize("cplx", [False, True])
def test_constant_image_is_fixed_point(oracle, cplx):
    f = np.full((9, 11), 0.7 + (0.2j if cplx else 0))
    u, _, _, _ = oracle.solve(oracle.TV, f, lambda_=3, max_iter=50)
    assert np.abs(u - f).max() < 1e-14
    u, v, _, _ = oracle.solve(oracle.TGV, f, lambda_=3, alpha0=2, alpha1=1, max_iter=50, want_v=True)
    assert np.abs(u - f).max() < 1e-14 and np.abs(v).max() < 1e-14


def test_max_iter_semantics(oracle):
    """k = 0..max_iter: max_iter + 1 iterations (tvoptimizer.cu:107)."""
    f = np.random.default_rng(0).standard_normal((6, 5))
    _, _, st, _ = oracle.solve(oracle.TV, f, max_iter=7)
    assert st["iters"] == 8


def test_rof_gap_nonnegative_and_decreasing(oracle):
    from primaldualtoolbox import synthetic as syn
    _, noisy = syn.tv_denoising_case(0, 64)
    _, _, _, hist = oracle.solve(oracle.TV, noisy.astype(np.float64), lambda_=10, max_iter=1999, gap_every=100)
    gaps = hist[:, 0] - hist[:, 1]
    assert np.all(gaps >= -1e-9)
    assert gaps[-1] < 0.2 * gaps[0]


def test_tgv_restored_gap_nonnegative(oracle):
    from primaldualtoolbox import synthetic as syn
    _, noisy = syn.tgv_denoising_case(0, 48)
    _, _, _, hist = oracle.solve(oracle.TGV, noisy.astype(np.complex128), lambda_=20, alpha0=2, alpha1=1,
                                 max_iter=1999, gap_every=200)
    gaps = hist[:, 0] - hist[:, 1]
    assert np.all(gaps >= -1e-9)
    assert gaps[-1] < 0.3 * gaps[0]


def test_mri_partial_gap_terms(oracle):
    """Partial gap = Fenchel-Young gaps + rho * stationarity residuals, each
    >= 0, and it decreases; reduction = 1 keeps alpha constant."""
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=2, n=24, coils=3, center=4)
    op = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"], case["mask"])
    f = case["kspace"].astype(np.complex128)
    a1 = 1e-2 * np.linalg.norm(f)
    _, _, st, hist = oracle.solve(oracle.TGV_MRI, f, op=op, alpha1=a1, alpha0=2 * a1, max_iter=1999,
                                  gap_every=100)
    P, D, fy, ru, rv = hist[:, 0], hist[:, 1], hist[:, 3], hist[:, 4], hist[:, 5]
    assert np.all(fy >= -1e-12) and np.all(ru >= 0) and np.all(rv >= 0)
    assert np.all(P - D >= -1e-12)
    assert (P - D)[-1] < 0.05 * (P - D)[0]


def test_mri_gap_reaches_tolerance(oracle):
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=5, n=24, coils=3, center=4)
    op = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"], case["mask"])
    f = case["kspace"].astype(np.complex128)
    a1 = 1e-2 * np.linalg.norm(f)
    _, _, st, _ = oracle.solve(oracle.TGV_MRI, f, op=op, alpha1=a1, alpha0=2 * a1, max_iter=20000,
                               gap_every=50, gap_tol=1e-3)
    assert st["rel_gap"] <= 1e-3 and st["iters"] < 20001 and st["iters"] % 50 == 0


def test_mri_cold_start_is_adjoint(oracle):
    """With max_iter = 0 and alpha continuation off, one iteration from u0 =
    A^* f: u1 = u0 + tau (div p1 - A^* r1), a check of the first iteration
    against quantities computed independently here."""
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=1, n=8, coils=2, center=2)
    op = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"], case["mask"])
    f = case["kspace"].astype(np.complex128)
    u0 = op.adjoint(f)
    # TvOptimizerWithOp, lambda = 1: r1 = (0 + s (A u0 - f)) / (1 + s); p1 = proj_1(s grad u0)
    s = t = 1 / 8
    r1 = s * (op.forward(u0) - f) / (1 + s)
    g = oracle.stencil(0, u0)
    a = s * g
    nrm = np.maximum(1.0, np.sqrt(np.abs(a[0]) ** 2 + np.abs(a[1]) ** 2))
    p1 = a / nrm
    u1 = u0 + t * (oracle.stencil(1, p1) - op.adjoint(r1))
    u, _, _, _ = oracle.solve(oracle.TV_OP, f, op=op, lambda_=1.0, max_iter=0)
    assert rel_l2(u, u1) < 1e-13


# ---------------------------------------------------------------------------
# planted mistakes: the checks above must fail on them
# ---------------------------------------------------------------------------
def test_planted_shift_mistake_is_caught(oracle, fft_gold):
    z = fft_gold["z9x9"]
    wrong = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(z), norm="ortho"))  # fftshift instead of ifftshift
    assert rel_l2(wrong, fft_gold["Z9x9"]) > 1e-3
    assert rel_l2(oracle.fft2c(z), fft_gold["Z9x9"]) < 1e-13


def test_planted_conjugation_mistake_is_caught(oracle, op_gold):
    key = "cart_10_10_5"
    g = dict(op_gold)
    g[key + "_S"] = np.conj(op_gold[key + "_S"])  # conj on the wrong side
    op_wrong = _op(oracle, g, key)
    assert rel_l2(op_wrong.forward(op_gold[key + "_x"]), op_gold[key + "_Ax"]) > 1e-3


def test_planted_boundary_mistake_is_caught(oracle):
    G = _mat(oracle, 0, (4, 5))
    Dv = _mat(oracle, 1, (2, 4, 5))
    Dv_wrong = Dv.copy()
    Dv_wrong[0, 0] = 0.0  # drop the first-column term p(0) of dm
    assert np.abs(Dv_wrong + G.T).max() > 0.5


def test_reference_mask_fixtures():
    """The reference's sampling masks (data/mask_cartesian_*.npy): 768 x 384,
    constant along FE, binary, 117 lines; random4 has a fully sampled centre
    run 175..206 and regular4 every 4th line plus centre 178..205."""
    m = np.load(os.path.join(GOLD, "masks.npz"))
    for name in ["random4", "regular4"]:
        assert m[f"{name}_shape"].tolist() == [768, 384]
        assert m[f"{name}_values"].tolist() == [0.0, 1.0]
        assert len(m[f"{name}_lines"]) == 117
    r = set(m["random4_lines"].tolist())
    assert set(range(175, 207)) <= r
    g = set(m["regular4_lines"].tolist())
    assert set(range(0, 384, 4)) <= g and set(range(178, 206)) <= g
