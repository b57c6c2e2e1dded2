"""MATLAB entry points and the reference's demo geometry on the GPU.

* primaldualtoolbox.matlab (MATLAB indexing, double in/out) against the
  classes in numpy layout: identical values, since both run the same library
  calls on the same single- or double-precision data;
* include/primaldualtoolbox/matlab.hpp through a C++ program with column-major
  arrays against the Python entries: identical values;
* the double-precision reconstruction entry against the FP64 oracle
  (rel. L2 <= 1e-9);
* the geometry of the reference's Cartesian demo (bin/mri_cartesian_test.py:
  768 x 384 read-out oversampled grid, its random4 mask of 117 PE lines,
  15 coils) in single precision against the FP64 oracle after 20 iterations
  (rel. L2 <= 1e-3).
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _case(n=24, n_pe=20, coils=3, seed=5):
    from primaldualtoolbox import synthetic as syn
    return syn.mri_case(seed=seed, n=n, coils=coils, n_pe=n_pe, center=6)


def _to_m(case, roos=True):
    """MATLAB-indexed double arrays of a case, and an image for the forward
    operator (the phantom, or a random image on the full grid without ROOS)."""
    from primaldualtoolbox.dataio import to_matlab
    s = to_matlab(case["sens"]).astype(np.complex128)
    f = to_matlab(case["kspace"]).astype(np.complex128)
    m = case["mask"].astype(np.float64)
    if roos:
        img = case["image"]
    else:
        rng = np.random.default_rng(1)
        img = rng.standard_normal(m.shape) + 1j * rng.standard_normal(m.shape)
    return s, f, m, np.ascontiguousarray(img, np.complex128)


@pytest.mark.parametrize("roos", [True, False])
def test_matlab_operators_equal_class_api(roos):
    import primaldualtoolbox as pdt
    from primaldualtoolbox import matlab as mx
    from primaldualtoolbox.dataio import to_matlab
    case = _case()
    s, f, m, img = _to_m(case, roos)
    cls = pdt.mri.MriCartesianRemoveROOSOperator if roos else pdt.mri.MriCartesianOperator
    op = cls()
    op.setCoilSens(case["sens"])
    op.setMask(case["mask"])
    fwd = (mx.gpuMriCartesianRemoveROOSFwd if roos else mx.gpuMriCartesianFwd)(s, img, m)
    adj = (mx.gpuMriCartesianRemoveROOSAdj if roos else mx.gpuMriCartesianAdj)(s, f, m)
    assert fwd.dtype == np.complex128 and fwd.shape == f.shape
    assert np.array_equal(fwd, to_matlab(op.forward(img.astype(np.complex64))))
    assert np.array_equal(adj, op.adjoint(case["kspace"]))
    print(f"roos={roos}: forward {fwd.shape}, adjoint {adj.shape} equal to the class API", flush=True)


def test_matlab_tgv_equals_class_api_and_oracle(oracle):
    import primaldualtoolbox as pdt
    from primaldualtoolbox import matlab as mx
    case = _case()
    s, f, m, _ = _to_m(case)
    a1 = 1e-3 * float(np.linalg.norm(case["kspace"]))
    prm = {"alpha0": 2 * a1, "alpha1": a1, "max_iter": 30, "reduction": 0.1}
    u = mx.gpuTgvMriOptimizer_cartesianRemoveROOS(s, f, m, prm)
    op = pdt.mri.MriCartesianRemoveROOSOperator(dtype=np.complex128)
    op.setCoilSens(case["sens"].astype(np.complex128))
    op.setMask(case["mask"].astype(np.float64))
    opt = pdt.mri.TgvMriOptimizer_2c3c(dtype=np.complex128)
    p = opt.getParameters()
    p.alpha0, p.alpha1, p.max_iter, p.reduction = 2 * a1, a1, 30, 0.1
    opt.setOperator(op)
    opt.setNoisyData(case["kspace"].astype(np.complex128))
    opt.solve()
    assert np.array_equal(u, opt.getResult())
    oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"].astype(np.complex128), case["mask"].astype(np.float64))
    ref, _, _, _ = oracle.solve(oracle.TGV_MRI, case["kspace"].astype(np.complex128), op=oop, alpha1=a1,
                                alpha0=2 * a1, reduction=0.1, max_iter=30)
    err = rel_l2(u, ref)
    print(f"gpuTgvMriOptimizer_cartesianRemoveROOS vs FP64 oracle: rel L2 {err:.3e}", flush=True)
    assert err < 1e-9


def test_matlab_hpp_program_equals_python_entries(tmp_path):
    from primaldualtoolbox import matlab as mx
    case = _case(n=20, n_pe=18, coils=4, seed=8)
    s, f, m, img = _to_m(case)
    nfe_k, npe, nc = s.shape
    libdir = os.path.join(ROOT, "primal-dual-toolbox_amd", "primaldualtoolbox")
    exe = str(tmp_path / "matlab_entry")
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{os.path.join(ROOT, 'include')}",
                    os.path.join(ROOT, "tests", "cpp", "matlab_entry.cpp"), "-o", exe, f"-L{libdir}", "-lpdt_hip",
                    f"-Wl,-rpath,{libdir}", "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    a1 = 1e-3 * float(np.linalg.norm(case["kspace"]))
    prm = {"alpha0": 2 * a1, "alpha1": a1, "max_iter": 15, "reduction": 1.0}
    col = lambda a: np.asfortranarray(a).ravel(order="F")  # noqa: E731  column-major memory
    with open(tmp_path / "in.bin", "wb") as fh:
        np.array([nfe_k, npe, nc, 1, prm["max_iter"]], np.int64).tofile(fh)
        np.array([prm["alpha0"], prm["alpha1"], prm["reduction"]], np.float64).tofile(fh)
        for a in (s, img, f, m):
            col(a).tofile(fh)
    r = subprocess.run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(tmp_path / "out.bin", np.complex128)
    n_k, n_i = nfe_k * npe * nc, img.size
    assert out.size == n_k + 2 * n_i
    unf = lambda v, shape: v.reshape(shape, order="F")  # noqa: E731
    assert np.array_equal(unf(out[:n_k], s.shape), mx.gpuMriCartesianRemoveROOSFwd(s, img, m))
    assert np.array_equal(unf(out[n_k:n_k + n_i], img.shape), mx.gpuMriCartesianRemoveROOSAdj(s, f, m))
    assert np.array_equal(unf(out[n_k + n_i:], img.shape), mx.gpuTgvMriOptimizer_cartesianRemoveROOS(s, f, m, prm))


def test_reference_demo_geometry_parity(oracle):
    """384 x 384 image on the 768 x 384 grid with the reference's random4 mask
    (117 of 384 PE lines), 15 coils: 20 iterations vs the FP64 oracle."""
    import primaldualtoolbox as pdt
    from primaldualtoolbox import synthetic as syn
    lines = np.load(os.path.join(ROOT, "tests", "golden", "masks.npz"))["random4_lines"]
    case = syn.mri_case(seed=11, n=384, coils=15, n_pe=384, lines=lines)
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    op = pdt.mri.MriCartesianRemoveROOSOperator()
    op.setCoilSens(case["sens"])
    op.setMask(case["mask"])
    opt = pdt.mri.TgvMriOptimizer_2c3c()
    p = opt.getParameters()
    p.alpha1, p.alpha0, p.max_iter = a1, 2 * a1, 19
    opt.setOperator(op)
    opt.setNoisyData(case["kspace"])
    opt.solve()
    oop = oracle.Op(oracle.OP_CARTESIAN_ROOS, case["sens"].astype(np.complex128), case["mask"].astype(np.float64))
    ref, _, _, _ = oracle.solve(oracle.TGV_MRI, case["kspace"].astype(np.complex128), op=oop, alpha1=a1,
                                alpha0=2 * a1, max_iter=19)
    err = rel_l2(opt.getResult(), ref)
    print(f"demo geometry 384x384 (768 FE, 117 lines, 15 coils), 20 iterations: rel L2 {err:.3e}", flush=True)
    assert err < 1e-3
