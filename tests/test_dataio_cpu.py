"""Data formats (primaldualtoolbox.dataio) and the argument checks of the
MATLAB entry points (primaldualtoolbox.matlab), on the CPU.

The layout rules are pinned by the reference's own mask files: its .mat and
.npy masks (data/mask_cartesian_{random4,regular4}) give the same PE-line
lists through ``load``/``from_matlab``/``mask_lines`` (golden/mat_masks.npz,
written by golden/make_mat_masks.py, against golden/masks.npz).
"""
import os

import numpy as np
import pytest

from primaldualtoolbox import dataio, synthetic

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reference_mat_masks_match_npy_masks():
    npy = np.load(os.path.join(GOLD, "masks.npz"))
    mat = np.load(os.path.join(GOLD, "mat_masks.npz"))
    for name in ("random4", "regular4"):
        assert mat[f"{name}_mask_shape"].tolist() == npy[f"{name}_shape"].tolist() == [768, 384]
        assert np.array_equal(mat[f"{name}_mask_lines"], npy[f"{name}_lines"])
    # regular4.mat also holds the fully sampled centre alone (mask_ref): 28 lines
    ref = mat["regular4_mask_ref_lines"]
    assert ref.tolist() == list(range(178, 206))
    assert set(ref.tolist()) <= set(npy["regular4_lines"].tolist())
    assert float(mat["random4_power"]) == 20.0 and abs(float(mat["random4_R"]) - 3.3) < 1e-12


def test_mask_lines_roundtrip_and_rejects():
    m = np.load(os.path.join(GOLD, "masks.npz"))
    lines = m["random4_lines"]
    mask = dataio.line_mask(lines, 768, 384)
    assert mask.dtype == np.float32 and mask.shape == (768, 384)
    assert np.array_equal(dataio.mask_lines(mask), lines)
    bad = mask.copy()
    bad[5, lines[0]] = 0  # one sample missing along FE
    with pytest.raises(ValueError, match="vary along FE"):
        dataio.mask_lines(bad)
    with pytest.raises(ValueError, match="only 0 and 1"):
        dataio.mask_lines(mask * 0.5)
    with pytest.raises(ValueError, match="out of range"):
        dataio.line_mask([0, 384], 768, 384)


def test_npy_npz_mat_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    k = (rng.standard_normal((3, 8, 6)) + 1j * rng.standard_normal((3, 8, 6))).astype(np.complex64)
    s = (rng.standard_normal((3, 8, 6)) + 1j * rng.standard_normal((3, 8, 6))).astype(np.complex64)
    m = dataio.line_mask([0, 2, 5], 8, 6)
    np.save(tmp_path / "k.npy", k)
    assert np.array_equal(dataio.load(str(tmp_path / "k.npy")), k)
    np.savez(tmp_path / "z.npz", a=k, b=m)
    assert np.array_equal(dataio.load(str(tmp_path / "z.npz"), "b"), m)
    with pytest.raises(ValueError, match="name one with key"):
        dataio.load(str(tmp_path / "z.npz"))
    with pytest.raises(KeyError):
        dataio.load(str(tmp_path / "z.npz"), "c")
    # .mat: MATLAB indexing on disk, (C, FE, PE) after from_matlab
    dataio.save_mat(str(tmp_path / "case.mat"), rawdata=k, sensitivities=s, mask=m)
    raw = dataio.load(str(tmp_path / "case.mat"), "rawdata")
    assert raw.shape == (8, 6, 3)
    assert raw[7, 5, 2] == k[2, 7, 5]
    case = dataio.load_cartesian(str(tmp_path / "case.mat"), str(tmp_path / "case.mat"), str(tmp_path / "case.mat"))
    assert np.array_equal(case["kspace"], k) and np.array_equal(case["sens"], s)
    assert np.array_equal(case["mask"], m) and case["lines"].tolist() == [0, 2, 5]
    assert np.array_equal(dataio.to_matlab(dataio.from_matlab(raw)), raw)
    with pytest.raises(ValueError, match="unknown file type"):
        dataio.load(str(tmp_path / "x.txt"))
    with pytest.raises(ValueError, match="mask"):
        dataio.load_cartesian(k, s, m[:, :5])


def test_mat73_refused(tmp_path):
    # the 128-byte header of a MATLAB 7.3 (HDF5) file: text, subsystem
    # offset, version 0x0200, endian indicator
    hdr = b"MATLAB 7.3 MAT-file, Platform: GLNXA64".ljust(116, b" ") + b"\0" * 8 + b"\x00\x02" + b"IM"
    p = tmp_path / "v73.mat"
    p.write_bytes(hdr + b"\0" * 512)
    with pytest.raises(ValueError, match="7.3"):
        dataio.load(str(p), "mask")


def test_roos_window_is_fe_idx():
    for n_fe in (640, 736, 768, 512, 10):
        w = dataio.roos_window(n_fe)
        assert w.start == synthetic.roos_offset(n_fe)  # mricartesianoperator.cu:341
        assert w.stop - w.start == n_fe // 2
    assert dataio.roos_window(768) == slice(193, 577)  # bin/mri_cartesian_test.py:45
    img = np.arange(768 * 3).reshape(768, 3)
    assert np.array_equal(dataio.crop_roos(img), img[193:577])


def test_matlab_entry_argument_checks():
    """Count and type checks of the mex sources (mexMriCartesianFwdOp.cpp:54-76,
    mexMriCartesianReconstructTgv.cpp:54-83), raised before any GPU work."""
    from primaldualtoolbox import matlab as mx
    s = np.zeros((8, 6, 2), np.complex128)
    f = np.zeros((8, 6, 2), np.complex128)
    m = np.ones((8, 6), np.float64)
    with pytest.raises(mx.MexError) as e:
        mx.gpuMriCartesianAdj(s, f)
    assert e.value.identifier == "MATLAB:gpuMriCartesianAdj:invalidNumInputs"
    assert e.value.message == "Three inputs required (sensitivities, rawdata, mask)"
    with pytest.raises(mx.MexError) as e:
        mx.gpuMriCartesianFwd(s.astype(np.complex64), f, m)
    assert e.value.message == "Input 1 (sensitivities) must be double complex!"
    with pytest.raises(mx.MexError) as e:
        mx.gpuMriCartesianRemoveROOSAdj(s, f.real.copy(), m)
    assert e.value.message == "Input 2 (rawdata) must be double complex!"
    with pytest.raises(mx.MexError) as e:
        mx.gpuMriCartesianRemoveROOSFwd(s, f, m.astype(np.float32))
    assert e.value.identifier == "MATLAB:gpuMriCartesianRemoveROOSFwd:wrongDataType"
    assert e.value.message == "Input 3 (mask) must be double!"
    with pytest.raises(mx.MexError) as e:
        mx.gpuTgvMriOptimizer_cartesian(s, f, m)
    assert e.value.message == "Four inputs required (sensitivities, rawdata, mask, tgv_params)"
    with pytest.raises(mx.MexError, match="Expect struct with tgv params!"):
        mx.gpuTgvMriOptimizer_cartesianRemoveROOS(s, f, m, 3.0)
    with pytest.raises(mx.MexError, match="missing"):
        mx.gpuTgvMriOptimizer_cartesian(s, f, m, {"alpha0": 1.0})
    with pytest.raises(mx.MexError, match="size mismatch"):
        mx.gpuTgvMriOptimizer_cartesian(s, f[:, :, :1].copy(), m,
                                        {"alpha0": 1.0, "alpha1": 1.0, "max_iter": 1, "reduction": 1.0})
