"""Writes mat_masks.npz: the sampling masks of the reference's MATLAB files
(data/mask_cartesian_{random4,regular4}.mat) read through
primaldualtoolbox.dataio, stored as their shapes and sampled PE-line lists,
plus the scalar fields of random4.mat.  The test suite compares them with the
line lists of the same masks' numpy files (masks.npz).  Needs the reference
tree mounted at /root/reference; the outputs are data only.

Usage: python tests/golden/make_mat_masks.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "primal-dual-toolbox_amd"))
from primaldualtoolbox import dataio  # noqa: E402

REF_DATA = "/root/reference/data"


def main():
    out = {}
    for name, keys in (("random4", ["mask"]), ("regular4", ["mask", "mask_ref"])):
        path = os.path.join(REF_DATA, f"mask_cartesian_{name}.mat")
        for k in keys:
            m = dataio.from_matlab(dataio.load(path, k))
            out[f"{name}_{k}_shape"] = np.array(m.shape, np.int64)
            out[f"{name}_{k}_lines"] = dataio.mask_lines(m).astype(np.int32)
    path = os.path.join(REF_DATA, "mask_cartesian_random4.mat")
    for k in ("R", "fullCenter", "power"):
        out[f"random4_{k}"] = np.array(float(np.asarray(dataio.load(path, k)).ravel()[0]))
    np.savez(os.path.join(HERE, "mat_masks.npz"), **out)
    print({k: (v.shape, v.dtype) for k, v in out.items()})


if __name__ == "__main__":
    main()
