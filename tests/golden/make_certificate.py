"""Writes tests/golden/certificate_24x24x3.npz: the TGV-MRI minimiser of a
small Cartesian case (24 x 24 image, 3 coils, ROOS grid 48 FE x 24 PE,
synthetic seed 2), computed by the FP64 oracle run for 200000 iterations
(relative partial gap ~1e-15, machine-level convergence).

The certificate test uses it as (u*, v*, P*) to check that the partial gap
at the first iterate with relative gap <= 1e-4 bounds P(ubar, vbar) - P* and
that u*, v* lie in the gap's balls around (ubar, vbar).  Run from the
repository root:  python tests/golden/make_certificate.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "primal-dual-toolbox_amd")]

import oracle as O  # noqa: E402
from primaldualtoolbox import synthetic as syn  # noqa: E402

ITERS = 200000
SEED, N, COILS, CENTER = 2, 24, 3, 4


def main():
    case = syn.mri_case(seed=SEED, n=N, coils=COILS, center=CENTER)
    f = case["kspace"].astype(np.complex128)
    a1 = 1e-2 * np.linalg.norm(f)
    op = O.Op(O.OP_CARTESIAN_ROOS, case["sens"], case["mask"])
    u, v, st, hist = O.solve(O.TGV_MRI, f, op=op, alpha1=a1, alpha0=2 * a1, max_iter=ITERS - 1,
                             gap_every=ITERS, want_v=True)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "certificate_24x24x3.npz")
    np.savez_compressed(out, u=u, v=v, alpha1=a1, alpha0=2 * a1, iters=ITERS, rel_gap=st["rel_gap"],
                        seed=SEED, n=N, coils=COILS, center=CENTER)
    print(f"wrote {out}: {ITERS} iterations, rel. gap {st['rel_gap']:.3e}")


if __name__ == "__main__":
    main()
