"""Writes tests/golden/endstate_320x320x15.npz: FP64 oracle solutions of the
320 x 320, 15-coil TGV-MRI workload at its stopping point, relative partial
gap 1e-4 (evaluated every 50 iterations), from the cold start u0 = A^* f.

Seeds: 0 (the benchmark slice) and 3000 to 3007 (held-out slices of the
warm-start network, trained on seeds 1000 to 2999).  alpha1 = 1e-5 ||f|| (the float of numpy's norm of the
complex64 k-space, as the benchmark computes it), alpha0 = 2 alpha1,
reduction 1.  Stored per seed: u (complex64), the stopping iteration and the
relative gap there.  The GPU tests compare their own solves to gap 1e-4,
cold and warm-started, with these.

Run from the repository root (about 5 minutes per seed on 8 cores; seeds
already in the file are kept, --force recomputes them):
    python tests/golden/make_endstate.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "primal-dual-toolbox_amd")]

import oracle as O  # noqa: E402
from primaldualtoolbox import synthetic as syn  # noqa: E402

SEEDS = (0,) + tuple(range(3000, 3008))
GAP_TOL, GAP_EVERY = 1e-4, 50


def main():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "endstate_320x320x15.npz")
    out = {}
    if os.path.exists(path) and "--force" not in sys.argv:
        with np.load(path) as z:
            out = {k: z[k] for k in z.files if k != "seeds"}
    for seed in SEEDS:
        if f"u_{seed}" in out:
            continue
        t0 = time.time()
        case = syn.mri320_case(seed)
        a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
        op = O.Op(O.OP_CARTESIAN_ROOS, case["sens"], case["mask"])
        u, _, st, _ = O.solve(O.TGV_MRI, case["kspace"].astype(np.complex128), op=op, alpha1=a1, alpha0=2 * a1,
                              max_iter=200000, gap_every=GAP_EVERY, gap_tol=GAP_TOL)
        out[f"u_{seed}"] = u.astype(np.complex64)
        out[f"iters_{seed}"] = st["iters"]
        out[f"rel_gap_{seed}"] = st["rel_gap"]
        out[f"alpha1_{seed}"] = a1
        print(f"seed {seed}: {st['iters']} iterations, rel. gap {st['rel_gap']:.4e}, {time.time() - t0:.0f} s",
              flush=True)
        np.savez_compressed(path, seeds=np.array([s for s in SEEDS if f"u_{s}" in out]), **out)
    np.savez_compressed(path, seeds=np.array(SEEDS), **out)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
