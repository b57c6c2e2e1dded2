#!/usr/bin/env python3
"""Benchmark: time-to-gap of the Cartesian TGV-MRI solver on one MI355X.

Workload (one "step" = one solve to relative primal-dual gap 1e-4, cold and
warm-started):
  320 x 320 complex64 image (ellipse phantom, quadratic ramp, polynomial
  phase), 15 coils (Gaussian bumps on a circle, smooth phase, sum |S|^2 = 1),
  read-out oversampled k-space {320 PE, 640 FE}, random 4x PE-line mask with
  24 fully sampled centre lines, complex noise at 30 dB SNR,
  MriCartesianRemoveROOSOperator, TgvMriOptimizer with alpha1 = 1e-5 ||f||,
  alpha0 = 2 alpha1, reduction 1, gap evaluated every 50 iterations.
  Cold start u0 = A^* f (v = p = q = r = 0), as the reference's solve();
  warm start (u0, v0) from the trained network (warmstart.py), its inference
  inside the timed region.

Prints one JSON line (see the "roofline" and "cpu_baseline" objects).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "primal-dual-toolbox_amd"))

HBM_PEAK_GBS = 8000.0
GAP_TOL = 1e-4
GAP_EVERY = 50
CONC_STREAMS = 4  # slices solved at once by pdt_solve_concurrent in the batch leg


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(seed: int = 0):
    from primaldualtoolbox import synthetic as syn
    case = syn.mri_case(seed=seed, n=320, coils=15, accel=4.0, center=24, snr_db=30.0, roos=True)
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    return case, a1


def build_solver(case, a1, max_iter=200000):
    import primaldualtoolbox as pdt
    op = pdt.mri.MriCartesianRemoveROOSOperator()
    op.setCoilSens(case["sens"])
    op.setMask(case["mask"])
    opt = pdt.mri.TgvMriOptimizer_2c3c()
    opt.setOperator(op)
    opt.setNoisyData(case["kspace"])
    prm = opt.getParameters()
    prm.alpha1, prm.alpha0, prm.reduction = a1, 2 * a1, 1.0
    prm.max_iter = max_iter
    prm.gap_tol, prm.gap_every = GAP_TOL, GAP_EVERY
    return op, opt


def cold_start(opt, u0_dev):
    """setInput0(A^* f) from a device-resident copy (v, p, q zeroed)."""
    from primaldualtoolbox import _capi as C
    C.check(C.lib().pdt_solver_set_input0(opt._h, u0_dev.h, None))


def host_info() -> dict:
    """The host the CPU baseline runs on: logical CPUs, physical cores and
    model (lscpu), and the OpenMP threads the oracle gets."""
    import subprocess
    info = {"nproc": os.cpu_count(), "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
        cores = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        info["physical_cores"] = len({ln for ln in cores.splitlines() if ln and not ln.startswith("#")})
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def default_cpu_threads() -> int:
    """OMP_NUM_THREADS where the host sets it (the CPU share of this GPU on
    a shared machine), else every physical core."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return int(host_info().get("physical_cores") or os.cpu_count() or 1)


def timed_cpu(run, runs: int = 5) -> tuple[float, list]:
    """Median seconds of `runs` timed calls after one untimed warm-up."""
    run()
    ts = [run() for _ in range(runs)]
    return float(np.median(ts)), ts


def cpu_baseline(case, a1, gpu_iters: int, sample_iters: int, threads: int):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.set_threads(threads)
    op = O.Op(O.OP_CARTESIAN_ROOS, case["sens"], case["mask"], prec=32)

    def run():
        _, _, st, _ = O.solve(O.TGV_MRI, case["kspace"], op=op, alpha1=a1, alpha0=2 * a1,
                              max_iter=sample_iters - 1, prec=32)
        return st["seconds"]
    med, ts = timed_cpu(run)
    per_iter = med / sample_iters
    host = host_info()
    host["omp_max_threads"] = O.max_threads()
    return {
        "value": per_iter * gpu_iters,
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": (f"FP32 OpenMP oracle on {threads} threads, the same 320x320x15 workload: median of 5 timed "
                   f"runs of {sample_iters} iterations after 1 warm-up ({per_iter * 1e3:.2f} ms/iteration, "
                   f"timed with steady_clock around the iteration loop), extrapolated to the GPU's "
                   f"{gpu_iters} iterations to rel. gap {GAP_TOL:g}"),
        "method": "extrapolated",
        "runs_s": ts,
        "host": host,
    }


def warm_start_steps(opt, case, steps: int, warmup: int):
    """Time-to-gap from the network's (u, v): warm_start (A^*f, RMS, network,
    set_input0) and solve, timed together.  None without trained weights."""
    from primaldualtoolbox import warmstart as ws
    if not os.path.exists(ws.DEFAULT_WEIGHTS):
        return None
    net = ws.WarmStartCNN.load()
    out = {"seconds": [], "iters": [], "rel_gap": [], "network_ms": []}
    for k in range(warmup + steps):
        opt.setNoisyData(case["kspace"])  # untimed: host->device upload of f, r = 0
        t0 = time.perf_counter()
        net.warm_start(opt)
        t1 = time.perf_counter()
        opt.solve()
        t2 = time.perf_counter()
        st = opt.getStats()
        if not st["converged"]:
            raise RuntimeError(f"warm-started solve did not reach rel. gap {GAP_TOL}: {st}")
        if k >= warmup:
            out["seconds"].append(t2 - t0)
            out["iters"].append(int(st["iters"]))
            out["rel_gap"].append(float(st["rel_gap"]))
            out["network_ms"].append((t1 - t0) * 1e3)
        log(f"[bench] warm {'step' if k >= warmup else 'warmup'}: {t2 - t0:.4f} s, {st['iters']} iterations")
    out["u"] = opt.getResult().copy()
    return out


def held_out_end_states(gold: str) -> list:
    """The held-out slices that have FP64 oracle end states in the fixture
    file (the network never saw them): cold and warm time-to-gap, iterations,
    and rel. L2 of each GPU end state vs the oracle's u at its own stop."""
    from primaldualtoolbox import _capi as C
    from primaldualtoolbox import synthetic as syn
    from primaldualtoolbox import warmstart as ws
    if not os.path.exists(ws.DEFAULT_WEIGHTS) or not os.path.exists(gold):
        return []
    net = ws.WarmStartCNN.load()
    out = []
    with np.load(gold) as z:
        seeds = sorted(int(k[6:]) for k in z.files if k.startswith("iters_") and int(k[6:]) != 0)
        for seed in seeds:
            case = syn.mri320_case(seed)
            op, opt = build_solver(case, float(z[f"alpha1_{seed}"]))
            # u0 = A^* f on the device, as for the headline's cold start
            f_dev = C.DeviceArray.from_numpy(case["kspace"], np.complex64)
            H, W = case["image"].shape
            u0_dev = C.DeviceArray.empty(C.PDT_C64, [W, H])
            C.check(C.lib().pdt_op_adjoint(op._h, f_dev.h, u0_dev.h))
            rec = {"seed": seed, "oracle_iterations": int(z[f"iters_{seed}"])}
            for start in ("cold", "warm"):
                best = None
                for _ in range(3):  # the first solve pays graph capture
                    opt.setNoisyData(case["kspace"])
                    t0 = time.perf_counter()
                    if start == "warm":
                        net.warm_start(opt)
                    else:
                        cold_start(opt, u0_dev)
                    opt.solve()
                    dt = time.perf_counter() - t0
                    best = dt if best is None else min(best, dt)
                st = opt.getStats()
                if not st["converged"]:
                    raise RuntimeError(f"held-out seed {seed} ({start}) did not reach rel. gap {GAP_TOL}: {st}")
                rec[start] = {"seconds": round(best, 5), "iterations": int(st["iters"]),
                              "rel_l2_vs_fp64_oracle": _rel(opt.getResult(), z[f"u_{seed}"])}
            log(f"[bench] held-out seed {seed}: {rec}")
            out.append(rec)
    return out


def extra_cnn(reps: int = 5) -> dict:
    """Warm-start network kernels alone: inference on one 320 x 320 slice and
    one Adam training step at batch 16 (random data), with MFMA TFLOP/s."""
    from primaldualtoolbox import _capi as C
    from primaldualtoolbox import warmstart as ws
    net = ws.WarmStartCNN(seed=0)
    ch = net.channels
    macs = sum(9 * a * b for a, b in zip(ch[:-1], ch[1:]))  # per pixel
    rng = np.random.default_rng(0)
    res = {"network": ch, "parameters": net.n_params}
    for name, n in (("inference_1x320x320", 1), ("train_step_16x320x320", 16)):
        x = C.DeviceArray.from_numpy(rng.standard_normal(n * 320 * 320 * 2).astype(np.float32))
        t = C.DeviceArray.from_numpy(rng.standard_normal(n * 320 * 320 * 6).astype(np.float32))
        y = C.DeviceArray.empty(C.PDT_F32, [n * 320 * 320 * 6])

        def run():
            if n == 1:
                C.check(C.lib().pdt_cnn_forward(net._h, x.h, y.h, 1, 320, 320))
            else:
                net.train_step(x, t, lr=1e-3, shape=(n, 320, 320))
        run()
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        dt = (time.perf_counter() - t0) / reps
        flops = 2.0 * macs * n * 320 * 320 * (1 if n == 1 else 3)  # training: forward + about 2x backward
        res[name] = {"ms": dt * 1e3, "TFLOPs": flops / dt / 1e12}
    return res


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def _rel(a, b) -> float:
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _timed_solves(opt, u0_dev, steps: int) -> float:
    """Median wall time of `steps` solves, each restarted from u0 on the device."""
    from primaldualtoolbox import _capi as C
    ts = []
    for _ in range(steps):
        C.check(C.lib().pdt_solver_set_input0(opt._h, u0_dev.h, None))
        t0 = time.perf_counter()
        opt.solve()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def _equivalent(one_iteration_bytes: float, us_per_iteration: float) -> dict:
    """The bytes the one-iteration kernels move per iteration (each plane
    read or written once per half-iteration) over the measured time per
    iteration: what a temporally blocked solve does in HBM-roofline terms
    (its own launches move those bytes once per k iterations)."""
    gbs = one_iteration_bytes / (us_per_iteration * 1e-6) / 1e9
    return {"one_iteration_MB": round(one_iteration_bytes / 1e6, 3), "GBs": round(gbs, 1),
            "frac_of_peak": round(gbs / HBM_PEAK_GBS, 3)}


def _kernels(prof) -> dict:
    """Per-launch duration (ms), algorithmic bytes (MB) and the GB/s they
    give, per kernel of one iteration (pdt_solver_profile)."""
    return {p["name"]: {"ms": round(p["ms"], 5), "alg_MB": round(p["bytes"] / 1e6, 3),
                        "GBs": round(p["bytes"] / (p["ms"] * 1e-3) / 1e9, 1)} for p in prof}


def extra_tv256(steps: int, cpu_threads: int, no_cpu: bool, autotune: bool = True) -> dict:
    """ROF denoising, 256^2 float32, lambda 10, 500 iterations.  Accuracy is
    the relative L2 distance to the FP64 GPU solver on the same input."""
    import primaldualtoolbox as pdt
    from primaldualtoolbox import _capi as C
    from primaldualtoolbox import synthetic as syn
    _, noisy = syn.tv_denoising_case(seed=0, n=256, sigma=0.1)
    res = {}
    tiles = None
    for name, dt in (("f32", np.float32), ("f64", np.float64)):
        opt = pdt.denoising.TvOptimizer_2f(dtype=dt)
        opt.setNoisyData(noisy.astype(dt))
        p = opt.getParameters()
        p.Lambda, p.max_iter, p.gap_every = 10, 499, 500
        f_dev = C.DeviceArray.from_numpy(noisy.astype(dt), dt)
        if autotune:
            t = pdt.denoising.autotune([opt], reps=20)
            tiles = t if name == "f32" else tiles
            opt.setInput0(noisy.astype(dt))
        opt.solve()
        res[name] = (opt, _timed_solves(opt, f_dev, steps), opt.getResult().copy())
    opt, s, u = res["f32"]
    st = opt.getStats()
    out = {"workload": "TV-L2 (ROF) 256x256 float32, lambda=10, 500 iterations",
           "seconds": s, "us_per_iteration": s * 1e6 / 500, "rel_gap": float(st["rel_gap"]),
           "rel_l2_vs_f64": _rel(u, res["f64"][2]), "seconds_f64": res["f64"][1], "tiles": tiles,
           "kernels": _kernels(opt.profileKernels(50)),
           # dual step: ubar r, p r/w; primal step: u r/w, f, p r, ubar w (4 bytes each)
           "equivalent": _equivalent((5 + 6) * 4.0 * 256 * 256, s * 1e6 / 500)}
    if not no_cpu:
        O = _oracle()
        O.set_threads(cpu_threads)
        med, ts = timed_cpu(lambda: O.solve(O.TV, noisy, lambda_=10, max_iter=499, gap_every=500,
                                            prec=32)[2]["seconds"])
        out["cpu_baseline"] = {"value": med, "unit": "s", "cores": cpu_threads, "kind": "port",
                               "sample": "the full 500 iterations in the FP32 OpenMP oracle, median of 5 after 1 warm-up",
                               "runs_s": ts}
    return out


def extra_tgv512(steps: int, cpu_threads: int, no_cpu: bool, autotune: bool = True) -> dict:
    """TGV^2 denoising, 512^2 complex64, lambda 20, alpha0 2, alpha1 1, 1000
    iterations, FP32 and FP16 dual storage; accuracy against the FP64 GPU
    solver on the same input."""
    import primaldualtoolbox as pdt
    from primaldualtoolbox import _capi as C
    from primaldualtoolbox import synthetic as syn
    _, noisy = syn.tgv_denoising_case(seed=0, n=512, sigma=0.05)
    res = {}
    tiles = {}
    for name, dt, fp16 in (("f32", np.complex64, 0), ("f16dual", np.complex64, 1), ("f64", np.complex128, 0)):
        opt = pdt.denoising.TgvOptimizer_2c(dtype=dt)
        opt.setNoisyData(noisy.astype(dt))
        p = opt.getParameters()
        p.Lambda, p.alpha0, p.alpha1, p.max_iter, p.dual_fp16, p.gap_every = 20, 2, 1, 999, fp16, 1000
        f_dev = C.DeviceArray.from_numpy(noisy.astype(dt), dt)
        if autotune:
            tiles[name] = pdt.denoising.autotune([opt], reps=20)
            opt.setNoisyData(noisy.astype(dt))
        opt.solve()
        res[name] = (opt, _timed_solves(opt, f_dev, steps), opt.getResult().copy())
    ref = res["f64"][2]
    out = {"workload": "TGV2-L2 512x512 complex64, lambda=20, alpha0=2, alpha1=1, 1000 iterations"}
    for name in ("f32", "f16dual"):
        opt, s, u = res[name]
        # primal step 10 pe + 5 pq, dual step 3 pe + 10 pq bytes per pixel
        pq = 4.0 if name == "f16dual" else 8.0
        out[name] = {"seconds": s, "us_per_iteration": s * 1e6 / 1000, "rel_l2_vs_f64": _rel(u, ref),
                     "rel_gap": float(opt.getStats()["rel_gap"]), "tiles": tiles.get(name),
                     "kernels": _kernels(opt.profileKernels(50)),
                     "equivalent": _equivalent((13 * 8.0 + 15 * pq) * 512 * 512, s * 1e6 / 1000)}
    out["seconds_f64"] = res["f64"][1]
    if not no_cpu:
        O = _oracle()
        O.set_threads(cpu_threads)
        n_it = 100
        med, ts = timed_cpu(lambda: O.solve(O.TGV, noisy, lambda_=20, alpha0=2, alpha1=1, max_iter=n_it - 1,
                                            gap_every=n_it, prec=32)[2]["seconds"])
        out["cpu_baseline"] = {"value": med * 1000 / n_it, "unit": "s", "cores": cpu_threads,
                               "kind": "port", "method": "extrapolated",
                               "sample": (f"{n_it} iterations in the FP32 OpenMP oracle (median of 5 after 1 "
                                          f"warm-up), scaled to 1000"), "runs_s": ts}
    return out


def extra_demo384(steps: int, cpu_threads: int, no_cpu: bool, autotune: bool = True, graph: bool = True) -> dict:
    """The geometry of the reference's Cartesian demo (bin/mri_cartesian_test.py):
    384x384 image on the 768 FE x 384 PE oversampled grid, the reference's
    random4 mask (117 PE lines, as stored in tests/golden/masks.npz), 15
    synthetic coils, alpha as in the headline workload.  Time-to-gap 1e-4 from
    the cold start, and the demo's fixed 1001 iterations (max_iter 1000)."""
    from primaldualtoolbox import _capi as C
    from primaldualtoolbox import synthetic as syn
    lines = np.load(os.path.join(ROOT, "tests", "golden", "masks.npz"))["random4_lines"]
    case = syn.mri_case(seed=0, n=384, coils=15, n_pe=384, lines=lines)
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    import primaldualtoolbox as pdt
    op, opt = build_solver(case, a1)
    opt.getParameters().use_graph = int(graph)
    u0 = C.DeviceArray.from_numpy(op.adjoint(case["kspace"]), np.complex64)
    tiles = pdt.mri.autotune([opt], reps=10) if autotune else None  # else the cached tiles
    opt.setNoisyData(case["kspace"])
    opt.solve()  # first solve: engine, graphs
    s = _timed_solves(opt, u0, steps)
    st = opt.getStats()
    if not st["converged"]:
        raise RuntimeError(f"demo geometry did not reach rel. gap {GAP_TOL}: {st}")
    it = int(st["iters"])
    p = opt.getParameters()
    p.max_iter, p.gap_tol, p.gap_every = 1000, 0.0, 0
    s_fixed = _timed_solves(opt, u0, steps)
    out = {"workload": ("TGV-MRI 384x384 complex64, 15 coils, ROOS grid 384 PE x 768 FE, reference random4 "
                        "mask (117 PE lines), 30 dB SNR, alpha1=1e-5||f||, alpha0=2 alpha1, cold start"),
           "seconds_to_gap": s, "iterations_to_gap": it, "us_per_iteration": s * 1e6 / it,
           "seconds_1001_iterations": s_fixed, "tiles": tiles,
           "kernels": _kernels(opt.profileKernels(50))}
    if not no_cpu:
        O = _oracle()
        O.set_threads(cpu_threads)
        n_it = 20
        oop = O.Op(O.OP_CARTESIAN_ROOS, case["sens"], case["mask"], prec=32)
        med, ts = timed_cpu(lambda: O.solve(O.TGV_MRI, case["kspace"], op=oop, alpha1=a1, alpha0=2 * a1,
                                            max_iter=n_it - 1, prec=32)[2]["seconds"])
        out["cpu_baseline"] = {"value": med * it / n_it, "unit": "s", "cores": cpu_threads,
                               "kind": "port", "method": "extrapolated",
                               "sample": (f"{n_it} iterations in the FP32 OpenMP oracle (median of 5 after 1 "
                                          f"warm-up), scaled to {it}"), "runs_s": ts}
    return out


def batch_fixed_iterations(count: int, workers: int, iters: int) -> None:
    """Profiling mode: the batched slices of extra_batch64 (same seeds and
    parameters) run `iters` iterations as one batched solve with direct
    launches and no gap evaluation, so a hardware-counter run sees each
    launch of a full-batch iteration (tiles from the tile cache or the
    defaults).  Prints log lines only."""
    import primaldualtoolbox as pdt
    from primaldualtoolbox import synthetic as syn
    cases = syn.batch_cases(range(1000, 1000 + count), workers=workers)
    keep, solvers = [], []
    for case in cases:
        op = pdt.mri.MriCartesianRemoveROOSOperator()
        op.setCoilSens(case["sens"])
        op.setMask(case["mask"])
        opt = pdt.mri.TgvMriOptimizer_2c3c()
        opt.setOperator(op)
        opt.setNoisyData(case["kspace"])
        a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
        prm = opt.getParameters()
        prm.alpha1, prm.alpha0, prm.reduction, prm.max_iter = a1, 2 * a1, 1.0, iters - 1
        prm.gap_every, prm.use_graph = 0, 0
        keep.append(op)
        solvers.append(opt)
    pdt.mri.batch_solve(solvers)
    log(f"[bench] batch: {count} slices, {iters} iterations with direct launches")


def extra_batch64(count: int, workers: int, cpu_threads: int, no_cpu: bool, autotune: bool = True,
                  graph: bool = True, runs: int = 3) -> dict:
    """`count` mixed Cartesian TGV-MRI slices (sizes 256^2..384x320, 8-20
    coils, random or regular masks) solved together to rel. gap 1e-4 in one
    batched solve (median of `runs` timed solves); slices/s, per-slice
    time-to-gap (alone, and to its stop inside the batch), algorithmic bytes
    per batch iteration and the per-kernel roofline of the batched launches
    (the working set, ~3 GB, is far beyond the 256 MB Infinity Cache)."""
    import primaldualtoolbox as pdt
    from primaldualtoolbox import _capi as C
    from primaldualtoolbox import synthetic as syn
    t0 = time.perf_counter()
    cases = syn.batch_cases(range(1000, 1000 + count), workers=workers)
    log(f"[bench] batch: generated {count} slices in {time.perf_counter() - t0:.1f} s")
    opts, u0s, a1s = [], [], []
    for case in cases:
        op = pdt.mri.MriCartesianRemoveROOSOperator()
        op.setCoilSens(case["sens"])
        op.setMask(case["mask"])
        opt = pdt.mri.TgvMriOptimizer_2c3c()
        opt.setOperator(op)
        opt.setNoisyData(case["kspace"])
        a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
        prm = opt.getParameters()
        prm.alpha1, prm.alpha0, prm.reduction, prm.max_iter = a1, 2 * a1, 1.0, 200000
        prm.gap_tol, prm.gap_every = GAP_TOL, GAP_EVERY
        prm.use_graph = int(graph)
        f_dev = C.DeviceArray.from_numpy(case["kspace"], np.complex64)
        H, W = case["image"].shape
        u0 = C.DeviceArray.empty(C.PDT_C64, [W, H])
        C.check(C.lib().pdt_op_adjoint(op._h, f_dev.h, u0.h))
        opts.append((op, opt))
        u0s.append(u0)
        a1s.append(a1)
    solvers = [o for _, o in opts]
    tune = pdt.mri.autotune(solvers, reps=5) if autotune else None  # else the cached tiles
    log(f"[bench] batch: tiles {tune}")

    def cold():
        for o, u0, case in zip(solvers, u0s, cases):
            o.setNoisyData(case["kspace"])
            C.check(C.lib().pdt_solver_set_input0(o._h, u0.h, None))

    cold()
    pdt.mri.batch_solve(solvers)  # warmup
    times = []
    for r in range(runs):
        cold()
        t0 = time.perf_counter()
        pdt.mri.batch_solve(solvers)
        times.append(time.perf_counter() - t0)
        log(f"[bench] batch run {r}: {times[-1]:.3f} s")
    batch_s = float(np.median(times))
    stats = [o.getStats() for o in solvers]
    its = [int(s["iters"]) for s in stats]
    stop_ms = [float(s["ms_to_stop"]) for s in stats]
    if not all(s["converged"] for s in stats):
        raise RuntimeError("batched solve: a slice did not reach the gap")
    # the launches of one batch iteration with every slice active
    cold()
    prof = pdt.mri.batch_profile(solvers, reps=10)
    per_iter_ms = sum(p["ms"] for p in prof)
    it_bytes = sum(p["bytes"] for p in prof)
    dom = max(prof, key=lambda p: p["ms"])
    # the same slices one at a time (single-slice engine per solver)
    cold()
    t0 = time.perf_counter()
    alone_ms = []
    for o in solvers:
        o.solve()
        alone_ms.append(float(o.getStats()["ms"]))
    seq_s = time.perf_counter() - t0
    log(f"[bench] batch: sequential {seq_s:.3f} s")
    # the same single-slice solves, CONC_STREAMS at a time on their own streams
    seq_u = [o.getResult() for o in solvers]
    conc_times = []
    for r in range(runs):
        cold()
        t0 = time.perf_counter()
        pdt.mri.solve_concurrent(solvers, streams=CONC_STREAMS)
        conc_times.append(time.perf_counter() - t0)
        log(f"[bench] batch: concurrent run {r}: {conc_times[-1]:.3f} s")
    conc_its = [int(o.getStats()["iters"]) for o in solvers]
    if conc_its != its or not all(np.array_equal(o.getResult(), u) for o, u in zip(solvers, seq_u)):
        raise RuntimeError("concurrent solves differ from the single solves")
    conc_s = float(np.median(conc_times))
    work = [it * c["meta"]["coils"] * 2 * c["meta"]["n"] * c["meta"]["n_pe"] for it, c in zip(its, cases)]

    def dist(x):
        return {"min": float(np.min(x)), "median": float(np.median(x)), "max": float(np.max(x))}
    out = {"workload": (f"{count} Cartesian TGV-MRI slices (seeds 1000..{999 + count}), sizes "
                        f"{sorted(set((c['meta']['n'], c['meta']['n_pe']) for c in cases))}, 8-20 coils, "
                        "rel. gap 1e-4 per slice, cold start"),
           "seconds": batch_s, "seconds_runs": times, "slices_per_s": count / batch_s,
           "sequential_seconds": seq_s, "sequential_slices_per_s": count / seq_s,
           "concurrent": {"streams": CONC_STREAMS, "seconds": conc_s, "seconds_runs": conc_times,
                          "slices_per_s": count / conc_s, "equal_to_single_solves": True},
           "best": ({"method": f"concurrent single-slice solves ({CONC_STREAMS} streams)", "slices_per_s": count / conc_s}
                    if conc_s < batch_s else {"method": "batched work lists", "slices_per_s": count / batch_s}),
           "iterations": {k: int(v) for k, v in dist(its).items()},
           "time_to_gap_alone_ms": dist(alone_ms),
           "time_to_stop_in_batch_ms": dist(stop_ms),
           "per_slice": [{"seed": c["meta"]["seed"], "n": c["meta"]["n"], "n_pe": c["meta"]["n_pe"],
                          "coils": c["meta"]["coils"], "iterations": it, "alone_ms": round(a, 3),
                          "in_batch_stop_ms": round(b, 1)}
                         for c, it, a, b in zip(cases, its, alone_ms, stop_ms)],
           "batch_iteration": {"ms": per_iter_ms, "alg_MB": it_bytes / 1e6,
                               "GBs": it_bytes / (per_iter_ms * 1e-3) / 1e9,
                               "kernels": _kernels(prof)},
           "roofline": {"kernel": dom["name"], "bound": "hbm",
                        "achieved": dom["bytes"] / (dom["ms"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": dom["bytes"] / (dom["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "tiles": tune}
    if not no_cpu:
        O = _oracle()
        import concurrent.futures as cf
        n_s, n_it = min(cpu_threads, count), 10

        def one(i):
            O.set_threads(1)
            c = cases[i]
            oop = O.Op(O.OP_CARTESIAN_ROOS, c["sens"], c["mask"], prec=32)
            _, _, ost, _ = O.solve(O.TGV_MRI, c["kspace"], op=oop, alpha1=a1s[i], alpha0=2 * a1s[i],
                                   max_iter=n_it - 1, prec=32)
            return ost["seconds"], n_it * c["meta"]["coils"] * 2 * c["meta"]["n"] * c["meta"]["n_pe"]

        with cf.ThreadPoolExecutor(max_workers=n_s) as ex:
            sample = list(ex.map(one, range(n_s)))
        per_unit = sum(t for t, _ in sample) / sum(u for _, u in sample)
        cpu_s = per_unit * sum(work) / n_s
        out["cpu_baseline"] = {"value": count / cpu_s, "unit": "slices/s", "cores": n_s, "kind": "port",
                               "method": "extrapolated",
                               "sample": (f"{n_it} iterations of each of the first {n_s} slices in the FP32 "
                                          f"oracle, one slice per thread; cost per coil-pixel-iteration "
                                          f"scaled to the {count} slices' GPU iteration counts and divided "
                                          f"over {n_s} threads"),
                               "host": host_info()}
    return out


def side_measurements(args, keys=None) -> dict:
    """The side measurements named in keys (all when None): cnn, tv_256,
    tgv_512, demo_384, batch (the batch of args.batch slices)."""
    runs = {"cnn": extra_cnn,
            "tv_256": lambda: extra_tv256(3, args.cpu_threads, args.no_cpu, not args.no_autotune),
            "tgv_512": lambda: extra_tgv512(3, args.cpu_threads, args.no_cpu, not args.no_autotune),
            "demo_384": lambda: extra_demo384(3, args.cpu_threads, args.no_cpu, not args.no_autotune,
                                              not args.no_graph),
            "batch": lambda: extra_batch64(args.batch, args.gen_workers, args.cpu_threads, args.no_cpu,
                                           not args.no_autotune, not args.no_graph)}
    unknown = set(keys or ()) - set(runs)
    if unknown:
        raise SystemExit(f"unknown side measurement(s): {sorted(unknown)}; known: {sorted(runs)}")
    extra = {}
    for key, fn in runs.items():
        if keys is not None and key not in keys:
            continue
        name = f"batch_{args.batch}" if key == "batch" else key
        log(f"[bench] side measurement {name}")
        try:
            extra[name] = fn()
        except Exception as e:  # reported, never hides the headline line
            extra[name] = {"error": f"{type(e).__name__}: {e}"}
        log(f"[bench] {name}: {extra[name]}")
    return extra


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-iters", type=int, default=40)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads of the CPU baseline (0: OMP_NUM_THREADS if set, else all physical cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=50)
    ap.add_argument("--fixed-iters", type=int, default=0,
                    help="profiling mode: run this many iterations per step (no gap stop), print no JSON")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the tile cache (PDT_TUNE_CACHE) or defaults instead of timing candidates")
    ap.add_argument("--batch-fixed-iters", type=int, default=0,
                    help="profiling mode: the --batch slices run this many iterations as one batched solve "
                         "with direct launches (no gap evaluation), print no JSON")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="profiling mode: run a plain device copy of this many MiB (PMC byte calibration)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the TV 256^2, TGV 512^2 and batched-slices side measurements")
    ap.add_argument("--batch", type=int, default=64, help="slices in the batched side measurement")
    ap.add_argument("--gen-workers", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-warm", action="store_true", help="report the cold start only")
    ap.add_argument("--no-graph", action="store_true",
                    help="side measurements launch their iterations directly instead of replaying graphs "
                         "(for profilers that do not follow graph launches)")
    ap.add_argument("--only", default="",
                    help="comma-separated side measurements to run alone (cnn, tv_256, tgv_512, demo_384, "
                         "batch), printed as one JSON line {\"extra\": ...} instead of the headline")
    args = ap.parse_args()
    if args.cpu_threads <= 0:
        args.cpu_threads = default_cpu_threads()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(slices would shard across GPUs without communication; only the one-GPU share is built)",
              file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.calibrate_copy} MiB x 5")
        return

    if args.batch_fixed_iters:
        batch_fixed_iterations(args.batch, args.gen_workers, args.batch_fixed_iters)
        return

    if args.only:
        print(json.dumps({"extra": side_measurements(args, args.only.split(","))}), flush=True)
        return

    from primaldualtoolbox import _capi as C

    log("[bench] generating workload (seed 0)")
    case, a1 = make_workload()
    op, opt = build_solver(case, a1)
    # u0 = A^* f on the device, kept resident for every cold start
    f_dev = C.DeviceArray.from_numpy(case["kspace"], np.complex64)
    H, W = case["image"].shape
    u0_dev = C.DeviceArray.empty(C.PDT_C64, [W, H])
    C.check(C.lib().pdt_op_adjoint(op._h, f_dev.h, u0_dev.h))

    import primaldualtoolbox as pdt
    tune = None if args.no_autotune else pdt.mri.autotune([opt], reps=20)
    log(f"[bench] autotuned tiles: {tune}")
    if args.fixed_iters:
        prm = opt.getParameters()
        prm.max_iter, prm.gap_tol = args.fixed_iters - 1, 0.0
        for k in range(max(args.steps, 1)):
            opt.setNoisyData(case["kspace"])
            cold_start(opt, u0_dev)
            opt.solve()
            log(f"[bench] fixed step {k}: {opt.getStats()}")
        return

    iters = []
    gaps = []
    loop_ms = []

    def one_step():
        cold_start(opt, u0_dev)
        opt.solve()
        st = opt.getStats()
        iters.append(int(st["iters"]))
        gaps.append(float(st["rel_gap"]))
        loop_ms.append(float(st["ms"]))
        if not st["converged"]:
            raise RuntimeError(f"solve did not reach rel. gap {GAP_TOL}: {st}")

    # r must restart at 0 for every cold solve: setNoisyData re-zeroes it.
    for w in range(args.warmup):
        opt.setNoisyData(case["kspace"])
        one_step()
        log(f"[bench] warmup {w}: {iters[-1]} iterations, rel gap {gaps[-1]:.3e}, loop {loop_ms[-1]:.1f} ms")
    iters.clear(), gaps.clear(), loop_ms.clear()
    step_s = []
    for k in range(args.steps):
        opt.setNoisyData(case["kspace"])  # untimed: host->device upload of f, r = 0
        t0 = time.perf_counter()
        one_step()
        step_s.append(time.perf_counter() - t0)
        log(f"[bench] step {k}: {step_s[-1]:.4f} s, {iters[-1]} iterations, rel gap {gaps[-1]:.3e}")
    cold_s = float(np.mean(step_s))
    u_cold = opt.getResult().copy()
    warm = None if args.no_warm else warm_start_steps(opt, case, args.steps, args.warmup)
    value = float(np.mean(warm["seconds"])) if warm else cold_s
    # the FP64 oracle's solution of the same slice at its own stop (gap 1e-4)
    oracle_end = None
    gold = os.path.join(ROOT, "tests", "golden", "endstate_320x320x15.npz")
    if os.path.exists(gold):
        with np.load(gold) as z:
            oracle_end = {"iterations": int(z["iters_0"]), "rel_l2_cold": _rel(u_cold, z["u_0"]),
                          "rel_l2_warm": _rel(warm["u"], z["u_0"]) if warm else None}
    held_out = [] if args.no_warm else held_out_end_states(gold)

    # per-kernel durations (hipEvents on the solver stream) and algorithmic bytes
    prof = opt.profileKernels(args.profile_reps)
    per_iter_ms = sum(p["ms"] for p in prof)
    iter_bytes = sum(p["bytes"] for p in prof)
    iter_gbs = iter_bytes / (per_iter_ms * 1e-3) / 1e9
    dom = max(prof, key=lambda p: p["ms"])
    achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_tgv_mri_320x320x15.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get(dom["name"])
    kernels = {p["name"]: {"ms": round(p["ms"], 5), "alg_MB": round(p["bytes"] / 1e6, 3),
                           "GBs": round(p["bytes"] / (p["ms"] * 1e-3) / 1e9, 1)} for p in prof}
    log(f"[bench] kernels: {kernels}")

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline: {args.cpu_iters} iterations on {args.cpu_threads} threads")
        cpu = cpu_baseline(case, a1, int(np.median(iters)), args.cpu_iters, args.cpu_threads)

    extra = None if args.no_extra else side_measurements(args)

    line = {
        "metric": "time-to-gap 1e-4 (s), 15-coil 320^2 Cartesian TGV, warm vs cold",
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": value * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "Cartesian TGV-MRI 320x320 complex64, 15 coils, ROOS grid 320 PE x 640 FE, random 4x "
                        "PE-line mask (24 centre lines), 30 dB SNR, alpha1=1e-5||f||, alpha0=2 alpha1, "
                        "rel. partial gap 1e-4 checked every 50 iterations; value = warm start (network "
                        "(u, v) from A^*f, inference timed) when trained weights exist, else cold start u0=A^*f",
            "start": "warm" if warm else "cold",
            "cold_start_s": cold_s,
            "warm_start_s": float(np.mean(warm["seconds"])) if warm else None,
            "warm_speedup": cold_s / float(np.mean(warm["seconds"])) if warm else None,
            "iterations_cold": int(np.median(iters)),
            "iterations_warm": int(np.median(warm["iters"])) if warm else None,
            "warm_network_ms": float(np.median(warm["network_ms"])) if warm else None,
            "warm_vs_cold_rel_l2_u": _rel(warm["u"], u_cold) if warm else None,
            "iterations_to_gap": int(np.median(iters)),
            "fp64_oracle_end_state": oracle_end,
            "held_out_vs_fp64_oracle": held_out,
            "final_rel_gap": float(np.median(gaps)),
            "device_loop_ms": float(np.median(loop_ms)),
            "us_per_iteration": float(np.median(loop_ms)) * 1e3 / max(1, int(np.median(iters))),
            "kernel_ms_per_iteration": per_iter_ms,
            "kernels": kernels,
            "tiles": tune,
        },
        "roofline": {
            "kernel": dom["name"],
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": ("static: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel, committed in "
                               "profiles/traffic_tgv_mri_320x320x15.json (not measured in this run)"
                               if traffic is not None else None),
            # every launch of an iteration together: algorithmic bytes of all
            # of them over their summed durations (a paired iteration moves
            # fewer bytes in two launches than three passes do)
            "iteration": {"launches": [p["name"] for p in prof], "bytes": iter_bytes,
                          "us": per_iter_ms * 1e3, "achieved": iter_gbs, "frac": iter_gbs / HBM_PEAK_GBS},
        },
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
