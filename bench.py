#!/usr/bin/env python
"""HSS2D N = 512^2 build + solve on one MI355X.

Workload (default): the Lippmann-Schwinger Helmholtz volume integral equation of the
reference, kernel w (den==0) + h^2 k^2 (i/4) b H0(k r) c (SL_H_2D, Common/Kernel_Eval.m:44-49,
Duan-Rokhlin diagonal w) with k = 20 on the uniform 512 x 512 grid of
HSS_tree_parameters('SL_H_2D', 20, 1, 8, 64, 1e-10, 3, 800, 0) (N = 262144, leaf 8x8,
depth 12, per-box operators), b = c = sqrt(q) for the reference's tanh-plateau contrast q
(Bump_function option 4, Common/Kernel_Eval.m:133,143), ID-selected skeletons at the two
bottom levels and 'bdry' skeletons above, 4-layer proxy rings, dense-block inverse, seeded
unit-norm complex Gaussian right-hand side.
`--kernel laplace` runs the SL_L_2D grid of HSS_tree_parameters('SL_L_2D', 0, 1, 8, 64,
1e-10, 2, 400, 0) instead; it is also reported as a side case.

One step runs two paths, each end to end from the parameters:
  HSS2D_build_tree_fsym (host index sets, FP64 skeleton IDs, GPU interpolation operators)
  + the exact grid operator A = Kernel_Eval(X, X, params) (FFT convolution)
  + HSS2D_invert_fsym + solve refined (Common/Iterative_Refinement.m) until the true
    relative residual ||b - A y|| / ||b|| <= 1e-10,
  mixed: complex64 interpolation operators (Tikhonov-regularised FP32 least squares with
         FP64 panel factorizations) and complex64 factorization (F^-1, E on MFMA), FP64
         refinement -> `value`
  fp64 : complex128 throughout (plain path).

Prints one JSON line.  `value` is the mixed path's build + solve seconds for the whole
job (time-like, lower is better).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The tree build queues its levels on four streams of its own, beside the null stream (the
# skeleton IDs, uploads) and the caller's; the HIP runtime maps a process's streams onto
# GPU_MAX_HW_QUEUES hardware queues (4 by default) and the kernels of streams sharing a queue
# run one after the other, so sixteen queues keep them apart.  Read when the runtime starts,
# so it is set before torch is imported.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "c128": 78.6, "c64": 157.3}  # MI355X dense MFMA
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of a kernel class, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs
# of `bench.py --no-cpu --no-hss1d --no-batched --no-laplace --mlp-cache F --profile-step`
# (one step of the default workload, the rank network loaded from a file written by an
# unprofiled run), averaged over the class's launches and scaled by what the counters read
# for a plain 1 GiB device copy (`bench.py --calibrate-copy`: FETCH_SIZE 512 MiB -> x2,
# WRITE_SIZE 1 GiB -> x1).  profiles/hss2d_512_helmholtz_traffic.md holds the runs.
MEASURED_TRAFFIC = {
    "gemm_c128_nn": 2 * 77905.1 * 1024 + 59022.3 * 1024,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--Ng", type=int, default=8)
    ap.add_argument("--np", dest="np_", type=int, default=64)
    ap.add_argument("--TI", type=int, default=0)
    ap.add_argument("--kernel", choices=("helmholtz", "laplace"), default="helmholtz")
    ap.add_argument("--kh", type=float, default=20.0)
    ap.add_argument("--proxy-layers", type=int, default=4)
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched HSS1D side case (256 contours)")
    ap.add_argument("--preset-all", action="store_true",
                    help="kh > 0: 'bdry' skeletons at every level instead of the reference's "
                         "ID-selected skeletons at the two bottom levels")
    ap.add_argument("--no-laplace", action="store_true", help="skip the Laplace side case")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--profile-step", action="store_true",
                    help="run one step of both paths only (for counter collection)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a plain 1 GiB device copy three times (scales the FETCH_SIZE / "
                         "WRITE_SIZE counters to bytes)")
    ap.add_argument("--cpu-stop", type=int, default=3,
                    help="CPU baseline: levels measured down to this one, above extrapolated")
    ap.add_argument("--no-hss1d", action="store_true", help="skip the HSS1D side case")
    ap.add_argument("--mlp-cache", default=None,
                    help="npz file of the trained rank network: loaded when present, else "
                         "written after training")
    ap.add_argument("--rank", choices=("id", "table", "mlp"), default="mlp",
                    help="skeleton IDs of the ID-selected levels: plain ID, or randomized IDs "
                         "with sketches sized by the closed-form guess / by the rank network "
                         "(trained before the timed steps on the plain-ID ranks of seeded "
                         "geometries)")
    return ap.parse_args()


def stats_snapshot(H):
    import ctypes as C
    L = H.lib()
    out = {}
    names = [f"{k}_{t}" for k in ("gemv", "gemvt", "symv", "gj_small", "gj_diag", "gen_block",
                                  "gen_parent", "cqr_panel")
             for t in ("f64", "f32", "c128", "c64")]
    names += [f"gemm_{t}_{a}{b}" for t in ("f64", "f32", "c128", "c64") for a in "ntch"
              for b in "nt"]
    names += ["fft_c128"]
    for name in names:
        n, fl, by, ms = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        L.hssc_stats_get(name.encode(), C.byref(n), C.byref(fl), C.byref(by), C.byref(ms))
        if n.value:
            out[name] = dict(launches=n.value, flops=fl.value, bytes=by.value, ms=ms.value)
    return out


def one_step(H, P, torch, b, tol, split=False, rank="id", mlp=None):
    """Runs each path end to end: the tree with its interpolation operators, the exact grid
    operator (FFT convolution, the refinement's residual operator), the inverse, and the solve
    refined until the true relative residual ||b - A y|| / ||b|| <= tol.
      mixed ("fp32"): FP32 / complex64 interpolation operators (skeleton IDs in FP64) and
                      FP32 / complex64 factorization;
      plain ("fp64"): everything in FP64 / complex128.
    `rank` / `mlp` select the skeleton IDs of the ID-selected levels (HSS2D_build_tree_fsym).
    Returns timings (s) and residual info per path.  Without `split` the tree, the grid
    operator and the inversion are called back to back with no host synchronisation, as a
    caller would: the tree build returns with its upper levels still queued, the grid operator
    is set up meanwhile, and the inversion drains the build before it starts.  With `split`
    every phase is synchronised on its own to report the breakdown."""
    sync = torch.cuda.synchronize
    res = {}
    for prec in ("fp32", "fp64"):
        P.raw.tree_precision = 1 if prec == "fp32" else 0
        sync()
        a = time.perf_counter()
        tree = H.HSS2D_build_tree_fsym(P, rank=rank, mlp=mlp)
        A = H.GridOperator(P)
        if split:
            sync()
        m = time.perf_counter()
        inv = H.HSS2D_invert_fsym(tree, precision=prec)
        sync()
        c = time.perf_counter()
        y = torch.empty_like(b)
        it, rel = H.solve_exact_device(inv, A, b, y, tol=tol, maxit=12)
        sync()
        d = time.perf_counter()
        res[prec] = dict(build=c - a, solve=d - c, iters=it, relres=rel, ids=tree.id_stats())
        if split:
            res[prec].update(tree=m - a, invert=c - m)
            # later solves with the same factorization replay the refinement's HIP graphs
            # (captured on the second): the time of one such solve
            for _ in range(3):
                sync()
                e = time.perf_counter()
                H.solve_exact_device(inv, A, b, y, tol=tol, maxit=12)
                sync()
            res[prec]["solve_replayed"] = time.perf_counter() - e
        del inv, A, tree
    P.raw.tree_precision = 0
    return res


def hss1d_case(H, torch, N=4096, reps=3):
    """HSS1D workload: second-kind Laplace double layer on the star contour
    r(t) = 1 + 0.3 cos(5t), N trapezoidal nodes, leaf 64, ID tol 1e-10, 64 proxy points;
    Gaussian RHS from seed 0.  Returns GPU times and the oracle's CPU time on the same
    problem (all of it: the CPU oracle finishes N = 4096 in about a second)."""
    import numpy as np
    from oracle import hss1d as o1
    curve = H.star_contour(N)
    f = np.random.default_rng(0).standard_normal(N)
    best = None
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        K = H.HSS1D_compress_nsym("green", curve, 64, 1e-10, nproxy=64)
        t1 = time.perf_counter()
        H.HSS1D_invert_nsym(K)
        t2 = time.perf_counter()
        x = H.HSS1D_applyinv_nsym(K, f)
        t3 = time.perf_counter()
        r = (t1 - t0, t2 - t1, t3 - t2)
        best = r if best is None or sum(r) < sum(best) else best
    resid = float(np.linalg.norm(H.HSS1D_apply_nsym(K, x) - f) / np.linalg.norm(f))
    C = o1.star_contour(N)
    c0 = time.perf_counter()
    Ho = o1.compress(C, 64, 1e-10, nproxy=64)
    c1 = time.perf_counter()
    inv = o1.invert(Ho)
    c2 = time.perf_counter()
    o1.applyinv(Ho, inv, f)
    c3 = time.perf_counter()
    return dict(workload=f"HSS1D DL star contour N={N} leaf 64 tol 1e-10 nproxy 64",
                compress_s=round(best[0], 4), invert_s=round(best[1], 4),
                solve_ms=round(1e3 * best[2], 3), total_s=round(sum(best), 4),
                relres_vs_hss=resid,
                cpu_oracle=dict(compress_s=round(c1 - c0, 3), invert_s=round(c2 - c1, 3),
                                solve_ms=round(1e3 * (c3 - c2), 2), total_s=round(c3 - c0, 3)))


def batched_case(H, torch, P=256, seed=2024):
    """Batched HSS1D: P star contours r(t) = 1 + a cos(m t), m in {3..8}, a in [0.1, 0.4], N
    log-uniform in [1024, 16384] (seeded), Laplace double layer, leaf 64, ID tol 1e-10, 64
    proxy points; one forest compressed, inverted and solved by shared per-level launches."""
    import numpy as np
    from hssc_directsolver_amd.rank_mlp import random_contour
    rng = np.random.default_rng(seed)
    Ns = [int(np.exp(rng.uniform(np.log(1024), np.log(16384)))) for _ in range(P)]
    curves = [random_contour(rng, N) for N in Ns]
    f = np.random.default_rng(0).standard_normal(sum(Ns))
    best = None
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        K = H.HSS1D_compress_batched(curves, 64, 1e-10, nproxy=64)
        t1 = time.perf_counter()
        H.HSS1D_invert_nsym(K)
        t2 = time.perf_counter()
        x = H.HSS1D_applyinv_nsym(K, f)
        t3 = time.perf_counter()
        r = (t1 - t0, t2 - t1, t3 - t2)
        best = r if best is None or sum(r) < sum(best) else best
    resid = float(np.linalg.norm(H.HSS1D_apply_nsym(K, x) - f) / np.linalg.norm(f))
    # mixed: FP32 factors + FP64 refinement to 1e-10 on the same compressed forest
    mixed = None
    for _ in range(2):
        torch.cuda.synchronize()
        a = time.perf_counter()
        H.HSS1D_invert_nsym(K, precision="fp32")
        b = time.perf_counter()
        xm, it, rel = H.HSS1D_solve(K, f, tol=1e-10, maxit=10)
        c = time.perf_counter()
        r = dict(invert_s=round(b - a, 4), solve_ms=round(1e3 * (c - b), 3), iters=it,
                 relres=rel, total_s=round(best[0] + (c - a), 4))
        mixed = r if mixed is None or r["total_s"] < mixed["total_s"] else mixed
    mixed["rel_vs_fp64"] = float(np.linalg.norm(xm - x) / np.linalg.norm(x))
    return dict(workload=f"batched HSS1D DL, {P} random star contours, total N={sum(Ns)}, "
                         f"leaf 64 tol 1e-10 nproxy 64",
                compress_s=round(best[0], 4), invert_s=round(best[1], 4),
                solve_ms=round(1e3 * best[2], 3), total_s=round(sum(best), 4),
                N_per_s=round(sum(Ns) / sum(best), 1), relres_vs_hss=resid, mixed=mixed)


def level_flops(n, k, M, cx):
    """Dense-block build flops of one box: T by Householder QR least squares
    (2Mk^2 + 4Mk(n-k)) and F^-1, R F^-1 R^T, E^-1 (2n^3 + 2n^2 k + 2nk^2 + 2k^3); x4 for
    complex arithmetic."""
    f = 4 if cx else 1
    return f * (2 * M * k * k + 4 * M * k * (n - k)), f * (2 * n**3 + 2 * n * n * k + 2 * n * k * k
                                                          + 2 * k**3)


def _cpu_run(o, H, P, args, p, stop, cx):
    """One oracle run on the TI tree (one box per level), levels depth..stop timed: build
    (least squares + inversion) and the solve's products with F^-1 and E (each applied twice
    per box in the two sweeps), scaled by the level's box count; the levels above stop are
    extrapolated at level stop's flop and byte rates."""
    import numpy as np
    tb, ti = {}, {}
    t0 = time.perf_counter()
    tree = o.build_tree(p, interform="dense", timings=tb, stop_level=stop,
                        preset_all=args.preset_all)
    inv = o.invert(tree, timings=ti, stop_level=stop)
    ts = {}
    rng = np.random.default_rng(0)
    for lev in range(tree.depth, stop - 1, -1):
        key = tree.key(2 ** lev)
        FI = inv.FI[key]
        E = inv.E.get(key)
        v = rng.standard_normal(FI.shape[0]) + (1j * rng.standard_normal(FI.shape[0]) if cx else 0)
        a = time.perf_counter()
        for _ in range(2):
            FI @ v
            if E is not None:
                E @ v[:E.shape[0]]
        ts[lev] = time.perf_counter() - a
    t1 = time.perf_counter()
    idx = H.HSS2D_build_index_sets(P)
    shapes = {l: idx.level_info(l) for l in range(idx.depth + 1)}
    fT, fI = level_flops(shapes[stop]["n"], shapes[stop]["k"], shapes[stop]["M"], cx)
    rT, rI = fT / tb[stop], fI / ti[stop]
    es = 16 if cx else 8
    bS = 2 * (shapes[stop]["n"] ** 2 + shapes[stop]["k"] ** 2) * es / ts[stop]
    build = solve = 0.0
    for l, s in shapes.items():
        if l >= stop:
            build += (tb[l] + ti[l]) * s["nbox"]
            solve += ts[l] * s["nbox"]
        else:
            a_, b_ = level_flops(s["n"], s["k"], s["M"], cx)
            build += (a_ / rT * (l > 0) + b_ / rI) * s["nbox"]
            solve += 2 * (s["n"] ** 2 + (s["k"] ** 2 if l > 0 else 0)) * es / bS * s["nbox"]
    return dict(build_s=round(build, 1), solve_s=round(solve, 3), measured_s=round(t1 - t0, 2),
                lsq_gflops=round(rT / 1e9, 1), inv_gflops=round(rI / 1e9, 1),
                solve_gbs=round(bS / 1e9, 2))


def cpu_baseline(H, P, args):
    """The oracle (numpy/scipy FP64 restatement of the reference) on the host cores, at all
    cores and at one thread: build and one solve (the FP64 path's first applyinv), measured
    on levels depth..stop of the TI tree of the same grid and scaled by box counts (the NTI
    path processes every box), the levels above extrapolated at the measured rates."""
    from oracle import hss2d as o
    from threadpoolctl import threadpool_limits
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cx = args.kernel == "helmholtz"
    if cx:
        from oracle.dr_weights import corr0
        h = 2.0 / (args.Ng * args.np_)
        p = o.tree_parameters("SL_H_2D", args.kh, 1, args.Ng, args.np_, 1e-10, 3, 800, 1,
                              corr0=corr0(h, args.kh))
        p["proxy_layers"] = args.proxy_layers
    else:
        p = o.tree_parameters("SL_L_2D", 0, 1, args.Ng, args.np_, 1e-10, 2, 400, 1)
    with threadpool_limits(limits=cores):
        allc = _cpu_run(o, H, P, args, p, args.cpu_stop, cx)
    with threadpool_limits(limits=1):
        one = _cpu_run(o, H, P, args, p, args.cpu_stop + 2, cx)
    return dict(value=round(allc["build_s"] + allc["solve_s"], 1), unit="s", cores=cores,
                kind="port",
                sample=(f"oracle (numpy/scipy FP64) on the TI tree of the same grid: levels "
                        f"{P.depth}..{args.cpu_stop} timed on one box each "
                        f"({allc['measured_s']} s at {cores} threads), scaled by boxes per level; "
                        f"levels above extrapolated at the measured rates; build + one solve"),
                all_cores=allc, one_thread=dict(
                    value=round(one["build_s"] + one["solve_s"], 1), levels_timed=(
                        f"{P.depth}..{args.cpu_stop + 2}"), **one))


def phase_rooflines(levels, split, stats, cx):
    """Per-phase figures of the synchronised step: the inversion and the interpolation
    operators against the MFMA peak of their precision (algorithmic flops: the symmetric
    sweep count sum 4 (n^3 + 2 n^2 k + 2 n k^2 + k^3) and the least squares' sum
    4 (2 M k^2 + 4 M k (n - k)) over boxes; the tree time includes the host index sets and the
    FP64 skeleton IDs), kernel-entry generation in entries/s and the solve kernels' HBM
    bytes/s, both from the instrumented step's HIP events."""
    f = 4 if cx else 1
    W_inv = W_T = 0.0
    for lev, (n, k, M) in levels.items():
        nb = 2 ** int(lev)
        W_inv += nb * f * (n ** 3 + 2 * n * n * k + 2 * n * k * k + k ** 3)
        if int(lev) > 0:
            W_T += nb * f * (2 * M * k * k + 4 * M * k * (n - k))
    out = {}
    for prec, peak in (("fp32", 157.3), ("fp64", 78.6)):
        sp = split[prec]
        ti, tt = sp["invert"], sp["tree"]
        out[prec] = dict(
            invert=dict(bound="mfma", flops=W_inv, achieved_tflops=round(W_inv / ti / 1e12, 2),
                        peak=peak, frac=round(W_inv / ti / 1e12 / peak, 3)),
            tree=dict(bound="mfma", flops=W_T, achieved_tflops=round(W_T / tt / 1e12, 2),
                      peak=peak, frac=round(W_T / tt / 1e12 / peak, 3)))
    gen = [v for k_, v in stats.items() if k_.startswith(("gen_block", "gen_parent"))]
    if gen:
        ent = sum(v["flops"] for v in gen)
        ms = sum(v["ms"] for v in gen)
        out["generation"] = dict(bound="valu", entries=ent,
                                 entries_per_s=round(ent / (ms * 1e-3), 1))
    sol = [v for k_, v in stats.items() if k_.startswith(("symv", "gemv"))]
    if sol:
        by = sum(v["bytes"] for v in sol)
        ms = sum(v["ms"] for v in sol)
        out["solve_kernels"] = dict(bound="hbm", bytes=by, achieved_gbs=round(by / (ms * 1e-3) / 1e9, 1),
                                    peak=PEAK_HBM_GBS, frac=round(by / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 3))
    return out


def level_shapes(H, P):
    """(n, k, M) of every level of the tree the workload builds."""
    t = H.HSS2D_build_index_sets(P)
    return {lev: t.level_info(lev) for lev in range(t.depth + 1)}


def make_problem(H, torch, kernel, args):
    """Params and the seeded unit-norm right-hand side of one workload."""
    gen = torch.Generator(device="cpu").manual_seed(0)
    if kernel == "helmholtz":
        P = H.HSS_tree_parameters("SL_H_2D", args.kh, 1, args.Ng, args.np_, 1e-10, 3, 800,
                                  args.TI, proxy_layers=args.proxy_layers,
                                  preset_all=args.preset_all)
        b = torch.complex(torch.randn(P.N, generator=gen, dtype=torch.float64),
                          torch.randn(P.N, generator=gen, dtype=torch.float64))
        wl = (f"HSS2D SL_H_2D Lippmann-Schwinger k={args.kh:g} {'TI' if args.TI else 'NTI'} "
              f"Ng={args.Ng} np={args.np_} N={P.N} acc=1e-10 layers=3 proxy_layers="
              f"{args.proxy_layers} dense blocks, tanh-plateau contrast (Bump_function "
              "option 4), "
              + ("'bdry' skeletons at every level" if args.preset_all else
                 "ID skeletons at the two bottom levels" + {
                     "id": "",
                     "table": " (randomized IDs, sketches sized by the closed-form rank guess)",
                     "mlp": " (randomized IDs, sketches sized by the rank network)"}[args.rank]))
    else:
        P = H.HSS_tree_parameters("SL_L_2D", 0, 1, args.Ng, args.np_, 1e-10, 2, 400, args.TI)
        b = torch.randn(P.N, generator=gen, dtype=torch.float64)
        wl = (f"HSS2D SL_L_2D {'TI' if args.TI else 'NTI'} Ng={args.Ng} np={args.np_} N={P.N} "
              f"acc=1e-10 layers=2 dense blocks")
    return P, (b / b.norm()).to("cuda:0"), wl


def summarize(runs, N, split=None):
    """Averages of the overlapped steps; `split` (one synchronised step) gives the breakdown of
    the build into the tree (with the compressed operator) and the inversion."""
    def path(prec):
        build = sum(r[prec]["build"] for r in runs) / len(runs)
        solve = sum(r[prec]["solve"] for r in runs) / len(runs)
        last = runs[-1][prec]
        out = {"build_s": round(build, 4), "solve_ms": round(1e3 * solve, 3),
               "iters": last["iters"], "relres": last["relres"],
               "total_s": round(build + solve, 4), "N_per_s": round(N / (build + solve), 1)}
        if split:
            out["synchronised"] = {"tree_s": round(split[prec]["tree"], 4),
                                   "invert_s": round(split[prec]["invert"], 4),
                                   "solve_ms": round(1e3 * split[prec]["solve"], 3),
                                   "solve_replayed_ms": round(1e3 * split[prec]["solve_replayed"], 3)}
        return out

    mixed, fp64 = path("fp32"), path("fp64")
    return {"mixed": mixed, "fp64": fp64,
            "speedup_mixed_vs_fp64": round(fp64["total_s"] / mixed["total_s"], 3)}


def roofline(stats):
    """Roofline object of the kernel class with the largest summed time in the
    instrumented step: algorithmic flops (or bytes) per launch / average launch time."""
    if not stats:
        return None
    name, s = max(stats.items(), key=lambda kv: kv[1]["ms"])
    prec = next(t for t in ("f64", "f32", "c128", "c64") if f"_{t}" in name)
    common = dict(kernel=name, traffic=MEASURED_TRAFFIC.get(name), launches=s["launches"],
                  avg_launch_us=round(1e3 * s["ms"] / s["launches"], 2),
                  algorithmic_bytes_per_launch=s["bytes"] / s["launches"])
    if name.startswith(("gemm", "gj_", "cqr_panel")):
        ach = s["flops"] / (s["ms"] * 1e-3) / 1e12
        return dict(bound="mfma", achieved=round(ach, 3), peak=PEAK_TFLOPS[prec], unit="TFLOP/s",
                    frac=round(ach / PEAK_TFLOPS[prec], 4), flops_per_launch=s["flops"] / s["launches"],
                    **common)
    ach = s["bytes"] / (s["ms"] * 1e-3) / 1e9
    return dict(bound="hbm", achieved=round(ach, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                frac=round(ach / PEAK_HBM_GBS, 4), bytes_per_launch=s["bytes"] / s["launches"],
                **common)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported "
                                   "(the tree would shard by subtrees, one per GPU)"}))
        sys.exit(2)
    import torch
    import hssc_directsolver_amd as H
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}))
        sys.exit(1)
    torch.cuda.set_device(0)
    if args.calibrate_copy:
        n = 1 << 28  # float32 elements: 1 GiB
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 4 * n, "copies": 3}), flush=True)
        return
    P, b, workload = make_problem(H, torch, args.kernel, args)
    N = P.N
    rank = args.rank if args.kernel == "helmholtz" and not args.preset_all else "id"
    mlp, rank_info = None, None
    if rank == "mlp":
        # the rank network, trained before the timed steps on the plain-ID ranks of 32 seeded
        # geometries (grids 24^2..640^2, kh 4..40, plateau or Gaussian contrast)
        from hssc_directsolver_amd.rank_mlp import RankMLP, collect_hss2d, train_rank_mlp
        if args.mlp_cache and os.path.exists(args.mlp_cache):
            mlp = RankMLP.load(args.mlp_cache)
            rank_info = dict(loaded=args.mlp_cache)
        else:
            t0 = time.perf_counter()
            Xr, tr = collect_hss2d(32, seed=7)
            mlp, loss = train_rank_mlp(Xr, tr, seed=0, epochs=2000)
            rank_info = dict(samples=len(tr), train_loss=round(loss, 6),
                             collect_and_train_s=round(time.perf_counter() - t0, 2))
            if args.mlp_cache:
                mlp.save(args.mlp_cache)
    step = lambda **kw: one_step(H, P, torch, b, args.tol, rank=rank, mlp=mlp, **kw)  # noqa: E731
    if args.profile_step:
        r = step()
        print(json.dumps({"profile_step": {k: round(v["build"] + v["solve"], 4)
                                           for k, v in r.items()}}), flush=True)
        return

    for _ in range(args.warmup):
        step()

    L = H.lib()
    torch.cuda.synchronize()
    T0 = time.perf_counter()
    runs = []
    for _ in range(args.steps):
        runs.append(step())
    torch.cuda.synchronize()
    T1 = time.perf_counter()
    split = step(split=True)
    if rank != "id":
        # the same synchronised step with plain skeleton IDs: what the rank buckets save
        plain = one_step(H, P, torch, b, args.tol, split=True)
        rank_info = dict(mode=rank, **(rank_info or {}),
                         ids=split["fp32"]["ids"], plain_ids=plain["fp32"]["ids"],
                         tree_s={"buckets": {p_: round(split[p_]["tree"], 4) for p_ in split},
                                 "plain": {p_: round(plain[p_]["tree"], 4) for p_ in plain}})
    # One more step of the same workload with per-kernel HIP-event instrumentation (same launches
    # and streams as the timed steps, so event times match a kernel trace of those steps).
    L.hssc_stats_reset()
    L.hssc_stats_enable(1)
    step()
    torch.cuda.synchronize()
    L.hssc_stats_enable(0)
    stats = stats_snapshot(H)
    summ = summarize(runs, N, split)

    side = None
    if args.kernel == "helmholtz" and not args.no_laplace:
        Pl, bl, wl = make_problem(H, torch, "laplace", args)
        one_step(H, Pl, torch, bl, args.tol)  # warm-up
        side = dict(workload=wl, **summarize([one_step(H, Pl, torch, bl, args.tol)], Pl.N))
    cpu = None if args.no_cpu else cpu_baseline(H, P, args)
    h1 = hss1d_case(H, torch) if not args.no_hss1d else None
    hb = batched_case(H, torch) if not args.no_batched else None
    cx = args.kernel == "helmholtz"
    levels = {lev: [v["n"], v["k"], v["M"]] for lev, v in level_shapes(H, P).items()}

    out = {
        "metric": "HSS2D N=512² build s + solve ms @1e-10 resid: mixed+batched vs FP64 GPU",
        "value": summ["mixed"]["total_s"],
        "unit": "s (build + solve to 1e-10 true residual, mixed FP32 tree and factorization + FP64 refinement)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * (T1 - T0) / args.steps, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c64 tree and factorization / c128 refinement" if cx
                  else "f32 tree and factorization / f64 refinement"),
        "data": "synthetic (seeded Gaussian RHS on the reference's uniform grid)",
        "config": {"workload": workload, "N": N},
        **summ,
        "roofline": roofline(stats),
        "phases": phase_rooflines(levels, split, stats, cx),
        "kernels": {k: dict(launches=v["launches"], ms=round(v["ms"], 3),
                            tflops=round(v["flops"] / (v["ms"] * 1e-3) / 1e12, 3) if v["ms"] else None,
                            gbs=round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] else None)
                    for k, v in stats.items()},
        "cpu_baseline": cpu,
        "rank_buckets": rank_info,
        "levels": levels,
        "laplace": side,
        "hss1d": h1,
        "hss1d_batched": hb,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
