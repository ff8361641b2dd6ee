"""The batched uneven Poisson stream: the mixed-precision batched solver against the plain FP64
GPU path on the same systems.

    python -m femd_amd.batch_bench [--count 256] [--nmin 16] [--nmax 512] [--seed 0]

Workload: `count` P2 problems -Laplace u = f on structured N x N triangulations of the unit square
(every side Dirichlet), N drawn log-uniform in [nmin, nmax] from the seed, f the 8-mode random
Fourier source of problem i's own seed.  All matrices and right-hand sides are assembled on the
GPU first (not timed in either path).  Then, with every vector resident in HBM:
  * plain FP64: one Jacobi-CG solve per system in turn (cg_jacobi's arithmetic, the reference's
    algorithm; tolerance 1e-10 relative), timed together;
  * batched mixed: all systems in one femd_batch solve (FP32 Jacobi-CG iterations shared by all
    systems, reliable FP64 updates per system) to --mixed-tol (default: the same 1e-10), timed.
Every batched solution is compared with the FP64 one (relative 2-norm error; the bar is 1e-8), and
one JSON line is printed: both solves/s figures, their ratio, the error, the systems' sizes and
iteration counts.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import sys
import time

import numpy as np


def problem_sizes(count, nmin, nmax, seed):
    rng = np.random.default_rng(seed)
    return [int(round(math.exp(u))) for u in rng.uniform(math.log(nmin), math.log(nmax), count)]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--nmin", type=int, default=16)
    ap.add_argument("--nmax", type=int, default=512)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-10, help="relative residual of the plain FP64 solves")
    ap.add_argument("--mixed-tol", type=float, default=None,
                    help="relative FP64 residual the batched solve stops at (default: --tol); its error against the "
                         "FP64 solutions is measured and reported")
    ap.add_argument("--inner-tol", type=float, default=0.3)
    ap.add_argument("--sweep", default="", help="comma-separated mixed-tol:inner-tol pairs to time as well")
    ap.add_argument("--no-graph", action="store_true", help="batched solves with plain launches instead of graph replays")
    ap.add_argument("--no-dict", action="store_true", help="batched solves read the matrix values, not value dictionaries")
    ap.add_argument("--original-order", action="store_true",
                    help="the inner CG keeps the original order of the unknowns instead of its own (first-met columns)")
    ap.add_argument("--parts", default="plain,fp64,mixed",
                    help="which solves to run (a profiling run may leave out the plain path; errors then are not measured)")
    args = ap.parse_args(argv)
    parts = set(args.parts.split(","))

    import femd_amd as fd
    from femd_amd import _lib as L
    from femd_amd import problems

    lib = L.load()
    L.check(lib.femd_init(0))
    lib.femd_value_dictionary(0 if args.no_dict else 1)
    lib.femd_inner_order(0 if args.original_order else 1)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)

    sizes = problem_sizes(args.count, args.nmin, args.nmax, args.seed)
    t0 = time.perf_counter()
    systems = []
    for i, N in enumerate(sizes):
        V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, N, N), 2, bc="dirichlet")
        f, _ = problems.fourier_source(args.seed * 100003 + i)
        a, Lf = problems.poisson_forms(V, f)
        systems.append((V, a.assemble(), Lf.assemble()))
    dims = [V.dim for V, _, _ in systems]
    nnzs = [A.nnz for _, A, _ in systems]
    log(f"setup: {len(systems)} systems, {sum(dims)} unknowns, {sum(nnzs)} stored entries, "
        f"N {min(sizes)}..{max(sizes)}, {time.perf_counter() - t0:.1f} s")

    def dev(nbytes):
        p = C.c_void_p()
        L.check(lib.femd_dev_alloc(max(nbytes, 8), C.byref(p)))
        return p.value

    # ---- plain FP64: one Jacobi-CG per system, vectors resident ----------------------------------
    bdev = [dev(8 * n) for n in dims]
    xdev = [dev(8 * n) for n in dims]
    precs = []
    for (V, A, b), bd in zip(systems, bdev):
        L.check(lib.femd_memcpy_htod(bd, b.ctypes.data, 8 * b.size))
        h, mp = C.c_void_p(), C.c_double()
        L.check(lib.femd_precond_create(A._h, L.PC_JACOBI, 0, 1.0, C.byref(h), C.byref(mp)))
        precs.append(h.value)
    zeros = np.zeros(max(dims))

    def plain_round():
        its = []
        for (V, A, b), bd, xd, M in zip(systems, bdev, xdev, precs):
            L.check(lib.femd_memcpy_htod(xd, zeros.ctypes.data, 8 * b.size))
            info = L.SolveInfo()
            L.check(lib.femd_cg_dev(A._h, M, bd, xd, max(1000, 10 * b.size), args.tol, C.byref(info), None))
            its.append(int(info.iterations))
        L.check(lib.femd_stream_sync(None))
        return its

    x64 = None
    t_plain, its64 = float("nan"), [0]
    if "plain" in parts:
        plain_round()                      # warm-up: tuning, graph and pool set-up
        t1 = time.perf_counter()
        its64 = plain_round()
        t_plain = time.perf_counter() - t1
        x64 = []
        for (V, A, b), xd in zip(systems, xdev):
            x = np.empty(b.size)
            L.check(lib.femd_memcpy_dtoh(x.ctypes.data, xd, 8 * b.size))
            x64.append(x)
        log(f"plain FP64: {t_plain:.3f} s, {sum(its64)} iterations in total (max {max(its64)}), "
            f"{sum(n * k for n, k in zip(dims, its64))} row-iterations")

    # ---- batched mixed precision ----------------------------------------------------------------
    B = fd.Batch([A for _, A, _ in systems])
    bcat = B.concat([b for _, _, b in systems])
    bd, xd = dev(8 * B.nrows), dev(8 * B.nrows)
    L.check(lib.femd_memcpy_htod(bd, bcat.ctypes.data, 8 * B.nrows))
    zcat = np.zeros(B.nrows)
    its32 = np.zeros(len(systems), dtype=np.int32)
    res = np.zeros(len(systems))

    mtol = args.mixed_tol if args.mixed_tol is not None else args.tol

    def batched_round(tol, itol, prec=L.FP32):
        L.check(lib.femd_memcpy_htod(xd, zcat.ctypes.data, 8 * B.nrows))
        info = L.SolveInfo()
        L.check(lib.femd_batch_solve_dev(B._h, bd, xd, tol, itol, 200, prec, 0 if args.no_graph else 1, its32.ctypes.data, res.ctypes.data,
                                         C.byref(info), None))
        L.check(lib.femd_stream_sync(None))
        return info

    def errors():
        if x64 is None:
            return [float("nan")]
        xb = np.empty(B.nrows)
        L.check(lib.femd_memcpy_dtoh(xb.ctypes.data, xd, 8 * B.nrows))
        return [float(np.linalg.norm(xs - x) / max(np.linalg.norm(x), 1e-300)) for xs, x in zip(B.split(xb), x64)]

    # the batched FP64 solve (same machinery, FP64 inner arithmetic)
    b64 = None
    if "fp64" in parts:
        batched_round(args.tol, args.tol, L.FP64)
        t3 = time.perf_counter()
        info64 = batched_round(args.tol, args.tol, L.FP64)
        t_b64 = time.perf_counter() - t3
        errs64 = errors()
        log(f"batched FP64: {t_b64:.3f} s, converged {bool(info64.converged)}, max error vs FP64 {max(errs64):.2e}, "
            f"inner iterations max {int(its32.max())}, {int(np.dot(dims, its32))} row-iterations")
        b64 = {"seconds": round(t_b64, 4), "solves_per_s": round(len(systems) / t_b64, 3),
               "speedup": round(t_plain / t_b64, 3), "converged": bool(info64.converged),
               "max_rel_err_vs_fp64": max(errs64), "inner_iterations_max": int(its32.max()),
               "inner_iterations_total": int(info64.iterations), "row_iterations": int(np.dot(dims, its32)),
               "value_dictionary": B.dictionary("fp64")}
    sweep = []
    for pair in [p for p in args.sweep.split(",") if p]:
        mt, it = (float(v) for v in pair.split(":"))
        batched_round(mt, it)
        t = time.perf_counter()
        inf = batched_round(mt, it)
        t = time.perf_counter() - t
        e = errors()
        sweep.append({"mixed_tol": mt, "inner_tol": it, "seconds": round(t, 4), "speedup": round(t_plain / t, 3),
                      "max_rel_err_vs_fp64": max(e), "inner_iterations_max": int(its32.max()),
                      "inner_iterations_total": int(inf.iterations), "rounds": int(inf.breakdown)})
        log("sweep", sweep[-1])
    t_batch, info, errs = float("nan"), L.SolveInfo(), [float("nan")]
    if "mixed" in parts:
        batched_round(mtol, args.inner_tol)
        t2 = time.perf_counter()
        info = batched_round(mtol, args.inner_tol)
        t_batch = time.perf_counter() - t2
        errs = errors()
        log(f"batched mixed: {t_batch:.3f} s, converged {bool(info.converged)}, refinement rounds {info.breakdown}, "
            f"max error vs FP64 {max(errs):.2e}, {int(np.dot(dims, its32))} row-iterations")
    out = {
        "workload": f"{len(systems)} P2 Poisson systems, structured N x N, N log-uniform in [{args.nmin}, {args.nmax}]",
        "seed": args.seed, "systems": len(systems), "unknowns": int(sum(dims)), "nnz": int(sum(nnzs)),
        "N": sizes, "tol": args.tol, "mixed_tol": mtol, "inner_tol": args.inner_tol,
        "plain_fp64": {"seconds": round(t_plain, 4), "solves_per_s": round(len(systems) / t_plain, 3),
                       "iterations_total": int(sum(its64)), "iterations_max": int(max(its64)),
                       "row_iterations": int(sum(n * k for n, k in zip(dims, its64)))},
        "batched_mixed": {"seconds": round(t_batch, 4), "solves_per_s": round(len(systems) / t_batch, 3),
                          "converged": bool(info.converged), "refinement_rounds": int(info.breakdown),
                          "inner_iterations_max": int(its32.max()), "worst_residual": float(info.residual),
                          "row_iterations": int(np.dot(dims, its32)),
                          "value_dictionary": B.dictionary("fp32") if "mixed" in parts else None},
        "speedup": round(t_plain / t_batch, 3), "max_rel_err_vs_fp64": max(errs),
        "batched_fp64": b64,
        "sweep": sweep,
    }
    print(json.dumps(out), flush=True)
    for M in precs:
        lib.femd_precond_free(M)
    for p in bdev + xdev + [bd, xd]:
        lib.femd_dev_free(p)


if __name__ == "__main__":
    main()
