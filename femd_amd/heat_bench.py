"""The heat configuration: u_t = Laplace u, P2 on a 256 x 256 structured triangulation, backward
Euler with dt = 1e-3 for 500 steps, u0 the random-Fourier field of the seed.

    python -m femd_amd.heat_bench [--N 256] [--steps 500] [--seed 0]

GPU paths advance the same u0 (every vector resident in HBM):
  * plain FP64: per step b = M u (FP64 product) and a Jacobi-CG solve warm-started at u to
    relative residual tol -- the reference's loop (example13.ipynb) with its cg_jacobi arithmetic;
  * graph stepper, FP32 and FP64 inner iterations: the whole step (FP64 b = M u, FP64 residual
    rounds of one scaled Jacobi-CG run) captured once into a hipGraph and replayed for every step,
    to the same tol.
The reference trajectory is the plain FP64 loop at the tighter tolerance ref_tol: per-step solve
errors add up over 500 steps (tol 1e-10 leaves two FP64 runs about 5e-6 apart, tol 1e-13 leaves the
FP32-inner stepper 1.3e-8 from the reference), so tol (5e-14) and ref_tol (1e-14) are set where that
sum stays below the 1e-8 bar for every path.  One JSON line: steps/s of every path, the
ratios to the plain one, inner iterations, and each final state's relative difference from the
reference (the bar is 1e-8).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--k", type=int, default=2)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=5e-14)
    ap.add_argument("--ref-tol", type=float, default=1e-14)
    ap.add_argument("--rounds", type=int, default=12)
    ap.add_argument("--inner-tol", type=float, default=1e-2)
    ap.add_argument("--max-inner", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--original-order", action="store_true",
                    help="the inner CG keeps the original order of the unknowns instead of its own (first-met columns)")
    args = ap.parse_args(argv)

    import femd_amd as fd
    from femd_amd import _lib as L
    from femd_amd import problems
    lib = L.load()
    L.check(lib.femd_init(0))
    lib.femd_inner_order(0 if args.original_order else 1)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    V, M, A, u0 = problems.heat_problem(args.N, args.k, dt=args.dt, seed=args.seed)
    n = V.dim
    log(f"setup: P{args.k} {args.N}x{args.N}, {n} unknowns, nnz {A.nnz}, {time.perf_counter() - t0:.1f} s")

    # ---- plain FP64 loop -------------------------------------------------------------------------
    def dev(nbytes):
        p = C.c_void_p()
        L.check(lib.femd_dev_alloc(max(nbytes, 8), C.byref(p)))
        return p.value
    ud, bd = dev(8 * n), dev(8 * n)
    h, mp = C.c_void_p(), C.c_double()
    L.check(lib.femd_precond_create(A._h, L.PC_JACOBI, 0, 1.0, C.byref(h), C.byref(mp)))
    Mj = h.value

    def plain(nsteps, tol):
        L.check(lib.femd_memcpy_htod(ud, u0.ctypes.data, 8 * n))
        its = 0
        for _ in range(nsteps):
            L.check(lib.femd_spmv_dev(M._h, ud, bd, L.FP64, None))
            info = L.SolveInfo()
            L.check(lib.femd_cg_dev(A._h, Mj, bd, ud, max(1000, 10 * n), tol, C.byref(info), None))
            its += int(info.iterations)
        L.check(lib.femd_stream_sync(None))
        out = np.empty(n)
        L.check(lib.femd_memcpy_dtoh(out.ctypes.data, ud, 8 * n))
        return its, out

    _, uref = plain(args.steps, args.ref_tol)
    plain(3, args.tol)
    t1 = time.perf_counter()
    its64, u64 = plain(args.steps, args.tol)
    t_plain = time.perf_counter() - t1
    rel = lambda u: float(np.linalg.norm(u - uref) / np.linalg.norm(uref))
    log(f"plain FP64: {args.steps} steps in {t_plain:.3f} s ({args.steps / t_plain:.1f} steps/s), {its64} CG iterations, "
        f"final state vs the reference {rel(u64):.2e}")
    out = {
        "workload": f"heat u_t = Laplace u, P{args.k} on {args.N}x{args.N}, backward Euler dt={args.dt}, "
                    f"{args.steps} steps, random-Fourier u0 (seed {args.seed})",
        "unknowns": n, "nnz": int(A.nnz), "steps": args.steps, "tol": args.tol, "rounds": args.rounds,
        "inner_tol": args.inner_tol, "max_inner": args.max_inner, "ref_tol": args.ref_tol,
        "plain_fp64": {"seconds": round(t_plain, 4), "steps_per_s": round(args.steps / t_plain, 2),
                       "cg_iterations": its64, "final_rel_diff_vs_reference": rel(u64)},
    }

    # ---- graph stepper -------------------------------------------------------------------------
    stepper = fd.BackwardEuler(M, A)
    kw = dict(tol=args.tol, rounds=args.rounds, inner_tol=args.inner_tol, max_inner=args.max_inner)
    for prec in ("fp32", "fp64"):
        stepper.run(u0, 3, precision=prec, **kw)
        t2 = time.perf_counter()
        u, info = stepper.run(u0, args.steps, precision=prec, **kw)
        t_graph = time.perf_counter() - t2
        t3 = time.perf_counter()
        u2, _ = stepper.run(u0, args.steps, precision=prec, graph=False, **kw)
        t_nograph = time.perf_counter() - t3
        log(f"graph stepper {prec}: {args.steps} steps in {t_graph:.3f} s ({args.steps / t_graph:.1f} steps/s), "
            f"{info['total_iterations']} inner iterations, worst step residual {info['residual']:.2e}, final state vs the "
            f"reference {rel(u):.2e}; without the graph {t_nograph:.3f} s")
        out[f"graph_stepper_{prec}"] = {
            "seconds": round(t_graph, 4), "steps_per_s": round(args.steps / t_graph, 2),
            "speedup_vs_plain": round(t_plain / t_graph, 3), "inner_iterations": info["total_iterations"],
            "every_step_converged": info["converged"], "failed_steps": info["failed_steps"], "worst_step_residual": info["residual"], "final_rel_diff_vs_reference": rel(u),
            "no_graph_seconds": round(t_nograph, 4), "bit_identical_without_graph": bool(np.array_equal(u, u2))}
    lib.femd_precond_free(Mj)
    for p in (ud, bd):
        lib.femd_dev_free(p)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
