"""GMRES / LGMRES on a nonsymmetric system: -eps Laplace u + beta . grad u = f, P2 on an N x N
structured triangulation of the unit square (eps = 0.01, beta = (1, 0.5), every side Dirichlet).

    python -m femd_amd.krylov_bench [--N 256] [--tol 1e-10] [--no-cpu]

The matrix and right-hand side are assembled once on the GPU.  Each leg -- GMRES(30) and
LGMRES(20, 2), preconditioned by ILU(0) under RCM and by Jacobi -- is solved on the GPU (the
device Arnoldi steps: MGS, Hessenberg column, Givens rotation and residual estimate on the device,
one host read per restart cycle) with its preconditioner set up inside the timed region, and by the
reference's own code (oracle/_ref) on the host cores from the same CSR arrays.  One JSON line:
per leg the GPU's seconds, iterations and residuals, the reference's, and the relative difference
of the two solutions.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
import warnings

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--eps", type=float, default=0.01)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--maxiter", type=int, default=20000)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args(argv)

    import femd_amd as fd
    from femd_amd import _lib as L
    from femd_amd import problems
    L.check(L.load().femd_init(0))
    V, a, Lf = problems.convection_problem(args.N, 2, eps=args.eps)
    A = a.assemble()
    b = Lf.assemble()
    n, nnz = A.shape[0], int(A.indices.size)
    print(f"[krylov] {n} unknowns, {nnz} entries, asymmetric: {not A.symmetric}", file=sys.stderr, flush=True)
    R = None
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if not args.no_cpu:
        import oracle
        R = oracle.reference() or oracle.oracle()
        R.set_threads(cores)
    legs = {}
    for pc in ("ilu0", "jacobi"):
        for lg in (False, True):
            name = f"{'lgmres(20,2)' if lg else 'gmres(30)'}+{pc}"
            restart = 20 if lg else 30
            f = fd.lgmres if lg else fd.gmres
            f(A, b, M=A.preconditioner(pc), restart=restart, tol=args.tol, maxiter=2, warn=False)   # warm-up
            t0 = time.perf_counter()
            M = A.preconditioner(pc)
            x, info = f(A, b, M=M, restart=restart, tol=args.tol, maxiter=args.maxiter, warn=False)
            t_gpu = time.perf_counter() - t0
            leg = {"gpu_s": round(t_gpu, 4), "gpu_iterations": info.iterations, "gpu_converged": bool(info.converged),
                   "gpu_residual": info.residual, "gpu_precond_residual": info.precond_residual,
                   "gpu_ms_per_iteration": round(1e3 * t_gpu / max(info.iterations, 1), 4)}
            if R is not None:
                t0 = time.perf_counter()
                xr, io = R.krylov(A.indptr, A.indices, A.data, b, M=pc, lgmres=lg, restart=restart, tol=args.tol,
                                  maxiter=args.maxiter)
                t_cpu = time.perf_counter() - t0
                leg.update({"cpu_s": round(t_cpu, 4), "cpu_iterations": io["iterations"],
                            "cpu_converged": io["converged"], "cpu_residual": io["residual"],
                            "rel_diff": float(np.linalg.norm(x - xr) / np.linalg.norm(xr)),
                            "speedup": round(t_cpu / t_gpu, 2)})
            legs[name] = leg
            print(f"[krylov] {name}: {leg}", file=sys.stderr, flush=True)
    print(json.dumps({"workload": f"-{args.eps:g} Laplace u + (1, 0.5) . grad u = f, P2 on {args.N}x{args.N}, "
                                  f"tol {args.tol:g}", "unknowns": n, "nnz": nnz, "cpu_cores": cores if R else None,
                      "legs": legs}), flush=True)


if __name__ == "__main__":
    main()
