"""Low-precision local matrices and FP64 iterative refinement: what 16-bit element matrices do to a
solve.

    python -m femd_amd.precision_bench [--problems c0,heat,kappa] [--max-outer 40]

For each problem the matrix is assembled exactly (FP64, the reference's order) and on the matrix
cores with FP64, f16 and bf16 operands (femd_amd.assembly_mode); the entries' error against the
exact matrix is reported row-relative (max over rows of max |dA_ij| / max |A_ij|).  Then the system
A x = b is solved by iterative refinement with the low-precision matrix A_lo as the correction
operator: r = b - A x in FP64 with the exact A, A_lo d = r solved by FP64 Jacobi-CG to 1e-12 (so
that only the matrix's perturbation limits the refinement), x += d, until ||r|| <= 1e-8 ||b|| or
max_outer corrections.  The contraction per correction, ||r_{k+1}|| / ||r_k||, is what
||I - A A_lo^{-1}|| comes to on the problem; the final solution is compared with the exact
matrix's.  One JSON line per problem and mode.
"""
from __future__ import annotations

import argparse
import json
import time
import warnings

import numpy as np


def _problems(names):
    import femd_amd as fd
    from femd_amd import problems
    out = []
    for name in names:
        if name == "c0":
            V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 32, 32), 2, bc="dirichlet")
            u, v = fd.TrialFunction(V), fd.TestFunction(V)
            a = fd.form((fd.dot(fd.grad(u), fd.grad(v)) + u * v) * fd.dx)
            f, _ = problems.fourier_source(0)
            out.append(("c0: P2 -Laplace u + u, 32x32", V, a, fd.form(f * v * fd.dx)))
        elif name == "heat":
            V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 128, 128), 2, bc="dirichlet")
            u, v = fd.TrialFunction(V), fd.TestFunction(V)
            a = fd.form((u * v + 1e-3 * fd.dot(fd.grad(u), fd.grad(v))) * fd.dx)
            f, _ = problems.fourier_source(2)
            out.append(("heat: P2 M + dt K (dt 1e-3), 128x128", V, a, fd.form(f * v * fd.dx)))
        elif name == "kappa":
            V, a, Lf, _ = problems.p3_kappa_problem(0.01)
            out.append(("kappa: P3 -div(kappa grad u), Delaunay h = 0.01", V, a, Lf))
        else:
            raise ValueError(f"unknown problem {name}")
    return out


def row_err(ip, got, want):
    rows = np.maximum.reduceat(np.abs(want), ip[:-1])
    return float((np.maximum.reduceat(np.abs(got - want), ip[:-1]) / np.where(rows > 0, rows, 1.0)).max())


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", default="c0,heat,kappa")
    ap.add_argument("--max-outer", type=int, default=40)
    ap.add_argument("--rtol", type=float, default=1e-8)
    args = ap.parse_args(argv)

    import femd_amd as fd
    from femd_amd import _lib as L
    L.check(L.load().femd_init(0))
    for label, V, a, Lf in _problems(args.problems.split(",")):
        A = a.assemble(reuse=False)
        b = Lf.assemble()
        nb = float(np.linalg.norm(b))
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            xe, ie = fd.cg(A, b, M=A.preconditioner("jacobi"), tol=1e-12, maxiter=200000)
        ip = A.indptr
        for mode in ("mfma", "mfma_f16", "mfma_bf16"):
            prev = fd.assembly_mode(mode)
            try:
                Alo = a.assemble(reuse=False)
            finally:
                fd.assembly_mode(prev)
            err = row_err(ip, Alo.data, A.data)
            rel_f = float(np.linalg.norm(Alo.data - A.data) / np.linalg.norm(A.data))
            Mlo = Alo.preconditioner("jacobi")
            x = np.zeros(A.shape[0])
            hist, inner, converged, t0 = [], 0, False, time.perf_counter()
            for k in range(args.max_outer + 1):
                r = b - A.matvec(x)
                rho = float(np.linalg.norm(r)) / nb
                hist.append(rho)
                if rho <= args.rtol:
                    converged = True
                    break
                if k == args.max_outer or not np.isfinite(rho) or (k >= 3 and rho > 10 * hist[0]):
                    break
                with warnings.catch_warnings():
                    warnings.simplefilter("ignore", RuntimeWarning)
                    d, info = fd.cg(Alo, r, M=Mlo, tol=1e-12, maxiter=200000)
                inner += info.iterations
                x = x + d
            secs = time.perf_counter() - t0
            steps = np.array(hist[1:]) / np.array(hist[:-1]) if len(hist) > 1 else np.array([])
            out = {"problem": label, "dim": int(A.shape[0]), "nnz": int(A.indices.size), "mode": mode,
                   "row_rel_err": err, "frobenius_rel_err": rel_f,
                   "refinement": {"converged": converged, "corrections": len(hist) - 1,
                                  "contraction_geomean": float(np.exp(np.mean(np.log(steps)))) if steps.size else None,
                                  "residuals": [float(f"{h:.3e}") for h in hist],
                                  "inner_cg_iterations": inner, "seconds": round(secs, 2),
                                  "rel_err_vs_exact_solution": float(np.linalg.norm(x - xe) / np.linalg.norm(xe))}}
            print(f"{label} [{mode}]: entries within {err:.1e} (row-relative), refinement "
                  f"{'converged' if converged else 'did not converge'} after {len(hist) - 1} corrections, "
                  f"contraction {out['refinement']['contraction_geomean']}", flush=True)
            print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
