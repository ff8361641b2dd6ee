"""ILU(0) and SSOR applications on the GPU, on the P3 workload (or a P2 convection system).

    python -m femd_amd.precond_bench [--problem p3|conv] [--reps 10] [--cg]

Builds the matrix, then ILU(0) under RCM and SSOR (omega 1.2) preconditioners, and times one
application z = M^-1 r with HIP events (reps applications back to back): the two planned sweeps,
the gather of r into the forward stream and the two scatters of the sweeps' results.  Reports the
application's time, its algorithmic bytes (12 per stored entry of the factor, 28 per unknown) as
GB/s, the sweeps' steps and dependency levels, the seconds per dependency level, and, with --cg,
one FP64 ILU(0)-RCM CG solve to 1e-8.  One JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
import warnings


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--problem", choices=("p3", "conv"), default="p3")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--cg", action="store_true")
    args = ap.parse_args(argv)

    import femd_amd as fd
    from femd_amd import _lib as L
    from femd_amd import problems
    from femd_amd.roofline import precond_apply_time
    L.check(L.load().femd_init(0))
    if args.problem == "p3":
        V, a, Lf, _ = problems.p3_kappa_problem()
        workload = "P3 -div(kappa grad u) = f, seeded Delaunay unit square (min angle 28), h = 0.00265"
    else:
        V, a, Lf = problems.convection_problem(256, 2)
        workload = "P2 -0.01 Laplace u + (1, 0.5) . grad u = f on 256x256"
    A = a.assemble(reuse=False)
    n, nnz = A.shape[0], int(A.indices.size)
    out = {"workload": workload, "unknowns": n, "nnz": nnz}
    for kind, omega in (("ilu0", 1.0), ("ssor", 1.2)):
        t0 = time.perf_counter()
        M = A.preconditioner(kind, omega=omega)
        t_make = time.perf_counter() - t0
        t = precond_apply_time(M, n, reps=args.reps)
        st = M.stats()
        B = 12 * nnz + 28 * n
        levels = st["levels_forward"] + st["levels_backward"]
        out[kind] = {"setup_s": round(t_make, 3), "apply_ms": round(t, 4), "algorithmic_bytes": B,
                     "gbs": round(B / (t * 1e-3) / 1e9, 2), "us_per_level": round(1e3 * t / max(levels, 1), 4),
                     **st}
        print(f"[precond] {kind}: setup {t_make:.2f} s, apply {t:.3f} ms, {out[kind]['gbs']} GB/s, "
              f"{out[kind]['us_per_level']} us per level", file=sys.stderr, flush=True)
        if args.cg and kind == "ilu0":
            b = Lf.assemble()
            t0 = time.perf_counter()
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", RuntimeWarning)
                _, info = fd.cg(A, b, M=M, tol=1e-8, maxiter=20000)
            out[kind]["cg"] = {"iterations": info.iterations, "residual": info.residual,
                               "seconds": round(time.perf_counter() - t0, 3)}
            print(f"[precond] ilu0 CG: {out[kind]['cg']}", file=sys.stderr, flush=True)
        del M
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
