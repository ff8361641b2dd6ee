"""The element-matrix kernels on the GPU: exact (FP64 VALU, the reference's operation order) against
the FP64 matrix cores, on the P3 workload or a P2 -Laplace u + u system.

    python -m femd_amd.assembly_bench [--problem p3|p2] [--reps 20]

Times the element kernel alone with HIP events (reps launches back to back, after one untimed),
then the whole matrix assembly (coefficient programs, element kernel, gather) in both modes, and
compares the two matrices row by row.  Reports the direct contraction's flops (2 nloc^2 nq per
term and cell) as TFLOP/s, the matrix-core flops issued (2048 per instruction) against the FP64
peak, and the kernel's least HBM bytes (element matrices written, coefficient values and cell
geometry read) as GB/s.  One JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--problem", choices=("p3", "p2"), default="p3")
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args(argv)

    import femd_amd as fd
    from femd_amd import _lib as L
    from femd_amd import problems
    lib = L.load()
    L.check(lib.femd_init(0))
    if args.problem == "p3":
        V, a_form, _, info = problems.p3_kappa_problem()
        ncells = int(info["triangles"].shape[0])
        workload = "P3 -div(kappa grad u), seeded Delaunay unit square (min angle 28), h = 0.00265"
    else:
        m = fd.rectangle_mesh(0, 0, 1, 1, 512, 512)
        V = fd.LagrangeSpace(m, 2, bc="dirichlet")
        u, v = fd.TrialFunction(V), fd.TestFunction(V)
        a_form = fd.form((fd.dot(fd.grad(u), fd.grad(v)) + u * v) * fd.dx)
        ncells = 2 * 512 * 512
        workload = "P2 -Laplace u + u on a 512x512 structured triangulation"
    nloc = (V.degree + 1) * (V.degree + 2) // 2
    coeffs, keep = a_form._coefficients()
    aa = L.i32([t.test.code() for t in a_form.terms])
    ab = L.i32([t.trial.code() for t in a_form.terms])
    nt, deg = len(a_form.terms), a_form.quad_degree
    nq_c = C.c_int()
    buf = [np.zeros(256) for _ in range(3)]
    L.check(lib.femd_quadrature(deg, C.byref(nq_c), *buf))
    nq = nq_c.value

    ms, fl = np.zeros(1), np.zeros(1)
    L.check(lib.femd_assembly_kernel_time(V._h, deg, nt, aa, ab, coeffs, 0, max(2, args.reps // 4), ms, fl))
    t_exact = float(ms[0])
    L.check(lib.femd_assembly_kernel_time(V._h, deg, nt, aa, ab, coeffs, 1, args.reps, ms, fl))
    t_mf, fl_mf = float(ms[0]), float(fl[0])

    def assembly(mode, reps=5):
        prev = fd.assembly_mode(mode)
        try:
            A = a_form.assemble(reuse=False)
            t0 = time.perf_counter()
            for _ in range(reps):
                A = a_form.assemble(reuse=False)
            return A, (time.perf_counter() - t0) / reps
        finally:
            fd.assembly_mode(prev)

    A_ex, t_asm_ex = assembly("exact")
    A_mf, t_asm_mf = assembly("mfma")
    ip = A_ex.indptr
    rows = np.maximum.reduceat(np.abs(A_ex.data), ip[:-1])
    err = float((np.maximum.reduceat(np.abs(A_mf.data - A_ex.data), ip[:-1]) / rows).max())

    codes = [np.asarray(p.arrays()[0]).reshape(-1, 4) for p in a_form.programs]
    progs = {c.tobytes() for c in codes if not (c.shape[0] == 1 and c[0, 0] == 0)}
    direct = 2.0 * nloc * nloc * nq * nt * ncells
    nbytes = 8 * nloc * nloc * ncells + 8 * nq * ncells * len(progs) + 48 * ncells
    out = {
        "workload": workload, "cells": ncells, "nloc": nloc, "quad_points": nq, "terms": nt,
        "exact_us": round(1e3 * t_exact, 2), "mfma_us": round(1e3 * t_mf, 2), "speedup": round(t_exact / t_mf, 2),
        "direct_flops": direct, "exact_tflops": round(direct / (t_exact * 1e-3) / 1e12, 2),
        "mfma_issued_flops": fl_mf, "mfma_issued_tflops": round(fl_mf / (t_mf * 1e-3) / 1e12, 2),
        "mfma_frac_of_fp64_peak": round(fl_mf / (t_mf * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
        "algorithmic_bytes": nbytes, "mfma_gbs": round(nbytes / (t_mf * 1e-3) / 1e9, 1),
        "mfma_hbm_frac": round(nbytes / (t_mf * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "assembly_ms": {"exact": round(1e3 * t_asm_ex, 3), "mfma": round(1e3 * t_asm_mf, 3)},
        "row_rel_err_mfma_vs_exact": err,
    }
    print(f"element kernel: exact {out['exact_us']} us ({out['exact_tflops']} TFLOP/s direct), matrix cores "
          f"{out['mfma_us']} us ({out['mfma_issued_tflops']} TFLOP/s issued, {out['mfma_gbs']} GB/s); assembly "
          f"{out['assembly_ms']['exact']} / {out['assembly_ms']['mfma']} ms; row-relative difference {err:.2e}", flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
