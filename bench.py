"""Benchmark of the 2D assemble -> SpMV -> Krylov path on one MI355X.

Workload (the P3 configuration): -div(kappa grad u) = f on the unit square, a seeded
unstructured Delaunay mesh refined to a minimum angle of 28 degrees (h = 0.00265: ~1.48M P3
unknowns, ~25M stored entries), kappa = exp(0.5 g) with g a sum of 16 seeded Fourier modes, f the
seeded 8-mode sine series, homogeneous Dirichlet sides.

One step = assembly of A and b in FP64 on the GPU (coefficient programs evaluated at the 25 / 20
quadrature points of every cell on the device) + the mixed-precision solve: FP32 Jacobi-PCG inner
iterations, FP64 residuals and updates (iterative refinement) until ||b - A x|| / ||b|| <= rtol,
1e-8 by default (the configuration's refinement target).  `value` is solves per second (1 / mean
step time, inputs resident in HBM).  The error of every step's solution is measured against the
FP64 solution of the reference algorithm (cg_jacobi, bit-identical to the reference on the CPU)
solved to 1e-12, computed once outside the timed region; it must stay within 1e-8.  Two more
steps at the stricter rtol --strict-rtol (1e-10) are reported beside it.

Also reported: the roofline of the dominant kernel, the FP32 inner CG's product (its
algorithmic bytes 8 nnz + 20 n + 4 over its average duration, sampled with HIP events on its stream
in one extra step after the timed ones); the plain SpMV kernels' HBM throughput (12 nnz + 20 n + 4 bytes in FP64,
8 nnz + 12 n + 4 in FP32), back to back and with the Infinity Cache flushed, on the matrix as
numbered and on its copy in the inner CG's order of the unknowns; the plain FP64 GPU paths to the
same rtol, each timed once with its assembly: Jacobi-CG, and ILU(0)-RCM CG (the configuration's
FP64 solver) with its factorisation and the throughput of one preconditioner application
(12 nnz + 28 n algorithmic bytes); the matrix assembly's throughput (8 nnz + 4 nloc ncells +
16 npoints + 12 ncells algorithmic bytes); p50 / p99 step latency; and the CPU baseline: the
reference's own code (oracle/_ref) on the host cores, the full assembly and full ILU(0)-RCM CG and
Jacobi-CG solves to the same rtol, measured, their iteration counts and solutions compared with the
GPU's FP64 solves, plus a sampled 1-thread figure.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--h 0.00265] [--rtol 1e-8] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
import warnings
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "solves/sec at 1e-8 rel err vs FP64; SpMV HBM GB/s; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6          # MI355X dense FP64, matrix cores and vector FMA alike


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--h", type=float, default=0.00265)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--strict-rtol", type=float, default=1e-10,
                    help="two extra steps at this rtol, reported beside the headline (0: none)")
    ap.add_argument("--inner-tol", type=float, default=0.5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-full-limit", type=float, default=150.0,
                    help="run a CPU reference solve in full when a short sample predicts it within this many "
                         "seconds (otherwise it is reported from the sample)")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--original-order", action="store_true",
                    help="the inner CG keeps the original order of the unknowns instead of its own (first-met columns)")
    ap.add_argument("--exact-assembly", action="store_true",
                    help="assemble the timed solves' matrix with the exact kernels instead of the FP64 matrix cores")
    ap.add_argument("--product-shape", default=None,
                    help="pin the launch shape of the inner CG's product (e.g. tile512x4096) instead of the tuned one")
    ap.add_argument("--no-live-traffic", action="store_true",
                    help="take roofline.traffic from the committed profile instead of two rocprofv3 --pmc passes")
    ap.add_argument("--no-graph", action="store_true",
                    help="every solve with plain launches (for profilers that cannot follow graph replays)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import femd_amd as fd
    from femd_amd import _lib as L
    from femd_amd import problems

    lib = L.load()
    L.check(lib.femd_init(0))
    lib.femd_inner_order(0 if args.original_order else 1)
    if args.no_graph:
        L.check(lib.femd_inner_timer(1))       # the sampled kernel timer runs solves without graphs

    V, a_form, L_form, pinfo = problems.p3_kappa_problem(args.h)
    if args.product_shape:
        L.check(lib.femd_spmv_select(a_form.assemble(reuse=False)._h, L.FP32, args.product_shape.encode()))
    P, T, min_ang = pinfo["points"], pinfo["triangles"], pinfo["min_angle"]
    ip, ix = V.pattern()
    n, nnz = V.dim, int(ix.size)
    log(f"setup: {P.shape[0]} vertices, {T.shape[0]} cells, min angle {min_ang:.1f} deg, dim {n}, nnz {nnz}, "
        f"quadrature degrees a {a_form.quad_degree} / L {L_form.quad_degree}, {pinfo['setup_s']:.1f} s")

    # device-resident right-hand side and solution vectors for the timed loop
    def dev(nbytes):
        p = C.c_void_p()
        L.check(lib.femd_dev_alloc(nbytes, C.byref(p)))
        return p.value

    b_dev, x_dev = dev(8 * n), dev(8 * n)
    coeffs_L, keep_L = L_form._coefficients()
    aL = L.i32([t.test.code() for t in L_form.terms])
    zeros = np.zeros(n)

    def step(rtol=args.rtol):
        prev = fd.assembly_mode("exact" if args.exact_assembly else "mfma")
        A = a_form.assemble(reuse=False)
        fd.assembly_mode(prev)
        L.check(lib.femd_assemble_vector_dev(V._h, L_form.quad_degree, len(L_form.terms), aL, coeffs_L, b_dev, None))
        L.check(lib.femd_memcpy_htod(x_dev, zeros.ctypes.data, 8 * n))
        info = L.SolveInfo()
        L.check(lib.femd_solve_refined_dev(A._h, b_dev, x_dev, rtol, args.inner_tol, 60, C.byref(info), None))
        L.check(lib.femd_stream_sync(None))
        return A, info

    # the FP64 solution of the reference algorithm (exact assembly, Jacobi-CG: bit-identical to the
    # CPU reference)
    step()
    A = a_form.assemble(reuse=False)
    b = L_form.assemble()
    with warnings.catch_warnings():
        # the recurrence reaches 1e-12 before the true residual does (FP64 attainable accuracy on
        # this matrix is about 4e-10); the iterate is still the reference's most accurate one
        warnings.simplefilter("ignore", RuntimeWarning)
        x64, i64 = fd.cg(A, b, M=A.preconditioner("jacobi"), tol=1e-12, maxiter=200000)
    log(f"FP64 Jacobi-CG reference solution: {i64.iterations} iterations, residual {i64.residual:.2e}")
    # the plain FP64 GPU path at the same tolerance: assembly + Jacobi-CG, timed once
    t1 = time.perf_counter()
    A64 = a_form.assemble(reuse=False)
    b64 = L_form.assemble()
    xj64, ip64 = fd.cg(A64, b64, M=A64.preconditioner("jacobi"), tol=args.rtol, maxiter=200000)
    t_fp64 = time.perf_counter() - t1
    log(f"plain FP64 GPU path (assembly + Jacobi-CG to {args.rtol:g}): {ip64.iterations} iterations, {t_fp64:.2f} s")

    # ---- the FP64 ILU(0)-RCM CG leg: assembly, factor, solve to rtol on the GPU --------------
    t1 = time.perf_counter()
    Ai = a_form.assemble(reuse=False)
    bi = L_form.assemble()
    t2 = time.perf_counter()
    Milu = Ai.preconditioner("ilu0", ordering="rcm")
    t3 = time.perf_counter()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        xilu, iilu = fd.cg(Ai, bi, M=Milu, tol=args.rtol, maxiter=20000)
    t4 = time.perf_counter()
    from femd_amd.roofline import precond_apply_time
    t_apply = precond_apply_time(Milu, n, reps=5)
    sw = Milu.stats()
    # one application reads the factor's values and columns (12 bytes per stored entry of L + U,
    # which has A's pattern) and per unknown r, the permutation both ways, the inverse pivot and the
    # intermediate y (28 bytes): the two sweeps' algorithmic bytes
    B_apply = 12 * nnz + 28 * n
    ilu = {"iterations": iilu.iterations, "converged": bool(iilu.converged), "residual": iilu.residual,
           "assembly_s": round(t2 - t1, 3), "factor_s": round(t3 - t2, 3), "solve_s": round(t4 - t3, 3),
           "seconds": round(t4 - t1, 3), "solves_per_s": round(1.0 / (t4 - t1), 4),
           "ms_per_iteration": round(1e3 * (t4 - t3) / max(iilu.iterations, 1), 3),
           "apply": {"ms": round(t_apply, 3), "algorithmic_bytes": B_apply,
                     "gbs": round(B_apply / (t_apply * 1e-3) / 1e9, 1),
                     "frac": round(B_apply / (t_apply * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                     "stream_bytes": sw["bytes_forward"] + sw["bytes_backward"], "planned": sw["planned"],
                     "levels": [sw["levels_forward"], sw["levels_backward"]],
                     "steps": [sw["steps_forward"], sw["steps_backward"]]}}
    del Milu, Ai
    log(f"FP64 ILU(0)-RCM CG on the GPU: {iilu.iterations} iterations, residual {iilu.residual:.2e}, assembly "
        f"{t2 - t1:.2f} s + factor {t3 - t2:.2f} s + solve {t4 - t3:.2f} s; apply {t_apply:.2f} ms "
        f"({ilu['apply']['gbs']:.0f} GB/s of {B_apply} algorithmic bytes)")

    # ---- assembly throughput: the matrix assembly kernels into the existing CSR values ---------
    coeffs_a, keep_a = a_form._coefficients()
    aa = L.i32([t.test.code() for t in a_form.terms])
    ab = L.i32([t.trial.code() for t in a_form.terms])
    asm = lambda: L.check(lib.femd_assemble_matrix(V._h, a_form.quad_degree, len(a_form.terms), aa, ab, coeffs_a, A64._h))
    asm()
    L.check(lib.femd_stream_sync(None))
    ta = time.perf_counter()
    for _ in range(5):
        asm()
    L.check(lib.femd_stream_sync(None))
    t_asm = (time.perf_counter() - ta) / 5
    # the element-matrix kernel alone: exact (the reference's operation order, FP64 VALU) and on the
    # FP64 matrix cores (one GEMM over the cells); the latter's matrix against the exact one
    nt_a = len(a_form.terms)
    t_el, f_el = np.zeros(1), np.zeros(1)
    L.check(lib.femd_assembly_kernel_time(V._h, a_form.quad_degree, nt_a, aa, ab, coeffs_a, 0, 5, t_el, f_el))
    t_exact = float(t_el[0])
    L.check(lib.femd_assembly_kernel_time(V._h, a_form.quad_degree, nt_a, aa, ab, coeffs_a, 1, 10, t_el, f_el))
    t_mf, fl_mf = float(t_el[0]), float(f_el[0])
    prev_mode = fd.assembly_mode("mfma")
    asm()
    L.check(lib.femd_stream_sync(None))
    ta = time.perf_counter()
    for _ in range(5):
        asm()
    L.check(lib.femd_stream_sync(None))
    t_asm_mf = (time.perf_counter() - ta) / 5
    A_mf = a_form.assemble(reuse=False)
    fd.assembly_mode(prev_mode)
    asm()
    A_ex = A64.data
    dif = np.abs(A_mf.data - A_ex)
    rows = np.maximum.reduceat(np.abs(A_ex), ip[:-1])
    mf_err = float((np.maximum.reduceat(dif, ip[:-1]) / rows).max())
    del A_mf, keep_a
    nloc = (V.degree + 1) * (V.degree + 2) // 2
    # the least an assembly moves: the values it writes (8 nnz), each cell's local-to-global map
    # (4 nloc per cell), the vertices (16 per point) and the cells' vertex indices (12 per cell)
    B_asm = 8 * nnz + 4 * nloc * int(T.shape[0]) + 16 * int(P.shape[0]) + 12 * int(T.shape[0])
    assembly = {"ms": round(1e3 * t_asm, 3), "algorithmic_bytes": B_asm, "gbs": round(B_asm / t_asm / 1e9, 1),
                "frac": round(B_asm / t_asm / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"matrix assembly: {1e3 * t_asm:.1f} ms, {assembly['gbs']:.0f} GB/s of {B_asm} algorithmic bytes")
    nq_c = C.c_int()
    xi_b, eta_b, w_b = np.zeros(256), np.zeros(256), np.zeros(256)
    L.check(lib.femd_quadrature(a_form.quad_degree, C.byref(nq_c), xi_b, eta_b, w_b))
    nq_a = nq_c.value
    codes = [np.asarray(pr.arrays()[0]).reshape(-1, 4) for pr in a_form.programs]
    nvar = sum(1 for c in codes if not (c.shape[0] == 1 and c[0, 0] == 0))
    nc = int(T.shape[0])
    # the direct contraction's flops (every term, point, row and column: a multiply and an add),
    # and the least the kernel moves: the element matrices it writes, the coefficient values at the
    # points of each non-constant term and the cell geometry (6 doubles)
    useful = 2.0 * nloc * nloc * nq_a * nt_a * nc
    B_el = 8 * nloc * nloc * nc + 8 * nq_a * nc * nvar + 48 * nc
    element = {"exact_us": round(1e3 * t_exact, 2), "mfma_us": round(1e3 * t_mf, 2),
               "speedup": round(t_exact / t_mf, 2),
               "mfma_issued_tflops": round(fl_mf / (t_mf * 1e-3) / 1e12, 2),
               "mfma_frac_of_fp64_peak": round(fl_mf / (t_mf * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
               "direct_flops": useful, "mfma_issued_flops": fl_mf,
               "exact_tflops": round(useful / (t_exact * 1e-3) / 1e12, 2),
               "algorithmic_bytes": B_el, "mfma_gbs": round(B_el / (t_mf * 1e-3) / 1e9, 1),
               "mfma_hbm_frac": round(B_el / (t_mf * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
               "assembly_ms_mfma": round(1e3 * t_asm_mf, 3), "row_rel_err_vs_exact": mf_err}
    log(f"element kernel: exact {1e3 * t_exact:.1f} us ({element['exact_tflops']:.1f} TFLOP/s direct), matrix cores "
        f"{1e3 * t_mf:.1f} us ({element['mfma_issued_tflops']:.1f} TFLOP/s issued, {element['mfma_gbs']:.0f} GB/s); "
        f"assembly {1e3 * t_asm_mf:.2f} ms in mfma mode, values within {mf_err:.1e} of the exact ones (row-relative)")

    for _ in range(args.warmup):
        step()
    times, errs, inner, outer = [], [], [], []
    x = np.empty(n)
    for k in range(args.steps):
        ts = time.perf_counter()
        A, info = step()
        te = time.perf_counter()
        times.append(te - ts)
        L.check(lib.femd_memcpy_dtoh(x.ctypes.data, x_dev, 8 * n))
        errs.append(float(np.linalg.norm(x - x64) / np.linalg.norm(x64)))
        inner.append(info.iterations)
        outer.append(info.breakdown)
        log(f"step {k}: {1e3 * times[-1]:.1f} ms, residual {info.residual:.2e}, rel err vs FP64 {errs[-1]:.2e}, "
            f"inner {info.iterations} / outer {info.breakdown}")
    ms = 1e3 * float(np.mean(times))
    # one more step, outside the timed ones, with the product kernel sampled: HIP events around
    # every 8th launch on the engine's stream (sampling needs plain launches, not graph replays)
    L.check(lib.femd_inner_timer(1))
    step()
    launches, tot = C.c_int64(), C.c_double()
    L.check(lib.femd_inner_timer_read(C.byref(launches), C.byref(tot)))
    inner_shape = C.create_string_buffer(64)
    L.check(lib.femd_inner_timer_shape(inner_shape, 64))
    inner_shape = inner_shape.value.decode()
    L.check(lib.femd_inner_timer(0 if not args.no_graph else 1))
    value = 1e3 / ms

    # ---- the same solve to the stricter tolerance (two steps, outside the timed ones) ---------
    strict = None
    if args.strict_rtol > 0:
        st_t, st_e, st_i = [], [], []
        for _ in range(2):
            ts = time.perf_counter()
            _, info = step(args.strict_rtol)
            st_t.append(time.perf_counter() - ts)
            L.check(lib.femd_memcpy_dtoh(x.ctypes.data, x_dev, 8 * n))
            st_e.append(float(np.linalg.norm(x - x64) / np.linalg.norm(x64)))
            st_i.append(info.iterations)
        strict = {"rtol": args.strict_rtol, "solves_per_s": round(1.0 / float(np.mean(st_t)), 4),
                  "rel_err_vs_fp64_max": max(st_e), "inner_iterations": st_i}
        log(f"at rtol {args.strict_rtol:g}: {1e3 * float(np.mean(st_t)):.1f} ms per solve, rel err {max(st_e):.2e}")

    # ---- roofline of the dominant kernel: the product of the FP32 inner CG -------------------
    # per launch it reads values + columns (8 nnz), row pointers (4 (n + 1)), the residual and the
    # old direction once (8 n) and writes the new direction p' = r + beta p and q = A p' (8 n):
    # 8 nnz + 20 n + 4 bytes (the partial of p'.q stays in registers)
    B_fused = 8 * nnz + 20 * n + 4
    t_fused = tot.value / max(launches.value, 1)
    gbs_fused = B_fused / (t_fused * 1e-3) / 1e9
    log(f"FP32 CG product: {launches.value} timed launches, {1e3 * t_fused:.1f} us average, {gbs_fused:.0f} GB/s")

    # ---- plain SpMV on this matrix: back to back, and with the Infinity Cache flushed --------
    from femd_amd.roofline import spmv_times
    t64, t32 = spmv_times(A, args.spmv_reps)
    c64, c32 = spmv_times(A, max(5, args.spmv_reps // 5), flush=True)
    variants = {}
    for prec, key in ((L.FP64, "fp64"), (L.FP32, "fp32")):
        nm, tm = C.create_string_buffer(64), C.c_double()
        L.check(lib.femd_spmv_tuning(A._h, prec, nm, 64, C.byref(tm)))
        variants[key] = nm.value.decode()
    B64, B32 = 12 * nnz + 20 * n + 4, 8 * nnz + 12 * n + 4
    gbs = lambda B, t: B / (t * 1e-3) / 1e9
    log(f"SpMV FP64 {t64 * 1e3:.1f} us {gbs(B64, t64):.0f} GB/s (flushed {c64 * 1e3:.1f} us {gbs(B64, c64):.0f} GB/s); "
        f"FP32 {t32 * 1e3:.1f} us {gbs(B32, t32):.0f} GB/s (flushed {c32 * 1e3:.1f} us {gbs(B32, c32):.0f} GB/s)")
    # the same kernels on the matrix in the inner CG's order of the unknowns (the SpMV its
    # iterations run); the products above keep the reference's order, whose row sums are its own
    from femd_amd.roofline import reordered
    Ar = reordered(A)
    o64, o32 = spmv_times(Ar, args.spmv_reps)
    q64, q32 = spmv_times(Ar, max(5, args.spmv_reps // 5), flush=True)
    del Ar
    log(f"SpMV in the inner CG's order: FP64 {o64 * 1e3:.1f} us {gbs(B64, o64):.0f} GB/s (flushed {gbs(B64, q64):.0f}); "
        f"FP32 {o32 * 1e3:.1f} us {gbs(B32, o32):.0f} GB/s (flushed {gbs(B32, q32):.0f})")
    traffic, traffic_source = (None, None) if args.no_live_traffic else live_traffic(inner_shape, log)
    if traffic is None:
        traffic = measured_traffic("k_cg_product")
        traffic_source = "profiles/hbm_traffic_p3_kappa_1.5M.md (committed)" if traffic is not None else None

    # ---- CPU baseline: the reference's own code, bounded sample ---------------------------
    cpu = None
    if not args.no_cpu:
        cpu, xref = cpu_baseline(P, T, a_form, L_form, {"ilu0": iilu.iterations, "jacobi": ip64.iterations},
                                 args.rtol, args.cpu_full_limit, log)
        # the GPU's FP64 solves against the reference's on the same inputs
        for key, xg, gi, leg in (("ilu0", xilu, iilu.iterations, "ilu0_rcm_cg"), ("jacobi", xj64, ip64.iterations, "jacobi_cg")):
            c = cpu["legs"][leg]
            if key in xref:
                c["gpu_iterations"] = gi
                c["iterations_equal"] = bool(gi == c["iterations"])
                c["gpu_rel_diff"] = float(np.linalg.norm(xg - xref[key]) / np.linalg.norm(xref[key]))
                log(f"{leg}: GPU {gi} / reference {c['iterations']} iterations, solutions differ by "
                    f"{c['gpu_rel_diff']:.2e} relative")
    del A64

    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64 assembly" if args.exact_assembly else "f64 MFMA assembly") + " + f32 CG + f64 refinement",
        "data": "synthetic",
        "config": {"workload": "P3 -div(kappa grad u) = f, seeded Delaunay unit square, FP64 "
                               + ("" if args.exact_assembly else "matrix-core ") + "assembly + FP32 Jacobi-CG with "
                               "FP64 iterative refinement",
                   "dofs": n, "nnz": nnz, "cells": int(T.shape[0]), "mesh_h": args.h, "min_angle_deg": round(min_ang, 2),
                   "rtol": args.rtol, "inner_tol": args.inner_tol, "parallelism": "single GPU"},
        "rel_err_vs_fp64_max": max(errs), "latency_ms_p50": round(1e3 * float(np.percentile(times, 50)), 3),
        "latency_ms_p99": round(1e3 * float(np.percentile(times, 99)), 3),
        "inner_iterations": inner, "outer_iterations": outer,
        "fp64_reference_cg": {"iterations": i64.iterations, "residual": i64.residual},
        "strict_rtol": strict,
        "plain_fp64_gpu": {"iterations": ip64.iterations, "seconds": round(t_fp64, 3),
                           "solves_per_s": round(1.0 / t_fp64, 4), "speedup_of_value": round(value * t_fp64, 2)},
        "fp64_ilu0_rcm_gpu": ilu,
        "assembly": assembly,
        "element_kernel": element,
        "spmv": {"fp64": {"us": round(t64 * 1e3, 2), "gbs": round(gbs(B64, t64), 1), "shape": variants["fp64"],
                          "flushed_us": round(c64 * 1e3, 2), "flushed_gbs": round(gbs(B64, c64), 1),
                          "flushed_frac": round(gbs(B64, c64) / HBM_PEAK_GBS, 4)},
                 "fp32": {"us": round(t32 * 1e3, 2), "gbs": round(gbs(B32, t32), 1), "shape": variants["fp32"],
                          "flushed_us": round(c32 * 1e3, 2), "flushed_gbs": round(gbs(B32, c32), 1),
                          "flushed_frac": round(gbs(B32, c32) / HBM_PEAK_GBS, 4)},
                 "inner_order": {"fp64_us": round(o64 * 1e3, 2), "fp64_gbs": round(gbs(B64, o64), 1),
                                 "fp64_flushed_frac": round(gbs(B64, q64) / HBM_PEAK_GBS, 4),
                                 "fp32_us": round(o32 * 1e3, 2), "fp32_gbs": round(gbs(B32, o32), 1),
                                 "fp32_flushed_frac": round(gbs(B32, q32) / HBM_PEAK_GBS, 4)},
                 "bytes_fp64": B64, "bytes_fp32": B32},
        "roofline": {"kernel": f"k_cg_product (product of the FP32 inner CG with the direction update, {inner_shape})",
                     "bound": "hbm",
                     "achieved": round(gbs_fused, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs_fused / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_source,
                     "algorithmic_bytes": B_fused, "avg_us": round(1e3 * t_fused, 2), "timed_launches": launches.value},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def live_traffic(shape, log, timeout=240):
    """HBM bytes per launch of the inner CG product, measured in this run: two rocprofv3 --pmc passes
    (FETCH_SIZE, then WRITE_SIZE, in KiB per dispatch) over a short mixed-precision solve of the same
    workload with the product's launch shape pinned, each scaled by what the counter reads for a
    64 MiB device copy (FETCH_SIZE counts about half of the bytes a copy is known to read).  Child
    processes after the timed region; (None, None) when the profiler is missing or fails."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    if not shutil.which("rocprofv3"):
        return None, None
    per = {}
    env = dict(os.environ, TMPDIR="/tmp")
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", d, "-o", counter, "--", sys.executable,
                   "-m", "femd_amd.roofline", "--product-only", "--shape", shape, "--cg-tol", "1e-3"]
            try:
                subprocess.run(cmd, cwd=ROOT, env=env, timeout=timeout, capture_output=True)
            except (subprocess.TimeoutExpired, OSError) as e:
                log(f"live traffic: {counter} pass failed ({e})")
                return None, None
            files = glob.glob(os.path.join(d, "**", f"{counter}_counter_collection.csv"), recursive=True)
            if not files:
                log(f"live traffic: no {counter} counters written")
                return None, None
            prod, copy = [], []
            with open(files[0]) as f:
                for r in csv.DictReader(f):
                    if r.get("Counter_Name") != counter:
                        continue
                    name, v = r.get("Kernel_Name", ""), float(r["Counter_Value"])
                    if "k_cg_product" in name:
                        prod.append(v)
                    elif "k_copy" in name and int(float(r.get("Grid_Size", 0))) == (64 << 20) // 8:
                        copy.append(v)
            if not prod or not copy:
                log(f"live traffic: {counter}: {len(prod)} product and {len(copy)} copy dispatches")
                return None, None
            scale = (64 << 20) / (sum(copy) / len(copy) * 1024.0)
            per[counter] = sum(prod) / len(prod) * 1024.0 * scale
            log(f"live traffic: {counter} {per[counter] / 1e6:.1f} MB per product launch ({len(prod)} dispatches, "
                f"copy calibration x{scale:.4f})")
    return int(per["FETCH_SIZE"] + per["WRITE_SIZE"]), "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, measured in this run"


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed hardware-counter summary (FETCH_SIZE and
    WRITE_SIZE per dispatch, scaled by a copy kernel of known size), or None."""
    path = ROOT / "profiles" / "hbm_traffic_p3_kappa_1.5M.md"
    if not path.exists():
        return None
    text = path.read_text()
    if "```json" not in text:
        return None
    blob = text.split("```json", 1)[1].split("```", 1)[0]
    try:
        return json.loads(blob).get(kernel)
    except ValueError:
        return None


def cpu_baseline(P, T, a_form, L_form, expected, rtol, limit, log):
    """The reference (oracle/_ref, the reference's own C++ compiled in place) on the host cores: the
    full assembly (cache, coefficient programs, element matrices, scatter), then its ILU(0)-RCM CG
    and its Jacobi-CG, each solved in full to rtol and timed when a sample of its iterations
    predicts the full solve within `limit` seconds (else reported from the sample, and marked so);
    then a 1-thread figure from a sample: the assembly and 20 ILU(0)-CG iterations at one thread,
    scaled to the full solve's iteration count.  Returns (the baseline dict, the solutions)."""
    import oracle
    R = oracle.reference()
    kind = "reference"
    if R is None:
        R = oracle.oracle()
        kind = "port"
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    from femd_amd.forms import Program
    Pr = R.from_arrays(P, T, k=3)

    def assemble():
        t0 = time.perf_counter()
        x, y, _ = Pr.cache_points(a_form.quad_degree)
        ck = R.integrand_eval(*Program(a_form.terms[0].coeff).arrays(), x, y)
        Av = Pr.assemble_matrix(a_form.quad_degree, [1, 2], [1, 2], np.stack([ck, ck]))
        xb, yb, _ = Pr.cache_points(L_form.quad_degree)
        cf = R.integrand_eval(*Program(L_form.terms[0].coeff).arrays(), xb, yb)
        b = Pr.assemble_vector(L_form.quad_degree, [0], cf[None, :])
        return Av, b, time.perf_counter() - t0

    R.set_threads(cores)
    Av, b, t_asm = assemble()
    ip, ix = Pr.pattern()
    log(f"CPU reference ({cores} threads): assembly {t_asm:.2f} s")
    legs, xs = {}, {}
    for key, leg, sample in (("ilu0", "ilu0_rcm_cg", 20), ("jacobi", "jacobi_cg", 200)):
        t1 = time.perf_counter()
        R.cg(ip, ix, Av, b, M=key, maxiter=sample)
        t_sample = time.perf_counter() - t1
        predicted = t_sample * expected[key] / sample
        if predicted <= limit:
            t1 = time.perf_counter()
            x, io = R.cg(ip, ix, Av, b, M=key, tol=rtol, maxiter=100000)
            t_solve = time.perf_counter() - t1
            xs[key] = x
            legs[leg] = {"measured": True, "iterations": int(io["iterations"]), "converged": bool(io["converged"]),
                         "residual": float(io["residual"]), "solve_s": round(t_solve, 3),
                         "seconds": round(t_asm + t_solve, 3)}
        else:
            legs[leg] = {"measured": False, "iterations": int(expected[key]), "sample_iterations": sample,
                         "sample_s": round(t_sample, 3), "solve_s": round(predicted, 3),
                         "seconds": round(t_asm + predicted, 3)}
        log(f"CPU reference {leg}: {legs[leg]}")
    # 1 thread: a sample
    R.set_threads(1)
    Av1, b1, t_asm1 = assemble()
    ts = []
    for it in (10, 30):
        t1 = time.perf_counter()
        R.cg(ip, ix, Av1, b1, M="ilu0", maxiter=it)
        ts.append(time.perf_counter() - t1)
    t_it1 = (ts[1] - ts[0]) / 20
    t_set1 = max(ts[0] - 10 * t_it1, 0.0)
    del Av1, b1
    R.set_threads(cores)
    it_ilu = legs["ilu0_rcm_cg"]["iterations"]
    one = {"cores": 1, "assembly_s": round(t_asm1, 3), "ilu0_factor_s": round(t_set1, 3),
           "ilu0_rcm_cg_ms_per_iteration": round(1e3 * t_it1, 3),
           "seconds": round(t_asm1 + t_set1 + it_ilu * t_it1, 3),
           "sample": f"the assembly and 10 + 30 ILU(0)-RCM CG iterations at one thread, scaled to {it_ilu} iterations"}
    log(f"CPU reference at 1 thread: assembly {t_asm1:.2f} s, ILU(0)-CG {1e3 * t_it1:.1f} ms per iteration "
        f"-> {one['seconds']:.1f} s per solve")
    best = min(legs, key=lambda k: legs[k]["seconds"])
    full = all(legs[k]["measured"] for k in legs)
    return ({"value": round(1.0 / legs[best]["seconds"], 6), "unit": "solves/s", "cores": cores, "kind": kind,
             "sample": (f"full solves, measured: the assembly ({t_asm:.2f} s) and the reference's ILU(0)-RCM CG and "
                        f"Jacobi-CG to rtol {rtol:g}; value is the faster, {best}" if full else
                        f"the assembly ({t_asm:.2f} s) and the reference's ILU(0)-RCM CG and Jacobi-CG to rtol "
                        f"{rtol:g}, those predicted beyond {limit:g} s from a sample of their iterations; value is "
                        f"the faster, {best}"),
             "seconds_per_solve": legs[best]["seconds"], "assembly_s": round(t_asm, 3), "legs": legs,
             "one_thread": one}, xs)


if __name__ == "__main__":
    main()
