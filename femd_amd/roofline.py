"""HBM roofline measurement of the SpMV kernels on one matrix.

    python -m femd_amd.roofline [--h 0.00265] [--reps 50]

Runs the CSR-stream SpMV (FP64 and FP32) `reps` times back to back on the P3 benchmark matrix,
then again with a 512 MiB copy before each product (so none is served from the Infinity Cache),
and a plain device copy of 64 MiB, timing with HIP events on the library stream; prints one JSON
line: kernel times, the tuned shapes, algorithmic bytes per launch (SpMV 12 nnz + 20 n + 4 in FP64,
8 nnz + 12 n + 4 in FP32; copy 2 x bytes) and GB/s.  The same program is what the hardware-counter runs profile, so
FETCH_SIZE / WRITE_SIZE per dispatch of the copy calibrate the counters against known bytes."""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys

import numpy as np

from . import _lib as L

HBM_PEAK_GBS = 8000.0


def _dev(lib, nbytes):
    p = C.c_void_p()
    L.check(lib.femd_dev_alloc(nbytes, C.byref(p)))
    return p.value


def _timed(lib, fn, reps, warm=3):
    for _ in range(warm):
        fn()
    L.check(lib.femd_event_timer_start(None))
    for _ in range(reps):
        fn()
    ms = C.c_double()
    L.check(lib.femd_event_timer_stop(None, C.byref(ms)))
    return ms.value / reps


def spmv_times(A, reps=50, flush=False):
    """Average kernel time (ms) of the FP64 and FP32 SpMV of A, from HIP events.  With flush, a
    512 MiB device copy runs before every product (outside the events), so the matrix cannot be
    served from the 256 MiB Infinity Cache left warm by the previous product."""
    lib = L.load()
    n, m = A.shape
    out = {}
    fl = _dev(lib, 1 << 30) if flush else None
    try:
        for prec in (L.FP64, L.FP32):
            xb, yb = _dev(lib, 8 * m), _dev(lib, 8 * n)
            xs = np.random.default_rng(0).standard_normal(m)
            if prec == L.FP32:
                xs = xs.astype(np.float32)
            L.check(lib.femd_memcpy_htod(xb, xs.ctypes.data, xs.nbytes))
            run = lambda: L.check(lib.femd_spmv_dev(A._h, xb, yb, prec, None))
            if not flush:
                out[prec] = _timed(lib, run, reps)
            else:
                run()
                tot = 0.0
                for _ in range(reps):
                    L.check(lib.femd_vec_copy_dev((1 << 29) // 8, fl, fl + (1 << 29), None))
                    tot += _timed(lib, run, 1, warm=0)
                out[prec] = tot / reps
            lib.femd_dev_free(xb)
            lib.femd_dev_free(yb)
    finally:
        if fl:
            lib.femd_dev_free(fl)
    return out[L.FP64], out[L.FP32]


def precond_apply_time(M, n, reps=10):
    """Average time (ms) of one application z = M^-1 r of a preconditioner on the device, from HIP
    events on the library stream, r a seeded random vector of n values."""
    lib = L.load()
    rb, zb = _dev(lib, 8 * n), _dev(lib, 8 * n)
    try:
        r = np.random.default_rng(0).standard_normal(n)
        L.check(lib.femd_memcpy_htod(rb, r.ctypes.data, r.nbytes))
        return _timed(lib, lambda: L.check(lib.femd_precond_apply_dev(M._h, rb, zb, None)), reps, warm=2)
    finally:
        lib.femd_dev_free(rb)
        lib.femd_dev_free(zb)


def reordered(A):
    """A copy of the square matrix A in the inner CG's order of the unknowns (each unknown numbered
    where a row-by-row scan first meets it as a column, those it never meets last), for timing the
    SpMV the inner iterations run.  Host-side (NumPy / SciPy) plumbing for measurement."""
    import scipy.sparse as sp
    from .sparse import SparseMatrix
    n = A.shape[0]
    ip, ix, v = A.indptr, A.indices, A.data
    cols, first = np.unique(ix, return_index=True)
    seen = np.zeros(n, dtype=bool)
    seen[cols] = True
    perm = np.concatenate([cols[np.argsort(first, kind="stable")], np.flatnonzero(~seen)])
    S = sp.csr_matrix((v, ix, ip), shape=(n, n))[perm][:, perm].tocsr()
    S.sort_indices()
    return SparseMatrix.from_arrays(n, n, S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data)


def copy_time(nbytes=64 << 20, reps=50):
    lib = L.load()
    a, b = _dev(lib, nbytes), _dev(lib, nbytes)
    n = nbytes // 8
    t = _timed(lib, lambda: L.check(lib.femd_vec_copy_dev(n, a, b, None)), reps)
    lib.femd_dev_free(a)
    lib.femd_dev_free(b)
    return t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--h", type=float, default=None)
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--shape", default=None,
                    help="pin the SpMV (and inner CG product) launch shape in both precisions, e.g. tile256x2048, "
                         "so that separate profiler runs measure the same kernels")
    ap.add_argument("--product-only", action="store_true",
                    help="skip the SpMV timings: a 64 MiB device copy (the counters' calibration) and the --cg-tol "
                         "solve only, for hardware-counter passes")
    ap.add_argument("--cg-tol", type=float, default=None,
                    help="also run one mixed-precision solve to this relative residual (the inner CG product "
                         "kernel, launched without graphs so that every launch is visible to a profiler)")
    args = ap.parse_args()
    from . import problems
    lib = L.load()
    L.check(lib.femd_init(0))
    V, a, Lf, info = problems.p3_kappa_problem(args.h or problems.C1_H)
    A = a.assemble()
    if args.shape:
        for prec in (L.FP64, L.FP32):
            L.check(lib.femd_spmv_select(A._h, prec, args.shape.encode()))
    ip, ix = V.pattern()
    n, nnz = V.dim, int(ix.size)
    if args.product_only:
        t64 = t32 = c64 = c32 = float("nan")
        tc = copy_time(reps=2)
    else:
        t64, t32 = spmv_times(A, args.reps)
        c64, c32 = spmv_times(A, max(5, args.reps // 5), flush=True)
        tc = copy_time(reps=args.reps)
    names = {}
    for prec in (L.FP64, L.FP32):
        buf = C.create_string_buffer(64)
        L.check(lib.femd_spmv_tuning(A._h, prec, buf, 64, None))
        names[prec] = buf.value.decode()
    B64, B32, Bc = 12 * nnz + 20 * n + 4, 8 * nnz + 12 * n + 4, 2 * (64 << 20)

    def row(t, B, name):
        return {"us": t * 1e3, "bytes": B, "gbs": B / t / 1e6, "frac": B / t / 1e6 / HBM_PEAK_GBS, "shape": name}

    res = {"n": n, "nnz": nnz,
           "spmv_fp64": row(t64, B64, names[L.FP64]), "spmv_fp32": row(t32, B32, names[L.FP32]),
           "spmv_fp64_flushed": row(c64, B64, names[L.FP64]), "spmv_fp32_flushed": row(c32, B32, names[L.FP32]),
           "copy_64MiB": {"us": tc * 1e3, "bytes": Bc, "gbs": Bc / tc / 1e6}}
    if args.cg_tol is not None:
        bvec = Lf.assemble()
        bd, xd = _dev(lib, 8 * n), _dev(lib, 8 * n)
        L.check(lib.femd_memcpy_htod(bd, bvec.ctypes.data, 8 * n))
        L.check(lib.femd_memcpy_htod(xd, np.zeros(n).ctypes.data, 8 * n))
        L.check(lib.femd_inner_timer(1))
        info = L.SolveInfo()
        L.check(lib.femd_solve_refined_dev(A._h, bd, xd, args.cg_tol, 0.3, 60, C.byref(info), None))
        L.check(lib.femd_stream_sync(None))
        k, tot = C.c_int64(), C.c_double()
        L.check(lib.femd_inner_timer_read(C.byref(k), C.byref(tot)))
        L.check(lib.femd_inner_timer(0))
        Bp = 8 * nnz + 20 * n + 4
        tp = tot.value / max(k.value, 1)
        res["cg_product"] = {"iterations": int(info.iterations), "timed_launches": k.value, "us": tp * 1e3, "bytes": Bp,
                             "gbs": Bp / tp / 1e6 if tp > 0 else None}
        lib.femd_dev_free(bd)
        lib.femd_dev_free(xd)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
