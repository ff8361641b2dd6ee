"""HBM roofline measurement of the SpMV kernels on one matrix.

    python -m femd_amd.roofline [--h 0.00265] [--reps 50]

Runs the CSR-stream SpMV (FP64 and FP32) `reps` times on the P3 benchmark matrix and a plain
device copy of 64 MiB, timing each loop with HIP events on the library stream, and prints one JSON
line: kernel times, algorithmic bytes per launch (SpMV 12 nnz + 20 n + 4 in FP64, 8 nnz + 12 n + 4
in FP32; copy 2 x bytes) and GB/s.  The same program is what the hardware-counter runs profile, so
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


def _timed(lib, fn, reps):
    for _ in range(3):
        fn()
    L.check(lib.femd_event_timer_start(None))
    for _ in range(reps):
        fn()
    ms = C.c_double()
    L.check(lib.femd_event_timer_stop(None, C.byref(ms)))
    return ms.value / reps


def spmv_times(A, reps=50):
    """Average kernel time (ms) of the FP64 and FP32 SpMV of A, from HIP events."""
    lib = L.load()
    n, m = A.shape
    out = {}
    for prec in (L.FP64, L.FP32):
        xb, yb = _dev(lib, 8 * m), _dev(lib, 8 * n)
        xs = np.random.default_rng(0).standard_normal(m)
        if prec == L.FP32:
            xs = xs.astype(np.float32)
        L.check(lib.femd_memcpy_htod(xb, xs.ctypes.data, xs.nbytes))
        out[prec] = _timed(lib, lambda: L.check(lib.femd_spmv_dev(A._h, xb, yb, prec, None)), reps)
        lib.femd_dev_free(xb)
        lib.femd_dev_free(yb)
    return out[L.FP64], out[L.FP32]


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
    args = ap.parse_args()
    from . import problems
    lib = L.load()
    L.check(lib.femd_init(0))
    V, a, Lf, info = problems.p3_kappa_problem(args.h or problems.C1_H)
    A = a.assemble()
    ip, ix = V.pattern()
    n, nnz = V.dim, int(ix.size)
    t64, t32 = spmv_times(A, args.reps)
    tc = copy_time(reps=args.reps)
    B64, B32, Bc = 12 * nnz + 20 * n + 4, 8 * nnz + 12 * n + 4, 2 * (64 << 20)
    res = {"n": n, "nnz": nnz,
           "spmv_fp64": {"us": t64 * 1e3, "bytes": B64, "gbs": B64 / t64 / 1e6, "frac": B64 / t64 / 1e6 / HBM_PEAK_GBS},
           "spmv_fp32": {"us": t32 * 1e3, "bytes": B32, "gbs": B32 / t32 / 1e6, "frac": B32 / t32 / 1e6 / HBM_PEAK_GBS},
           "copy_64MiB": {"us": tc * 1e3, "bytes": Bc, "gbs": Bc / tc / 1e6}}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
