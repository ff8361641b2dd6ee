"""Time stepping with a fixed matrix pair on the GPU (the time-stepping pattern).

The reference steps in Python (example13.ipynb: ``b = M @ u`` then a solve with ``A = M + dt K``,
every step).  Here the whole step -- the FP64 product ``M u`` and FP64 residual rounds of one scaled
Jacobi-CG run (FP32 or FP64 iterations, reliable updates between rounds) warm-started at the
previous ``u`` -- is one launch sequence, captured once into a hipGraph and replayed for every step:

    stepper = BackwardEuler(M, A)                  # SparseMatrix pair on one pattern size
    u, info = stepper.run(u0, nsteps=500, tol=1e-12)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .sparse import SparseMatrix


class BackwardEuler:
    """u_{k+1} solves A u_{k+1} = M u_k (backward Euler for M u' + K u = 0 with A = M + dt K)."""

    def __init__(self, M: SparseMatrix, A: SparseMatrix):
        if not isinstance(M, SparseMatrix) or not isinstance(A, SparseMatrix):
            raise TypeError("BackwardEuler: M and A must be SparseMatrix")
        if M.shape != A.shape or M.shape[0] != M.shape[1]:
            raise ValueError(f"BackwardEuler: M {M.shape} and A {A.shape} must be square and of one size")
        self.M, self.A = M, A

    def run(self, u0, nsteps, *, tol=1e-10, rounds=8, inner_tol=1e-2, max_inner=64, precision="fp32", graph=True):
        """Advance u0 by nsteps steps.  Each step runs up to `rounds` FP64 residual rounds, each a
        segment of at most max_inner CG iterations in `precision` ("fp32" or "fp64") that ends when
        its residual has dropped by inner_tol or the step's residual meets tol; later rounds continue
        the same CG (reliable updates).  Returns (u, info) with info a dict: per-step inner
        iterations, their total, the largest relative residual any step ended with, the steps
        whose residual the reference's CG verdict would not accept (krylov/cg.hpp:222-237: tol, or
        10 tol within the iteration budget) and whether there were none."""
        n = self.A.shape[0]
        u = np.array(u0, dtype=np.float64).reshape(-1)
        if u.size != n:
            raise ValueError(f"BackwardEuler: u0 has {u.size} values, the system {n}")
        prec = {"fp32": L.FP32, "fp64": L.FP64}.get(precision)
        if prec is None:
            raise ValueError(f"BackwardEuler: precision is 'fp32' or 'fp64', not {precision!r}")
        its = np.zeros(max(int(nsteps), 1), dtype=np.int32)
        info = L.SolveInfo()
        failed = C.c_int(0)
        L.check(L.load().femd_heat_run(self.M._h, self.A._h, u, int(nsteps), float(tol), int(rounds), float(inner_tol),
                                       int(max_inner), prec, int(bool(graph)), its.ctypes.data, C.byref(failed),
                                       C.byref(info)))
        return u, dict(converged=bool(info.converged), iterations=its[:int(nsteps)], total_iterations=int(info.iterations),
                       steps=int(info.breakdown), residual=float(info.residual), failed_steps=int(failed.value))
