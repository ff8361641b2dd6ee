"""fd.cg, fd.gmres, fd.lgmres and KrylovInfo (python/femd/krylov.py:31-208) on the GPU.

    x, info = cg(A, b, M=A.preconditioner("ilu0"))

A is a SparseMatrix; M is None or a Preconditioner.  cg's tol bounds the true relative residual
||b - A x|| / ||b||; gmres / lgmres are left preconditioned and tol bounds
||M^{-1} r|| / ||M^{-1} b||.  Non-convergence and breakdown are RuntimeWarnings, as there."""
from __future__ import annotations

import ctypes as C
import warnings
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .sparse import Preconditioner, SparseMatrix


@dataclass
class KrylovInfo:
    converged: bool
    iterations: int
    residual: float            # ||b - A x|| / ||b||
    precond_residual: float    # ||M^{-1}(b - A x)|| / ||M^{-1} b||, what tol bounds (gmres)
    method: str

    def __repr__(self):
        state = "converged" if self.converged else "NOT converged"
        return (f"KrylovInfo({self.method} {state} in {self.iterations} iterations, residual {self.residual:.2e}, "
                f"preconditioned {self.precond_residual:.2e})")


def _vector(b):
    v = np.asarray(getattr(b, "vector", b))
    if np.iscomplexobj(v):
        raise TypeError("the Krylov solvers are real; solve the real and imaginary parts separately")
    return np.ascontiguousarray(v, dtype=np.float64).reshape(-1)


def _operator(A, n):
    if not isinstance(A, SparseMatrix):
        raise TypeError(f"krylov: A must be a SparseMatrix on this path, got {type(A).__name__}")
    if A.shape != (n, n):
        raise ValueError(f"krylov: A is {A.shape[0]} x {A.shape[1]}, b has length {n}")
    return A


def _preconditioner(M, n):
    if M is None:
        return None
    if isinstance(M, str):
        raise TypeError("krylov: M is a Preconditioner (A.preconditioner(kind)) or None")
    if not isinstance(M, Preconditioner):
        raise TypeError(f"krylov: cannot use {type(M).__name__} as a preconditioner on this path")
    if M.size != n:
        raise ValueError(f"krylov: the preconditioner has size {M.size}, b has length {n}")
    return M


def cg(A, b, M=None, x0=None, *, maxiter=None, tol=1e-10, warn=True):
    """Preconditioned CG (krylov.py:178-202).  Returns (x, KrylovInfo)."""
    bv = _vector(b)
    n = bv.size
    A = _operator(A, n)
    M = _preconditioner(M, n)
    x = np.zeros(n) if x0 is None else _vector(x0).copy()
    if x.shape != (n,):
        raise ValueError(f"cg: x0 has length {x.size}, b has length {n}")
    maxiter = int(maxiter if maxiter is not None else max(1000, 10 * n))
    info = L.SolveInfo()
    L.check(L.load().femd_cg(A._h, M._h if M is not None else None, bv if n else np.zeros(1), x if n else np.zeros(1),
                             maxiter, float(tol), C.byref(info)))
    ki = KrylovInfo(bool(info.converged), int(info.iterations), float(info.residual), float(info.residual), "cg")
    if warn and info.breakdown:
        warnings.warn("cg broke down (p^T A p <= 0 or r^T M^{-1} r <= 0): A or the preconditioner is not "
                      "symmetric positive definite; use gmres", RuntimeWarning, stacklevel=2)
    elif warn and not ki.converged:
        warnings.warn(f"cg did not converge in {ki.iterations} iterations (residual {ki.residual:.2e}, "
                      f"tol {tol:.1e}); raise maxiter or improve the preconditioner", RuntimeWarning, stacklevel=2)
    ki.breakdown = bool(info.breakdown)
    return x, ki


def _krylov(A, b, M, x0, restart, maxiter, tol, lgmres, k_aug, warn):
    bv = _vector(b)
    n = bv.size
    A = _operator(A, n)
    M = _preconditioner(M, n)
    x = np.zeros(n) if x0 is None else _vector(x0).copy()
    if x.shape != (n,):
        raise ValueError(f"gmres: x0 has length {x.size}, b has length {n}")
    restart = int(min(restart, max(n, 1)))
    maxiter = int(maxiter if maxiter is not None else max(1000, 10 * restart))
    info = L.SolveInfo()
    L.check(L.load().femd_krylov(A._h, M._h if M is not None else None, bv, x, int(bool(lgmres)), restart, maxiter,
                                 float(tol), int(k_aug), C.byref(info)))
    ki = KrylovInfo(bool(info.converged), int(info.iterations), float(info.residual), float(info.precond_residual),
                    "lgmres" if lgmres else "gmres")
    if warn and not ki.converged:
        warnings.warn(f"{ki.method} did not converge in {ki.iterations} iterations "
                      f"(preconditioned residual {ki.precond_residual:.2e}, tol {tol:.1e}); "
                      "raise maxiter or restart, or improve the preconditioner", RuntimeWarning, stacklevel=3)
    return x, ki


def gmres(A, b, M=None, x0=None, *, restart=30, maxiter=None, tol=1e-10, warn=True):
    """GMRES(restart), left preconditioned by M (krylov.py:148-163).  Returns (x, KrylovInfo)."""
    return _krylov(A, b, M, x0, restart, maxiter, tol, False, 0, warn)


def lgmres(A, b, M=None, x0=None, *, restart=20, k_aug=2, maxiter=None, tol=1e-10, warn=True):
    """LGMRES(restart, k_aug) (krylov.py:205-208).  Returns (x, KrylovInfo)."""
    return _krylov(A, b, M, x0, restart, maxiter, tol, True, k_aug, warn)


def solve_refined(A, b, x0=None, *, tol=1e-8, inner_tol=1e-4, max_outer=50):
    """FP64 iterative refinement around an FP32 Jacobi-PCG (no counterpart in the reference):
    returns (x, KrylovInfo) with info.iterations the inner FP32 iterations."""
    bv = _vector(b)
    n = bv.size
    A = _operator(A, n)
    x = np.zeros(n) if x0 is None else _vector(x0).copy()
    info = L.SolveInfo()
    L.check(L.load().femd_solve_refined(A._h, bv, x, float(tol), float(inner_tol), int(max_outer), C.byref(info)))
    ki = KrylovInfo(bool(info.converged), int(info.iterations), float(info.residual), float(info.residual), "refined-cg32")
    ki.outer = int(info.breakdown)
    return x, ki
