"""fd.solve(a, L, backend=...) on a 2D mesh (python/femd/linalg.py:710-755), Krylov backends."""
from __future__ import annotations

import numpy as np

from .forms import Form, form


def solve(a, L, backend="cg", dirichlet=None, **options):
    """Solve a(u, v) = L(v) for all v in V: A = a.assemble(), b = L.assemble(), then
    A.solver(backend).solve(b) (linalg.py:740-752).  backend is "cg", "gmres" or "lgmres"
    (options: M= (default "ilu0"), tol=, maxiter=).  Non-homogeneous Dirichlet data (the lift)
    is outside this path."""
    a = a if isinstance(a, Form) else form(a)
    L = L if isinstance(L, Form) else form(L)
    if a.rank != 2 or L.rank != 1:
        raise TypeError(f"solve(a, L): a must be bilinear (rank 2) and L linear (rank 1), got ranks {a.rank} and {L.rank}")
    V = a.test_space
    if L.test_space is not V:
        raise ValueError("solve(a, L): a and L must have the same test space")
    if a.trial_space is not V:
        raise TypeError("solve(a, L): the form must be square (trial space = test space)")
    if dirichlet is not None and not (isinstance(dirichlet, (int, float)) and dirichlet == 0):
        raise NotImplementedError("solve(dirichlet=...): non-homogeneous Dirichlet data is outside this path")
    if str(backend).lower() == "auto":
        backend = "cg"
    A = a.assemble()
    b = np.array(L.assemble(), dtype=np.float64)
    S = A.solver(backend, **options)
    return S.solve(b)
