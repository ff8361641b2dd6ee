"""fd.solve(a, L, backend=...) on a 2D mesh (python/femd/linalg.py:710-755), Krylov backends."""
from __future__ import annotations

import numpy as np

from .forms import Form, form


def solve(a, L, backend="cg", dirichlet=None, **options):
    """Solve a(u, v) = L(v) for all v in V (linalg.py:710-755): A = a.assemble(), b = L.assemble(),
    then A.solver(backend).solve(b).  backend is "cg", "gmres" or "lgmres" (options: M= (default
    "ilu0"), tol=, maxiter=).  With Dirichlet data -- dirichlet= (a number, a callable g(x, y), an
    expression in fd.x, fd.y, or a dict {marker: ...}, as V.lift takes) -- the solution is
    u = u_0 + u_g with u_g = V.lift(dirichlet) and u_0 in V solving a(u_0, v) = L(v) - a(u_g, v); it
    is returned as the raw vector of V.unconstrained.  Without data in the call, the data given with the
    space (LagrangeSpace(m, k, dirichlet={marker: data}) or data=) are used; without any, the adapted
    vector of V."""
    a = a if isinstance(a, Form) else form(a)
    L = L if isinstance(L, Form) else form(L)
    if a.rank != 2 or L.rank != 1:
        raise TypeError(f"solve(a, L): a must be bilinear (rank 2) and L linear (rank 1), got ranks {a.rank} and {L.rank}")
    V = a.test_space
    if L.test_space is not V:
        raise ValueError("solve(a, L): a and L must have the same test space")
    if a.trial_space is not V:
        raise TypeError("solve(a, L): the form must be square (trial space = test space)")
    if str(backend).lower() == "auto":
        backend = "cg"
    if dirichlet is None:
        # the data given with the space, when the call gives none (linalg.py:697-706)
        dirichlet = getattr(V, "boundary_data", None)
    A = a.assemble()
    b = np.array(L.assemble(), dtype=np.float64)
    S = A.solver(backend, **options)
    if dirichlet is None or V.n_constraints == 0:
        if dirichlet is not None and not (isinstance(dirichlet, (int, float)) and dirichlet == 0):
            raise ValueError("solve(dirichlet=...): the space has no Dirichlet condition built in; create it "
                             "with dirichlet=[markers] or bc='dirichlet'")
        return S.solve(b)
    ug = V.lift(dirichlet)
    b -= a.action(ug)
    return V.prolongate(S.solve(b)) + ug
