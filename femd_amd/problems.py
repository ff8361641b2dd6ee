"""The synthetic workloads of the benchmark configurations, built from recorded seeds.

* ``fourier_source(seed)``      f(x, y) = sum_{k<=8} a_k sin(kx pi x) sin(ky pi y), a_k ~ N(0, 1),
                                kx, ky ~ U{1..6} (numpy.random.default_rng(seed));
* ``kappa_field(seed)``         kappa = exp(0.5 g), g = sum_{m<=16} c_m cos(2 pi (w_m . x) + phi_m),
                                c_m ~ N(0, 1)/4, w_m integer with |w| <= 8, phi_m ~ U[0, 2 pi);
* ``delaunay_square(h, seed)``  a seeded unstructured Delaunay mesh of the unit square: boundary
                                nodes every h, interior nodes on a jittered triangular lattice,
                                triangulated by Qhull and sorted by centroid rows (the reference's
                                mesher has no seed, so the mesh arrives as arrays, as
                                mesh_from_arrays takes them);
* ``poisson_problem(...)``      the forms of -div(kappa grad u) + c u = f with Dirichlet sides.

Each returns form-language expressions (femd_amd.forms) or arrays; nothing here runs on the GPU.
"""
from __future__ import annotations

import math

import numpy as np

from . import forms as F


def fourier_source(seed: int = 0, nmodes: int = 8, kmax: int = 6):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(nmodes)
    kx = rng.integers(1, kmax + 1, nmodes)
    ky = rng.integers(1, kmax + 1, nmodes)
    terms = [F.Const(float(a[k])) * F.sin(F.Const(float(kx[k] * math.pi)) * F.x) * F.sin(F.Const(float(ky[k] * math.pi)) * F.y)
             for k in range(nmodes)]
    return F.Sum(terms), dict(a=a, kx=kx, ky=ky)


def kappa_field(seed: int = 1, nmodes: int = 16, wmax: int = 8):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal(nmodes) / math.sqrt(nmodes)
    w = []
    while len(w) < nmodes:
        v = rng.integers(-wmax, wmax + 1, 2)
        if 0 < v[0] ** 2 + v[1] ** 2 <= wmax ** 2:
            w.append(v)
    w = np.array(w)
    phi = rng.uniform(0.0, 2.0 * math.pi, nmodes)
    g = F.Sum([F.Const(float(c[m])) * F.cos(F.Const(float(2 * math.pi * w[m, 0])) * F.x
                                            + F.Const(float(2 * math.pi * w[m, 1])) * F.y + F.Const(float(phi[m])))
               for m in range(nmodes)])
    return F.exp(F.Const(0.5) * g), dict(c=c, w=w, phi=phi)


def _triangles(P, h):
    """The Delaunay triangles of P (Qhull), without the slivers of collinear boundary nodes."""
    from scipy.spatial import Delaunay
    tri = Delaunay(P).simplices.astype(np.int64)
    a, b, c = P[tri[:, 0]], P[tri[:, 1]], P[tri[:, 2]]
    area = 0.5 * ((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))
    return tri[np.abs(area) > 1e-12 * h * h]


def _angles(P, tri):
    """The three angles (degrees) of every triangle."""
    a, b, c = P[tri[:, 0]], P[tri[:, 1]], P[tri[:, 2]]

    def at(u, v, w):
        e1, e2 = v - u, w - u
        cs = (e1 * e2).sum(1) / (np.linalg.norm(e1, axis=1) * np.linalg.norm(e2, axis=1))
        return np.degrees(np.arccos(np.clip(cs, -1.0, 1.0)))
    return np.stack([at(a, b, c), at(b, c, a), at(c, a, b)], 1)


def _circumcentres(P, tri):
    """Circumcentres (m, 2) and circumradii (m,) of triangles."""
    a, b, c = P[tri[:, 0]], P[tri[:, 1]], P[tri[:, 2]]
    d = 2.0 * (a[:, 0] * (b[:, 1] - c[:, 1]) + b[:, 0] * (c[:, 1] - a[:, 1]) + c[:, 0] * (a[:, 1] - b[:, 1]))
    sa, sb, sc = (a * a).sum(1), (b * b).sum(1), (c * c).sum(1)
    cc = np.stack([(sa * (b[:, 1] - c[:, 1]) + sb * (c[:, 1] - a[:, 1]) + sc * (a[:, 1] - b[:, 1])) / d,
                   (sa * (c[:, 0] - b[:, 0]) + sb * (a[:, 0] - c[:, 0]) + sc * (b[:, 0] - a[:, 0])) / d], 1)
    return cc, np.linalg.norm(cc - a, axis=1)


def _refine_square(P, h, min_angle, max_rounds=200):
    """Delaunay refinement of a point set of the unit square whose boundary nodes lie on its sides
    (Ruppert's scheme, which the reference's mesher follows: triangulate/refine.hpp).  Every round
    each triangle with an angle below min_angle proposes its circumcentre -- or, when that point
    lies outside the square or inside the diametral circle of a boundary segment, the segment's
    midpoint -- and the proposals are inserted as an independent set: worst triangle first, a
    proposal is dropped for this round when it lies within its own radius (circumradius, or half
    the segment) of an accepted one or within an accepted one's radius of it.  The proposals of one
    round then change disjoint cavities, as one-at-a-time insertion would.  Deterministic."""
    from scipy.spatial import cKDTree
    for _ in range(max_rounds):
        tri = _triangles(P, h)
        worst = _angles(P, tri).min(axis=1)
        bad = np.flatnonzero(worst < min_angle)
        if bad.size == 0:
            return P
        bad = bad[np.argsort(worst[bad], kind="stable")]
        x, rad = _circumcentres(P, tri[bad])
        split = np.full(bad.size, -1, dtype=np.int64)
        # the boundary segments: consecutive boundary nodes along each side; a proposal that is
        # outside the square or encroaches a segment becomes that segment's midpoint
        for si, (axis, val) in enumerate(((1, 0.0), (0, 1.0), (1, 1.0), (0, 0.0))):
            along = np.sort(P[np.abs(P[:, axis] - val) <= 1e-14, 1 - axis])
            t = np.clip(x[:, 1 - axis], 0.0, 1.0)
            j = np.clip(np.searchsorted(along, t, side="right") - 1, 0, along.size - 2)
            lo, hi = along[j], along[j + 1]
            mid = np.empty_like(x)
            mid[:, axis], mid[:, 1 - axis] = val, 0.5 * (lo + hi)
            outside = x[:, axis] < 0.0 if val == 0.0 else x[:, axis] > 1.0
            hit = (split < 0) & (outside | (np.linalg.norm(x - mid, axis=1) < 0.5 * (hi - lo)))
            split[hit] = si * (1 << 40) + j[hit]
            x[hit], rad[hit] = mid[hit], 0.5 * (hi - lo)[hit]
        # one proposal per split segment (its worst triangle's)
        keep = np.ones(bad.size, dtype=bool)
        _, first = np.unique(split[split >= 0], return_index=True)
        keep[split >= 0] = False
        keep[np.flatnonzero(split >= 0)[first]] = True
        x, rad = x[keep], rad[keep]
        tree = cKDTree(x)
        blocked = np.zeros(x.shape[0], dtype=bool)
        accepted = np.zeros(x.shape[0], dtype=bool)
        near = tree.query_ball_point(x, rad)
        for k in range(x.shape[0]):
            if blocked[k] or accepted[near[k]].any():
                continue
            accepted[k] = True
            blocked[near[k]] = True
        P = np.concatenate([P, x[accepted]])
    raise RuntimeError(f"mesh refinement did not reach {min_angle} degrees in {max_rounds} rounds")


def delaunay_square(h: float, seed: int = 0, jitter: float = 0.12, min_angle_deg: float = None):
    """(points (n, 2), triangles (m, 3), min angle in degrees) of the unit square: boundary nodes
    every h, a seeded jittered lattice inside, triangulated by Qhull and, with min_angle_deg,
    refined until no angle is smaller (Ruppert's scheme)."""
    rng = np.random.default_rng(seed)
    nb = max(2, int(round(1.0 / h)))
    t = np.linspace(0.0, 1.0, nb + 1)[:-1]
    bnd = np.concatenate([np.stack([t, np.zeros_like(t)], 1), np.stack([np.ones_like(t), t], 1),
                          np.stack([1.0 - t, np.ones_like(t)], 1), np.stack([np.zeros_like(t), 1.0 - t], 1)])
    dy = h * math.sqrt(3.0) / 2.0
    rows = int(math.floor(1.0 / dy))
    pts = []
    for j in range(1, rows + 1):
        yj = j * dy
        off = 0.5 * h if j % 2 else 0.0
        xs = np.arange(off, 1.0, h)
        pts.append(np.stack([xs, np.full_like(xs, yj)], 1))
    P = np.concatenate(pts)
    P = P + rng.uniform(-jitter * h, jitter * h, P.shape)
    inside = (P[:, 0] > 0.45 * h) & (P[:, 0] < 1 - 0.45 * h) & (P[:, 1] > 0.45 * h) & (P[:, 1] < 1 - 0.45 * h)
    P = np.concatenate([bnd, P[inside]])
    if min_angle_deg:
        P = _refine_square(P, h, float(min_angle_deg))
    tri = _triangles(P, h)
    cen = P[tri].mean(axis=1)
    order = np.lexsort((cen[:, 0], np.floor(cen[:, 1] / (4 * h))))
    tri = tri[order]
    # renumber vertices in first-use order (locality of the vertex numbering)
    first = np.full(P.shape[0], -1, dtype=np.int64)
    used, at = np.unique(tri.reshape(-1), return_index=True)
    newid = used[np.argsort(at, kind="stable")]
    first[newid] = np.arange(newid.size)
    P = P[newid]
    tri = first[tri]
    return P, tri.astype(np.int32), min_angle(P, tri)


def min_angle(P, tri) -> float:
    return float(_angles(P, tri).min())


def poisson_forms(V, f_expr, kappa=None, reaction=0.0):
    """a = (kappa grad u . grad v + reaction u v) dx, L = f v dx."""
    u, v = F.TrialFunction(V), F.TestFunction(V)
    g = F.dot(F.grad(u), F.grad(v))
    a_expr = g if kappa is None else kappa * g
    if reaction:
        a_expr = a_expr + (reaction * u * v if reaction != 1.0 else u * v)
    return F.form(a_expr * F.dx), F.form(f_expr * v * F.dx)


C1_H = 0.00265


C1_MIN_ANGLE = 28.0


def p3_kappa_problem(h: float = C1_H, seed: int = 0, min_angle_deg: float = C1_MIN_ANGLE):
    """The P3 configuration: seeded Delaunay mesh of the unit square refined to min_angle_deg, P3
    with every side Dirichlet, a = kappa grad u . grad v dx, L = f v dx.  Returns (space, a, L, info)."""
    import time
    from .mesh import mesh_from_arrays
    from .spaces import LagrangeSpace
    t0 = time.perf_counter()
    P, T, ma = delaunay_square(h, seed=seed, min_angle_deg=min_angle_deg)
    V = LagrangeSpace(mesh_from_arrays(P, T), 3, bc="dirichlet")
    kap, _ = kappa_field(1)
    f, _ = fourier_source(0)
    a, L = poisson_forms(V, f, kappa=kap)
    return V, a, L, dict(points=P, triangles=T, min_angle=ma, setup_s=time.perf_counter() - t0)


def fourier_values(info, xs, ys):
    """The field of fourier_source (its info dict) evaluated with NumPy at points."""
    out = np.zeros_like(np.asarray(xs, dtype=np.float64))
    for a, kx, ky in zip(info["a"], info["kx"], info["ky"]):
        out = out + a * np.sin(kx * math.pi * xs) * np.sin(ky * math.pi * ys)
    return out


def heat_problem(N: int = 256, k: int = 2, dt: float = 1e-3, seed: int = 0):
    """The heat configuration: u_t = Laplace u on an N x N structured triangulation, P_k with every
    side Dirichlet, backward Euler with step dt: M = (u v) dx, A = (u v + dt grad u . grad v) dx,
    u0 the random-Fourier field of the seed interpolated.  Returns (space, M, A, u0)."""
    from .mesh import rectangle_mesh
    from .spaces import LagrangeSpace
    V = LagrangeSpace(rectangle_mesh(0, 0, 1, 1, N, N), k, bc="dirichlet")
    u, v = F.TrialFunction(V), F.TestFunction(V)
    M = F.form(u * v * F.dx).assemble()
    A = F.form((u * v + F.Const(float(dt)) * F.dot(F.grad(u), F.grad(v))) * F.dx).assemble()
    _, info = fourier_source(seed)
    u0 = V.interpolate(lambda xs, ys: fourier_values(info, xs, ys))
    return V, M, A, u0


def convection_problem(N: int = 256, k: int = 2, eps: float = 0.01, beta=(1.0, 0.5), seed: int = 0):
    """A nonsymmetric system for GMRES / LGMRES: -eps Laplace u + beta . grad u = f on an N x N
    structured triangulation, P_k with every side Dirichlet, f the seeded random-Fourier series.
    a = (eps grad u . grad v + (beta . grad u) v) dx, L = f v dx.  Returns (space, a, L)."""
    from .mesh import rectangle_mesh
    from .spaces import LagrangeSpace
    V = LagrangeSpace(rectangle_mesh(0, 0, 1, 1, N, N), k, bc="dirichlet")
    u, v = F.TrialFunction(V), F.TestFunction(V)
    b = F.as_vector([F.Const(float(beta[0])), F.Const(float(beta[1]))])
    a = F.form((F.Const(float(eps)) * F.dot(F.grad(u), F.grad(v)) + F.dot(b, F.grad(u)) * v) * F.dx)
    f, _ = fourier_source(seed)
    return V, a, F.form(f * v * F.dx)

