"""fd.LagrangeSpace on a 2D mesh (python/femd/spaces2d.py:84-279, spaces.py:435-442): the
continuous P_k space with the reference's numbering and Dirichlet elimination, held by the
native library (fe/space_2d.hpp:390-526, fe/lagrange_space_2d.hpp:32-67)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .mesh import Mesh2D


def _markers(m, what):
    if m is None:
        return []
    if isinstance(m, (int, np.integer)):
        return [int(m)]
    out = []
    for v in m:
        if isinstance(v, bool) or int(v) != v:
            raise TypeError(f"{what}: markers are integers, got {v!r}")
        out.append(int(v))
    return out


class LagrangeSpace2D:
    """P_k (k >= 1) on a triangular Mesh2D, or Q_k (k = 1..4) on a QuadMesh
    (fe/lagrange_space_quad.hpp:112-210: tensor Lagrange on the reference square, bilinear map).

    bc:        "free" (default) or "dirichlet" (every exterior side eliminated).
    dirichlet: a marker or list of markers whose sides are eliminated; combines with bc.  A dict
               {marker: data} names the sides and gives their data at once (spaces2d.py:115-119).
    nodes:     "equispaced" (default) or "lobatto": Warburton's warp-and-blend points on triangles,
               Gauss-Lobatto tensor points on quadrilaterals (fe/simplex_basis.hpp:150-285,
               lagrange_space_quad.hpp:31-50); the same space with another basis, identical to the
               equispaced one, to rounding, for k <= 2.
    data:      Dirichlet data kept with the space (a number, a callable g(x, y), an expression in
               fd.x, fd.y, or a dict {marker: ...}, as lift() takes); lift() and fd.solve use it
               when they are given none (spaces2d.py:161-172)."""

    family = "lagrange"
    tdim = 2
    nverts = 3

    def __init__(self, mesh: Mesh2D, degree: int, bc="free", dirichlet=None, nodes="equispaced", data=None):
        if isinstance(dirichlet, dict):
            if data is not None:
                raise TypeError(f"{type(self).__name__}: give the data either in dirichlet={{marker: data}} or in "
                                "data=, not both")
            data, dirichlet = dict(dirichlet), sorted(dirichlet)
        if not isinstance(mesh, Mesh2D):
            raise TypeError(f"LagrangeSpace2D: mesh must be a Mesh2D, got {type(mesh).__name__}")
        if isinstance(degree, bool) or int(degree) != degree:
            raise TypeError(f"degree must be an integer, got {degree!r}")
        k = int(degree)
        if k < 1:
            raise ValueError(f"LagrangeSpace2D: degree must be >= 1, got {k}")
        b = "free" if bc is None else str(bc).lower()
        if b not in ("free", "dirichlet"):
            raise ValueError(f"LagrangeSpace2D: bc is 'free' or 'dirichlet' (every exterior side), got {bc!r}")
        marks = sorted(set(_markers(dirichlet, "dirichlet")))
        present = set(np.asarray(mesh.segment_markers).tolist())
        missing = [m for m in marks if m not in present]
        if missing:
            raise ValueError(f"LagrangeSpace2D: no side of the mesh carries marker(s) {missing}; "
                             f"the markers present are {sorted(present)}")
        nd = str(nodes).lower()
        if nd not in ("equispaced", "lobatto"):
            raise ValueError(f"LagrangeSpace2D: nodes is 'equispaced' or 'lobatto', got {nodes!r}")
        h = C.c_void_p()
        d = L.i32(marks) if marks else np.zeros(1, dtype=np.int32)
        L.check(L.load().femd_space_create(mesh._h, k, len(marks), d, int(b == "dirichlet"),
                                           int(nd == "equispaced"), C.byref(h)))
        self._h = h.value
        self.mesh = mesh
        self.nverts = mesh.nverts
        self.cell_name = "quadrilateral" if self.nverts == 4 else "triangle"
        self.degree = k
        self.dirichlet_markers = marks
        self.all_exterior = b == "dirichlet"
        self.node_set = nd
        s = np.zeros(6, dtype=np.int32)
        L.check(L.load().femd_space_sizes(self._h, s))
        self.ncells, self.nloc, self.raw_dim, self.dim, self.nedges, self.npoints = map(int, s)
        self._num = None
        self._pattern = None
        self._data = None
        if data is not None:
            if self.n_constraints == 0:
                raise ValueError(f"{type(self).__name__}: data= gives Dirichlet data, and this space has no Dirichlet "
                                 "side; name them with dirichlet=[markers] or bc='dirichlet'")
            # evaluated once now, so that a marker without a Dirichlet side or a callable that does
            # not evaluate fails here rather than at the first solve
            self._lift(data)
            self._data = data

    @property
    def boundary_data(self):
        """The Dirichlet data given with the space (dirichlet={marker: data} or data=), or None."""
        return self._data

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                lib = L.load()
                for f in getattr(self, "_facet_sets", {}).values():
                    lib.femd_facets_free(f)
                lib.femd_space_free(h)
            except Exception:
                pass

    def facet_set(self, degree, markers=None):
        """The library's ds point set of this space (cached): exterior facets with a marker in
        `markers`, every exterior facet for None, Gauss-Legendre points exact to `degree`
        (Cache2D on facets, fe/cache_2d.hpp:124-183; boundary_cache, spaces2d.py:265-270)."""
        cache = self.__dict__.setdefault("_facet_sets", {})
        key = (int(degree), None if markers is None else tuple(sorted({int(m) for m in markers})))
        h = cache.get(key)
        if h is None:
            m = L.i32(key[1]) if key[1] else np.zeros(1, dtype=np.int32)
            out = C.c_void_p()
            L.check(L.load().femd_facets_create(self._h, key[0], len(key[1]) if key[1] else 0, m, int(key[1] is None),
                                                C.byref(out)))
            h = cache[key] = out.value
        return h

    def boundary_points(self, degree, markers=None):
        """(x, y, w, nx, ny) per (facet, point) of facet_set(degree, markers)."""
        h = self.facet_set(degree, markers)
        sz = np.zeros(2, dtype=np.int32)
        L.check(L.load().femd_facets_sizes(h, sz))
        out = [np.empty(int(sz[0]) * int(sz[1])) for _ in range(5)]
        L.check(L.load().femd_facets_points(h, *out))
        return tuple(out)

    @property
    def n_constraints(self) -> int:
        return self.raw_dim - self.dim

    def _numbering(self):
        if self._num is None:
            cd = np.empty(self.ncells * self.nloc, dtype=np.int32)
            cr = np.empty_like(cd)
            a2r = np.empty(max(self.dim, 1), dtype=np.int32)
            xy = np.empty(2 * self.raw_dim)
            L.check(L.load().femd_space_numbering(self._h, cd, cr, a2r, xy))
            self._num = (cd.reshape(self.ncells, self.nloc), cr.reshape(self.ncells, self.nloc), a2r[: self.dim],
                         xy.reshape(-1, 2))
        return self._num

    def cell_dofs(self) -> np.ndarray:
        """(ncells, nloc) adapted indices, -1 where the node is eliminated."""
        return self._numbering()[0]

    def cell_raw_dofs(self) -> np.ndarray:
        return self._numbering()[1]

    def adapted_to_raw(self) -> np.ndarray:
        return self._numbering()[2]

    def raw_coordinates(self) -> np.ndarray:
        return self._numbering()[3]

    def dof_coordinates(self) -> np.ndarray:
        """(dim, 2) coordinates of the degrees of freedom."""
        return self.raw_coordinates()[self.adapted_to_raw()]

    def node_coordinates(self) -> np.ndarray:
        return self.raw_coordinates()

    def prolongate(self, c) -> np.ndarray:
        """Adapted -> raw coefficients (zeros on the eliminated nodes)."""
        r = np.zeros(self.raw_dim)
        r[self.adapted_to_raw()] = np.asarray(c, dtype=np.float64)
        return r

    def restrict(self, raw) -> np.ndarray:
        return np.asarray(raw, dtype=np.float64)[self.adapted_to_raw()]

    # ---- Dirichlet data (spaces2d.py:212-220, 239-248, 373-414) -------------------------------
    @property
    def unconstrained(self):
        """The same mesh, degree and nodes with nothing eliminated (created once); self if nothing is."""
        if self.n_constraints == 0:
            return self
        U = self.__dict__.get("_unconstrained")
        if U is None:
            U = type(self)(self.mesh, self.degree, nodes=self.node_set)
            self._unconstrained = U
        return U

    def constrained(self) -> np.ndarray:
        """Raw indices of the eliminated nodes."""
        mask = np.ones(self.raw_dim, dtype=bool)
        mask[self.adapted_to_raw()] = False
        return np.flatnonzero(mask)

    def facets(self):
        """The facet list of fe/space_2d.hpp:432-448 -- every exterior edge and every interior edge
        that is a marked segment (once), cells ascending, local edges ascending: arrays cell,
        local edge, marker (0 for none), exterior."""
        m = self.mesh
        T = m.triangles
        nv = T.shape[1]
        va = np.arange(nv) if nv == 4 else (np.arange(3) + 1) % 3
        vb = (np.arange(nv) + 1) % 4 if nv == 4 else (np.arange(3) + 2) % 3
        A, B = T[:, va], T[:, vb]
        lo, hi = np.minimum(A, B).astype(np.int64), np.maximum(A, B).astype(np.int64)
        key = (lo * m.npoints + hi).ravel()
        seg = m.segments.astype(np.int64)
        skey = np.minimum(seg[:, 0], seg[:, 1]) * m.npoints + np.maximum(seg[:, 0], seg[:, 1])
        mark = np.full(key.size, -1, dtype=np.int64)
        if skey.size:
            order = np.argsort(skey, kind="stable")
            sk = skey[order]
            pos = np.minimum(np.searchsorted(sk, key), sk.size - 1)
            hit = sk[pos] == key
            mark[hit] = np.asarray(m.segment_markers)[order][pos[hit]]
        ext = (m.neighbours < 0).ravel()
        keep = ext | (mark >= 0)
        first = np.zeros(key.size, dtype=bool)                       # an interior segment once, first met
        _, idx = np.unique(key, return_index=True)
        first[idx] = True
        keep &= ext | first
        sel = np.flatnonzero(keep)
        return sel // nv, sel % nv, np.where(mark[sel] >= 0, mark[sel], 0), ext[sel]

    def boundary_nodes(self, marker=None) -> np.ndarray:
        """Raw indices of the nodes on the sides with this marker (every exterior side for None):
        the edges' vertices and edge nodes."""
        cell, local, mark, ext = self.facets()
        sel = ext if marker is None else np.isin(mark, _markers(marker, "marker"))
        nv, k = self.nverts, self.degree
        cr = self.cell_raw_dofs()
        va = local[sel] if nv == 4 else (local[sel] + 1) % 3
        vb = (local[sel] + 1) % 4 if nv == 4 else (local[sel] + 2) % 3
        parts = [cr[cell[sel], va], cr[cell[sel], vb]]
        if k > 1:
            l = nv + local[sel][:, None] * (k - 1) + np.arange(k - 1)[None, :]
            parts.append(cr[cell[sel][:, None], l].ravel())
        return np.unique(np.concatenate(parts)).astype(np.int64)

    def lift(self, g=None):
        """The raw vector (a function of V.unconstrained) carrying Dirichlet data on the eliminated
        nodes: a number or a callable g(x, y) / expression in fd.x, fd.y for every Dirichlet side,
        or a dict {marker: ...}; sides left out get zero data, and where two sides meet the corner
        takes the data of the larger marker (spaces2d.py:373-414).  Without data, the data given
        with the space (boundary_data), if any."""
        if g is None and self._data is not None:
            g = self._data
        return self._lift(g)

    def _lift(self, g=None):
        """lift() with exactly the data given: None is zero; the space's own data are not used."""
        raw = np.zeros(self.raw_dim)
        cons = self.constrained()
        if g is None or cons.size == 0:
            if g is not None and cons.size == 0 and not _is_zero(g):
                raise ValueError("lift(): this space has no Dirichlet condition built in; natural data enter "
                                 "through ds(marker)")
            return raw
        xy = self.raw_coordinates()
        if isinstance(g, dict):
            allowed = set(self.dirichlet_markers)
            if self.all_exterior:
                allowed |= set(np.asarray(self.mesh.segment_markers).tolist()) | {0}
            for mk in sorted(g):
                if int(mk) not in allowed:
                    raise ValueError(f"lift(): marker {mk} carries no Dirichlet condition in this space "
                                     f"(Dirichlet markers {sorted(allowed)}); natural data enter through ds({mk})")
                nodes = self.boundary_nodes(int(mk))
                nodes = nodes[np.isin(nodes, cons)]
                raw[nodes] = _values(g[mk], xy[nodes])
        else:
            raw[cons] = _values(g, xy[cons])
        return raw

    def interpolate(self, f) -> np.ndarray:
        """Nodal values of a callable f(x, y) at the degrees of freedom (the coefficients)."""
        xy = self.dof_coordinates()
        return np.asarray(f(xy[:, 0], xy[:, 1]), dtype=np.float64) * np.ones(self.dim)

    def pattern(self, trial=None):
        """The CSR pattern (indptr, indices) of matrices with test and trial space self (cached)."""
        if trial is not None and trial is not self:
            raise NotImplementedError("pattern: only the square pattern of a space with itself")
        if self._pattern is None:
            nnz = C.c_int64()
            L.check(L.load().femd_pattern(self._h, C.byref(nnz), None, None))
            ip = np.empty(self.dim + 1, dtype=np.int32)
            ix = np.empty(max(nnz.value, 1), dtype=np.int32)
            L.check(L.load().femd_pattern(self._h, None, ip.ctypes.data, ix.ctypes.data))
            self._pattern = (ip, ix[: nnz.value])
        return self._pattern

    @property
    def bc_description(self) -> str:
        parts = []
        if self.all_exterior:
            parts.append("Dirichlet on every exterior side")
        if self.dirichlet_markers:
            parts.append("Dirichlet on marker(s) " + ", ".join(map(str, self.dirichlet_markers)))
        return "; ".join(parts) or "free (natural on every side)"

    def __repr__(self):
        name = "LagrangeSpaceQ" if self.nverts == 4 else "LagrangeSpace2D"
        return f"{name}(degree={self.degree}, ncells={self.ncells}, bc={self.bc_description}, dim={self.dim})"


def _is_zero(g):
    return isinstance(g, (int, float)) and g == 0


def _values(g, xy):
    """Boundary data at the nodes xy: an expression in fd.x, fd.y and Constants, a callable g(x, y),
    or a number (spaces2d.py:540-555)."""
    from .forms import Expr, evaluate
    if isinstance(g, Expr):
        return np.broadcast_to(np.asarray(evaluate(g, xy[:, 0], xy[:, 1]), dtype=np.float64), (xy.shape[0],))
    if callable(g):
        return np.broadcast_to(np.asarray(g(xy[:, 0], xy[:, 1]), dtype=np.float64), (xy.shape[0],))
    return float(g)


def LagrangeSpace(mesh, degree, bc="free", dirichlet=None, nodes="equispaced", data=None):
    """fd.LagrangeSpace on a Mesh2D (P_k) or a QuadMesh (Q_k) (the 2D branches of spaces.py:435-442)."""
    if not isinstance(mesh, Mesh2D):
        raise TypeError("LagrangeSpace: this package builds the 2D (Mesh2D) spaces only")
    return LagrangeSpace2D(mesh, degree, bc=bc, dirichlet=dirichlet, nodes=nodes, data=data)
