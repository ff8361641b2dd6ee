"""Meshes: femd.Mesh2D, rectangle_mesh and mesh_from_arrays (python/femd/mesh2d.py:687-717 and
the mesh_from_arrays wrapper), held by the native library (mesh/structured.hpp:99-351)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

_DIAG = {"right": 0, "/": 0, "left": 1, "\\": 1, "alternate": 2, "alternating": 2, "right/left": 2,
         "crossed": 3, "crisscross": 3, "criss-cross": 3, "union-jack": 3, "unionjack": 3}


class Mesh2D:
    """A triangulation: counter-clockwise triangles, neighbour links (across the edge opposite
    each vertex), boundary segments with markers."""

    nverts = 3

    def __init__(self, handle, report=None):
        self._h = handle
        self.report = report or {}
        s = np.zeros(3, dtype=np.int32)
        L.check(L.load().femd_mesh_sizes(self._h, s))
        self.npoints, self.ncells, self.nsegments = map(int, s)
        self.ntriangles = self.ncells if self.nverts == 3 else 0
        self._arrays = None

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                L.load().femd_mesh_free(h)
            except Exception:
                pass

    def _fetch(self):
        if self._arrays is None:
            pts = np.empty(2 * self.npoints)
            nv = self.nverts
            tri = np.empty(max(nv * self.ncells, 1), dtype=np.int32)
            nbr = np.empty_like(tri)
            seg = np.empty(max(2 * self.nsegments, 1), dtype=np.int32)
            mk = np.empty(max(self.nsegments, 1), dtype=np.int32)
            L.check(L.load().femd_mesh_arrays(self._h, pts, tri, nbr, seg, mk))
            self._arrays = (pts.reshape(-1, 2), tri[: nv * self.ncells].reshape(-1, nv), nbr[: nv * self.ncells].reshape(-1, nv),
                            seg[: 2 * self.nsegments].reshape(-1, 2), mk[: self.nsegments])
        return self._arrays

    @property
    def points(self) -> np.ndarray:
        return self._fetch()[0]

    @property
    def triangles(self) -> np.ndarray:
        return self._fetch()[1]

    @property
    def neighbours(self) -> np.ndarray:
        return self._fetch()[2]

    @property
    def segments(self) -> np.ndarray:
        return self._fetch()[3]

    @property
    def segment_markers(self) -> np.ndarray:
        return self._fetch()[4]

    def __repr__(self):
        return f"Mesh2D({self.npoints} points, {self.ntriangles} triangles, {self.nsegments} segments)"


def _markers4(markers):
    if isinstance(markers, (int, np.integer)):
        return [int(markers)] * 4
    m = [int(v) for v in markers]
    if len(m) != 4:
        raise ValueError("markers: one int or four (bottom, right, top, left)")
    return m


def rectangle_mesh(x0, y0, x1, y1, nx, ny, *, diagonal="right", markers=(1, 2, 3, 4)) -> Mesh2D:
    """Structured mesh of [x0, x1] x [y0, y1] with nx x ny cells (mesh2d.py:687-717): grid vertex
    (i, j) is point j(nx+1) + i, each cell split by `diagonal` ("right", "left", "alternate",
    "crossed"), sides marked bottom, right, top, left."""
    d = _DIAG.get(str(diagonal).lower())
    if d is None:
        raise ValueError(f"unknown diagonal {diagonal!r}: use 'right', 'left', 'alternate' or 'crossed'")
    if int(nx) != nx or int(ny) != ny:
        raise TypeError("nx and ny must be integers")
    h = C.c_void_p()
    L.check(L.load().femd_mesh_rectangle(float(x0), float(y0), float(x1), float(y1), int(nx), int(ny), d,
                                         L.i32(_markers4(markers)), C.byref(h)))
    return Mesh2D(h.value, {"structured": (int(nx), int(ny), str(diagonal).lower())})


def mesh_from_arrays(points, triangles, segments=None, segment_markers=None, marker=1) -> Mesh2D:
    """A mesh from arrays (mesh_from_arrays, structured.hpp:99-221): clockwise triangles are
    re-oriented, neighbours linked, the listed segments kept with their markers and every other
    boundary edge added with `marker`.  The triangle order is kept."""
    pts = L.f64(points).reshape(-1)
    tri = L.i32(triangles).reshape(-1)
    if pts.size % 2 or tri.size % 3:
        raise ValueError("mesh_from_arrays: points is (n, 2) and triangles (m, 3)")
    seg = L.i32(segments if segments is not None else np.zeros((0, 2))).reshape(-1)
    mk = None if segment_markers is None else L.i32(segment_markers)
    if mk is not None and mk.size * 2 != seg.size:
        raise ValueError("mesh_from_arrays: segment_markers must have one entry per segment")
    h = C.c_void_p()
    L.check(L.load().femd_mesh_from_arrays(pts.size // 2, pts, tri.size // 3, tri, seg.size // 2,
                                           seg if seg.size else np.zeros(1, dtype=np.int32),
                                           mk.ctypes.data if mk is not None else None, int(marker), C.byref(h)))
    return Mesh2D(h.value)


class QuadMesh(Mesh2D):
    """A mesh of convex quadrilaterals (mesh/quad_mesh.hpp, python/femd/quadmesh.py:21-128):
    counter-clockwise quads, neighbour i across the edge from vertex i to vertex i+1, boundary
    segments with markers."""

    nverts = 4

    @property
    def nquads(self) -> int:
        return self.ncells

    @property
    def quads(self) -> np.ndarray:
        return self._fetch()[1]

    @property
    def triangles(self):
        raise AttributeError("a QuadMesh has quads, not triangles")

    def __repr__(self):
        return f"QuadMesh({self.npoints} points, {self.nquads} quads, {self.nsegments} segments)"


def rectangle_quad_mesh(x0, y0, x1, y1, nx, ny, *, markers=(1, 2, 3, 4)) -> QuadMesh:
    """Structured mesh of [x0, x1] x [y0, y1] with nx x ny equal rectangles (quadmesh.py:161-170):
    grid vertex (i, j) is point j(nx+1) + i, quad (i, j) is quad j nx + i with vertices (i, j),
    (i+1, j), (i+1, j+1), (i, j+1); sides marked bottom, right, top, left."""
    if int(nx) != nx or int(ny) != ny:
        raise TypeError("nx and ny must be integers")
    h = C.c_void_p()
    L.check(L.load().femd_quad_mesh_rectangle(float(x0), float(y0), float(x1), float(y1), int(nx), int(ny),
                                              L.i32(_markers4(markers)), C.byref(h)))
    return QuadMesh(h.value, {"structured": (int(nx), int(ny), "rectangle")})


def quad_mesh_from_arrays(points, quads, segments=None, segment_markers=None, marker=1) -> QuadMesh:
    """A QuadMesh from arrays (quadmesh.py:130-143): clockwise quads are re-oriented, a degenerate
    or non-convex quad is refused, the listed segments kept with their markers and every other
    boundary edge added with `marker`."""
    pts = L.f64(points).reshape(-1)
    q = L.i32(quads).reshape(-1)
    if pts.size % 2 or q.size % 4:
        raise ValueError("quad_mesh_from_arrays: points is (n, 2) and quads (m, 4)")
    seg = L.i32(segments if segments is not None else np.zeros((0, 2))).reshape(-1)
    mk = None if segment_markers is None else L.i32(segment_markers)
    if mk is not None and mk.size * 2 != seg.size:
        raise ValueError("quad_mesh_from_arrays: segment_markers must have one entry per segment")
    h = C.c_void_p()
    L.check(L.load().femd_quad_mesh_from_arrays(pts.size // 2, pts, q.size // 4, q, seg.size // 2,
                                                seg if seg.size else np.zeros(1, dtype=np.int32),
                                                mk.ctypes.data if mk is not None else None, int(marker), C.byref(h)))
    return QuadMesh(h.value)
