"""The product's host-side setup and its C ABI, on the CPU (no kernel launches):
the library loads and exports every entry point of include/femd_hip.h; meshes, the P_k
numbering, the pattern, quadrature and RCM agree bit for bit with the oracle; bad input is
rejected with the reference's exception types."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import femd_amd as fd
from femd_amd import _lib, problems

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "femd_hip.h").read_text()
    declared = sorted(set(re.findall(r"\b(femd_[a-z0-9_]+)\s*\(", header)))
    assert len(declared) > 40
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_lib.exported_symbols()) <= set(declared)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("diag", ["right", "left", "alternate", "crossed"])
def test_rectangle_numbering_vs_oracle(orc, k, diag):
    m = fd.rectangle_mesh(0, 0, 1, 1, 5, 3, diagonal=diag)
    for kw, okw in ((dict(bc="dirichlet"), dict(all_exterior=True)),
                    (dict(dirichlet=[1, 4]), dict(dirichlet=[1, 4], all_exterior=False)),
                    (dict(), dict(all_exterior=False))):
        V = fd.LagrangeSpace(m, k, **kw)
        P = orc.rectangle(5, 3, k, diag, **okw)
        num = P.numbering()
        assert np.array_equal(V.cell_dofs(), num["cell_dof"])
        assert np.array_equal(V.cell_raw_dofs(), num["cell_raw"])
        assert np.array_equal(V.adapted_to_raw(), num["ad2raw"])
        assert np.array_equal(V.raw_coordinates(), num["raw_xy"], equal_nan=True)
        for a, b in zip(V.pattern(), P.pattern()):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_lobatto_numbering_vs_oracle(orc, k):
    """nodes="lobatto": warp-and-blend nodes on triangles, Gauss-Lobatto tensor nodes on quads;
    the numbering is that of the equispaced space and the node coordinates match bit for bit."""
    pts, q = _distorted_quads(4, 5, k)
    cases = [(fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 5, 3, diagonal="crossed"), k, bc="dirichlet", nodes="lobatto"),
              orc.rectangle(5, 3, k, "crossed", equispaced=False))]
    if k <= 4:
        cases.append((fd.LagrangeSpace(fd.quad_mesh_from_arrays(pts, q), k, nodes="lobatto"),
                      orc.quad_from_arrays(pts, q, k=k, all_exterior=False, equispaced=False)))
    for V, P in cases:
        assert V.node_set == "lobatto"
        num = P.numbering()
        assert np.array_equal(V.cell_dofs(), num["cell_dof"])
        assert np.array_equal(V.raw_coordinates(), num["raw_xy"], equal_nan=True)
    with pytest.raises(ValueError):
        fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 2, 2), 3, nodes="chebyshev")


def _distorted_quads(nx, ny, seed):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.linspace(0, 1, nx + 1), np.linspace(0, 1, ny + 1))
    pts = np.stack([xs.ravel(), ys.ravel()], 1)
    inner = (pts[:, 0] > 0) & (pts[:, 0] < 1) & (pts[:, 1] > 0) & (pts[:, 1] < 1)
    pts[inner] += rng.uniform(-0.25, 0.25, (inner.sum(), 2)) / max(nx, ny)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    v = (j * (nx + 1) + i).ravel()
    q = np.stack([v, v + 1, v + nx + 2, v + nx + 1], 1)
    q[::3] = q[::3, ::-1]
    return pts, q


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_quad_numbering_vs_oracle(orc, k):
    m = fd.rectangle_quad_mesh(0, 0, 1, 1, 6, 4)
    assert m.nverts == 4 and m.nquads == 24 and m.quads.shape == (24, 4)
    pts, q = _distorted_quads(5, 6, k)
    mq = fd.quad_mesh_from_arrays(pts, q)
    cases = [(fd.LagrangeSpace(m, k, bc="dirichlet"), orc.quad_rectangle(6, 4, k)),
             (fd.LagrangeSpace(m, k, dirichlet=[2, 3]), orc.quad_rectangle(6, 4, k, dirichlet=[2, 3], all_exterior=False)),
             (fd.LagrangeSpace(mq, k), orc.quad_from_arrays(pts, q, k=k, all_exterior=False))]
    for V, P in cases:
        num = P.numbering()
        assert np.array_equal(V.cell_dofs(), num["cell_dof"])
        assert np.array_equal(V.cell_raw_dofs(), num["cell_raw"])
        assert np.array_equal(V.adapted_to_raw(), num["ad2raw"])
        assert np.array_equal(V.raw_coordinates(), num["raw_xy"], equal_nan=True)
        for a, b in zip(V.pattern(), P.pattern()):
            assert np.array_equal(a, b)
    po = orc.quad_from_arrays(pts, q, k=1)
    opts, oq, onbr, oseg, omk = orc.mesh_arrays(po)
    assert np.array_equal(mq.quads, oq) and np.array_equal(mq.neighbours, onbr)
    assert np.array_equal(mq.segments, oseg) and np.array_equal(mq.segment_markers, omk)
    n = __import__("ctypes").c_int()
    xi, eta, w = np.empty(64), np.empty(64), np.empty(64)
    _lib.check(_lib.load().femd_quadrature_square(7, __import__("ctypes").byref(n), xi, eta, w))
    for a, b in zip((xi[: n.value], eta[: n.value], w[: n.value]), orc.quad_quadrature(7)):
        assert np.array_equal(a, b)


def test_quad_mesh_errors():
    with pytest.raises(ValueError):                          # a bow-tie is not convex
        fd.quad_mesh_from_arrays([[0, 0], [1, 0], [0, 1], [1, 1]], [[0, 1, 2, 3]])
    with pytest.raises(ValueError):
        fd.LagrangeSpace(fd.rectangle_quad_mesh(0, 0, 1, 1, 2, 2), 5)     # Q5 is not built


def test_mesh_arrays_vs_oracle(orc):
    P, T, _ = problems.delaunay_square(0.15, seed=4)
    T = T.copy()
    T[::3] = T[::3][:, [0, 2, 1]]
    m = fd.mesh_from_arrays(P, T)
    po = orc.from_arrays(P, T, k=2)
    pts, tri, nbr, seg, mk = orc.mesh_arrays(po)
    assert np.array_equal(m.triangles, tri) and np.array_equal(m.neighbours, nbr)
    assert np.array_equal(m.segments, seg) and np.array_equal(m.segment_markers, mk)
    V = fd.LagrangeSpace(m, 3, bc="dirichlet")
    p3 = orc.from_arrays(P, T, k=3)
    assert np.array_equal(V.cell_dofs(), p3.numbering()["cell_dof"])
    for a, b in zip(V.pattern(), p3.pattern()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("h,seed", [(0.1, 0), (0.04, 3), (0.015, 7)])
def test_refined_delaunay_square(h, seed):
    """The quality mesh of the unit square: every angle at least 28 degrees, a conforming
    triangulation that tiles the square (areas sum to 1, every interior edge shared by two
    triangles, boundary edges only on the sides, one per boundary node), Delaunay (no vertex
    strictly inside any circumcircle, by brute force), and reproducible from its seed."""
    P, T, ma = problems.delaunay_square(h, seed=seed, min_angle_deg=28)
    assert ma >= 28.0 and ma == problems.min_angle(P, T)
    a, b, c = P[T[:, 0]], P[T[:, 1]], P[T[:, 2]]
    area = 0.5 * np.abs((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))
    assert abs(area.sum() - 1.0) < 1e-12 and area.min() > 0
    E = np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), 1)
    edges, count = np.unique(E, axis=0, return_counts=True)
    assert count.max() == 2
    side = lambda q: (np.abs(q[:, 0]) < 1e-14) | (np.abs(q[:, 0] - 1) < 1e-14) | (np.abs(q[:, 1]) < 1e-14) | (np.abs(q[:, 1] - 1) < 1e-14)
    bnd = edges[count == 1]
    assert side(P[bnd[:, 0]]).all() and side(P[bnd[:, 1]]).all() and bnd.shape[0] == side(P).sum()
    assert np.unique(T).size == P.shape[0]
    cc, rad = problems._circumcentres(P, T)
    for lo in range(0, T.shape[0], 512):
        d = np.linalg.norm(P[None, :, :] - cc[lo:lo + 512, None, :], axis=2)
        assert (d >= rad[lo:lo + 512, None] * (1 - 1e-9)).all()
    P2, T2, _ = problems.delaunay_square(h, seed=seed, min_angle_deg=28)
    assert np.array_equal(P, P2) and np.array_equal(T, T2)
    # the unrefined lattice of the same seed is worse: refinement did the work
    assert problems.delaunay_square(h, seed=seed)[2] < 28.0


def test_user_segments_and_markers(orc):
    pts = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [0.5, 0.5]], dtype=float)
    tris = np.array([[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]])
    m = fd.mesh_from_arrays(pts, tris, segments=[[2, 1], [0, 4]], segment_markers=[7, 9], marker=5)
    po = orc.from_arrays(pts, tris, segments=[[2, 1], [0, 4]], markers=[7, 9], k=1)
    _, _, _, seg, mk = orc.mesh_arrays(po)
    # the oracle numbers unlisted boundary edges with marker 1; the product with `marker`
    assert np.array_equal(m.segments, seg)
    assert list(m.segment_markers[:2]) == [7, 9] and set(m.segment_markers[2:]) == {5}
    V = fd.LagrangeSpace(m, 2, dirichlet=[9])                 # an interior segment as a Dirichlet side
    P = orc.from_arrays(pts, tris, segments=[[2, 1], [0, 4]], markers=[7, 9], k=2, dirichlet=[9], all_exterior=False)
    assert np.array_equal(V.cell_dofs(), P.numbering()["cell_dof"])


def test_quadrature_and_rcm_vs_oracle(orc):
    lib = _lib.load()
    for d in range(0, 13):
        xi, eta, w = np.empty(64), np.empty(64), np.empty(64)
        nq = ctypes.c_int()
        _lib.check(lib.femd_quadrature(d, ctypes.byref(nq), xi, eta, w))
        oxi, oeta, ow = orc.quadrature(d)
        n = nq.value
        assert np.array_equal(xi[:n], oxi) and np.array_equal(eta[:n], oeta) and np.array_equal(w[:n], ow)
    V = fd.LagrangeSpace(fd.mesh_from_arrays(*problems.delaunay_square(0.1, seed=1)[:2]), 3, bc="dirichlet")
    ip, ix = V.pattern()
    p = np.empty(V.dim, dtype=np.int32)
    _lib.check(lib.femd_rcm(V.dim, ip, ix, p))
    assert np.array_equal(p, orc.rcm(ip, ix))


def test_errors_are_the_references():
    with pytest.raises(ValueError):
        fd.rectangle_mesh(0, 0, -1, 1, 3, 3)
    with pytest.raises(ValueError):
        fd.rectangle_mesh(0, 0, 1, 1, 3, 3, diagonal="zigzag")
    m = fd.rectangle_mesh(0, 0, 1, 1, 2, 2)
    with pytest.raises(ValueError):
        fd.LagrangeSpace(m, 0)
    with pytest.raises(ValueError):
        fd.LagrangeSpace(m, 2, dirichlet=[42])
    with pytest.raises(ValueError):
        fd.mesh_from_arrays([[0, 0], [1, 0], [2, 0]], [[0, 1, 2]])            # collinear
    with pytest.raises(ValueError):
        fd.mesh_from_arrays([[0, 0], [1, 0], [0, 1]], [[0, 1, 5]])            # index out of range


def test_form_canonicalization_and_degree():
    V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 2, 2), 2, bc="dirichlet")
    u, v = fd.TrialFunction(V), fd.TestFunction(V)
    F = fd.form((fd.dot(fd.grad(u), fd.grad(v)) + u * v) * fd.dx)
    assert [(t.test.code(), t.trial.code()) for t in F.terms] == [(1, 1), (2, 2), (0, 0)]
    assert F.quad_degree == 4                                                   # 7-point rule
    f, _ = problems.fourier_source(0)
    assert fd.form(f * v * fd.dx).quad_degree == 5                              # 3 + 2
    V3 = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 2, 2), 3, bc="dirichlet")
    u3, v3 = fd.TrialFunction(V3), fd.TestFunction(V3)
    kap, _ = problems.kappa_field(1)
    assert fd.form(kap * fd.dot(fd.grad(u3), fd.grad(v3)) * fd.dx).quad_degree == 8   # 4 + 4: 25 points
    assert fd.form(f * v3 * fd.dx).quad_degree == 7
    assert fd.form((fd.dot(fd.grad(u), fd.grad(v))) * fd.dx(9)).quad_degree == 9
    with pytest.raises(TypeError):
        fd.form(u * fd.dx)
    with pytest.raises(NotImplementedError):
        fd.form(u * v * fd.dS)
    # ds terms: their own degree (forms.py:2487-2500) and point set per marker set, in the form's order
    nv = fd.FacetNormal()
    F = fd.form(u * v * fd.ds(3) + fd.dot(fd.grad(u), fd.grad(v)) * fd.dx + fd.x * u * v * fd.ds((1, 3)))
    assert F.quad_degree == 2 and F.facet_degree == 5
    assert list(F.groups) == [("ds", (3,)), ("dx",), ("ds", (1, 3))]
    assert fd.form(nv[0] * v * fd.ds).facet_degree == 2
    with pytest.raises(TypeError):                                   # the normal exists on facets only
        fd.form(nv[0] * v * fd.dx)
    with pytest.raises(ValueError):                                  # no side carries marker 7
        fd.form(v * fd.ds(7))
    with pytest.raises(TypeError):
        fd.ds("left")


def test_program_matches_reference_instruction_order():
    """The program of a*sin(b*x) emits CONST a, CONST b, X, MUL, SIN, MUL (forms.py:1870-1900)."""
    e = fd.Const(2.0) * fd.sin(fd.Const(3.0) * fd.x)
    from femd_amd.forms import Program
    code, nreg, consts = Program(e).arrays()
    assert code.tolist() == [[0, 0, 0, 0], [0, 1, 1, 0], [1, 2, 0, 0], [7, 3, 1, 2], [9, 4, 3, 0], [7, 5, 0, 4]]
    assert nreg == 6 and consts.tolist() == [2.0, 3.0]


def test_program_inputs_for_known_fields():
    """A known field emits INPUT with the slot of its (vector, derivative), slots in order of first
    use; the same node met again reuses its register, another node of the same slot gets its own
    INPUT of that slot (forms.py:1870-1880)."""
    from femd_amd.forms import Program
    m = fd.rectangle_mesh(0, 0, 1, 1, 3, 3)
    V = fd.LagrangeSpace(m, 2)
    w, z = fd.Function(V, "w"), fd.Function(V, "z")
    dw = fd.Dx(w, 0)
    e = w * w + dw * fd.Dx(w, 0) + z
    P = Program(e)
    code, nreg, consts = P.arrays()
    assert [c[0] for c in code.tolist()] == [5, 7, 5, 5, 7, 6, 5, 6]
    assert [c[2] for c in code.tolist() if c[0] == 5] == [0, 1, 1, 2]
    assert [(f.owner.name, f.deriv) for f in P.inputs] == [("w", (0, 0)), ("w", (1, 0)), ("z", (0, 0))]
    assert [f[2] for f in P.fields()] == [0, 1, 0]
    # a form with a known field is assembled anew each time; its degree counts the field's
    u, v = fd.TrialFunction(V), fd.TestFunction(V)
    a = fd.form(w * fd.dot(fd.grad(u), fd.grad(v)) * fd.dx)
    assert not a.constant and a.quad_degree == 4
    assert fd.form(fd.sin(w) * u * v * fd.dx).quad_degree == 4 + 3
    with pytest.raises(ValueError):
        fd.Function(V, "bad", np.zeros(V.dim + 1))
    with pytest.raises(NotImplementedError):
        fd.form(w * u * v * fd.ds)


def _interpret(code, consts, x, y):
    """A numpy interpreter of the instruction set (forms/integrand.hpp:300-342)."""
    reg = {}
    f = {9: np.sin, 10: np.cos, 11: np.exp, 12: np.log, 13: np.tanh, 14: np.sqrt, 15: np.sinh, 16: np.cosh,
         18: np.abs, 19: np.sign}
    res = 0
    for op, o, a, b in code:
        if op == 0: v = np.full_like(x, consts[a])
        elif op == 1: v = x.copy()
        elif op == 2: v = y.copy()
        elif op == 6: v = reg[a] + reg[b]
        elif op == 7: v = reg[a] * reg[b]
        elif op == 8: v = reg[a] ** consts[b]
        elif op == 17: t = np.exp(-np.abs(reg[a])); v = 2 * t / (1 + t * t)
        else: v = f[op](reg[a])
        reg[o] = v
        res = o
    return reg[res]


def test_register_compaction_keeps_the_arithmetic():
    from femd_amd.forms import Program
    rng = np.random.default_rng(0)
    x, y = rng.random(200), rng.random(200)
    f, _ = problems.fourier_source(0)
    kap, _ = problems.kappa_field(1)
    z = fd.x * fd.x + fd.sin(fd.x) * fd.sin(fd.x)
    for e in (f, kap, z, fd.Const(3.0)):
        code, nreg, consts = Program(e).arrays()
        c, keep = _lib.coefficient(prog=(code, nreg, consts))
        out = np.empty(code.size, dtype=np.int32)
        nr = ctypes.c_int()
        _lib.check(_lib.load().femd_program_compact(ctypes.byref(c), out, ctypes.byref(nr)))
        comp = out.reshape(-1, 4)
        assert nr.value <= min(nreg, 16)
        assert np.array_equal(_interpret(comp, consts, x, y), _interpret(code, consts, x, y))
        assert np.array_equal(comp[:, 0], code[:, 0])                      # same operations, same order


@pytest.mark.parametrize("k", [1, 2, 3])
def test_boundary_facets_vs_oracle(orc, k):
    """The library's ds point sets (host side of the facet set) bit for bit with the oracle's, and
    the Dirichlet-data helpers: boundary nodes, constrained nodes, the lift."""
    pts, q = _distorted_quads(5, 4, k)
    P, T, _ = problems.delaunay_square(0.2, seed=1)
    cases = [(fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 5, 4, diagonal="alternate"), k, dirichlet=[4]),
              orc.rectangle(5, 4, k, "alternate", dirichlet=[4], all_exterior=False)),
             (fd.LagrangeSpace(fd.quad_mesh_from_arrays(pts, q), k), orc.quad_from_arrays(pts, q, k=k, all_exterior=False)),
             (fd.LagrangeSpace(fd.mesh_from_arrays(P, T), k, bc="dirichlet"), orc.from_arrays(P, T, k=k))]
    for V, Po in cases:
        for mk in (None, [1], [2, 3]):
            for deg in (1, 4):
                for a, b in zip(V.boundary_points(deg, mk), orc.facet_points(Po, deg, mk)):
                    assert np.array_equal(a, b)
    V = cases[0][0]
    xy = V.raw_coordinates()
    assert np.array_equal(V.boundary_nodes(1), np.flatnonzero(xy[:, 1] == 0.0))
    assert np.array_equal(V.boundary_nodes(4), np.flatnonzero(xy[:, 0] == 0.0))
    assert np.array_equal(V.constrained(), V.boundary_nodes(4))
    g = V.lift(lambda x, y: 1 + y)
    assert np.array_equal(np.flatnonzero(g), V.constrained()) and np.array_equal(g[V.constrained()], 1 + xy[V.constrained(), 1])
    assert np.array_equal(V.lift({4: 2.5})[V.constrained()], np.full(V.n_constraints, 2.5))
    assert np.array_equal(V.lift(fd.x + 2 * fd.y)[V.constrained()], 2 * xy[V.constrained(), 1])
    U = V.unconstrained
    assert U.dim == V.raw_dim and U.n_constraints == 0 and U.unconstrained is U
    with pytest.raises(ValueError):
        V.lift({1: 1.0})                         # side 1 is natural here
    with pytest.raises(ValueError):
        U.lift(1.0)                              # nothing eliminated


def _first_met(ip, ix, segments):
    """The inner CG's order restated from its definition: per segment, the unknowns in the order a
    row-by-row scan of the segment's entries first meets them as columns, then the unmet ones."""
    perm = []
    for r0, r1 in zip(segments[:-1], segments[1:]):
        seen = []
        met = set()
        for j in ix[ip[r0]:ip[r1]]:
            if j not in met:
                met.add(int(j))
                seen.append(int(j))
        perm += seen + [j for j in range(r0, r1) if j not in met]
    return np.array(perm, dtype=np.int64)


def test_inner_order_pattern_first_met():
    """The host construction of the inner CG's order and of P A P^T: the permutation against its
    definition, the reordered pattern and the entry map against SciPy's permutation of a matrix
    whose values are the entry numbers; an isolated unknown goes last in its segment; a segment
    whose rows reach outside it is refused."""
    import scipy.sparse as sp
    lib = _lib.load()
    rng = np.random.default_rng(11)
    sizes = [7, 12, 1, 5, 9]
    blocks = []
    for k, m in enumerate(sizes):
        a = (rng.random((m, m)) < 0.3).astype(float)
        a = a + a.T + np.eye(m)
        if k == 3:
            a[2, :] = 0.0
            a[:, 2] = 0.0                     # unknown 2 of this segment: no entries at all
        blocks.append(sp.csr_matrix(a))
    S = sp.block_diag(blocks, format="csr")
    S.sort_indices()
    n, nnz = S.shape[0], S.nnz
    ip, ix = S.indptr.astype(np.int32), S.indices.astype(np.int32)
    segments = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    perm = np.empty(n, np.int32)
    ip2 = np.empty(n + 1, np.int32)
    ix2 = np.empty(max(nnz, 1), np.int32)
    emap = np.empty(max(nnz, 1), np.int32)
    _lib.check(lib.femd_inner_order_pattern(n, ip, ix, len(sizes), segments.ctypes.data, perm, ip2, ix2, emap))
    want = _first_met(ip, ix, segments)
    assert np.array_equal(perm, want)
    assert sorted(perm.tolist()) == list(range(n))
    iso = segments[3] + 2
    assert perm[segments[4] - 1] == iso
    E = sp.csr_matrix((np.arange(nnz, dtype=np.float64), ix, ip), shape=(n, n))
    Q = E[want][:, want].tocsr()
    Q.sort_indices()
    assert np.array_equal(ip2, Q.indptr) and np.array_equal(ix2[:nnz], Q.indices)
    assert np.array_equal(emap[:nnz], Q.data.astype(np.int64))
    # one segment (no segment list): the whole matrix
    _lib.check(lib.femd_inner_order_pattern(n, ip, ix, 0, None, perm, ip2, ix2, emap))
    assert np.array_equal(perm, _first_met(ip, ix, [0, n]))
    # rows of one segment with a column in another
    bad = S.tolil()
    bad[0, n - 1] = 1.0
    bad = bad.tocsr()
    bad.sort_indices()
    with pytest.raises(ValueError):
        _lib.check(lib.femd_inner_order_pattern(n, bad.indptr.astype(np.int32), bad.indices.astype(np.int32), len(sizes),
                                                segments.ctypes.data, perm, ip2, np.empty(bad.nnz, np.int32),
                                                np.empty(bad.nnz, np.int32)))


def test_dirichlet_data_kept_with_the_space():
    """LagrangeSpace(m, k, dirichlet={marker: data}) names the Dirichlet sides and keeps their data
    (spaces2d.py:115-119, 161-172): boundary_data holds it, lift() without data uses it, and the
    eliminated nodes of the marked side carry it; data= does the same; giving both, or data for a
    space with no Dirichlet side, is an error."""
    m = fd.rectangle_mesh(0, 0, 1, 1, 4, 4)
    V = fd.LagrangeSpace(m, 2, dirichlet={1: 5.0})
    assert V.dirichlet_markers == [1] and V.boundary_data == {1: 5.0}
    ug = V.lift()
    assert np.array_equal(ug, V.lift({1: 5.0}))
    cons = V.constrained()
    assert cons.size == 9 and np.all(ug[cons] == 5.0) and not np.delete(ug, cons).any()
    W = fd.LagrangeSpace(m, 2, dirichlet=[1, 3], data=lambda x, y: x + 2 * y)
    xy = W.raw_coordinates()
    assert np.array_equal(W.lift()[W.constrained()], (xy[:, 0] + 2 * xy[:, 1])[W.constrained()])
    with pytest.raises(TypeError):
        fd.LagrangeSpace(m, 2, dirichlet={1: 5.0}, data=1.0)
    with pytest.raises(ValueError):
        fd.LagrangeSpace(m, 2, data=1.0)
    with pytest.raises(ValueError):
        fd.LagrangeSpace(m, 2, dirichlet=[1], data={3: 1.0})


def _kappa_p3(orc, N):
    """A P3 matrix with a variable coefficient on an N x N grid (oracle assembly) and a right-hand side."""
    P = orc.rectangle(N, N, 3, "crossed")
    ip, ix = P.pattern()
    deg = 8
    xq, yq, _ = P.cache_points(deg)
    kap = np.exp(0.5 * np.sin(6 * xq) * np.cos(5 * yq))
    A = P.assemble_matrix(deg, [1, 2], [1, 2], np.vstack([kap, kap]))
    b = np.random.default_rng(N).standard_normal(ip.size - 1)
    return ip, ix, A, b


@pytest.mark.parametrize("kind", ["ilu0", "ssor"])
@pytest.mark.parametrize("ordering", ["natural", "rcm"])
def test_sweep_plan_matches_oracle(orc, kind, ordering):
    """The one-workgroup sweep program that femd_precond_create plans -- dependency levels cut into
    steps, LDS slots for the z values, the streamed ring and its producer schedule -- run on the host
    step by step (femd_sweep_host_apply) gives the oracle's z = M^{-1} r bit for bit, for ILU(0) and
    SSOR(1.3) in both orderings, on a P3 pattern whose sweeps need several steps per level."""
    import ctypes as C
    import scipy.sparse as sp
    ip, ix, A, r = _kappa_p3(orc, 14)
    n = ip.size - 1
    omega = 1.3 if kind == "ssor" else 1.0
    want, _ = orc.precond_apply(ip, ix, A, r, kind, ordering, omega)
    if ordering == "rcm":
        perm = orc.rcm(ip, ix)
    else:
        perm = np.arange(n, dtype=np.int32)
    if kind == "ilu0":
        _, ip2, ix2, vals, _ = orc.ilu0_factor(ip, ix, A, ordering)
    else:
        B = sp.csr_matrix((A, ix, ip), shape=(n, n))[perm][:, perm].tocsr()
        B.sort_indices()
        ip2, ix2, vals = B.indptr.astype(np.int32), B.indices.astype(np.int32), B.data.copy()
    zp = np.empty(n)
    stats = np.zeros(6, dtype=np.int32)
    lib = _lib.load()
    _lib.check(lib.femd_sweep_host_apply(n, _lib.i32(ip2), _lib.i32(ix2), np.ascontiguousarray(vals),
                                         _lib.PC_ILU0 if kind == "ilu0" else _lib.PC_SSOR, omega,
                                         np.ascontiguousarray(r[perm]), zp, stats))
    got = np.empty(n)
    got[perm] = zp
    assert np.array_equal(got, want)
    assert stats[0] >= stats[2] > 1 and stats[1] >= stats[3] > 1


def test_sweep_plan_refuses_bad_input():
    lib = _lib.load()
    ip = np.array([0, 1, 2], dtype=np.int32)
    ix = np.array([1, 0], dtype=np.int32)                    # no diagonal
    z = np.empty(2)
    st = np.zeros(6, dtype=np.int32)
    rc = lib.femd_sweep_host_apply(2, ip, ix, np.ones(2), _lib.PC_ILU0, 1.0, np.ones(2), z, st)
    assert rc == _lib.EZEROPIVOT
    rc = lib.femd_sweep_host_apply(2, ip, ix, np.ones(2), _lib.PC_JACOBI, 1.0, np.ones(2), z, st)
    assert rc == _lib.EINVAL
