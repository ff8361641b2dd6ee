"""Known fields inside coefficients (fd.Function), on the GPU against the oracle.

A field is evaluated at the quadrature points in the order of Cache2D::eval_multi
(fe/cache_2d.hpp:245-288) and read by the program's INPUT instructions (forms/integrand.hpp:39):
the field values are bit for bit the oracle's, and so is every matrix whose coefficient programs
are arithmetic (the GPU's sin / cos / exp are not glibc's: within 1e-14 there).  A Picard
iteration for -div((1 + w^2) grad u) = f, each step assembling with the previous solution,
follows the oracle's bit for bit."""
import numpy as np
import pytest

import femd_amd as fd
from femd_amd import _lib as L
from femd_amd import problems

pytestmark = pytest.mark.gpu


def _name(s):
    print(f"\n[case] {s}", flush=True)


def _field_values(V, deg, W, vec, code):
    """The field (vec on space W) at V's points through a one-instruction INPUT program."""
    prog = (np.array([[5, 0, 0, 0]], dtype=np.int32), 1, np.zeros(0))
    arr, keep = L.coefficient_array([(prog, None, [(W._h, vec, code)])])
    out = np.empty(V.ncells * _nq(V, deg))
    L.check(L.load().femd_eval_coefficient(V._h, deg, arr, out))
    return out


def _nq(V, deg):
    import ctypes as C
    n = C.c_int()
    b = [np.zeros(512) for _ in range(3)]
    fn = L.load().femd_quadrature_square if V.nverts == 4 else L.load().femd_quadrature
    L.check(fn(deg, C.byref(n), *b))
    return n.value


def _oracle_coefficients(form, P, fields):
    """Each term's coefficient at the points, computed by the oracle: its fields by eval_field on
    their own oracle problems (fields: Function -> oracle problem), then the program with inputs."""
    from femd_amd.forms import Program
    deg = form.quad_degree
    x, y, _ = P.cache_points(deg)
    out = []
    for prog in form.programs:
        code, nreg, consts = prog.arrays()
        inp = [fields[f.owner].eval_field(deg, f.owner.vector, [f.code()])[0] for f in prog.inputs]
        out.append(P.o.integrand_eval_in(code, nreg, consts, x, y, np.array(inp)) if inp
                   else P.o.integrand_eval(code, nreg, consts, x, y))
    return np.stack(out)


@pytest.mark.parametrize("k,kw,bc", [(1, 1, "free"), (2, 2, "dirichlet"), (3, 1, "dirichlet"), (2, 3, "free")])
def test_field_values_bit_exact(orc, k, kw, bc):
    _name(f"field values: P{kw} field at P{k} points, {bc}")
    for mesh_kind in ("structured", "unstructured"):
        if mesh_kind == "structured":
            m = fd.rectangle_mesh(0, 0, 1, 1, 6, 5, diagonal="alternate")
            P, Pw = orc.rectangle(6, 5, k, "alternate", all_exterior=bc == "dirichlet"), \
                orc.rectangle(6, 5, kw, "alternate", all_exterior=bc == "dirichlet")
        else:
            Pts, T, _ = problems.delaunay_square(0.12, seed=k + kw)
            m = fd.mesh_from_arrays(Pts, T)
            P = orc.from_arrays(Pts, T, k=k, all_exterior=bc == "dirichlet")
            Pw = orc.from_arrays(Pts, T, k=kw, all_exterior=bc == "dirichlet")
        V = fd.LagrangeSpace(m, k, bc=bc)
        W = fd.LagrangeSpace(m, kw, bc=bc)
        c = np.random.default_rng(k * 7 + kw).standard_normal(W.dim)
        deg = 2 * k + 1
        want = Pw.eval_field(deg, c, [0, 1, 2])
        for code in (0, 1, 2):
            got = _field_values(V, deg, W, c, code)
            assert np.array_equal(got, want[code]), (mesh_kind, code)


def test_field_values_quads_bit_exact(orc):
    _name("field values: Q2 on a distorted quad mesh")
    rng = np.random.default_rng(4)
    nx, ny = 6, 5
    pts = np.array([(i / nx + (0 if i in (0, nx) else 0.2 * (rng.random() - 0.5) / nx),
                     j / ny + (0 if j in (0, ny) else 0.2 * (rng.random() - 0.5) / ny))
                    for j in range(ny + 1) for i in range(nx + 1)])
    q = np.array([[j * (nx + 1) + i, j * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i]
                  for j in range(ny) for i in range(nx)], dtype=np.int32)
    V = fd.LagrangeSpace(fd.quad_mesh_from_arrays(pts, q), 2)
    P = orc.quad_from_arrays(pts, q, k=2, all_exterior=False)
    c = rng.standard_normal(V.dim)
    want = P.eval_field(5, c, [0, 1, 2])
    for code in (0, 1, 2):
        assert np.array_equal(_field_values(V, 5, V, c, code), want[code])


def test_nonlinear_forms_against_oracle(orc):
    """(1 + w^2) grad u . grad v + (dw/dx) u v (arithmetic: bit for bit) and sin(w) u v, f(w) v
    (transcendental: within 1e-14), w a P1 field in a P2 form on an unstructured mesh."""
    _name("forms with known fields")
    Pts, T, _ = problems.delaunay_square(0.08, seed=7)
    m = fd.mesh_from_arrays(Pts, T)
    V, W = fd.LagrangeSpace(m, 2, bc="dirichlet"), fd.LagrangeSpace(m, 1)
    P, Pw = orc.from_arrays(Pts, T, k=2), orc.from_arrays(Pts, T, k=1, all_exterior=False)
    u, v = fd.TrialFunction(V), fd.TestFunction(V)
    w = fd.Function(W, "w", np.random.default_rng(0).standard_normal(W.dim))
    fields = {w: Pw}
    a = fd.form((1 + w * w) * fd.dot(fd.grad(u), fd.grad(v)) * fd.dx + fd.Dx(w, 0) * u * v * fd.dx)
    A = a.assemble()
    co = _oracle_coefficients(a, P, fields)
    terms = [(t.test.code(), t.trial.code()) for t in a.terms]
    Ao = P.assemble_matrix(a.quad_degree, [t[0] for t in terms], [t[1] for t in terms], co)
    assert np.array_equal(A.data, Ao)
    s = fd.form(fd.sin(w) * u * v * fd.dx)
    S = s.assemble()
    So = P.assemble_matrix(s.quad_degree, [0], [0], _oracle_coefficients(s, P, fields))
    assert np.abs(S.data - So).max() <= 1e-14 * np.abs(So).max()
    L_ = fd.form(fd.exp(w) * fd.Dx(w, 1) * v * fd.dx)
    b = L_.assemble()
    bo = P.assemble_vector(L_.quad_degree, [0], _oracle_coefficients(L_, P, fields))
    assert np.abs(b - bo).max() <= 1e-14 * np.abs(bo).max()
    # the vector is read at every assembly
    w.vector[:] = 0.0
    assert np.array_equal(a.assemble().data, P.assemble_matrix(a.quad_degree, [1, 2, 0], [1, 2, 0],
                                                              _oracle_coefficients(a, P, fields)))


def test_picard_iteration_bit_exact(orc):
    """-div((1 + 4 w^2) grad u) = 10 with u = 0 on the boundary, w the previous iterate: five Picard
    steps, each assembling with w on the GPU and solving with Jacobi-CG, against the oracle doing
    the same; assembly, iteration counts and iterates bit for bit."""
    _name("Picard iteration with a known field")
    m = fd.rectangle_mesh(0, 0, 1, 1, 16, 16)
    V = fd.LagrangeSpace(m, 2, bc="dirichlet")
    P = orc.rectangle(16, 16, 2)
    ip, ix = P.pattern()
    u, v = fd.TrialFunction(V), fd.TestFunction(V)
    w = fd.Function(V, "w")
    a = fd.form((1 + 4.0 * w * w) * fd.dot(fd.grad(u), fd.grad(v)) * fd.dx)
    b = fd.form(10.0 * v * fd.dx).assemble()
    wo = np.zeros(V.dim)
    for it in range(5):
        A = a.assemble()
        x, y, _ = P.cache_points(a.quad_degree)
        code, nreg, consts = a.programs[0].arrays()
        co = P.o.integrand_eval_in(code, nreg, consts, x, y, P.eval_field(a.quad_degree, wo, [0]))
        Ao = P.assemble_matrix(a.quad_degree, [1, 2], [1, 2], np.stack([co, co]))
        assert np.array_equal(A.data, Ao), it
        xg, info = fd.cg(A, b, M=A.preconditioner("jacobi"), tol=1e-12)
        xo, io = orc.cg(ip, ix, Ao, b, M="jacobi", tol=1e-12)
        assert info.iterations == io["iterations"] and np.array_equal(xg, xo), it
        change = np.linalg.norm(xg - w.vector) / np.linalg.norm(xg)
        print(f"  step {it}: {info.iterations} CG iterations, change {change:.2e}")
        w.vector = xg
        wo = xo
    assert change < 0.05
