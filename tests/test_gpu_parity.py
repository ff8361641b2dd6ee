"""Parity of the HIP path with the oracle, through the C ABI, on the GPU.

Bit for bit: element matrices and vectors summed in the reference's element order, the SpMV,
the ordered dot products, every CG variant (iterates, iteration counts, residuals), the
Jacobi / SSOR / ILU(0) applications and factors.  Within tolerances: coefficient programs
(the GPU's sin / cos / exp are not glibc's: 4 ulp), GMRES / LGMRES (their dot products are tree
reductions; the reference's own are OpenMP reductions that change bits with the thread count),
and the FP32-inner refined solve (1e-8 against the FP64 oracle solution)."""
import ctypes as C
import math
import warnings

import numpy as np
import pytest

import femd_amd as fd
from femd_amd import _lib as L
from femd_amd import problems

pytestmark = pytest.mark.gpu


def _name(s):
    print(f"\n[case] {s}", flush=True)


def _oracle_c0(orc, n=32, k=2, diag="right"):
    P = orc.rectangle(n, n, k, diag)
    ip, ix = P.pattern()
    x, y, w = P.cache_points(4)
    A = P.assemble_matrix(4, [1, 2, 0], [1, 2, 0], np.ones((3, x.size)))
    f, _ = problems.fourier_source(0)
    code, nreg, consts = fd.forms.Program(f).arrays()
    xb, yb, _ = P.cache_points(5)
    cb = orc.integrand_eval(code, nreg, consts, xb, yb)
    b = P.assemble_vector(5, [0], cb[None, :])
    return P, ip, ix, A, b, cb


def _assemble_values(V, deg, a, b, coeffs):
    """Assemble through the C ABI with host coefficient arrays (one per term)."""
    arr, keep = L.coefficient_array([(None, c) for c in coeffs])
    K = fd.SparseMatrix._on_space(V)
    L.check(L.load().femd_assemble_matrix(V._h, deg, len(a), L.i32(a), L.i32(b), arr, K._h))
    return K


def _vector_values(V, deg, a, coeffs):
    arr, keep = L.coefficient_array([(None, c) for c in coeffs])
    f = np.empty(V.dim)
    L.check(L.load().femd_assemble_vector(V._h, deg, len(a), L.i32(a), arr, f))
    return f


@pytest.mark.parametrize("k,deg,diag", [(1, 2, "right"), (2, 4, "alternate"), (3, 8, "crossed"), (4, 9, "left"), (5, 10, "right")])
def test_assembly_bit_exact(orc, k, deg, diag):
    _name(f"assembly k={k} deg={deg} diag={diag}")
    m = fd.rectangle_mesh(0, 0, 1, 1, 7, 6, diagonal=diag)
    V = fd.LagrangeSpace(m, k, bc="dirichlet")
    P = orc.rectangle(7, 6, k, diag)
    npt = P.ncells * len(orc.quadrature(deg)[2])
    rng = np.random.default_rng(k)
    c = rng.standard_normal((4, npt))
    c[2, ::5] = 0.0
    a, b = [1, 2, 0, 1], [1, 2, 0, 2]
    K = _assemble_values(V, deg, a, b, c)
    assert np.array_equal(K.data, P.assemble_matrix(deg, a, b, c))
    assert np.array_equal(_vector_values(V, deg, [0, 1, 2], c[:3]), P.assemble_vector(deg, [0, 1, 2], c[:3]))


def test_form_assembly_c0(orc):
    _name("form assembly C0 (device coefficient programs)")
    m = fd.rectangle_mesh(0, 0, 1, 1, 32, 32)
    V = fd.LagrangeSpace(m, 2, bc="dirichlet")
    u, v = fd.TrialFunction(V), fd.TestFunction(V)
    A = fd.form((fd.dot(fd.grad(u), fd.grad(v)) + u * v) * fd.dx).assemble()
    f, _ = problems.fourier_source(0)
    b = fd.form(f * v * fd.dx).assemble()
    P, ip, ix, Ao, bo, cb = _oracle_c0(orc)
    assert np.array_equal(A.indptr, ip) and np.array_equal(A.indices, ix)
    assert np.array_equal(A.data, Ao)                        # constant coefficients: bit for bit
    assert A.symmetric
    assert np.abs(b - bo).max() <= 1e-14 * np.abs(bo).max()
    # the device program against the oracle's libm at the same points: a few ulp
    cg = np.empty_like(cb)
    code, nreg, consts = fd.forms.Program(f).arrays()
    arr, keep = L.coefficient_array([((code, nreg, consts), None)])
    L.check(L.load().femd_eval_coefficient(V._h, 5, arr, cg))
    assert np.abs(cg - cb).max() <= 4e-15 * max(1.0, np.abs(cb).max())


def test_kappa_program_p3(orc):
    _name("kappa program P3")
    V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 6, 6), 3, bc="dirichlet")
    P = orc.rectangle(6, 6, 3)
    kap, _ = problems.kappa_field(1)
    code, nreg, consts = fd.forms.Program(kap).arrays()
    x, y, _ = P.cache_points(8)
    co = orc.integrand_eval(code, nreg, consts, x, y)
    cg = np.empty_like(co)
    arr, keep = L.coefficient_array([((code, nreg, consts), None)])
    L.check(L.load().femd_eval_coefficient(V._h, 8, arr, cg))
    assert np.abs(cg / co - 1).max() <= 1e-14


def test_spmv_and_dot_bit_exact(orc):
    _name("spmv + ddot")
    P, ip, ix, A, b, _ = _oracle_c0(orc, 40, 3)
    M = fd.SparseMatrix.from_arrays(ip.size - 1, ip.size - 1, ip, ix, A)
    rng = np.random.default_rng(0)
    for _ in range(3):
        x = rng.standard_normal(ip.size - 1)
        assert np.array_equal(M.matvec(x), orc.spmv(ip, ix, A, x))
    lib = L.load()
    for n in (1, 7, 8, 9, 4095, 4096, 4097, 8191, 12289, 100003, 1 << 20):
        u, v = rng.standard_normal(n), rng.standard_normal(n)
        out = C.c_double()
        L.check(lib.femd_ddot(n, u, v, C.byref(out)))
        assert out.value == orc.ddot(u, v), n


def test_spmv_long_rows_and_empty_rows(orc):
    _name("spmv long/empty rows")
    rng = np.random.default_rng(5)
    n = 300
    rows = []
    for i in range(n):
        if i % 50 == 0:
            cols = np.arange(n)                                 # dense rows
        elif i % 7 == 0:
            cols = np.array([], dtype=int)                      # empty rows
        else:
            cols = np.unique(rng.integers(0, n, 12))
        rows.append(cols)
    ip = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    ix = np.concatenate(rows).astype(np.int32)
    v = rng.standard_normal(ix.size)
    # a row longer than the staging buffer
    big = 5000
    ip2 = np.concatenate([ip, [ip[-1] + big]]).astype(np.int32)
    ix2 = np.concatenate([ix, np.sort(rng.choice(n + 1, size=min(big, n + 1), replace=False))[:big]]).astype(np.int32)
    ip2[-1] = ip[-1] + (ix2.size - ix.size)
    v2 = rng.standard_normal(ix2.size)
    M = fd.SparseMatrix.from_arrays(n + 1, n + 1, ip2, ix2, v2)
    x = rng.standard_normal(n + 1)
    assert np.array_equal(M.matvec(x), orc.spmv(ip2, ix2, v2, x))


@pytest.mark.parametrize("M", [None, "jacobi", "ilu0", "ssor"])
@pytest.mark.parametrize("ordering", ["natural", "rcm"])
def test_cg_bit_exact_c0(orc, M, ordering):
    _name(f"cg C0 M={M} ordering={ordering}")
    P, ip, ix, A, b, _ = _oracle_c0(orc)
    Ag = fd.SparseMatrix.from_arrays(ip.size - 1, ip.size - 1, ip, ix, A, symmetric=True)
    pc = None if M is None else Ag.preconditioner(M, omega=1.0, ordering=ordering)
    x, info = fd.cg(Ag, b, M=pc)
    xo, io = orc.cg(ip, ix, A, b, M=M, ordering=ordering)
    assert info.iterations == io["iterations"] and info.converged == io["converged"]
    assert info.residual == io["residual"]
    assert np.array_equal(x, xo)


def test_cg_unstructured_p3_ilu_rcm(orc):
    _name("cg unstructured P3 ilu0 rcm + kappa")
    Pts, T, _ = problems.delaunay_square(0.08, seed=2)
    m = fd.mesh_from_arrays(Pts, T)
    V = fd.LagrangeSpace(m, 3, bc="dirichlet")
    Po = orc.from_arrays(Pts, T, k=3)
    kap, _ = problems.kappa_field(1)
    code, nreg, consts = fd.forms.Program(kap).arrays()
    x, y, _ = Po.cache_points(8)
    ck = orc.integrand_eval(code, nreg, consts, x, y)
    ip, ix = Po.pattern()
    Ao = Po.assemble_matrix(8, [1, 2], [1, 2], np.stack([ck, ck]))
    K = _assemble_values(V, 8, [1, 2], [1, 2], [ck, ck])
    assert np.array_equal(K.data, Ao)
    f, _ = problems.fourier_source(0)
    cf = orc.integrand_eval(*fd.forms.Program(f).arrays(), *Po.cache_points(7)[:2])
    b = Po.assemble_vector(7, [0], cf[None, :])
    pc = K.preconditioner("ilu0", ordering="rcm")
    xg, info = fd.cg(K, b, M=pc)
    xo, io = orc.cg(ip, ix, Ao, b, M="ilu0", ordering="rcm")
    assert info.iterations == io["iterations"] and np.array_equal(xg, xo)
    _, mp = orc.precond_apply(ip, ix, Ao, b, "ilu0", "rcm")
    assert pc.min_pivot == mp


@pytest.mark.parametrize("kind,ordering", [("jacobi", "natural"), ("ilu0", "natural"), ("ilu0", "rcm"), ("ssor", "rcm")])
def test_precond_apply_bit_exact(orc, kind, ordering):
    _name(f"precond {kind} {ordering}")
    P, ip, ix, A, b, _ = _oracle_c0(orc, 20, 3)
    Ag = fd.SparseMatrix.from_arrays(ip.size - 1, ip.size - 1, ip, ix, A)
    om = 1.4 if kind == "ssor" else 1.0
    z = Ag.preconditioner(kind, omega=om, ordering=ordering).solve(b)
    zo, _ = orc.precond_apply(ip, ix, A, b, kind, ordering, om)
    assert np.array_equal(z, zo)


@pytest.mark.parametrize("kind,ordering", [("ilu0", "rcm"), ("ilu0", "natural"), ("ssor", "rcm"), ("ssor", "natural")])
def test_planned_sweeps_bit_exact_unstructured_p3(orc, kind, ordering):
    """The one-workgroup planned sweeps on a P3 kappa matrix of an unstructured mesh (~26k rows,
    about a thousand dependency levels, several steps per level): the application, and a whole CG
    solve (iterates and iteration count), equal the oracle's bit for bit."""
    _name(f"planned sweeps P3 unstructured {kind} {ordering}")
    Pts, T, _ = problems.delaunay_square(0.02, seed=4)
    Po = orc.from_arrays(Pts, T, k=3)
    kap, _ = problems.kappa_field(1)
    x, y, _ = Po.cache_points(8)
    ck = orc.integrand_eval(*fd.forms.Program(kap).arrays(), x, y)
    ip, ix = Po.pattern()
    Ao = Po.assemble_matrix(8, [1, 2], [1, 2], np.stack([ck, ck]))
    n = ip.size - 1
    Ag = fd.SparseMatrix.from_arrays(n, n, ip, ix, Ao, symmetric=True)
    om = 1.2 if kind == "ssor" else 1.0
    pc = Ag.preconditioner(kind, omega=om, ordering=ordering)
    st = pc.stats()
    assert st["planned"] and st["levels_forward"] > 100 and st["steps_forward"] >= st["levels_forward"], st
    r = np.random.default_rng(5).standard_normal(n)
    zo, _ = orc.precond_apply(ip, ix, Ao, r, kind, ordering, om)
    assert np.array_equal(pc.solve(r), zo)
    assert np.array_equal(pc.solve(2.0 * r), orc.precond_apply(ip, ix, Ao, 2.0 * r, kind, ordering, om)[0])
    b = r[::-1].copy()
    xg, info = fd.cg(Ag, b, M=pc, tol=1e-9)
    xo, io = orc.cg(ip, ix, Ao, b, M=kind, ordering=ordering, omega=om, tol=1e-9)
    assert info.iterations == io["iterations"] and info.residual == io["residual"]
    assert np.array_equal(xg, xo)


def test_ilu_factor_bit_exact(orc):
    _name("ilu factor values")
    P, ip, ix, A, b, _ = _oracle_c0(orc, 16, 2)
    perm, ip2, ix2, lu, mp = orc.ilu0_factor(ip, ix, A, "rcm")
    Ag = fd.SparseMatrix.from_arrays(ip.size - 1, ip.size - 1, ip, ix, A)
    assert np.array_equal(Ag.ordering("rcm"), perm)
    assert Ag.preconditioner("ilu0", ordering="rcm").min_pivot == mp


def _nonsymmetric_c0(orc, n):
    P, ip, ix, A, b, _ = _oracle_c0(orc, n)
    rows = np.repeat(np.arange(ip.size - 1), np.diff(ip))
    A2 = A + 0.05 * np.where(ix > rows, 1.0, np.where(ix < rows, -1.0, 0.0))
    return ip, ix, A2, b


@pytest.mark.parametrize("lgmres", [False, True])
@pytest.mark.parametrize("M", [None, "jacobi", "ilu0", "ssor"])
def test_gmres_bit_exact_below_threshold(orc, lgmres, M):
    """Below 4096 unknowns the reference runs GMRES / LGMRES on one thread (krylov/adapters.hpp:50-68):
    the GPU path's serial sums give its iterates bit for bit -- the same iteration count, solution,
    true and preconditioned residuals -- over several restart cycles, and when max_iter stops it."""
    _name(f"gmres bit-exact lgmres={lgmres} M={M}")
    ip, ix, A2, b = _nonsymmetric_c0(orc, 14)
    Ag = fd.SparseMatrix.from_arrays(ip.size - 1, ip.size - 1, ip, ix, A2)
    om = 1.3 if M == "ssor" else 1.0
    pc = None if M is None else Ag.preconditioner(M, omega=om)
    f = fd.lgmres if lgmres else fd.gmres
    for restart, maxiter in ((20, None), (7, 23)):
        kw = {} if maxiter is None else dict(maxiter=maxiter)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            x, info = f(Ag, b, M=pc, restart=restart, **kw)
        xo, io = orc.krylov(ip, ix, A2, b, M=M, omega=om, lgmres=lgmres, restart=restart, **({} if maxiter is None else dict(maxiter=maxiter)))
        assert (info.converged, info.iterations) == (io["converged"], io["iterations"]), (restart, info, io)
        assert info.residual == io["residual"] and info.precond_residual == io["precond_residual"]
        assert np.array_equal(x, xo)
    assert io["iterations"] == 23 and not io["converged"]


def test_gmres_breakdown_bit_exact(orc):
    """An exact Krylov space: on 2 I the first Arnoldi vector spans it (h_10 = 0), GMRES returns after
    one step with the unrotated column (csnewton.hpp:211-214), LGMRES ends its cycle there."""
    _name("gmres breakdown")
    n = 50
    ip = np.arange(n + 1, dtype=np.int32)
    ix = np.arange(n, dtype=np.int32)
    v = np.full(n, 2.0)
    b = np.linspace(1.0, 2.0, n)
    Ag = fd.SparseMatrix.from_arrays(n, n, ip, ix, v)
    for lg, f in ((False, fd.gmres), (True, fd.lgmres)):
        x, info = f(Ag, b, M=None, restart=20)
        xo, io = orc.krylov(ip, ix, v, b, M=None, lgmres=lg, restart=20)
        assert (info.converged, info.iterations) == (io["converged"], io["iterations"]) and np.array_equal(x, xo)


@pytest.mark.parametrize("lgmres", [False, True])
def test_gmres_above_threshold(orc, lgmres):
    """Above 4096 unknowns the reference's dot products are OpenMP reductions whose bits depend on the
    thread count; the GPU's are fixed-order tree sums.  Solved to tol 1e-14, the GPU solution is
    within 1e-12 of the oracle's (relative), with the iteration count within 2%."""
    _name(f"gmres above the threshold lgmres={lgmres}")
    ip, ix, A2, b = _nonsymmetric_c0(orc, 36)
    n = ip.size - 1
    assert n >= 4096
    Ag = fd.SparseMatrix.from_arrays(n, n, ip, ix, A2)
    pc = Ag.preconditioner("ilu0")
    f = fd.lgmres if lgmres else fd.gmres
    x, info = f(Ag, b, M=pc, restart=30, tol=1e-14)
    xo, io = orc.krylov(ip, ix, A2, b, M="ilu0", lgmres=lgmres, restart=30, tol=1e-14)
    assert info.converged and io["converged"]
    assert abs(info.iterations - io["iterations"]) <= max(2, io["iterations"] // 50), (info, io)
    assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.norm(xo)
    assert info.precond_residual <= 1e-14 * (1 + 1e-8)


def test_refined_fp32_reaches_1e8(orc):
    _name("refined fp32 P3 unstructured")
    Pts, T, _ = problems.delaunay_square(0.05, seed=3)
    m = fd.mesh_from_arrays(Pts, T)
    V = fd.LagrangeSpace(m, 3, bc="dirichlet")
    kap, _ = problems.kappa_field(1)
    f, _ = problems.fourier_source(0)
    a, Lf = problems.poisson_forms(V, f, kappa=kap)
    A, b = a.assemble(), Lf.assemble()
    x, info = fd.solve_refined(A, b, tol=1e-10, inner_tol=1e-4)
    assert info.converged
    ip, ix = V.pattern()
    xo, io = orc.cg(ip, ix, A.data, b, M="ilu0", tol=1e-13)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-8


def test_breakdown_and_errors(orc):
    _name("cg breakdown + zero diagonal + maxiter")
    n = 10
    ip = np.arange(n + 1, dtype=np.int32)
    ix = np.arange(n, dtype=np.int32)
    v = np.ones(n); v[3] = -1.0
    A = fd.SparseMatrix.from_arrays(n, n, ip, ix, v)
    b = np.ones(n)
    with pytest.warns(RuntimeWarning):
        x, info = fd.cg(A, b)
    xo, io = orc.cg(ip, ix, v, b)
    assert info.breakdown == io["breakdown"] and info.iterations == io["iterations"]
    v0 = v.copy(); v0[5] = 0.0
    Z = fd.SparseMatrix.from_arrays(n, n, ip, ix, v0)
    with pytest.raises(ValueError):
        Z.preconditioner("jacobi")
    with pytest.raises(ValueError):
        Z.preconditioner("ilu0")
    P, ipc, ixc, Ac, bc, _ = _oracle_c0(orc, 12)
    Ag = fd.SparseMatrix.from_arrays(ipc.size - 1, ipc.size - 1, ipc, ixc, Ac)
    with pytest.warns(RuntimeWarning):
        x, info = fd.cg(Ag, bc, maxiter=3)
    xo, io = orc.cg(ipc, ixc, Ac, bc, maxiter=3)
    assert not info.converged and info.iterations == 3 and np.array_equal(x, xo)
    x, info = fd.cg(Ag, np.zeros_like(bc))
    assert info.converged and not x.any()


# ------------------------------------------------------------------- Q_k on quadrilaterals
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


@pytest.mark.parametrize("k,deg", [(1, 2), (2, 4), (3, 7), (4, 9)])
def test_quad_assembly_bit_exact(orc, k, deg):
    _name(f"Q{k} assembly deg={deg}, rectangle and distorted meshes")
    pts, q = _distorted_quads(7, 6, k)
    cases = [(fd.LagrangeSpace(fd.rectangle_quad_mesh(0, 0, 1, 1, 7, 6), k, bc="dirichlet"), orc.quad_rectangle(7, 6, k)),
             (fd.LagrangeSpace(fd.quad_mesh_from_arrays(pts, q), k), orc.quad_from_arrays(pts, q, k=k, all_exterior=False))]
    for V, P in cases:
        npt = P.ncells * len(orc.quad_quadrature(deg)[2])
        rng = np.random.default_rng(k)
        c = rng.standard_normal((4, npt))
        c[2, ::5] = 0.0
        a, b = [1, 2, 0, 1], [1, 2, 0, 2]
        K = _assemble_values(V, deg, a, b, c)
        assert np.array_equal(K.data, P.assemble_matrix(deg, a, b, c))
        assert np.array_equal(_vector_values(V, deg, [0, 1, 2], c[:3]), P.assemble_vector(deg, [0, 1, 2], c[:3]))


def test_quad_forms_and_solve(orc):
    _name("Q2 forms with device coefficients on a distorted mesh, then Jacobi-CG")
    pts, q = _distorted_quads(12, 10, 5)
    V = fd.LagrangeSpace(fd.quad_mesh_from_arrays(pts, q), 2, bc="dirichlet")
    P = orc.quad_from_arrays(pts, q, k=2)
    kap, _ = problems.kappa_field(1)
    f, _ = problems.fourier_source(3)
    a, L_ = problems.poisson_forms(V, f, kappa=kap)
    A, b = a.assemble(), L_.assemble()
    code, nreg, consts = fd.forms.Program(kap).arrays()
    x, y, _ = P.cache_points(a.quad_degree)
    ck = orc.integrand_eval(code, nreg, consts, x, y)
    Ao = P.assemble_matrix(a.quad_degree, [1, 2], [1, 2], np.stack([ck, ck]))
    assert np.abs(A.data - Ao).max() <= 1e-13 * np.abs(Ao).max()
    # fed the oracle's coefficient values the assembly is bit for bit
    K = _assemble_values(V, a.quad_degree, [1, 2], [1, 2], [ck, ck])
    assert np.array_equal(K.data, Ao)
    ip, ix = V.pattern()
    xg, info = fd.cg(K, b, M=K.preconditioner("jacobi"), tol=1e-10)
    xo, io = orc.cg(ip, ix, Ao, b, M="jacobi", tol=1e-10)
    assert info.iterations == io["iterations"] and np.array_equal(xg, xo)


@pytest.mark.parametrize("k,quad", [(3, False), (5, False), (3, True), (4, True)])
def test_lobatto_assembly_bit_exact(orc, k, quad):
    _name(f"nodes=lobatto assembly k={k} quad={quad}")
    if quad:
        pts, q = _distorted_quads(6, 5, k)
        V = fd.LagrangeSpace(fd.quad_mesh_from_arrays(pts, q), k, bc="dirichlet", nodes="lobatto")
        P = orc.quad_from_arrays(pts, q, k=k, equispaced=False)
        rule = orc.quad_quadrature
    else:
        V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 6, 5, diagonal="alternate"), k, bc="dirichlet", nodes="lobatto")
        P = orc.rectangle(6, 5, k, "alternate", equispaced=False)
        rule = orc.quadrature
    deg = 2 * k
    c = np.random.default_rng(k).standard_normal((3, P.ncells * len(rule(deg)[2])))
    a, b = [1, 2, 0], [1, 2, 0]
    K = _assemble_values(V, deg, a, b, c)
    assert np.array_equal(K.data, P.assemble_matrix(deg, a, b, c))
    assert np.array_equal(_vector_values(V, deg, [0, 1], c[:2]), P.assemble_vector(deg, [0, 1], c[:2]))


# ------------------------------------------------------------------- boundary facets and lifts
def _facet_set(V, deg, markers=None):
    return V.facet_set(deg, markers)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_facet_groups_bit_exact(orc, k):
    """A form's ds and dx groups through the C ABI, added in the form's order: bit for bit with the
    oracle (pinned to the reference's facet cache and group combination)."""
    _name(f"ds/dx groups k={k}")
    lib = L.load()
    pts, q = _distorted_quads(5, 4, k)
    cases = [(fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 5, 4, diagonal="alternate"), k, dirichlet=[4]),
              orc.rectangle(5, 4, k, "alternate", dirichlet=[4], all_exterior=False), orc.quadrature),
             (fd.LagrangeSpace(fd.quad_mesh_from_arrays(pts, q), k), orc.quad_from_arrays(pts, q, k=k, all_exterior=False),
              orc.quad_quadrature)]
    rng = np.random.default_rng(k)
    for V, P, rule in cases:
        npf = lambda d, mk: orc.facet_points(P, d, mk)[0].size
        gs = ("ds", 2 * k + 1, [1, 3], [0, 1], [0, 0], rng.standard_normal((2, npf(2 * k + 1, [1, 3]))))
        gd = ("dx", 2 * k, None, [1, 2, 0], [1, 2, 0], rng.standard_normal((3, P.ncells * len(rule(2 * k)[2]))))
        ga = ("ds", 2, None, [0], [0], rng.standard_normal((1, npf(2, None))))
        gs[5][0, ::4] = 0.0
        for order in ([gs, gd, ga], [gd, gs, ga]):
            K = fd.SparseMatrix._on_space(V)
            f = np.empty(V.dim)
            for g, (kind, deg, mk, a, b, c) in enumerate(order):
                arr, keep = L.coefficient_array([(None, row) for row in c])
                if kind == "dx":
                    fm = lib.femd_assemble_matrix if g == 0 else lib.femd_assemble_matrix_add
                    fv = lib.femd_assemble_vector if g == 0 else lib.femd_assemble_vector_add
                    L.check(fm(V._h, deg, len(a), L.i32(a), L.i32(b), arr, K._h))
                    L.check(fv(V._h, deg, len(a), L.i32(a), arr, f))
                else:
                    F = _facet_set(V, deg, mk)
                    L.check(lib.femd_assemble_matrix_facets(F, len(a), L.i32(a), L.i32(b), arr, K._h, int(g > 0)))
                    L.check(lib.femd_assemble_vector_facets(F, len(a), L.i32(a), arr, f, int(g > 0)))
            assert np.array_equal(K.data, orc.assemble_matrix_groups(P, order))
            assert np.array_equal(f, orc.assemble_vector_groups(P, [(g[0], g[1], g[2], g[3], None, g[5]) for g in order]))


def test_facet_forms_with_normals(orc):
    _name("ds forms: normal programs on the GPU, Robin matrix")
    V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 6, 6), 2, dirichlet=[4])
    P = orc.rectangle(6, 6, 2, dirichlet=[4], all_exterior=False)
    u, v, n = fd.TrialFunction(V), fd.TestFunction(V), fd.n
    g = fd.sin(fd.Const(2.0) * fd.x) * n[0] + fd.y * n[1]
    Lf = fd.form(g * v * fd.ds((2, 3)) + fd.x * v * fd.dx)
    fg = Lf.assemble()
    x, y, w, nx, ny = orc.facet_points(P, Lf.facet_degree, [2, 3])
    cg_ = orc.integrand_eval_n(*fd.forms.Program(g).arrays(), x, y, nx, ny)
    cd = orc.integrand_eval(*fd.forms.Program(fd.x).arrays(), *P.cache_points(Lf.quad_degree)[:2])
    fo = orc.assemble_vector_groups(P, [("ds", Lf.facet_degree, [2, 3], [0], None, cg_[None, :]),
                                        ("dx", Lf.quad_degree, None, [0], None, cd[None, :])])
    assert np.abs(fg - fo).max() <= 1e-14 * np.abs(fo).max()
    # the GPU's program at the facet points against the oracle's libm
    out = np.empty(x.size)
    arr, keep = L.coefficient_array([(fd.forms.Program(g).arrays(), None)])
    L.check(L.load().femd_eval_coefficient_facets(V.facet_set(Lf.facet_degree, (2, 3)), arr, out))
    assert np.abs(out - cg_).max() <= 4e-15
    # Robin: grad.grad + alpha u v on sides 1, 2 (a constant coefficient: bit for bit)
    A = fd.form(fd.dot(fd.grad(u), fd.grad(v)) * fd.dx + fd.Const(3.0) * u * v * fd.ds((1, 2))).assemble()
    npd = P.ncells * len(orc.quadrature(2)[2])                      # grad.grad of P2: degree 2
    nf = orc.facet_points(P, 4, [1, 2])[0].size                     # u v of P2 on the sides: degree 4
    Ao = orc.assemble_matrix_groups(P, [("dx", 2, None, [1, 2], [1, 2], np.ones((2, npd))),
                                        ("ds", 4, [1, 2], [0], [0], np.full((1, nf), 3.0))])
    assert np.array_equal(A.data, Ao) and A.symmetric


@pytest.mark.parametrize("k", [1, 2])
def test_solve_dirichlet_lift(orc, k):
    """fd.solve with Dirichlet data: the patch test (u in P_k is reproduced to the solver's
    tolerance) with every side Dirichlet, and with Neumann data g = grad(u).n on three sides
    through ds; and against the same lift solved with the oracle."""
    _name(f"solve with dirichlet= k={k}")
    Pts, T, _ = problems.delaunay_square(0.15, seed=3)
    ue = (lambda x, y: 1 + 2 * x - y) if k == 1 else (lambda x, y: 1 + 2 * x - y + 0.5 * x * y)
    V = fd.LagrangeSpace(fd.mesh_from_arrays(Pts, T), k, bc="dirichlet")
    u, v = fd.TrialFunction(V), fd.TestFunction(V)
    a = fd.dot(fd.grad(u), fd.grad(v)) * fd.dx
    L0 = fd.Const(0.0) * v * fd.dx
    uh = fd.solve(a, L0, backend="cg", M="jacobi", tol=1e-13, dirichlet=ue)
    xy = V.raw_coordinates()
    assert uh.size == V.raw_dim and np.abs(uh - ue(xy[:, 0], xy[:, 1])).max() < 1e-9
    # Dirichlet on the left side only, g = grad(u).n on the others (P1: grad u = (2, -1))
    if k == 1:
        W = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 8, 8), 1, dirichlet=[4])
        u, v, n = fd.TrialFunction(W), fd.TestFunction(W), fd.n
        aw = fd.dot(fd.grad(u), fd.grad(v)) * fd.dx
        Lw = (fd.Const(2.0) * n[0] - n[1]) * v * fd.ds((1, 2, 3))
        uw = fd.solve(aw, Lw, backend="cg", M="jacobi", tol=1e-13, dirichlet={4: ue})
        xw = W.raw_coordinates()
        assert np.abs(uw - ue(xw[:, 0], xw[:, 1])).max() < 1e-9
        # the same data kept with the space: fd.solve takes them when the call gives none
        Wd = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 8, 8), 1, dirichlet={4: ue})
        u, v = fd.TrialFunction(Wd), fd.TestFunction(Wd)
        ud = fd.solve(fd.dot(fd.grad(u), fd.grad(v)) * fd.dx, (fd.Const(2.0) * n[0] - n[1]) * v * fd.ds((1, 2, 3)),
                      backend="cg", M="jacobi", tol=1e-13)
        assert np.array_equal(ud, uw)
    # the same lift through the oracle: unconstrained matrix, b_f - (A_U u_g)_f, Jacobi-CG
    f, _ = problems.fourier_source(2)
    V2 = V
    u2, v2 = fd.TrialFunction(V2), fd.TestFunction(V2)
    a2, L2 = fd.dot(fd.grad(u2), fd.grad(v2)) * fd.dx, f * v2 * fd.dx
    got = fd.solve(a2, L2, backend="cg", M="jacobi", tol=1e-12, dirichlet=ue)
    Po, Pu = orc.from_arrays(Pts, T, k=k), orc.from_arrays(Pts, T, k=k, all_exterior=False)
    deg_a, deg_l = fd.form(a2).quad_degree, fd.form(L2).quad_degree
    Ao = Po.assemble_matrix(deg_a, [1, 2], [1, 2], np.ones((2, Po.ncells * len(orc.quadrature(deg_a)[2]))))
    cf = orc.integrand_eval(*fd.forms.Program(f).arrays(), *Po.cache_points(deg_l)[:2])
    bo = Po.assemble_vector(deg_l, [0], cf[None, :])
    Au = Pu.assemble_matrix(deg_a, [1, 2], [1, 2], np.ones((2, Pu.ncells * len(orc.quadrature(deg_a)[2]))))
    ipu, ixu = Pu.pattern()
    ug = V2.lift(ue)
    num = Po.numbering()
    bo = bo - orc.spmv(ipu, ixu, Au, ug)[num["ad2raw"]]
    ip, ix = Po.pattern()
    xo, _ = orc.cg(ip, ix, Ao, bo, M="jacobi", tol=1e-12)
    want = V2.prolongate(xo) + ug
    assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want)


def test_asymmetry_matches_definition():
    """CSR asymmetry (csr.hpp:631-644): max |a_ij - a_ji| / max |a_ij|, a missing transpose counting
    as 0, exact maxima; the transpose positions are kept with the pattern, so a second call after
    new values reuses them.  Non-square matrices give 1."""
    import scipy.sparse as sp
    print("\n[case] asymmetry of random square patterns, new values, symmetric and non-square", flush=True)
    rng = np.random.default_rng(21)

    def ref(S):
        S = S.tocsr()
        S.sort_indices()
        big = np.abs(S.data).max() if S.nnz else 0.0
        T = S.T.tocsr()
        d = 0.0
        for i in range(S.shape[0]):
            for p in range(S.indptr[i], S.indptr[i + 1]):
                j = S.indices[p]
                d = max(d, abs(S.data[p] - T[i, j]))     # T[i, j] = a_ji, 0 when not stored
        return d / big if big > 0 else 0.0

    n = 300
    S = sp.random(n, n, density=0.03, random_state=1, format="csr") + sp.eye(n)
    S.sort_indices()
    A = fd.SparseMatrix.from_arrays(n, n, S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.copy(),
                                    symmetric=False)
    assert A.asymmetry() == ref(S)
    S2 = S.copy()
    S2.data = rng.standard_normal(S.nnz)
    A.set_data(S2.data)
    assert A.asymmetry() == ref(S2)
    Y = (S + S.T).tocsr()
    Y.sort_indices()
    B = fd.SparseMatrix.from_arrays(n, n, Y.indptr.astype(np.int32), Y.indices.astype(np.int32), Y.data.copy(),
                                    symmetric=False)
    assert B.asymmetry() == 0.0
    R = sp.random(20, 30, density=0.2, random_state=2, format="csr")
    R.sort_indices()
    C = fd.SparseMatrix.from_arrays(20, 30, R.indptr.astype(np.int32), R.indices.astype(np.int32), R.data.copy(),
                                    symmetric=False)
    assert C.asymmetry() == 1.0
