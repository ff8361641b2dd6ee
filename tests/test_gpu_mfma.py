"""The matrix-core assembly mode (assembly_mode "mfma") against the oracle, through the C ABI.

The FP64 MFMA kernel forms the same sums as forms/assembly_2d.hpp:102-126 in another order (one
GEMM over the cells of D_e[k] R[k][ij]), so its values are held to the FP64 bar, 1e-12 of each
row's largest entry, instead of bit for bit; the exact mode stays bit-identical before and after
the switch is used, and groups the kernel is not built for fall back to the exact kernel."""
import numpy as np
import pytest

import femd_amd as fd
from femd_amd import _lib as L
from femd_amd import problems

pytestmark = pytest.mark.gpu

BAR = 1e-12


def _name(s):
    print(f"\n[case] {s}", flush=True)


class _mfma:
    def __enter__(self):
        self.prev = fd.assembly_mode("mfma")

    def __exit__(self, *exc):
        fd.assembly_mode(self.prev)


def _assemble(V, deg, a, b, coeffs):
    arr, keep = L.coefficient_array([(None, c) for c in coeffs])
    K = fd.SparseMatrix._on_space(V)
    L.check(L.load().femd_assemble_matrix(V._h, deg, len(a), L.i32(a), L.i32(b), arr, K._h))
    return K


def _runs_on_matrix_cores(V, deg, a, b, coeffs):
    """Whether the group runs on the matrix-core kernel (not the exact kernel's fallback)."""
    arr, keep = L.coefficient_array([(None, c) for c in coeffs])
    ms, fl = np.zeros(1), np.zeros(1)
    rc = L.load().femd_assembly_kernel_time(V._h, deg, len(a), L.i32(a), L.i32(b), arr, 1, 1, ms, fl)
    return rc == 0 and fl[0] > 0


def _row_err(ip, got, want):
    """max over rows of max |got - want| / max |want| (rows of zeros compare absolutely)."""
    worst = 0.0
    d = np.abs(got - want)
    for i in range(ip.size - 1):
        lo, hi = ip[i], ip[i + 1]
        if hi == lo:
            continue
        scale = np.abs(want[lo:hi]).max()
        worst = max(worst, d[lo:hi].max() / (scale if scale > 0 else 1.0))
    return worst


FORMS = [
    ("stiffness kappa", [1, 2], [1, 2], lambda c: [c[0], c[0]]),
    ("nonsymmetric mix", [1, 2, 0, 1], [1, 2, 0, 2], lambda c: list(c[:4])),
    ("symmetric cross terms", [1, 2, 0, 1, 2], [2, 1, 0, 1, 2], lambda c: [c[1], c[1], c[2], c[3], c[3]]),
    ("mass + convection", [0, 0, 0], [0, 1, 2], lambda c: list(c[:3])),
]


@pytest.mark.parametrize("k,deg,diag", [(1, 2, "right"), (2, 4, "alternate"), (3, 8, "crossed"), (3, 6, "left")])
@pytest.mark.parametrize("form", FORMS, ids=[f[0] for f in FORMS])
def test_mfma_triangles(orc, k, deg, diag, form):
    label, a, b, pick = form
    _name(f"mfma P{k} deg={deg} {diag} {label}")
    m = fd.rectangle_mesh(0, 0, 1, 1, 9, 7, diagonal=diag)
    V = fd.LagrangeSpace(m, k, bc="dirichlet")
    P = orc.rectangle(9, 7, k, diag)
    npt = P.ncells * len(orc.quadrature(deg)[2])
    c = np.random.default_rng(k + deg).standard_normal((4, npt))
    coeffs = pick(c)
    want = P.assemble_matrix(deg, a, b, np.stack(coeffs))
    assert _runs_on_matrix_cores(V, deg, a, b, coeffs)
    with _mfma():
        got = _assemble(V, deg, a, b, coeffs).data
    ip, _ = P.pattern()
    err = _row_err(ip, got, want)
    print(f"  row-relative error {err:.2e}")
    assert err <= BAR
    # the exact mode is untouched by the switch
    assert np.array_equal(_assemble(V, deg, a, b, coeffs).data, want)


@pytest.mark.parametrize("k,deg", [(1, 2), (2, 4), (3, 7)])
def test_mfma_quads(orc, k, deg):
    _name(f"mfma Q{k} deg={deg}, rectangle and distorted meshes")
    rng = np.random.default_rng(10 + k)
    pts = []
    nx, ny = 7, 6
    for j in range(ny + 1):
        for i in range(nx + 1):
            dx = 0.0 if i in (0, nx) else 0.25 * (rng.random() - 0.5) / nx
            dy = 0.0 if j in (0, ny) else 0.25 * (rng.random() - 0.5) / ny
            pts.append((i / nx + dx, j / ny + dy))
    pts = np.array(pts)
    q = np.array([[j * (nx + 1) + i, j * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i]
                  for j in range(ny) for i in range(nx)], dtype=np.int32)
    cases = [(fd.LagrangeSpace(fd.rectangle_quad_mesh(0, 0, 1, 1, nx, ny), k, bc="dirichlet"), orc.quad_rectangle(nx, ny, k)),
             (fd.LagrangeSpace(fd.quad_mesh_from_arrays(pts, q), k), orc.quad_from_arrays(pts, q, k=k, all_exterior=False))]
    for V, P in cases:
        npt = P.ncells * len(orc.quad_quadrature(deg)[2])
        c = rng.standard_normal((4, npt))
        for label, a, b, pick in FORMS:
            coeffs = pick(c)
            want = P.assemble_matrix(deg, a, b, np.stack(coeffs))
            assert _runs_on_matrix_cores(V, deg, a, b, coeffs), label
            with _mfma():
                got = _assemble(V, deg, a, b, coeffs).data
            ip, _ = P.pattern()
            err = _row_err(ip, got, want)
            print(f"  {label}: row-relative error {err:.2e}")
            assert err <= BAR, label


def test_mfma_unstructured_p3_solve(orc):
    """C1's form (kappa-Poisson, P3, degree 8) on an unstructured mesh: values within the bar and
    the ILU(0)-RCM CG solution of the MFMA matrix within 1e-12 of the oracle's at tol 1e-13."""
    _name("mfma unstructured P3 kappa-Poisson + ILU(0) CG")
    Pts, T, _ = problems.delaunay_square(0.06, seed=5)
    V = fd.LagrangeSpace(fd.mesh_from_arrays(Pts, T), 3, bc="dirichlet")
    Po = orc.from_arrays(Pts, T, k=3)
    kap, _ = problems.kappa_field(1)
    x, y, _ = Po.cache_points(8)
    ck = orc.integrand_eval(*fd.forms.Program(kap).arrays(), x, y)
    ip, ix = Po.pattern()
    Ao = Po.assemble_matrix(8, [1, 2], [1, 2], np.stack([ck, ck]))
    with _mfma():
        K = _assemble(V, 8, [1, 2], [1, 2], [ck, ck])
        # through the form language too: the device coefficient program, then the MFMA kernel
        a_form, L_form = problems.poisson_forms(V, problems.fourier_source(0)[0], kappa=kap)
        Kf = a_form.assemble()
    err = _row_err(ip, K.data, Ao)
    print(f"  row-relative error {err:.2e}, form path {_row_err(ip, Kf.data, Ao):.2e}")
    assert err <= BAR
    assert _row_err(ip, Kf.data, Ao) <= BAR
    b = L_form.assemble()
    xg, info = fd.cg(K, b, M=K.preconditioner("ilu0", ordering="rcm"), tol=1e-13)
    xo, io = orc.cg(ip, ix, Ao, b, M="ilu0", ordering="rcm", tol=1e-13)
    rel = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    print(f"  iterations {info.iterations} / oracle {io['iterations']}, solution {rel:.2e}")
    assert info.converged and rel <= 1e-12


def test_mfma_fallback_and_mode_switch(orc):
    """A group whose rule has more points than the kernel's steps hold (degree 10: 36 > 32) runs
    on the exact kernel even in mfma mode, bit for bit; the switch rejects unknown modes."""
    _name("mfma fallback P5 + mode switch")
    m = fd.rectangle_mesh(0, 0, 1, 1, 4, 4)
    V = fd.LagrangeSpace(m, 5, bc="dirichlet")
    P = orc.rectangle(4, 4, 5)
    deg = 10
    npt = P.ncells * len(orc.quadrature(deg)[2])
    c = np.random.default_rng(3).standard_normal((4, npt))
    a, b = [0, 1, 2, 1], [0, 2, 1, 1]
    want = P.assemble_matrix(deg, a, b, c)
    assert fd.assembly_mode() == "exact"
    assert not _runs_on_matrix_cores(V, deg, a, b, list(c))
    with _mfma():
        assert fd.assembly_mode() == "mfma"
        got = _assemble(V, deg, a, b, list(c)).data
    assert fd.assembly_mode() == "exact"
    assert np.array_equal(got, want)
    with pytest.raises(ValueError):
        fd.assembly_mode("bf16")


def test_mfma_16bit_local_matrices(orc):
    """Modes mfma_f16 / mfma_bf16 round the GEMM's operands to 16 bits and sum in FP32: the
    entries carry the 16-bit rounding (f16 2^-11, bf16 2^-8 relative per operand) and no more;
    refinement with the f16 matrix as the correction operator converges on a well-conditioned
    system (P2 -Laplace u + u on 32 x 32)."""
    _name("mfma f16 / bf16 local matrices")
    Pts, T, _ = problems.delaunay_square(0.06, seed=5)
    V = fd.LagrangeSpace(fd.mesh_from_arrays(Pts, T), 3, bc="dirichlet")
    Po = orc.from_arrays(Pts, T, k=3)
    kap, _ = problems.kappa_field(1)
    x, y, _ = Po.cache_points(8)
    ck = orc.integrand_eval(*fd.forms.Program(kap).arrays(), x, y)
    ip, ix = Po.pattern()
    Ao = Po.assemble_matrix(8, [1, 2], [1, 2], np.stack([ck, ck]))
    bars = {"mfma_f16": (1e-6, 5e-3), "mfma_bf16": (1e-5, 5e-2)}
    for mode, (lo, hi) in bars.items():
        prev = fd.assembly_mode(mode)
        try:
            K = _assemble(V, 8, [1, 2], [1, 2], [ck, ck])
        finally:
            fd.assembly_mode(prev)
        err = _row_err(ip, K.data, Ao)
        print(f"  {mode}: row-relative error {err:.2e}")
        assert lo <= err <= hi, (mode, err)
        assert K.asymmetry() == 0.0                # i <= j computed, written to both places
    from femd_amd import precision_bench
    m = fd.rectangle_mesh(0, 0, 1, 1, 32, 32)
    V2 = fd.LagrangeSpace(m, 2, bc="dirichlet")
    u, v = fd.TrialFunction(V2), fd.TestFunction(V2)
    a = fd.form((fd.dot(fd.grad(u), fd.grad(v)) + u * v) * fd.dx)
    A = a.assemble()
    b = fd.form(problems.fourier_source(0)[0] * v * fd.dx).assemble()
    prev = fd.assembly_mode("mfma_f16")
    try:
        Alo = a.assemble(reuse=False)
    finally:
        fd.assembly_mode(prev)
    xr = np.zeros(A.shape[0])
    hist = []
    for _ in range(12):
        r = b - A.matvec(xr)
        hist.append(np.linalg.norm(r) / np.linalg.norm(b))
        if hist[-1] <= 1e-10:
            break
        d, _info = fd.cg(Alo, r, M=Alo.preconditioner("jacobi"), tol=1e-13)
        xr = xr + d
    print(f"  f16 refinement residuals {[f'{h:.1e}' for h in hist]}")
    assert hist[-1] <= 1e-10 and len(hist) <= 8
    assert precision_bench.row_err(A.indptr, Alo.data, A.data) < 5e-3
