"""femd_cg_dev_noalloc, the CG entry point that allocates nothing: through the C ABI, its solution,
iteration count and residuals equal femd_cg_dev's bit for bit, on two calls in a row (the second on
a smaller system, so the shared work space holds the first call's contents), for no preconditioner,
Jacobi and ILU(0) in RCM order; bad sizes and a matrix with no SpMV launch shape are refused."""
import ctypes as C

import numpy as np
import pytest

import femd_amd as fd
from femd_amd import _lib as L
from femd_amd import problems

pytestmark = pytest.mark.gpu


def _system(N, seed):
    V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, N, N), 2, bc="dirichlet")
    f, _ = problems.fourier_source(seed)
    a, Lf = problems.poisson_forms(V, f, kappa=problems.kappa_field(seed)[0])
    return a.assemble(), Lf.assemble()


def _dev(host):
    p = C.c_void_p()
    L.check(L.load().femd_dev_alloc(max(8 * host.size, 8), C.byref(p)))
    L.check(L.load().femd_memcpy_htod(p.value, host.ctypes.data, 8 * host.size))
    return p.value


def _get(p, n):
    out = np.empty(n)
    L.check(L.load().femd_memcpy_dtoh(out.ctypes.data, p, 8 * n))
    return out


def _solve(entry, A, M, b, tol):
    lib = L.load()
    bd, xd = _dev(b), _dev(np.zeros_like(b))
    info = L.SolveInfo()
    L.check(getattr(lib, entry)(A._h, M, bd, xd, 20000, tol, C.byref(info), None))
    x = _get(xd, b.size)
    lib.femd_dev_free(bd)
    lib.femd_dev_free(xd)
    return x, (info.converged, info.iterations, info.breakdown, info.residual, info.precond_residual)


def _precond(A, kind):
    if kind is None:
        return None
    h = C.c_void_p()
    L.check(L.load().femd_precond_create(A._h, {"jacobi": L.PC_JACOBI, "ilu0": L.PC_ILU0}[kind],
                                         L.ORDER_RCM, 1.0, C.byref(h), None))
    return h.value


@pytest.mark.parametrize("kind", [None, "jacobi", "ilu0"])
def test_noalloc_cg_equals_cg_dev(kind):
    print(f"\n[case] femd_cg_dev_noalloc vs femd_cg_dev, preconditioner {kind}, P2 30x30 then 12x12", flush=True)
    for N, seed in ((30, 1), (12, 2)):
        A, b = _system(N, seed)
        M = _precond(A, kind)
        want, iw = _solve("femd_cg_dev", A, M, b, 1e-10)        # also builds A's FP64 SpMV launch shape
        got, ig = _solve("femd_cg_dev_noalloc", A, M, b, 1e-10)
        assert ig == iw and iw[0] == 1
        assert np.array_equal(got, want)
        if M:
            L.load().femd_precond_free(M)


def test_noalloc_cg_refuses_what_it_cannot_hold():
    print("\n[case] femd_cg_dev_noalloc: no launch shape yet, too many rows", flush=True)
    lib = L.load()
    A, b = _system(8, 3)
    bd, xd = _dev(b), _dev(np.zeros_like(b))
    info = L.SolveInfo()
    rc = lib.femd_cg_dev_noalloc(A._h, None, bd, xd, 100, 1e-10, C.byref(info), None)
    assert rc == L.EINVAL and b"femd_spmv_tuning" in lib.femd_last_error()
    name, ms = C.create_string_buffer(64), C.c_double()
    L.check(lib.femd_spmv_tuning(A._h, L.FP64, name, 64, C.byref(ms)))
    L.check(lib.femd_cg_dev_noalloc(A._h, None, bd, xd, 1000, 1e-10, C.byref(info), None))
    assert info.converged
    n = (1 << 20) + 1
    big = C.c_void_p()
    ip = np.arange(n + 1, dtype=np.int32)
    L.check(lib.femd_csr_create(n, n, ip, ip[:-1].copy(), np.ones(n), C.byref(big)))
    rc = lib.femd_cg_dev_noalloc(big.value, None, bd, xd, 100, 1e-10, C.byref(info), None)
    assert rc == L.EINVAL and b"1048576" in lib.femd_last_error()
    lib.femd_csr_free(big.value)
    lib.femd_dev_free(bd)
    lib.femd_dev_free(xd)
