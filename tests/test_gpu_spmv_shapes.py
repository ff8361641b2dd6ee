"""Every SpMV launch shape of the autotuner's space, on the GPU, through the C ABI.

FP64: bit-identical to the oracle's CSRMatrix::apply order.  FP32: within float rounding of the
exact product of the float values.  The matrices reach each path of the tile: row blocks that
start at any offset, nnz not a multiple of four (the guarded tail group), empty rows, rows longer
than the staging buffer, a single row.  The fused product of the FP32 inner CG is checked through
the refined solve with each shape pinned."""
import ctypes as C

import numpy as np
import pytest

import femd_amd as fd
from femd_amd import _lib as L
from femd_amd import problems

pytestmark = pytest.mark.gpu


def _shapes():
    lib = L.load()
    buf = C.create_string_buffer(64)
    out = []
    for i in range(lib.femd_spmv_shape_count()):
        L.check(lib.femd_spmv_shape_name(i, buf, 64))
        out.append(buf.value.decode())
    return out


def _csr(rng, lens, m):
    ip = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ix = np.concatenate([np.sort(rng.choice(m, size=k, replace=False)) if k else np.zeros(0, int) for k in lens]).astype(np.int32)
    return ip, ix, rng.standard_normal(ix.size)


def _cases(orc):
    rng = np.random.default_rng(11)
    out = {}
    out["one entry"] = (np.array([0, 1], np.int32), np.array([0], np.int32), np.array([2.5]), 1)
    lens = rng.integers(0, 41, 5003)
    lens[::97] = 0
    out["ragged 5003"] = (*_csr(rng, lens, 5003), 5003)
    lens = rng.integers(5, 30, 700)
    lens[[3, 350, 699]] = [3000, 9000, 5000]                 # beyond the staging buffer of some or all shapes
    lens[[100, 200]] = [300, 600]                            # within it, with hundreds of distinct columns
    out["long rows"] = (*_csr(rng, lens, 12000), 12000)
    P = orc.rectangle(32, 32, 2)
    ip, ix = P.pattern()
    A = P.assemble_matrix(4, [1, 2, 0], [1, 2, 0], np.ones((3, P.cache_points(4)[0].size)))
    out["P2 32x32"] = (ip, ix, A, ip.size - 1)
    return out


def _dev_product32(lib, M, x32):
    n, m = M.shape
    xb, yb = C.c_void_p(), C.c_void_p()
    L.check(lib.femd_dev_alloc(max(4 * m, 4), C.byref(xb)))
    L.check(lib.femd_dev_alloc(max(4 * n, 4), C.byref(yb)))
    try:
        L.check(lib.femd_memcpy_htod(xb, x32.ctypes.data, x32.nbytes))
        L.check(lib.femd_spmv_dev(M._h, xb, yb, L.FP32, None))
        y = np.empty(n, dtype=np.float32)
        L.check(lib.femd_memcpy_dtoh(y.ctypes.data, yb, y.nbytes))
        return y
    finally:
        lib.femd_dev_free(xb)
        lib.femd_dev_free(yb)


def test_every_shape_fp64_exact_fp32_rounded(orc):
    lib = L.load()
    names = _shapes()
    assert len(names) >= 3
    rng = np.random.default_rng(3)
    for label, (ip, ix, v, m) in _cases(orc).items():
        n = ip.size - 1
        M = fd.SparseMatrix.from_arrays(n, m, ip, ix, v, symmetric=False)
        x = rng.standard_normal(m)
        y_ref = orc.spmv(ip, ix, v, x)
        x32 = x.astype(np.float32)
        v32 = v.astype(np.float32).astype(np.float64)
        rows = np.repeat(np.arange(n), np.diff(ip))
        exact32 = np.bincount(rows, weights=v32 * x32[ix], minlength=n)
        scale = np.bincount(rows, weights=np.abs(v32 * x32[ix]), minlength=n)
        for name in names:
            print(f"\n[case] spmv shape {name} on {label}", flush=True)
            if lib.femd_spmv_select(M._h, L.FP64, name.encode()) == 0:
                assert np.array_equal(M.matvec(x), y_ref), (name, label)
            L.check(lib.femd_spmv_select(M._h, L.FP32, name.encode()))
            y32 = _dev_product32(lib, M, x32).astype(np.float64)
            assert np.all(np.abs(y32 - exact32) <= 2e-6 * scale + 1e-30), (name, label)
        L.check(lib.femd_spmv_select(M._h, L.FP64, None))
        L.check(lib.femd_spmv_select(M._h, L.FP32, None))
    with pytest.raises(ValueError):
        L.check(lib.femd_spmv_select(M._h, L.FP64, b"no-such-shape"))


def test_fused_inner_cg_every_shape(orc):
    lib = L.load()
    Pts, T, _ = problems.delaunay_square(0.06, seed=5)
    V = fd.LagrangeSpace(fd.mesh_from_arrays(Pts, T), 3, bc="dirichlet")
    kap, _ = problems.kappa_field(2)
    f, _ = problems.fourier_source(1)
    a, Lf = problems.poisson_forms(V, f, kappa=kap)
    A, b = a.assemble(), Lf.assemble()
    ip, ix = V.pattern()
    xo, _ = orc.cg(ip, ix, A.data, b, M="jacobi", tol=1e-13, maxiter=100000)
    for name in _shapes():
        print(f"\n[case] refined solve with FP32 shape {name}", flush=True)
        L.check(lib.femd_spmv_select(A._h, L.FP32, name.encode()))
        # the inner CG works on its reordered copy of the pattern, which takes the selected shape
        L.check(lib.femd_inner_timer(1))
        x, info = fd.solve_refined(A, b, tol=1e-11, inner_tol=1e-4)
        used = C.create_string_buffer(64)
        L.check(lib.femd_inner_timer_shape(used, 64))
        L.check(lib.femd_inner_timer(0))
        assert used.value.decode() == name
        assert info.converged, name
        assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-8, name
        # run to run: the reductions have a fixed order
        x2, _ = fd.solve_refined(A, b, tol=1e-11, inner_tol=1e-4)
        assert np.array_equal(x, x2), name
    L.check(lib.femd_spmv_select(A._h, L.FP32, None))
