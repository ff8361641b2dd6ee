"""The inner CG's set-up built on the GPU: the first-met order and P A P^T's pattern equal the host
construction entry for entry (segmented, with isolated unknowns and empty rows, on finite-element
patterns), a segment reaching outside itself is refused, and solves whose order and tile column
lists were built on the GPU are bit-identical to solves set up on the host."""
import numpy as np
import pytest
import scipy.sparse as sp

import femd_amd as fd
from femd_amd import _lib as L
from femd_amd import problems

pytestmark = pytest.mark.gpu


def _name(s):
    print(f"\n[case] {s}", flush=True)


def _order(fn, ip, ix, segments):
    n, nnz = ip.size - 1, int(ip[-1])
    perm = np.full(n, -7, np.int32)
    ip2 = np.full(n + 1, -7, np.int32)
    ix2 = np.full(max(nnz, 1), -7, np.int32)
    emap = np.full(max(nnz, 1), -7, np.int32)
    seg = None if segments is None else np.asarray(segments, np.int64)
    L.check(fn(n, ip, ix, 0 if seg is None else seg.size - 1, None if seg is None else seg.ctypes.data, perm, ip2, ix2,
               emap))
    return perm, ip2, ix2[:nnz], emap[:nnz]


def _same_as_host(ip, ix, segments):
    lib = L.load()
    ip, ix = np.ascontiguousarray(ip, np.int32), np.ascontiguousarray(ix if ix.size else np.zeros(1), np.int32)
    host = _order(lib.femd_inner_order_pattern, ip, ix, segments)
    dev = _order(lib.femd_inner_order_pattern_gpu, ip, ix, segments)
    for h, d in zip(host, dev):
        assert np.array_equal(h, d)
    return dev


def _fe_pattern(V):
    ip, ix = V.pattern()
    return np.asarray(ip, np.int32), np.asarray(ix, np.int32)


def test_order_equals_host_construction():
    _name("first-met order on the GPU against the host's")
    rng = np.random.default_rng(3)
    sizes = [7, 12, 1, 5, 9, 300]
    blocks = []
    for k, m in enumerate(sizes):
        a = (rng.random((m, m)) < min(0.3, 8.0 / m)).astype(float)
        a = a + a.T + np.eye(m)
        if k == 3:
            a[2, :] = 0.0
            a[:, 2] = 0.0                     # an unknown with no entries: last in its segment
        if k == 5:
            a[[4, 50, 51], :] = 0.0           # empty rows whose columns are met elsewhere
        blocks.append(sp.csr_matrix(a))
    S = sp.block_diag(blocks, format="csr")
    S.sort_indices()
    seg = np.concatenate([[0], np.cumsum(sizes)])
    perm, ip2, ix2, emap = _same_as_host(S.indptr, S.indices, seg)
    E = sp.csr_matrix((np.arange(S.nnz, dtype=np.float64), S.indices, S.indptr), shape=S.shape)
    Q = E[perm][:, perm].tocsr()
    Q.sort_indices()
    assert np.array_equal(ip2, Q.indptr) and np.array_equal(ix2, Q.indices)
    assert np.array_equal(emap, Q.data.astype(np.int64))
    _same_as_host(S.indptr, S.indices, None)
    # finite-element patterns: P3 on a Delaunay mesh, and three spaces as one segmented pattern
    Pts, T, _ = problems.delaunay_square(0.03, seed=9)
    V3 = fd.LagrangeSpace(fd.mesh_from_arrays(Pts, T), 3, bc="dirichlet")
    ip, ix = _fe_pattern(V3)
    _same_as_host(ip, ix, None)
    parts = [_fe_pattern(fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, N, N), k)) for N, k in ((9, 2), (30, 1), (6, 3))]
    mats = [sp.csr_matrix((np.ones(p[1].size), p[1], p[0])) for p in parts]
    B = sp.block_diag(mats, format="csr")
    B.sort_indices()
    _same_as_host(B.indptr, B.indices, np.concatenate([[0], np.cumsum([m.shape[0] for m in mats])]))
    # no rows at all
    _same_as_host(np.zeros(1, np.int32), np.zeros(0, np.int32), None)


def test_order_refuses_rows_outside_their_segment():
    _name("first-met order on the GPU: a row outside its segment")
    S = sp.block_diag([sp.eye(4), sp.eye(5)], format="lil")
    S[1, 7] = 1.0
    S = S.tocsr()
    S.sort_indices()
    seg = np.array([0, 4, 9], np.int64)
    with pytest.raises(ValueError):
        _order(L.load().femd_inner_order_pattern_gpu, S.indptr.astype(np.int32), S.indices.astype(np.int32), seg)


def _systems():
    out = []
    for N, k, seed in ((14, 2, 7), (9, 3, 8), (21, 1, 9)):
        V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, N, N), k, bc="dirichlet")
        f, _ = problems.fourier_source(seed)
        a, Lf = problems.poisson_forms(V, f, kappa=problems.kappa_field(seed)[0])
        out.append((V, a.assemble(), Lf.assemble()))
    return out


def test_solves_set_up_on_gpu_are_bit_identical():
    """Batch (FP32 and FP64 inner) and a refined single solve on an unstructured P3 system (every
    SpMV launch shape planned for the autotuner): new patterns in each mode, the same bits; the
    server, whose groups take the inner order only with the device set-up, within 1e-8."""
    _name("batch / refined / server solves: device set-up against host set-up")
    prev, prev_order = fd.device_setup(), fd.inner_order(True)
    Pts, T, _ = problems.delaunay_square(0.02, seed=4)
    try:
        runs = {}
        for mode in (False, True):
            fd.device_setup(mode)
            systems = _systems()
            mats, bs = [A for _, A, _ in systems], [b for _, _, b in systems]
            r = {}
            r["batch32"] = fd.Batch(mats).solve(bs, tol=1e-11, inner_tol=1e-4)[0]
            r["batch64"] = fd.Batch(mats).solve(bs, tol=1e-11, inner_tol=1e-4, precision="fp64")[0]
            V = fd.LagrangeSpace(fd.mesh_from_arrays(Pts, T), 3, bc="dirichlet")
            f, _ = problems.fourier_source(2)
            a, Lf = problems.poisson_forms(V, f, kappa=problems.kappa_field(2)[0])
            A, b = a.assemble(), Lf.assemble()
            x, info = fd.solve_refined(A, b, tol=1e-10, inner_tol=1e-3)
            assert info.converged
            r["refined"] = [x]
            r["spmv"] = [A.matvec(b)]
            runs[mode] = r
        for key in runs[True]:
            for h, d in zip(runs[False][key], runs[True][key]):
                assert np.array_equal(h, d), key
        # the server's groups work in the inner order once it is built on the device
        S = fd.Server(hbm_budget=1 << 30, max_systems=8, tol=1e-11, inner_tol=1e-4)
        for t, (Am, bm) in enumerate(zip(mats, bs)):
            S.submit(t, Am, bm)
        done = S.run()
        for t, x64 in enumerate(runs[True]["batch64"]):
            assert done[t].converged and np.linalg.norm(done[t].x - x64) <= 1e-8 * np.linalg.norm(x64)
    finally:
        fd.device_setup(prev)
        fd.inner_order(prev_order)


def _lists(V, where):
    import ctypes as C
    lib = L.load()
    nnz, ncon = C.c_int64(), C.c_int64()
    L.check(lib.femd_pattern_lists(V._h, where, C.byref(nnz), C.byref(ncon), *([None] * 6)))
    nloc = V.nloc if hasattr(V, "nloc") else None
    nv = int((np.asarray(V.cell_dofs()) >= 0).sum()) if hasattr(V, "cell_dofs") else None
    arrs = [np.full(V.dim + 1, -7, np.int32), np.full(max(nnz.value, 1), -7, np.int32),
            np.full(nnz.value + 1, -7, np.int32), np.full(max(ncon.value, 1), -7, np.int32),
            np.full(V.dim + 1, -7, np.int32), np.full(max(nv if nv is not None else 1, 1), -7, np.int32)]
    L.check(lib.femd_pattern_lists(V._h, where, None, None, *[a.ctypes.data for a in arrs]))
    return nnz.value, ncon.value, arrs


@pytest.mark.parametrize("case", ["p1", "p2_dirichlet", "p3_delaunay", "q2_markers", "p2_alternate"])
def test_pattern_and_gather_lists_equal_host(case):
    """The pattern and its gather lists built on the GPU (radix sorts of the element entries by
    (row, column) and of the cell dofs by dof) equal the host builder's, array for array; assembly
    on a space whose pattern the GPU built is bit-identical to assembly on a host-built one."""
    _name(f"pattern and gather lists on the GPU: {case}")
    if case == "p1":
        V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 23, 17), 1)
    elif case == "p2_dirichlet":
        V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 30, 30), 2, bc="dirichlet")
    elif case == "p3_delaunay":
        Pts, T, _ = problems.delaunay_square(0.03, seed=6)
        V = fd.LagrangeSpace(fd.mesh_from_arrays(Pts, T), 3, bc="dirichlet")
    elif case == "q2_markers":
        V = fd.LagrangeSpace(fd.quad_mesh_from_arrays(*_quads(9, 7)), 2, bc="dirichlet")
    else:
        V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, 12, 9, diagonal="alternate"), 2, dirichlet=[2])
    nnz_h, ncon_h, host = _lists(V, 0)
    nnz_d, ncon_d, dev = _lists(V, 1)
    assert nnz_h == nnz_d and ncon_h == ncon_d
    for k, (h, d) in enumerate(zip(host, dev)):
        assert np.array_equal(h, d), k
    ip, ix = V.pattern()
    assert np.array_equal(ip, host[0]) and np.array_equal(ix, host[1][:nnz_h])


def _quads(nx, ny):
    rng = np.random.default_rng(8)
    pts = np.array([(i / nx + (0 if i in (0, nx) else 0.2 * (rng.random() - 0.5) / nx),
                     j / ny + (0 if j in (0, ny) else 0.2 * (rng.random() - 0.5) / ny))
                    for j in range(ny + 1) for i in range(nx + 1)])
    q = np.array([[j * (nx + 1) + i, j * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i]
                  for j in range(ny) for i in range(nx)], dtype=np.int32)
    return pts, q


def test_assembly_on_gpu_built_pattern_is_bit_identical():
    _name("assembly on GPU-built and host-built patterns")
    prev = fd.device_setup()
    Pts, T, _ = problems.delaunay_square(0.04, seed=3)
    try:
        out = {}
        for mode in (False, True):
            fd.device_setup(mode)
            V = fd.LagrangeSpace(fd.mesh_from_arrays(Pts, T), 2, bc="dirichlet")
            f, _ = problems.fourier_source(1)
            u, v = fd.TrialFunction(V), fd.TestFunction(V)
            a = fd.form(problems.kappa_field(1)[0] * fd.dot(fd.grad(u), fd.grad(v)) * fd.dx + u * v * fd.ds)
            A = a.assemble()
            b = fd.form(f * v * fd.dx).assemble()
            out[mode] = (A.indptr.copy(), A.indices.copy(), A.data.copy(), b)
        for h, d in zip(out[False], out[True]):
            assert np.array_equal(h, d)
    finally:
        fd.device_setup(prev)
