"""The batched solver on the GPU: uneven systems in one launch sequence, each system's result the
same alone and in the mix, and within 1e-8 of the FP64 oracle solution."""
import numpy as np
import pytest

import femd_amd as fd
from femd_amd import problems

pytestmark = pytest.mark.gpu


def _system(N, k, seed):
    V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, N, N), k, bc="dirichlet")
    f, _ = problems.fourier_source(seed)
    a, L = problems.poisson_forms(V, f)
    return V, a.assemble(), L.assemble()


def test_batch_matches_oracle_and_is_mix_independent(orc):
    print("\n[case] batch of 7 uneven P2/P1/P3 systems", flush=True)
    specs = [(6, 2, 0), (17, 2, 1), (33, 1, 2), (9, 3, 3), (40, 2, 4), (5, 2, 5), (25, 3, 6)]
    systems = [_system(*s) for s in specs]
    mats = [A for _, A, _ in systems]
    bs = [b for _, _, b in systems]
    B = fd.Batch(mats)
    assert B.nrows == sum(V.dim for V, _, _ in systems)
    xs, info = B.solve(bs, tol=1e-11, inner_tol=1e-4)
    assert info["converged"] and info["worst_residual"] <= 1e-11
    for (V, A, b), x in zip(systems, xs):
        ip, ix = V.pattern()
        xo, io = orc.cg(ip, ix, A.data, b, M="jacobi", tol=1e-13)
        assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-8
    # the same system alone: bit-identical (every reduction is per system)
    for j in (1, 3):
        alone, _ = fd.Batch([mats[j]]).solve([bs[j]], tol=1e-11, inner_tol=1e-4)
        assert np.array_equal(alone[0], xs[j])
    # graph replay and plain launches agree bit for bit
    xs2, _ = B.solve(bs, tol=1e-11, inner_tol=1e-4, graph=False)
    for a, b in zip(xs, xs2):
        assert np.array_equal(a, b)


def test_batch_update_values_and_zero_rhs():
    print("\n[case] batch update + zero rhs", flush=True)
    (V1, A1, b1), (V2, A2, b2) = _system(8, 2, 0), _system(12, 2, 1)
    B = fd.Batch([A1, A2])
    xs, info = B.solve([b1, np.zeros_like(b2)], tol=1e-10)
    assert info["converged"] and not xs[1].any()
    A1.set_data(2.0 * A1.data)
    B.update_values()
    ys, info = B.solve([b1, b2], tol=1e-10)
    assert np.linalg.norm(2.0 * ys[0] - xs[0]) <= 1e-8 * np.linalg.norm(xs[0])


def test_batch_fp64_inner(orc):
    print("\n[case] batch with FP64 inner iterations", flush=True)
    specs = [(10, 2, 7), (23, 2, 8), (7, 3, 9)]
    systems = [_system(*s) for s in specs]
    B = fd.Batch([A for _, A, _ in systems])
    xs, info = B.solve([b for _, _, b in systems], tol=1e-11, inner_tol=1e-11, precision="fp64")
    assert info["converged"] and info["worst_residual"] <= 1e-11
    for (V, A, b), x in zip(systems, xs):
        ip, ix = V.pattern()
        xo, _ = orc.cg(ip, ix, A.data, b, M="jacobi", tol=1e-13)
        assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-9
    # FP64 inner iterations take about as many steps as the FP64 reference algorithm
    for (V, A, b), it in zip(systems, info["iterations"]):
        ip, ix = V.pattern()
        _, io = orc.cg(ip, ix, A.data, b, M="jacobi", tol=1e-11)
        assert it <= 1.25 * io["iterations"] + 5
    with pytest.raises(ValueError):
        B.solve([b for _, _, b in systems], precision="fp16")


def test_inner_order_switch(orc):
    """The inner CG's own order of the unknowns and the original one: both within 1e-8 of the FP64
    oracle, alone and in the mix the same bits in either, and switching the order back gives the
    first result again (the scaled values are rebuilt for the order in use)."""
    print("\n[case] inner order on / off / on, batch of 3 and a refined solve", flush=True)
    systems = [_system(*s) for s in [(14, 2, 7), (9, 3, 8), (21, 1, 9)]]
    mats = [A for _, A, _ in systems]
    bs = [b for _, _, b in systems]
    refs = []
    for V, A, b in systems:
        ip, ix = V.pattern()
        refs.append(orc.cg(ip, ix, A.data, b, M="jacobi", tol=1e-13)[0])
    prev = fd.inner_order(True)
    try:
        B = fd.Batch(mats)
        runs = {}
        for mode in (True, False, True):
            fd.inner_order(mode)
            xs, info = B.solve(bs, tol=1e-11, inner_tol=1e-4)
            assert info["converged"], mode
            for x, xo in zip(xs, refs):
                assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-8, mode
            alone, _ = fd.Batch([mats[1]]).solve([bs[1]], tol=1e-11, inner_tol=1e-4)
            assert np.array_equal(alone[0], xs[1]), mode
            if mode in runs:
                for a, b in zip(runs[mode], xs):
                    assert np.array_equal(a, b)
            runs[mode] = xs
        # a single refined solve on one matrix, both orders
        V, A, b = systems[0]
        for mode in (False, True):
            fd.inner_order(mode)
            x, info = fd.solve_refined(A, b, tol=1e-11, inner_tol=1e-4)
            assert info.converged and np.linalg.norm(x - refs[0]) / np.linalg.norm(refs[0]) <= 1e-8, mode
    finally:
        fd.inner_order(prev)


def test_large_system_alone_and_in_a_mix():
    """A system large enough for the single-system update path (alpha formed in the update from
    the tile partials, each workgroup taking a run of chunks): alone, its iterates are those it has
    next to a small system (where alpha is formed by its own kernel), bit for bit."""
    print("\n[case] 1.2M-unknown P1 system alone and with a small one", flush=True)
    V, A, b = _system(1100, 1, 11)
    _, As, bs = _system(9, 2, 12)
    alone, info1 = fd.Batch([A]).solve([b], tol=1e-9, inner_tol=0.3)
    mix, info2 = fd.Batch([A, As]).solve([b, bs], tol=1e-9, inner_tol=0.3)
    assert info1["converged"] and info2["converged"]
    assert np.array_equal(alone[0], mix[0])
    x, info = fd.solve_refined(A, b, tol=1e-9, inner_tol=0.3)
    assert info.converged and np.array_equal(x, alone[0])


@pytest.mark.parametrize("blocks", [5, 35000])
def test_inner_cg_breakdown_ends_the_solve(blocks):
    """An indefinite matrix with a positive diagonal (2x2 blocks [[1, 2], [2, 1]]) and b = (1, -1)
    per block: b.Ab = -2 per block, so the first inner iteration breaks down (p.q <= 0).  Every
    update workgroup of the system sees it; the last to finish pauses the system (one arrival
    counter for 5 blocks, two levels for 35000: 69 update workgroups), the host restarts it, and
    the solve ends unconverged after max_outer rounds instead of looping or returning NaN."""
    print(f"\n[case] inner CG breakdown, {blocks} blocks of [[1, 2], [2, 1]]", flush=True)
    n = 2 * blocks
    ip = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    ix = np.empty(2 * n, dtype=np.int32)
    ix[0::4], ix[1::4] = np.arange(0, n, 2), np.arange(1, n, 2)
    ix[2::4], ix[3::4] = np.arange(0, n, 2), np.arange(1, n, 2)
    v = np.tile([1.0, 2.0, 2.0, 1.0], blocks)
    A = fd.SparseMatrix.from_arrays(n, n, ip, ix, v)
    b = np.tile([1.0, -1.0], blocks)
    x, info = fd.solve_refined(A, b, tol=1e-8, inner_tol=0.3, max_outer=4)
    assert not info.converged
    assert np.all(np.isfinite(x))
    # the same system again: the arrival counters were left cleared
    x2, info2 = fd.solve_refined(A, b, tol=1e-8, inner_tol=0.3, max_outer=4)
    assert not info2.converged and np.array_equal(x, x2) and info2.iterations == info.iterations
