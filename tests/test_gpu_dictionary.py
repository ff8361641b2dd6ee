"""Per-tile value dictionaries in the inner CG: where a matrix allows one (at most 256 distinct
values per SpMV tile) the product reads 1-byte positions into an LDS table instead of the values.
A tile with more distinct values reads the values themselves; a matrix where such tiles hold more
than a quarter of the entries keeps the plain format.  The table holds the values bit for bit, so
every solver gives bit-identical results with the dictionary on and off."""
import numpy as np
import pytest

import femd_amd as fd
from femd_amd import problems

pytestmark = pytest.mark.gpu


def _system(N, k, seed, kappa=None):
    V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, N, N), k, bc="dirichlet")
    f, _ = problems.fourier_source(seed)
    a, L = problems.poisson_forms(V, f, kappa=kappa)
    return V, a.assemble(), L.assemble()


@pytest.fixture
def dictionary_switch():
    prev = fd.value_dictionary()
    yield
    fd.value_dictionary(prev)


def test_batch_dictionary_bit_identical(dictionary_switch):
    print("\n[case] batch of 4 P1/P2/P3 systems, dictionary on and off, fp32 and fp64", flush=True)
    # every SpMV tile of these holds at most 200 distinct scaled values in FP64 (about 80-190;
    # rounding in the element geometry makes equal entries differ in their last bits)
    systems = [_system(*s) for s in [(12, 2, 10), (31, 2, 11), (8, 3, 12), (20, 1, 13)]]
    mats, bs = [A for _, A, _ in systems], [b for _, _, b in systems]
    B = fd.Batch(mats)
    runs = {"fp32": dict(tol=1e-10, inner_tol=0.3), "fp64": dict(tol=1e-10, inner_tol=0.5)}
    with_dict = {}
    fd.value_dictionary(True)
    for prec, kw in runs.items():
        xs, info = B.solve(bs, precision=prec, **kw)
        assert info["converged"]
        d = B.dictionary(prec)
        assert d is not None and 0 < d[1] <= 256, d
        with_dict[prec] = (xs, info)
    fd.value_dictionary(False)
    for prec, kw in runs.items():
        xs, info = B.solve(bs, precision=prec, **kw)
        assert B.dictionary(prec) is None
        xd, infod = with_dict[prec]
        assert all(np.array_equal(a, b) for a, b in zip(xs, xd))
        assert np.array_equal(info["iterations"], infod["iterations"])


def test_refined_and_stepper_dictionary_bit_identical(dictionary_switch):
    print("\n[case] refined P2 solve and heat stepper, dictionary on and off", flush=True)
    V, A, b = _system(40, 2, 14)
    V2, M, A2, u0 = problems.heat_problem(24, 2, dt=1e-3, seed=3)
    out = []
    for on in (True, False):
        fd.value_dictionary(on)
        x, info = fd.solve_refined(A, b, tol=1e-10, inner_tol=0.3)
        assert info.converged
        u, hinfo = fd.BackwardEuler(M, A2).run(u0, 10, tol=1e-11, rounds=3, inner_tol=1e-4)
        out.append((x, info.iterations, u, hinfo["iterations"]))
    (x1, i1, u1, h1), (x2, i2, u2, h2) = out
    assert np.array_equal(x1, x2) and i1 == i2
    assert np.array_equal(u1, u2) and np.array_equal(h1, h2)


def test_too_many_values_keep_plain_format(dictionary_switch):
    print("\n[case] variable-coefficient P3, and P3 10x10 in FP64: too many distinct values", flush=True)
    kap, _ = problems.kappa_field(1)
    V, A, b = _system(24, 3, 15, kappa=kap)
    fd.value_dictionary(True)
    B = fd.Batch([A])
    xs, info = B.solve([b], tol=1e-10)
    assert info["converged"]
    assert B.dictionary("fp32") is None
    # FP64 tiles of this matrix hold up to about 310 distinct scaled values, FP32 rounding merges
    # them: whatever the format, the FP64 solve gives the same bits as without dictionaries
    V, A, b = _system(10, 3, 16)
    B = fd.Batch([A])
    x64, _ = B.solve([b], tol=1e-10, precision="fp64", inner_tol=0.5)
    fd.value_dictionary(False)
    y64, _ = B.solve([b], tol=1e-10, precision="fp64", inner_tol=0.5)
    fd.value_dictionary(True)
    assert np.array_equal(x64[0], y64[0])
    x32, _ = B.solve([b], tol=1e-10)
    assert B.dictionary("fp32") is not None
    assert np.linalg.norm(x32[0] - x64[0]) <= 1e-8 * np.linalg.norm(x64[0])


def test_tiles_with_many_values_read_the_values(dictionary_switch):
    """A batch whose variable-coefficient member has tiles of more than 256 distinct values: those
    tiles read the values themselves, the others their dictionaries, and every result has the same
    bits as with dictionaries off."""
    print("\n[case] batch: constant-coefficient P2 systems and a small variable-coefficient P1", flush=True)
    kap, _ = problems.kappa_field(2)
    systems = [_system(34, 2, 17), _system(26, 2, 18), _system(12, 1, 19, kappa=kap)]
    mats, bs = [A for _, A, _ in systems], [b for _, _, b in systems]
    B = fd.Batch(mats)
    for prec, kw in {"fp32": dict(tol=1e-10, inner_tol=0.3), "fp64": dict(tol=1e-10, inner_tol=0.5)}.items():
        fd.value_dictionary(True)
        xs, info = B.solve(bs, precision=prec, **kw)
        assert info["converged"]
        d = B.dictionary(prec)
        assert d is not None and d[1] > 256, d
        fd.value_dictionary(False)
        ys, infoy = B.solve(bs, precision=prec, **kw)
        assert B.dictionary(prec) is None
        assert all(np.array_equal(a, c) for a, c in zip(xs, ys)), prec
        assert np.array_equal(info["iterations"], infoy["iterations"])
