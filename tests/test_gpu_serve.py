"""The serving loop on the GPU: requests of uneven sizes and kinds admitted by an HBM budget with
continuous batching; every result the same bit for bit alone and in the mix, and within 1e-8 of
the FP64 oracle solution."""
import numpy as np
import pytest

import femd_amd as fd
from femd_amd import problems

pytestmark = pytest.mark.gpu


def _request(kind, N, k, seed):
    V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, N, N), k, bc="dirichlet")
    f, _ = problems.fourier_source(seed)
    if kind == "heat":
        u, v = fd.TrialFunction(V), fd.TestFunction(V)
        A = fd.form((u * v + fd.Const(1e-3) * fd.dot(fd.grad(u), fd.grad(v))) * fd.dx).assemble()
        b = fd.form(f * v * fd.dx).assemble()
        return V, A, b
    kap = problems.kappa_field(seed)[0] if kind == "kappa" else None
    a, L = problems.poisson_forms(V, f, kappa=kap)
    return V, a.assemble(), L.assemble()


SPECS = [("poisson", 12, 2, 1), ("kappa", 20, 1, 2), ("heat", 9, 3, 3), ("poisson", 30, 2, 4), ("kappa", 7, 3, 5),
         ("poisson", 40, 1, 6), ("heat", 25, 2, 7), ("poisson", 5, 2, 8), ("kappa", 16, 2, 9), ("poisson", 18, 3, 10)]


def test_serve_mix_is_alone_and_matches_oracle(orc):
    print("\n[case] serving: 10 uneven requests in two waves, then each alone", flush=True)
    reqs = [_request(*s) for s in SPECS]
    S = fd.Server(hbm_budget=1 << 30, max_systems=64)
    for t in range(5):
        S.submit(t, reqs[t][1], reqs[t][2])
    mix = {}
    early = 0

    def take(rs):
        # each result as step() hands it out, while the rest of its group keeps running: its
        # output must already hold the solution (true residual at the tolerance), not the guess
        nonlocal early
        running = S.status()["running"]
        for r in rs:
            V, A, b = reqs[r.tag]
            res = np.linalg.norm(b - A.tocsr() @ r.x) / np.linalg.norm(b)
            assert res <= 1.01e-10, (r.tag, res)
            early += running > 0
            mix[r.tag] = r

    for _ in range(3):                    # the first wave runs before the second arrives
        take(S.step())
    assert S.status()["running"] > 0
    for t in range(5, len(reqs)):
        S.submit(t, reqs[t][1], reqs[t][2])
    while S.busy():
        take(S.step())
    assert early > 0
    assert sorted(mix) == list(range(len(reqs)))
    st = S.status()
    assert st["running"] == 0 and st["queued"] == 0 and st["bytes_in_use"] == 0
    for t, (V, A, b) in enumerate(reqs):
        r = mix[t]
        assert r.converged and r.residual <= 1e-10
        alone = fd.Server(hbm_budget=1 << 30, max_systems=64)
        alone.submit(0, A, b)
        a = alone.run()[0]
        assert np.array_equal(a.x, r.x) and a.iterations == r.iterations
        ip, ix = V.pattern()
        xo, _ = orc.cg(ip, ix, A.data, b, M="jacobi", tol=1e-13)
        assert np.linalg.norm(r.x - xo) / np.linalg.norm(xo) <= 1e-8


def test_serve_budget_admission():
    print("\n[case] serving: a budget that holds two requests at a time", flush=True)
    reqs = [_request("poisson", 24, 2, s) for s in range(6)]
    probe = fd.Server(hbm_budget=1 << 40)
    need = probe.submit(0, reqs[0][1], reqs[0][2])
    probe.run()
    S = fd.Server(hbm_budget=int(2.5 * need), max_systems=64)
    for t, (_, A, b) in enumerate(reqs):
        S.submit(t, A, b)
    done, most = {}, 0
    while S.busy():
        for r in S.step():
            done[r.tag] = r
        st = S.status()
        most = max(most, st["running"])
        assert st["bytes_in_use"] <= int(2.5 * need)
    assert most == 2 and len(done) == len(reqs) and all(r.converged for r in done.values())
    with pytest.raises(ValueError):
        fd.Server(hbm_budget=need // 2).submit(0, reqs[0][1], reqs[0][2])


def test_serve_edge_cases():
    print("\n[case] serving: zero right-hand side, converged guess, slot limit 1, fp64", flush=True)
    V, A, b = _request("poisson", 10, 2, 11)
    S = fd.Server(hbm_budget=1 << 30, max_systems=1)
    S.submit(0, A, np.zeros_like(b))
    S.submit(1, A, b)
    out = S.run()
    assert not out[0].x.any() and out[0].iterations == 0 and out[0].converged
    # the solution as the guess: nothing to do
    S.submit(2, A, b, out[1].x)
    again = S.run()[2]
    assert again.iterations == 0 and np.array_equal(again.x, out[1].x)
    S64 = fd.Server(hbm_budget=1 << 30, precision="fp64", tol=1e-11, inner_tol=0.5)
    S64.submit(0, A, b)
    r = S64.run()[0]
    assert r.converged and np.linalg.norm(r.x - out[1].x) <= 1e-8 * np.linalg.norm(r.x)
