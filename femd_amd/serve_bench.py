"""The serving stream: Poisson / heat requests of uneven sizes arriving as a Poisson process,
served with continuous batching under an HBM budget.

    python -m femd_amd.serve_bench [--count 200] [--rate 20] [--budget-gib 64] [--seed 0]

Workload: `count` requests, arrival times of a Poisson process of rate `rate` per second; each
request draws (from the seed) its degree k in {1, 2, 3}, its size -- unknowns log-uniform in
[dof_min, dof_max], a structured N x N triangulation of the unit square with N chosen to hit it --
and its kind: Poisson -Laplace u = f, variable-coefficient Poisson -div(kappa grad u) = f
(kappa = exp(0.5 g), g 16 seeded Fourier modes), or one backward-Euler heat step
(M + dt K) u = M u0 with dt = 1e-3; f and u0 random-Fourier fields of the request's own seed;
every side Dirichlet.  Matrices and right-hand sides are assembled on the GPU before the stream
starts (reported as setup, not part of any latency).  During the stream each request is submitted
at its arrival time and its latency runs from arrival to its solution back on the host.  Then
the same requests are solved one after another by the plain FP64 GPU path (Jacobi-CG, tolerance
tol) for the throughput baseline and the error reference.  One JSON line: p50 / p99 latency,
solves/s of the stream, of the plain path, the ratio, and the largest error against FP64.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import sys
import time

import numpy as np

KINDS = ("poisson", "kappa", "heat")


def draw_requests(count, rate, dof_min, dof_max, seed):
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count)) if rate > 0 else np.zeros(count)
    out = []
    for i in range(count):
        k = int(rng.integers(1, 4))
        dofs = math.exp(rng.uniform(math.log(dof_min), math.log(dof_max)))
        N = max(2, int(round((math.sqrt(dofs) + 1.0) / k)))
        kind = KINDS[int(rng.integers(0, 3))]
        out.append(dict(arrival=float(arrivals[i]), k=k, N=N, kind=kind, seed=seed * 1000003 + i))
    return out


def build(req):
    import femd_amd as fd
    from femd_amd import problems
    V = fd.LagrangeSpace(fd.rectangle_mesh(0, 0, 1, 1, req["N"], req["N"]), req["k"], bc="dirichlet")
    f, _ = problems.fourier_source(req["seed"])
    u, v = fd.TrialFunction(V), fd.TestFunction(V)
    if req["kind"] == "heat":
        A = fd.form((u * v + fd.Const(1e-3) * fd.dot(fd.grad(u), fd.grad(v))) * fd.dx).assemble()
        b = fd.form(f * v * fd.dx).assemble()
    else:
        kap = problems.kappa_field(req["seed"])[0] if req["kind"] == "kappa" else None
        a, L = problems.poisson_forms(V, f, kappa=kap)
        A, b = a.assemble(), L.assemble()
    return V, A, b


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=200)
    ap.add_argument("--rate", type=float, default=20.0, help="arrivals per second (0: all at once)")
    ap.add_argument("--dof-min", type=float, default=1e3)
    ap.add_argument("--dof-max", type=float, default=2e6)
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--max-systems", type=int, default=256)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--inner-tol", type=float, default=0.3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 baseline and the error check")
    ap.add_argument("--original-order", action="store_true",
                    help="the inner CG keeps the original order of the unknowns instead of its own (first-met columns)")
    args = ap.parse_args(argv)

    import femd_amd as fd
    from femd_amd import _lib as L
    lib = L.load()
    L.check(lib.femd_init(0))
    lib.femd_inner_order(0 if args.original_order else 1)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)

    reqs = draw_requests(args.count, args.rate, args.dof_min, args.dof_max, args.seed)
    t0 = time.perf_counter()
    systems = [build(r) for r in reqs]
    dims = [V.dim for V, _, _ in systems]
    nnzs = [A.nnz for _, A, _ in systems]
    t_setup = time.perf_counter() - t0
    log(f"setup: {len(reqs)} requests, {sum(dims)} unknowns, {sum(nnzs)} stored entries, "
        f"sizes {min(dims)}..{max(dims)}, {t_setup:.1f} s")

    # warm-up: one request of every degree and kind through a server (code objects, pools)
    warm = fd.Server(hbm_budget=int(args.budget_gib * (1 << 30)), max_systems=args.max_systems,
                     precision=args.precision, tol=args.tol, inner_tol=args.inner_tol)
    for i, (V, A, b) in enumerate(systems[:6]):
        warm.submit(i, A, b)
    warm.run()
    del warm

    S = fd.Server(hbm_budget=int(args.budget_gib * (1 << 30)), max_systems=args.max_systems, precision=args.precision,
                  tol=args.tol, inner_tol=args.inner_tol)
    arrivals = [r["arrival"] for r in reqs]
    n = len(reqs)
    lat, xs, its, conv = {}, {}, {}, {}
    most_groups = most_running = 0
    i = 0
    t0 = time.perf_counter()
    while len(lat) < n:
        now = time.perf_counter() - t0
        while i < n and arrivals[i] <= now:
            S.submit(i, systems[i][1], systems[i][2])
            i += 1
        if S.busy():
            for r in S.step():
                lat[r.tag] = time.perf_counter() - t0 - arrivals[r.tag]
                xs[r.tag], its[r.tag], conv[r.tag] = r.x, r.iterations, r.converged
            st = S.status()
            most_groups, most_running = max(most_groups, st["groups"]), max(most_running, st["running"])
        elif i < n:
            time.sleep(max(0.0, arrivals[i] - (time.perf_counter() - t0)))
    t_end = time.perf_counter() - t0
    makespan = t_end - arrivals[0]
    lats = np.array([lat[t] for t in range(n)])
    log(f"stream: {n} requests in {makespan:.3f} s ({n / makespan:.2f} solves/s), latency p50 "
        f"{1e3 * np.percentile(lats, 50):.1f} ms, p99 {1e3 * np.percentile(lats, 99):.1f} ms, "
        f"up to {most_groups} groups / {most_running} systems at once, all converged {all(conv.values())}")

    plain = None
    if not args.no_plain:
        # the plain FP64 GPU path: one Jacobi-CG solve per request in turn, vectors resident
        def dev(nbytes):
            p = C.c_void_p()
            L.check(lib.femd_dev_alloc(max(nbytes, 8), C.byref(p)))
            return p.value
        bdev, xdev, precs = [], [], []
        for V, A, b in systems:
            bd, xd = dev(8 * b.size), dev(8 * b.size)
            L.check(lib.femd_memcpy_htod(bd, b.ctypes.data, 8 * b.size))
            h, mp = C.c_void_p(), C.c_double()
            L.check(lib.femd_precond_create(A._h, L.PC_JACOBI, 0, 1.0, C.byref(h), C.byref(mp)))
            bdev.append(bd)
            xdev.append(xd)
            precs.append(h.value)
        zeros = np.zeros(max(dims))

        def plain_round():
            total = 0
            for (V, A, b), bd, xd, M in zip(systems, bdev, xdev, precs):
                L.check(lib.femd_memcpy_htod(xd, zeros.ctypes.data, 8 * b.size))
                info = L.SolveInfo()
                L.check(lib.femd_cg_dev(A._h, M, bd, xd, max(1000, 10 * b.size), args.tol, C.byref(info), None))
                total += int(info.iterations)
            L.check(lib.femd_stream_sync(None))
            return total
        t1 = time.perf_counter()
        it_plain = plain_round()
        t_plain = time.perf_counter() - t1
        errs = []
        for t, ((V, A, b), xd) in enumerate(zip(systems, xdev)):
            x64 = np.empty(b.size)
            L.check(lib.femd_memcpy_dtoh(x64.ctypes.data, xd, 8 * b.size))
            errs.append(float(np.linalg.norm(xs[t] - x64) / max(np.linalg.norm(x64), 1e-300)))
        for p in bdev + xdev:
            lib.femd_dev_free(p)
        for M in precs:
            lib.femd_precond_free(M)
        plain = {"seconds": round(t_plain, 4), "solves_per_s": round(n / t_plain, 3), "iterations_total": it_plain,
                 "max_rel_err_vs_fp64": max(errs)}
        log(f"plain FP64: {t_plain:.3f} s ({n / t_plain:.2f} solves/s), max error of the served solutions vs FP64 "
            f"{max(errs):.2e}")

    out = {
        "workload": f"{n} Poisson / kappa-Poisson / heat requests, P1-P3, unknowns log-uniform in "
                    f"[{args.dof_min:.0f}, {args.dof_max:.0f}], Poisson arrivals at {args.rate}/s",
        "seed": args.seed, "requests": n, "unknowns": int(sum(dims)), "nnz": int(sum(nnzs)),
        "setup_s": round(t_setup, 2), "budget_gib": args.budget_gib, "precision": args.precision, "tol": args.tol,
        "inner_tol": args.inner_tol,
        "stream": {"makespan_s": round(makespan, 4), "solves_per_s": round(n / makespan, 3),
                   "latency_ms_p50": round(1e3 * float(np.percentile(lats, 50)), 2),
                   "latency_ms_p99": round(1e3 * float(np.percentile(lats, 99)), 2),
                   "latency_ms_mean": round(1e3 * float(lats.mean()), 2),
                   "max_groups": most_groups, "max_systems": most_running,
                   "inner_iterations_total": int(sum(its.values())), "all_converged": all(conv.values())},
        "plain_fp64": plain,
        "speedup_vs_plain": round((n / makespan) / plain["solves_per_s"], 3) if plain else None,
        "kinds": {k: sum(1 for r in reqs if r["kind"] == k) for k in KINDS},
        "degrees": {str(k): sum(1 for r in reqs if r["k"] == k) for k in (1, 2, 3)},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
