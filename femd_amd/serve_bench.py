"""The serving stream: Poisson / heat requests of uneven sizes on structured and unstructured meshes,
arriving as a Poisson process, assembled on arrival and served with continuous batching under an
HBM budget.

    python -m femd_amd.serve_bench [--count 200] [--rate 20] [--budget-gib 64] [--seed 0]
                                   [--space-at-arrival] [--assembly exact|mfma]

Workload: `count` requests, arrival times of a Poisson process of rate `rate` per second; each
request draws (from the seed) its degree k in {1, 2, 3}, its size (unknowns log-uniform in
[dof_min, dof_max]), its mesh -- with probability `unstructured` one of six seeded Delaunay meshes
of the unit square (h = 0.04 ... 0.00125, the one whose P_k space comes nearest the size), else a
structured N x N triangulation with N chosen to hit the size -- and its kind: Poisson
-Laplace u = f, variable-coefficient Poisson -div(kappa grad u) = f (kappa = exp(0.5 g), g 16
seeded Fourier modes), or one backward-Euler heat step (M + dt K) u = f with dt = 1e-3; f a
random-Fourier field of the request's own seed; every side Dirichlet.

Set-up generates the Delaunay meshes (Qhull) and, unless --space-at-arrival, each request's
Lagrange space with its sparsity pattern, as a server holds the geometries it serves.  In the
stream each request is taken at its arrival time: its matrix and right-hand side are assembled on
the GPU (with --space-at-arrival its space and pattern are built first), it is submitted, and its
latency runs from arrival to its solution back on the host, so assembly is inside every latency.
Then the same requests are assembled and solved one after another by the plain FP64 GPU path
(Jacobi-CG, tolerance tol) for the throughput baseline and the error reference.  One JSON line:
p50 / p99 latency, solves/s of the stream, of the plain path, the ratio, the assembly time spent
inside the stream, and the largest error against FP64.
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
MESH_HS = (0.04, 0.02, 0.01, 0.005, 0.0025, 0.00125)


def _delaunay_dofs(h, k):
    """About how many unknowns P_k has on delaunay_square(h): k^2 per vertex, a vertex per
    sqrt(3)/2 h^2 of area."""
    return k * k * 2.0 / (math.sqrt(3.0) * h * h)


def draw_requests(count, rate, dof_min, dof_max, seed, unstructured=0.5):
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0, count)           # drawn at every rate, so that the requests do not depend on it
    arrivals = np.cumsum(gaps) / rate if rate > 0 else np.zeros(count)
    out = []
    for i in range(count):
        k = int(rng.integers(1, 4))
        dofs = math.exp(rng.uniform(math.log(dof_min), math.log(dof_max)))
        kind = KINDS[int(rng.integers(0, 3))]
        r = dict(arrival=float(arrivals[i]), k=k, kind=kind, seed=seed * 1000003 + i)
        if rng.random() < unstructured:
            r["h"] = min(MESH_HS, key=lambda h: abs(math.log(_delaunay_dofs(h, k) / dofs)))
        else:
            r["N"] = max(2, int(round((math.sqrt(dofs) + 1.0) / k)))
        out.append(r)
    return out


def make_meshes(reqs, seed):
    """The Delaunay meshes the requests name, one per h (seeded by h's place in MESH_HS)."""
    import femd_amd as fd
    from femd_amd import problems
    out = {}
    for h in sorted({r["h"] for r in reqs if "h" in r}, reverse=True):
        P, T, _ = problems.delaunay_square(h, seed=seed * 31 + MESH_HS.index(h))
        out[h] = fd.mesh_from_arrays(P, T)
    return out


def describe(req, meshes):
    """The request's space and its bilinear and linear forms; the pattern is built here too (the
    matrix made on the space is dropped), so that assembly finds it."""
    import femd_amd as fd
    from femd_amd import problems
    from femd_amd.sparse import SparseMatrix
    mesh = meshes[req["h"]] if "h" in req else fd.rectangle_mesh(0, 0, 1, 1, req["N"], req["N"])
    V = fd.LagrangeSpace(mesh, req["k"], bc="dirichlet")
    SparseMatrix._on_space(V)
    f, _ = problems.fourier_source(req["seed"])
    u, v = fd.TrialFunction(V), fd.TestFunction(V)
    if req["kind"] == "heat":
        a = fd.form((u * v + fd.Const(1e-3) * fd.dot(fd.grad(u), fd.grad(v))) * fd.dx)
        Lf = fd.form(f * v * fd.dx)
    else:
        kap = problems.kappa_field(req["seed"])[0] if req["kind"] == "kappa" else None
        a, Lf = problems.poisson_forms(V, f, kappa=kap)
    return V, a, Lf


def assemble(a, Lf):
    return a.assemble(reuse=False), Lf.assemble(reuse=False)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=200)
    ap.add_argument("--rate", type=float, default=20.0, help="arrivals per second (0: all at once)")
    ap.add_argument("--dof-min", type=float, default=1e3)
    ap.add_argument("--dof-max", type=float, default=2e6)
    ap.add_argument("--unstructured", type=float, default=0.5, help="share of requests on Delaunay meshes")
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--max-systems", type=int, default=256)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--inner-tol", type=float, default=0.3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--assembly", choices=("exact", "mfma"), default="exact",
                    help="element matrices in the reference's operation order or on the FP64 matrix cores")
    ap.add_argument("--space-at-arrival", action="store_true",
                    help="build each request's space and pattern at its arrival (inside its latency)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 baseline and the error check")
    ap.add_argument("--original-order", action="store_true",
                    help="the inner CG keeps the original order of the unknowns instead of its own (first-met columns)")
    args = ap.parse_args(argv)

    import femd_amd as fd
    from femd_amd import _lib as L
    lib = L.load()
    L.check(lib.femd_init(0))
    lib.femd_inner_order(0 if args.original_order else 1)
    fd.assembly_mode(args.assembly)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)

    reqs = draw_requests(args.count, args.rate, args.dof_min, args.dof_max, args.seed, args.unstructured)
    n = len(reqs)
    t0 = time.perf_counter()
    meshes = make_meshes(reqs, args.seed)
    t_mesh = time.perf_counter() - t0
    spaces = None if args.space_at_arrival else [describe(r, meshes) for r in reqs]
    t_setup = time.perf_counter() - t0
    log(f"setup: {len(meshes)} Delaunay meshes in {t_mesh:.1f} s"
        + ("" if spaces is None else f", {n} spaces and patterns, {t_setup:.1f} s in all"))

    # warm-up: requests of every degree and kind through a server (code objects, pools)
    warm = fd.Server(hbm_budget=int(args.budget_gib * (1 << 30)), max_systems=args.max_systems,
                     precision=args.precision, tol=args.tol, inner_tol=args.inner_tol)
    seen = set()
    for i, r in enumerate(reqs):
        key = (r["k"], r["kind"], "h" in r)
        if key in seen:
            continue
        seen.add(key)
        _, a, Lf = spaces[i] if spaces is not None else describe(r, meshes)
        warm.submit(i, *assemble(a, Lf))
    warm.run()
    del warm

    S = fd.Server(hbm_budget=int(args.budget_gib * (1 << 30)), max_systems=args.max_systems, precision=args.precision,
                  tol=args.tol, inner_tol=args.inner_tol)
    arrivals = [r["arrival"] for r in reqs]
    lat, xs, its, conv = {}, {}, {}, {}
    dims, nnzs, t_asm = [0] * n, [0] * n, [0.0] * n
    most_groups = most_running = 0
    i = 0
    t0 = time.perf_counter()
    while len(lat) < n:
        now = time.perf_counter() - t0
        while i < n and arrivals[i] <= now:
            ta = time.perf_counter()
            _, a, Lf = spaces[i] if spaces is not None else describe(reqs[i], meshes)
            A, b = assemble(a, Lf)
            t_asm[i] = time.perf_counter() - ta
            dims[i], nnzs[i] = b.size, A.nnz
            S.submit(i, A, b)
            del A
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
    asm = np.array(t_asm)
    log(f"stream: {n} requests ({sum(dims)} unknowns, {sum(nnzs)} stored entries, sizes {min(dims)}..{max(dims)}) "
        f"in {makespan:.3f} s ({n / makespan:.2f} solves/s), latency p50 {1e3 * np.percentile(lats, 50):.1f} ms, "
        f"p99 {1e3 * np.percentile(lats, 99):.1f} ms, assembly {asm.sum():.3f} s in all (median "
        f"{1e3 * np.median(asm):.1f} ms), up to {most_groups} groups / {most_running} systems at once, "
        f"all converged {all(conv.values())}")

    plain = None
    if not args.no_plain:
        # the plain FP64 GPU path: each request assembled, then one Jacobi-CG solve, in turn
        def dev(nbytes):
            p = C.c_void_p()
            L.check(lib.femd_dev_alloc(max(nbytes, 8), C.byref(p)))
            return p.value
        errs, it_plain, t_plain_asm = [], 0, 0.0
        t1 = time.perf_counter()
        for t, r in enumerate(reqs):
            ta = time.perf_counter()
            _, a, Lf = spaces[t] if spaces is not None else describe(r, meshes)
            A, b = assemble(a, Lf)
            t_plain_asm += time.perf_counter() - ta
            bd, xd = dev(8 * b.size), dev(8 * b.size)
            L.check(lib.femd_memcpy_htod(bd, b.ctypes.data, 8 * b.size))
            L.check(lib.femd_memcpy_htod(xd, np.zeros(b.size).ctypes.data, 8 * b.size))
            h, mp = C.c_void_p(), C.c_double()
            L.check(lib.femd_precond_create(A._h, L.PC_JACOBI, 0, 1.0, C.byref(h), C.byref(mp)))
            info = L.SolveInfo()
            L.check(lib.femd_cg_dev(A._h, h.value, bd, xd, max(1000, 10 * b.size), args.tol, C.byref(info), None))
            it_plain += int(info.iterations)
            x64 = np.empty(b.size)
            L.check(lib.femd_memcpy_dtoh(x64.ctypes.data, xd, 8 * b.size))
            errs.append(float(np.linalg.norm(xs[t] - x64) / max(np.linalg.norm(x64), 1e-300)))
            lib.femd_precond_free(h.value)
            lib.femd_dev_free(bd)
            lib.femd_dev_free(xd)
            del A
        t_plain = time.perf_counter() - t1
        plain = {"seconds": round(t_plain, 4), "assembly_s": round(t_plain_asm, 4),
                 "solves_per_s": round(n / t_plain, 3), "iterations_total": it_plain,
                 "max_rel_err_vs_fp64": max(errs)}
        log(f"plain FP64: {t_plain:.3f} s ({n / t_plain:.2f} solves/s, assembly {t_plain_asm:.3f} s of it), "
            f"max error of the served solutions vs FP64 {max(errs):.2e}")

    out = {
        "workload": f"{n} Poisson / kappa-Poisson / heat requests, P1-P3, unknowns log-uniform in "
                    f"[{args.dof_min:.0f}, {args.dof_max:.0f}], {args.unstructured:.0%} on Delaunay meshes, "
                    f"Poisson arrivals at {args.rate}/s, assembled on arrival"
                    + (", spaces built on arrival" if spaces is None else ""),
        "seed": args.seed, "requests": n, "unknowns": int(sum(dims)), "nnz": int(sum(nnzs)),
        "setup_s": round(t_setup, 2), "budget_gib": args.budget_gib, "precision": args.precision, "tol": args.tol,
        "inner_tol": args.inner_tol, "assembly": args.assembly, "space_at_arrival": spaces is None,
        "stream": {"makespan_s": round(makespan, 4), "solves_per_s": round(n / makespan, 3),
                   "latency_ms_p50": round(1e3 * float(np.percentile(lats, 50)), 2),
                   "latency_ms_p99": round(1e3 * float(np.percentile(lats, 99)), 2),
                   "latency_ms_mean": round(1e3 * float(lats.mean()), 2),
                   "assembly_s_total": round(float(asm.sum()), 4),
                   "assembly_ms_p50": round(1e3 * float(np.median(asm)), 2),
                   "assembly_ms_max": round(1e3 * float(asm.max()), 2),
                   "max_groups": most_groups, "max_systems": most_running,
                   "inner_iterations_total": int(sum(its.values())), "all_converged": all(conv.values())},
        "plain_fp64": plain,
        "speedup_vs_plain": round((n / makespan) / plain["solves_per_s"], 3) if plain else None,
        "kinds": {k: sum(1 for r in reqs if r["kind"] == k) for k in KINDS},
        "degrees": {str(k): sum(1 for r in reqs if r["k"] == k) for k in (1, 2, 3)},
        "meshes": {"structured": sum(1 for r in reqs if "N" in r), "delaunay": sum(1 for r in reqs if "h" in r)},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
