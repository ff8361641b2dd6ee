"""The serving loop: independent solve requests of uneven sizes, admitted by an HBM budget and solved
with continuous batching on the device (no reference counterpart; the reference solves one system
per call).

    S = Server(hbm_budget=8 << 30, max_systems=64, tol=1e-10)
    S.submit(tag, A, b)              # A: an assembled SparseMatrix; b, x0: host vectors
    while S.busy():
        for r in S.step():           # one round of every running group; finished requests
            r.tag, r.x, r.iterations, r.residual, r.converged

Each step queues one round of iterations of every running group, admits the waiting requests that
fit the budget as a new group (which then overlaps with the running ones), retires every system
that has converged and frees groups whose systems have all finished.  A request's result is
bit-identical whatever else it is admitted with.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .sparse import SparseMatrix


@dataclass
class Finished:
    tag: int
    x: np.ndarray
    iterations: int
    residual: float
    converged: bool


class _DevVec:
    def __init__(self, host: np.ndarray):
        self.n = host.size
        p = C.c_void_p()
        L.check(L.load().femd_dev_alloc(max(8 * self.n, 8), C.byref(p)))
        self.p = p.value
        if self.n:
            L.check(L.load().femd_memcpy_htod(self.p, host.ctypes.data, 8 * self.n))

    def get(self) -> np.ndarray:
        out = np.empty(self.n)
        if self.n:
            L.check(L.load().femd_memcpy_dtoh(out.ctypes.data, self.p, 8 * self.n))
        return out

    def free(self):
        if self.p:
            L.load().femd_dev_free(self.p)
            self.p = None


class Server:
    def __init__(self, hbm_budget: int, max_systems: int = 64, *, precision: str = "fp32", tol: float = 1e-10,
                 inner_tol: float = 0.3, max_outer: int = 200):
        prec = {"fp32": L.FP32, "fp64": L.FP64}.get(precision)
        if prec is None:
            raise ValueError(f"Server: precision is 'fp32' or 'fp64', not {precision!r}")
        h = C.c_void_p()
        L.check(L.load().femd_server_create(int(hbm_budget), int(max_systems), prec, float(tol), float(inner_tol),
                                            int(max_outer), C.byref(h)))
        self._h = h.value
        self._live = {}                    # tag -> (A, b_dev, x_dev)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                L.load().femd_server_free(h)
            except Exception:
                pass
        for _, bd, xd in getattr(self, "_live", {}).values():
            bd.free()
            xd.free()

    def submit(self, tag: int, A: SparseMatrix, b, x0=None) -> int:
        """Queue A x = b (x0: the initial guess, zero by default); returns the device bytes the request
        holds while admitted.  Raises ValueError when it can never fit the budget."""
        if tag in self._live:
            raise ValueError(f"Server.submit: tag {tag} is already in use")
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).reshape(-1)
        if b.size != A.shape[0] or x.size != A.shape[0]:
            raise ValueError("Server.submit: b and x0 must have one entry per row of A")
        bd, xd = _DevVec(b), _DevVec(x)
        nb = C.c_int64()
        try:
            L.check(L.load().femd_server_submit(self._h, int(tag), A._h, bd.p, xd.p, C.byref(nb)))
        except Exception:
            bd.free()
            xd.free()
            raise
        self._live[tag] = (A, bd, xd)
        return nb.value

    def step(self, max_done: int = 1024):
        """One round; returns the requests that finished in it."""
        tags = np.zeros(max_done, dtype=np.int64)
        its = np.zeros(max_done, dtype=np.int32)
        res = np.zeros(max_done)
        conv = np.zeros(max_done, dtype=np.int32)
        nd = C.c_int()
        L.check(L.load().femd_server_step(self._h, max_done, tags.ctypes.data, its.ctypes.data, res.ctypes.data,
                                          conv.ctypes.data, C.byref(nd)))
        out = []
        for k in range(nd.value):
            _, bd, xd = self._live.pop(int(tags[k]))
            out.append(Finished(int(tags[k]), xd.get(), int(its[k]), float(res[k]), bool(conv[k])))
            bd.free()
            xd.free()
        return out

    def status(self) -> dict:
        q, r, g, b = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        L.check(L.load().femd_server_status(self._h, C.byref(q), C.byref(r), C.byref(g), C.byref(b)))
        return dict(queued=q.value, running=r.value, groups=g.value, bytes_in_use=b.value)

    def busy(self) -> bool:
        return bool(self._live)

    def run(self):
        """Step until every submitted request has finished; returns them by tag."""
        done = {}
        while self.busy():
            for r in self.step():
                done[r.tag] = r
        return done
