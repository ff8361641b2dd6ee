"""Batched solves of many independent systems (the batching pattern; no reference counterpart).

    B = Batch([A1, A2, ...])                       # SparseMatrix members, any sizes
    xs, info = B.solve([b1, b2, ...], tol=1e-10)   # FP32 Jacobi-CG + reliable FP64 updates per system

All systems iterate in one launch sequence on the GPU (block-diagonal concatenation, per-system
scalars and convergence flags, the iteration replayed from a captured hipGraph)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .sparse import SparseMatrix


class Batch:
    def __init__(self, matrices):
        mats = list(matrices)
        if not mats or not all(isinstance(A, SparseMatrix) for A in mats):
            raise TypeError("Batch: a non-empty list of SparseMatrix")
        self._members = mats                               # kept alive for update_values()
        arr = (C.c_void_p * len(mats))(*[A._h for A in mats])
        h = C.c_void_p()
        L.check(L.load().femd_batch_create(len(mats), arr, C.byref(h)))
        self._h = h.value
        nr, nz = C.c_int64(), C.c_int64()
        off = np.zeros(len(mats) + 1, dtype=np.int64)
        L.check(L.load().femd_batch_info(self._h, C.byref(nr), C.byref(nz), off.ctypes.data))
        self.nrows, self.nnz, self.offsets = nr.value, nz.value, off

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                L.load().femd_batch_free(h)
            except Exception:
                pass

    def __len__(self):
        return len(self._members)

    def update_values(self):
        """Re-read the members' values (after they were re-assembled)."""
        L.check(L.load().femd_batch_update_values(self._h))

    def concat(self, vectors) -> np.ndarray:
        v = np.concatenate([np.asarray(b, dtype=np.float64).reshape(-1) for b in vectors])
        if v.size != self.nrows:
            raise ValueError(f"Batch: {v.size} values for {self.nrows} rows")
        return v

    def split(self, v):
        return [v[self.offsets[p]:self.offsets[p + 1]] for p in range(len(self))]

    def solve(self, bs, x0=None, *, tol=1e-10, inner_tol=0.3, max_outer=200, graph=True, precision="fp32"):
        """Returns (list of solutions, dict with per-system iterations and relative residuals).
        precision: "fp32" (FP32 inner CG with reliable FP64 updates) or "fp64" (FP64 inner CG)."""
        prec = {"fp32": L.FP32, "fp64": L.FP64}.get(precision)
        if prec is None:
            raise ValueError(f"Batch.solve: precision is 'fp32' or 'fp64', not {precision!r}")
        b = self.concat(bs)
        x = np.zeros(self.nrows) if x0 is None else self.concat(x0)
        its = np.zeros(len(self), dtype=np.int32)
        res = np.zeros(len(self))
        info = L.SolveInfo()
        L.check(L.load().femd_batch_solve(self._h, b, x, float(tol), float(inner_tol), int(max_outer), prec, int(bool(graph)),
                                          its.ctypes.data, res.ctypes.data, C.byref(info)))
        return self.split(x), dict(converged=bool(info.converged), iterations=its, residuals=res,
                                   refinement_rounds=int(info.breakdown), worst_residual=float(info.residual))

    def dictionary(self, precision="fp32"):
        """(distinct values over all tiles, most in one tile) of the value dictionary the last solve
        in `precision` read its matrix through, or None when it read the values themselves."""
        prec = {"fp32": L.FP32, "fp64": L.FP64}[precision]
        d, m = C.c_int64(), C.c_int()
        L.check(L.load().femd_batch_dictionary(self._h, prec, C.byref(d), C.byref(m)))
        return None if d.value < 0 else (d.value, m.value)


def inner_order(enable=None) -> bool:
    """Whether the inner CG works in its own order of the unknowns (each system's unknowns in the
    order its rows first meet them as columns, which keeps an SpMV tile's gathers close together);
    with `enable` set it first switches it on or off.  Returns the previous setting."""
    return bool(L.load().femd_inner_order(-1 if enable is None else int(bool(enable))))


def device_setup(enable=None) -> bool:
    """Whether the set-up (a space's sparsity pattern and gather lists, the inner CG's order of the
    unknowns with its reordered pattern, and the SpMV tiles' column lists) is built on the GPU
    rather than on the host; both give the same arrays, so results are bit-identical either way.  With `enable` set it first switches; returns the
    previous setting."""
    return bool(L.load().femd_device_setup(-1 if enable is None else int(bool(enable))))


def value_dictionary(enable=None) -> bool:
    """Whether the inner CG may read matrices through lossless per-tile value dictionaries (results
    are bit-identical either way); with `enable` set it first switches them on or off.  Returns the
    previous setting."""
    return bool(L.load().femd_value_dictionary(-1 if enable is None else int(bool(enable))))
