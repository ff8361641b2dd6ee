"""SparseMatrix and Preconditioner (python/femd/sparse.py:40-301), their store on the GPU.

A SparseMatrix holds FEMd's CSR layout (int32 indptr / indices, float64 data) in device memory;
indptr, indices and data come back as NumPy copies.  The products, preconditioners and Krylov
solvers run on the GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class SparseMatrix:
    """An assembled matrix in CSR form (sparse.py:40-185)."""

    def __init__(self, handle, space=None, row_space=None):
        self._h = handle
        self.space = space
        self.row_space = row_space if row_space is not None else space
        s = np.zeros(2, dtype=np.int32)
        nnz = C.c_int64()
        L.check(L.load().femd_csr_info(self._h, s, C.byref(nnz)))
        self._shape = (int(s[0]), int(s[1]))
        self._nnz = int(nnz.value)
        self.symmetric = False

    @classmethod
    def _on_space(cls, V):
        h = C.c_void_p()
        L.check(L.load().femd_csr_create_space(V._h, C.byref(h)))
        return cls(h.value, V, V)

    @classmethod
    def from_arrays(cls, nrows, ncols, indptr, indices, data, symmetric=None):
        ip, ix, v = L.i32(indptr), L.i32(indices), L.f64(data)
        if ip.size != nrows + 1 or ix.size != ip[-1] or v.size != ip[-1]:
            raise ValueError("CSRMatrix: indptr has nrows + 1 entries, indices and data indptr[-1]")
        h = C.c_void_p()
        L.check(L.load().femd_csr_create(int(nrows), int(ncols), ip, ix if ix.size else np.zeros(1, dtype=np.int32),
                                         v if v.size else np.zeros(1), C.byref(h)))
        M = cls(h.value)
        M.symmetric = M.is_symmetric() if symmetric is None else bool(symmetric)
        return M

    @classmethod
    def from_scipy(cls, A, space=None, symmetric=None):
        import scipy.sparse as sp
        A = sp.csr_matrix(A, dtype=np.float64)
        A.sum_duplicates()
        A.sort_indices()
        M = cls.from_arrays(A.shape[0], A.shape[1], A.indptr, A.indices, A.data, symmetric)
        M.space = M.row_space = space
        return M

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                L.load().femd_csr_free(h)
            except Exception:
                pass

    # ---- structure ---------------------------------------------------------------------
    @property
    def shape(self): return self._shape

    @property
    def nnz(self): return self._nnz

    def _download(self, what):
        ip = np.empty(self._shape[0] + 1, dtype=np.int32) if "p" in what else None
        ix = np.empty(max(self._nnz, 1), dtype=np.int32) if "i" in what else None
        v = np.empty(max(self._nnz, 1)) if "v" in what else None
        L.check(L.load().femd_csr_download(self._h, ip.ctypes.data if ip is not None else None,
                                           ix.ctypes.data if ix is not None else None,
                                           v.ctypes.data if v is not None else None))
        return ip, (ix[: self._nnz] if ix is not None else None), (v[: self._nnz] if v is not None else None)

    @property
    def indptr(self) -> np.ndarray: return self._download("p")[0]

    @property
    def indices(self) -> np.ndarray: return self._download("i")[1]

    @property
    def data(self) -> np.ndarray:
        """A copy of the stored values (assign through set_data)."""
        return self._download("v")[2]

    def set_data(self, values):
        v = L.f64(values)
        if v.size != self._nnz:
            raise ValueError(f"set_data: {v.size} values for {self._nnz} entries")
        L.check(L.load().femd_csr_upload_values(self._h, v if v.size else np.zeros(1)))

    def asymmetry(self) -> float:
        out = C.c_double()
        L.check(L.load().femd_csr_asymmetry(self._h, C.byref(out)))
        return out.value

    def is_symmetric(self, tol: float = 1e-12) -> bool:
        return self._shape[0] == self._shape[1] and self.asymmetry() <= tol

    def tocsr(self):
        import scipy.sparse as sp
        ip, ix, v = self._download("piv")
        A = sp.csr_matrix((v, ix, ip), shape=self._shape, copy=False)
        A.has_sorted_indices = True
        return A

    def toarray(self) -> np.ndarray:
        return self.tocsr().toarray()

    def diagonal(self) -> np.ndarray:
        d = np.empty(min(self._shape))
        L.check(L.load().femd_csr_diagonal(self._h, d))
        return d

    # ---- orderings ---------------------------------------------------------------------
    def ordering(self, kind: str = "rcm") -> np.ndarray:
        """p[k] = the unknown that becomes number k: "rcm" or "natural" (sparse.py:107-125)."""
        n = self._square("ordering")
        k = str(kind).lower()
        if k == "natural":
            return np.arange(n, dtype=np.int64)
        if k != "rcm":
            raise ValueError(f"ordering: kind is 'rcm' or 'natural' here, got {kind!r}")
        ip, ix, _ = self._download("pi")
        p = np.empty(n, dtype=np.int32)
        L.check(L.load().femd_rcm(n, ip, ix if ix.size else np.zeros(1, dtype=np.int32), p))
        return p.astype(np.int64)

    def bandwidth(self, p=None) -> int:
        n = self._square("bandwidth")
        ip, ix, _ = self._download("pi")
        rows = np.repeat(np.arange(n), np.diff(ip))
        if p is None:
            return int(np.abs(rows - ix).max(initial=0))
        pinv = np.empty(n, dtype=np.int64)
        pinv[np.asarray(p)] = np.arange(n)
        return int(np.abs(pinv[rows] - pinv[ix]).max(initial=0))

    def _square(self, what) -> int:
        if self._shape[0] != self._shape[1]:
            raise ValueError(f"{what}: the matrix must be square, it is {self._shape[0]} x {self._shape[1]}")
        return self._shape[0]

    # ---- products ----------------------------------------------------------------------
    def matvec(self, x) -> np.ndarray:
        v = L.f64(getattr(x, "vector", x)).reshape(-1)
        if v.size != self._shape[1]:
            raise ValueError(f"matvec: x has {v.size} entries, the matrix {self._shape[1]} columns")
        y = np.empty(self._shape[0])
        L.check(L.load().femd_spmv(self._h, v if v.size else np.zeros(1), y if y.size else np.zeros(1)))
        return y

    def __matmul__(self, x):
        return self.matvec(x)

    # ---- solvers -----------------------------------------------------------------------
    def preconditioner(self, kind: str = "ilu0", omega: float = 1.0, ordering: str = "rcm") -> "Preconditioner":
        return Preconditioner(self, kind, omega, ordering)

    def solver(self, backend: str = "cg", **options) -> "SparseSolver":
        return SparseSolver(self, backend, **options)

    def __repr__(self):
        sym = ", symmetric" if self.symmetric else ""
        return f"SparseMatrix({self._shape[0]} x {self._shape[1]}, nnz={self._nnz}{sym})"


class Preconditioner:
    """z = M^{-1} r on the GPU: kind "jacobi", "ssor" (omega) or "ilu0" (sparse.py:251-297).
    ordering="rcm" (the default) builds SSOR and ILU(0) on P A P^T with the reverse Cuthill-McKee
    order; Jacobi ignores it."""

    def __init__(self, A: SparseMatrix, kind: str = "ilu0", omega: float = 1.0, ordering: str = "rcm"):
        k = str(kind).lower().replace("(", "").replace(")", "")
        o = str(ordering).lower()
        if o not in ("rcm", "natural"):
            raise ValueError(f"preconditioner ordering must be 'rcm' or 'natural', got {ordering!r}")
        if A.shape[0] != A.shape[1]:
            raise ValueError("a preconditioner needs a square matrix")
        if k == "jacobi":
            kk = L.PC_JACOBI
        elif k in ("ssor", "sgs"):
            kk = L.PC_SSOR
            omega = float(omega) if k == "ssor" else 1.0
        elif k in ("ilu0", "ilu", "ic0"):
            kk = L.PC_ILU0
        else:
            raise ValueError(f"preconditioner kind must be 'jacobi', 'ssor' or 'ilu0', got {kind!r}")
        h = C.c_void_p()
        mp = C.c_double(0.0)
        L.check(L.load().femd_precond_create(A._h, kk, L.ORDER_RCM if o == "rcm" else L.ORDER_NATURAL, float(omega),
                                             C.byref(h), C.byref(mp)))
        self._h = h.value
        self._A = A
        self.kind = {L.PC_JACOBI: "jacobi", L.PC_SSOR: "ssor", L.PC_ILU0: "ilu0"}[kk]
        self.ordering = "natural" if kk == L.PC_JACOBI else o
        self.space = A.space
        self.size = A.shape[0]
        self._min_pivot = mp.value if kk == L.PC_ILU0 else None

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                L.load().femd_precond_free(h)
            except Exception:
                pass

    @property
    def min_pivot(self):
        """Smallest ILU(0) pivot (ILU(0) only)."""
        return self._min_pivot

    def stats(self) -> dict:
        """How the triangular sweeps run (no reference counterpart): planned (one workgroup's program
        per sweep) or on the level-set kernels, steps and dependency levels per sweep, stream bytes
        and the stored entries the two sweeps read."""
        out = np.zeros(8, dtype=np.int64)
        L.check(L.load().femd_precond_stats(self._h, out.ctypes.data))
        keys = ("planned", "steps_forward", "steps_backward", "levels_forward", "levels_backward", "bytes_forward",
                "bytes_backward", "entries")
        d = {k: int(v) for k, v in zip(keys, out)}
        d["planned"] = bool(d["planned"])
        return d

    def solve(self, r):
        v = L.f64(getattr(r, "vector", r)).reshape(-1)
        z = np.empty_like(v)
        L.check(L.load().femd_precond_apply(self._h, v, z))
        return z

    __call__ = solve

    def __repr__(self):
        return f"Preconditioner({self.kind}, {self.ordering}, n={self.size})"


class SparseSolver:
    """A.solver(backend): the Krylov backends "cg", "gmres", "lgmres" (sparse.py:433-462), with
    M= ("ilu0" by default, or a Preconditioner, or None), tol=, maxiter=.  The direct backends of
    the reference (cholesky, ldlt, lu, superlu, dense) are outside this path."""

    def __init__(self, A: SparseMatrix, backend: str = "cg", **options):
        b = str(backend).lower()
        if b not in ("cg", "gmres", "lgmres"):
            raise NotImplementedError(f"backend {backend!r}: only the Krylov backends 'cg', 'gmres', 'lgmres' run here")
        self.A, self.backend = A, b
        M = options.pop("M", "ilu0")
        if isinstance(M, str):
            M = A.preconditioner(M)
        self._M = M
        self._opts = options
        self.info = None

    @property
    def size(self): return self.A.shape[0]

    def solve(self, b):
        from . import krylov
        f = {"cg": krylov.cg, "gmres": krylov.gmres, "lgmres": krylov.lgmres}[self.backend]
        x, self.info = f(self.A, b, M=self._M, **self._opts)
        return x
