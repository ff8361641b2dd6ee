"""The oracle: CPU checkers of the HIP product, for tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py only.  The product (femd_amd) never imports this package.

Two libraries sit behind one Python interface:

* ``Oracle``  -- oracle/lib/liboracle.so, the C++ restatement in oracle/femd_oracle.cpp;
* ``Ref``     -- oracle/_ref/libfemd_ref.so, the reference's own header-only C++ compiled in
                 place by ``make -C oracle ref`` (only where the reference sources exist).

Both are built by ``make -C oracle`` (see oracle/Makefile).  Every array crosses as a NumPy
array: int32 for indices, float64 for values.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "lib" / "liboracle.so"
REF_SO = HERE / "_ref" / "libfemd_ref.so"

_P = C.c_void_p
_I = C.c_int
_L = C.c_long
_D = C.c_double
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

DIAG = {"right": 0, "left": 1, "alternate": 2, "crossed": 3}


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class _Lib:
    prefix = ""

    def __init__(self, path):
        self.lib = C.CDLL(str(path))
        self._declare()

    def _fn(self, name, res, *args):
        f = getattr(self.lib, self.prefix + name)
        f.restype, f.argtypes = res, list(args)
        return f

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self._err())
        return rc

    def _err(self):
        f = getattr(self.lib, self.prefix + "last_error")
        f.restype = C.c_char_p
        return f().decode()

    def ref_nodes(self, k, equispaced=True, quad=False):
        """(xi, eta) of the local nodes of the P_k (or, quad=True, Q_k) reference element."""
        n = (k + 1) ** 2 if quad else (k + 1) * (k + 2) // 2
        xi, eta = np.empty(n), np.empty(n)
        self._check(self._nodes(int(k), int(equispaced), int(quad), xi, eta))
        return xi, eta

    def lobatto_points(self, p):
        out = np.empty(p + 1)
        self._check(self._lgl(int(p), out))
        return out

    @staticmethod
    def _marks(markers):
        all_ = markers is None
        m = _i32(list(markers)) if markers is not None and len(markers) else np.zeros(1, dtype=np.int32)
        return (0 if all_ else len(markers)), m, int(all_)

    def facet_points(self, P, degree, markers=None):
        """(x, y, w, nx, ny) of the ds point set: markers None = every exterior facet."""
        nm, m, all_ = self._marks(markers)
        sz = np.zeros(2, dtype=np.int32)
        self._check(self._fsz(P.h, int(degree), nm, m, all_, sz))
        out = [np.empty(int(sz[0]) * int(sz[1])) for _ in range(5)]
        self._check(self._fpts(P.h, int(degree), nm, m, all_, *out))
        return tuple(out)

    def _groups(self, P, groups):
        """groups: [(kind 'dx' | 'ds', degree, markers or None, a, b or None, coeffs (nterms x npts))]."""
        kind = _i32([0 if g[0] == "dx" else 1 for g in groups])
        deg = _i32([g[1] for g in groups])
        nmark, marks, all_ = [], [], []
        for g in groups:
            if g[0] == "ds" and g[2] is not None:
                nmark.append(len(g[2])); marks.extend(g[2]); all_.append(0)
            else:
                nmark.append(0); all_.append(1 if g[0] == "ds" else 0)
        nterms = _i32([len(g[3]) for g in groups])
        a = _i32([x for g in groups for x in g[3]])
        b = _i32([x for g in groups for x in (g[4] if g[4] is not None else [0] * len(g[3]))])
        co = _f64(np.concatenate([np.asarray(g[5], dtype=np.float64).reshape(-1) for g in groups]))
        return (len(groups), kind, deg, _i32(nmark), _i32(marks) if marks else np.zeros(1, dtype=np.int32), _i32(all_), nterms,
                a, b, co)

    def assemble_matrix_groups(self, P, groups):
        n, kind, deg, nmark, marks, all_, nterms, a, b, co = self._groups(P, groups)
        out = np.empty(int(self._pnnz(P.h)))
        self._check(self._amg(P.h, n, kind, deg, nmark, marks, all_, nterms, a, b, co, out))
        return out

    def assemble_vector_groups(self, P, groups):
        n, kind, deg, nmark, marks, all_, nterms, a, b, co = self._groups(P, groups)
        out = np.empty(P.dim)
        self._check(self._avg(P.h, n, kind, deg, nmark, marks, all_, nterms, a, co, out))
        return out

    # -- shared CSR / Krylov entry points (same signatures in both libraries) ------------
    def _declare(self):
        self._spmv = self._fn("spmv", _I, _I, _ip, _ip, _dp, _dp, _dp)
        self._ddot = self._fn("ddot", _D, _L, _dp, _dp)
        self._rcm = self._fn("rcm", _I, _I, _ip, _ip, _ip)
        self._pc = self._fn("precond_apply", _I, _I, _ip, _ip, _dp, _I, _I, _D, _dp, _dp, C.POINTER(_D))
        self._cg = self._fn("cg", _I, _I, _ip, _ip, _dp, _I, _I, _D, _dp, _dp, _I, _D, _ip, C.POINTER(_D))
        self._kr = self._fn("krylov", _I, _I, _ip, _ip, _dp, _I, _I, _D, _dp, _dp, _I, _I, _I, _D, _I, _ip, _dp)
        self._ieval = None
        self._threads = self._fn("set_threads", None, _I)

    def set_threads(self, n):
        self._threads(int(n))

    def spmv(self, ip, ix, v, x):
        ip, ix, v, x = _i32(ip), _i32(ix), _f64(v), _f64(x)
        y = np.empty(ip.size - 1)
        self._check(self._spmv(ip.size - 1, ip, ix, v, x, y))
        return y

    def ddot(self, x, y):
        x, y = _f64(x), _f64(y)
        return float(self._ddot(x.size, x, y))

    def rcm(self, ip, ix):
        ip, ix = _i32(ip), _i32(ix)
        p = np.empty(ip.size - 1, dtype=np.int32)
        self._check(self._rcm(ip.size - 1, ip, ix, p))
        return p

    def precond_apply(self, ip, ix, v, r, kind="ilu0", ordering="rcm", omega=1.0):
        ip, ix, v, r = _i32(ip), _i32(ix), _f64(v), _f64(r)
        z = np.empty_like(r)
        mp = _D(0.0)
        k = {"jacobi": 1, "ilu0": 2, "ssor": 3}[kind]
        self._check(self._pc(ip.size - 1, ip, ix, v, k, int(ordering == "rcm"), float(omega), r, z, C.byref(mp)))
        return z, mp.value

    def cg(self, ip, ix, v, b, x0=None, M=None, ordering="rcm", omega=1.0, maxiter=None, tol=1e-10):
        """fd.cg semantics: M in (None, "jacobi", "ilu0", "ssor").  Returns x, dict."""
        ip, ix, v, b = _i32(ip), _i32(ix), _f64(v), _f64(b)
        n = b.size
        x = np.zeros(n) if x0 is None else _f64(x0).copy()
        maxiter = max(1000, 10 * n) if maxiter is None else int(maxiter)
        k = {None: 0, "jacobi": 1, "ilu0": 2, "ssor": 3}[M]
        info = np.zeros(3, dtype=np.int32)
        res = _D(0.0)
        self._check(self._cg(n, ip, ix, v, k, int(ordering == "rcm"), float(omega), b, x, maxiter, float(tol), info, C.byref(res)))
        return x, dict(converged=bool(info[0]), iterations=int(info[1]), breakdown=bool(info[2]), residual=res.value)

    def krylov(self, ip, ix, v, b, x0=None, M=None, ordering="rcm", omega=1.0, lgmres=False, restart=30,
               maxiter=None, tol=1e-10, k_aug=2):
        ip, ix, v, b = _i32(ip), _i32(ix), _f64(v), _f64(b)
        n = b.size
        x = np.zeros(n) if x0 is None else _f64(x0).copy()
        restart = int(min(restart, max(n, 1)))
        maxiter = max(1000, 10 * restart) if maxiter is None else int(maxiter)
        k = {None: 0, "jacobi": 1, "ilu0": 2, "ssor": 3}[M]
        info = np.zeros(2, dtype=np.int32)
        res = np.zeros(2)
        self._check(self._kr(n, ip, ix, v, k, int(ordering == "rcm"), float(omega), b, x, int(lgmres), restart, maxiter,
                             float(tol), int(k_aug), info, res))
        return x, dict(converged=bool(info[0]), iterations=int(info[1]), residual=float(res[0]),
                       precond_residual=float(res[1]))


class OracleProblem:
    """A mesh plus a P_k space held by one of the libraries."""

    def __init__(self, owner, handle, free, quad=False):
        self.o, self.h, self._free, self.quad = owner, handle, free, quad
        self.nv = 4 if quad else 3
        s = self.sizes()
        self.ncells, self.nloc, self.raw_dim, self.dim = s[0], s[1], s[2], s[3]

    def __del__(self):
        try:
            if self.h:
                self._free(self.h)
        except Exception:
            pass

    def sizes(self):
        return self.o._sizes(self.h)

    def numbering(self):
        return self.o._numbering(self)

    def pattern(self):
        return self.o._pattern(self)

    def cache_points(self, degree):
        return self.o._cache_points(self, degree)

    def assemble_matrix(self, degree, a, b, coeffs):
        return self.o._assemble_matrix(self, degree, a, b, coeffs)

    def assemble_vector(self, degree, a, coeffs):
        return self.o._assemble_vector(self, degree, a, coeffs)

    def eval_field(self, degree, c, codes):
        """A known field with coefficients c at the cell points: (len(codes), ncells * nq)."""
        return self.o._eval_field(self, degree, c, codes)


class Oracle(_Lib):
    prefix = "orc_"

    def __init__(self, path=ORACLE_SO):
        super().__init__(path)
        self._prect = self._fn("problem_rectangle", _P, _D, _D, _D, _D, _I, _I, _I, _I, _I, _ip, _I, _I)
        self._parr = self._fn("problem_arrays", _P, _I, _dp, _I, _ip, _I, _ip, C.c_void_p, _I, _I, _ip, _I, _I)
        self._pfree = self._fn("problem_free", None, _P)
        self._psizes = self._fn("problem_sizes", _I, _P, _ip)
        self._pnum = self._fn("numbering", _I, _P, _ip, _ip, _ip, _dp)
        self._pnnz = self._fn("pattern_nnz", _L, _P)
        self._ppat = self._fn("pattern", _I, _P, _ip, _ip)
        self._quad = self._fn("quadrature", _I, _I, C.POINTER(_I), _dp, _dp, _dp)
        self._basis = self._fn("basis_eval", _I, _I, _I, _D, _D, _dp)
        self._cpts = self._fn("cache_points", _I, _P, _I, _dp, _dp, _dp)
        self._amat = self._fn("assemble_matrix", _I, _P, _I, _I, _ip, _ip, _dp, _dp)
        self._avec = self._fn("assemble_vector", _I, _P, _I, _I, _ip, _dp, _dp)
        self._ie = self._fn("integrand_eval", _I, _I, _ip, _I, _dp, _L, _dp, _dp, _dp)
        self._ilu = self._fn("ilu0_factor", _I, _I, _ip, _ip, _dp, _I, _ip, _ip, _ip, _dp, C.POINTER(_D))
        self._marr = self._fn("mesh_arrays", _I, _P, _dp, _ip, _ip, _ip, _ip)
        self._pqrect = self._fn("problem_quad_rectangle", _P, _D, _D, _D, _D, _I, _I, _I, _I, _ip, _I, _I)
        self._pqarr = self._fn("problem_quad_arrays", _P, _I, _dp, _I, _ip, _I, _ip, C.c_void_p, _I, _I, _ip, _I, _I)
        self._qquad = self._fn("quadrature_quad", _I, _I, C.POINTER(_I), _dp, _dp, _dp)
        self._qbasis = self._fn("basis_eval_quad", _I, _I, _I, _D, _D, _dp)
        self._nodes = self._fn("ref_nodes", _I, _I, _I, _I, _dp, _dp)
        self._lgl = self._fn("lobatto_points", _I, _I, _dp)
        self._fsz = self._fn("facet_sizes", _I, _P, _I, _I, _ip, _I, _ip)
        self._fpts = self._fn("facet_points", _I, _P, _I, _I, _ip, _I, _dp, _dp, _dp, _dp, _dp)
        self._amg = self._fn("assemble_matrix_groups", _I, _P, _I, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp)
        self._avg = self._fn("assemble_vector_groups", _I, _P, _I, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp)
        self._ien = self._fn("integrand_eval_n", _I, _I, _ip, _I, _dp, _L, _dp, _dp, _dp, _dp, _dp)
        self._iei = self._fn("integrand_eval_in", _I, _I, _ip, _I, _dp, _L, _dp, _dp, _I, _dp, _dp)
        self._evf = self._fn("eval_field", _I, _P, _I, _dp, _I, _ip, _dp)

    def quad_rectangle(self, nx, ny, k, dirichlet=(), all_exterior=True, x0=0.0, y0=0.0, x1=1.0, y1=1.0, equispaced=True):
        d = _i32(list(dirichlet)) if len(dirichlet) else np.zeros(1, dtype=np.int32)
        h = self._pqrect(x0, y0, x1, y1, nx, ny, k, len(dirichlet), d, int(all_exterior), int(equispaced))
        if not h:
            raise RuntimeError(self._err())
        return OracleProblem(self, h, self._pfree, quad=True)

    def quad_from_arrays(self, points, quads, segments=None, markers=None, k=1, dirichlet=(), all_exterior=True,
                         equispaced=True):
        pts, q = _f64(points).reshape(-1), _i32(quads).reshape(-1)
        segs = _i32(segments).reshape(-1) if segments is not None else np.zeros(0, dtype=np.int32)
        mk = _i32(markers) if markers is not None else None
        d = _i32(list(dirichlet)) if len(dirichlet) else np.zeros(1, dtype=np.int32)
        h = self._pqarr(pts.size // 2, pts, q.size // 4, q, segs.size // 2, segs, mk.ctypes.data if mk is not None else None,
                        k, len(dirichlet), d, int(all_exterior), int(equispaced))
        if not h:
            raise RuntimeError(self._err())
        return OracleProblem(self, h, self._pfree, quad=True)

    def quad_quadrature(self, degree):
        xi, eta, w = np.empty(4096), np.empty(4096), np.empty(4096)
        nq = _I(0)
        self._check(self._qquad(int(degree), C.byref(nq), xi, eta, w))
        return xi[: nq.value].copy(), eta[: nq.value].copy(), w[: nq.value].copy()

    def quad_basis_eval(self, k, xi, eta, equispaced=True):
        n = (k + 1) ** 2
        out = np.empty(3 * n)
        self._check(self._qbasis(int(k), int(equispaced), float(xi), float(eta), out))
        return out.reshape(3, n)

    def rectangle(self, nx, ny, k, diagonal="right", dirichlet=(), all_exterior=True, x0=0.0, y0=0.0, x1=1.0, y1=1.0,
                  equispaced=True):
        d = _i32(list(dirichlet)) if len(dirichlet) else np.zeros(1, dtype=np.int32)
        h = self._prect(x0, y0, x1, y1, nx, ny, DIAG[diagonal], k, len(dirichlet), d, int(all_exterior), int(equispaced))
        if not h:
            raise RuntimeError(self._err())
        return OracleProblem(self, h, self._pfree)

    def from_arrays(self, points, triangles, segments=None, markers=None, k=1, dirichlet=(), all_exterior=True,
                    equispaced=True):
        pts, tris = _f64(points).reshape(-1), _i32(triangles).reshape(-1)
        segs = _i32(segments).reshape(-1) if segments is not None else np.zeros(0, dtype=np.int32)
        mk = _i32(markers) if markers is not None else None
        d = _i32(list(dirichlet)) if len(dirichlet) else np.zeros(1, dtype=np.int32)
        h = self._parr(pts.size // 2, pts, tris.size // 3, tris, segs.size // 2, segs,
                       mk.ctypes.data if mk is not None else None, k, len(dirichlet), d, int(all_exterior), int(equispaced))
        if not h:
            raise RuntimeError(self._err())
        return OracleProblem(self, h, self._pfree)

    def _sizes(self, h):
        out = np.zeros(7, dtype=np.int32)
        self._check(self._psizes(h, out))
        return out

    def mesh_arrays(self, P):
        s = P.sizes()
        pts, tris, nbr = np.empty(2 * s[5]), np.empty(P.nv * s[0], dtype=np.int32), np.empty(P.nv * s[0], dtype=np.int32)
        segs, mk = np.empty(2 * s[6], dtype=np.int32), np.empty(max(s[6], 1), dtype=np.int32)
        self._check(self._marr(P.h, pts, tris, nbr, segs, mk))
        return pts.reshape(-1, 2), tris.reshape(-1, P.nv), nbr.reshape(-1, P.nv), segs.reshape(-1, 2), mk[: s[6]]

    def _numbering(self, P):
        cd = np.empty(P.ncells * P.nloc, dtype=np.int32)
        cr = np.empty_like(cd)
        a2r = np.empty(max(P.dim, 1), dtype=np.int32)
        xy = np.empty(2 * P.raw_dim)
        self._check(self._pnum(P.h, cd, cr, a2r, xy))
        return dict(cell_dof=cd.reshape(P.ncells, P.nloc), cell_raw=cr.reshape(P.ncells, P.nloc),
                    ad2raw=a2r[: P.dim], raw_xy=xy.reshape(-1, 2))

    def _pattern(self, P):
        nnz = int(self._pnnz(P.h))
        ip, ix = np.empty(P.dim + 1, dtype=np.int32), np.empty(max(nnz, 1), dtype=np.int32)
        self._check(self._ppat(P.h, ip, ix))
        return ip, ix[:nnz]

    def quadrature(self, degree):
        xi, eta, w = np.empty(256), np.empty(256), np.empty(256)
        nq = _I(0)
        self._check(self._quad(int(degree), C.byref(nq), xi, eta, w))
        return xi[: nq.value].copy(), eta[: nq.value].copy(), w[: nq.value].copy()

    def basis_eval(self, k, xi, eta, equispaced=True):
        n = (k + 1) * (k + 2) // 2
        out = np.empty(3 * n)
        self._check(self._basis(int(k), int(equispaced), float(xi), float(eta), out))
        return out.reshape(3, n)

    def _npts(self, P, degree):
        return P.ncells * len((self.quad_quadrature if P.quad else self.quadrature)(degree)[2])

    def _cache_points(self, P, degree):
        m = self._npts(P, degree)
        x, y, w = np.empty(m), np.empty(m), np.empty(m)
        self._check(self._cpts(P.h, int(degree), x, y, w))
        return x, y, w

    def _assemble_matrix(self, P, degree, a, b, coeffs):
        nnz = int(self._pnnz(P.h))
        out = np.empty(nnz)
        c = _f64(coeffs).reshape(-1)
        self._check(self._amat(P.h, int(degree), len(a), _i32(a), _i32(b), c, out))
        return out

    def _assemble_vector(self, P, degree, a, coeffs):
        out = np.empty(P.dim)
        c = _f64(coeffs).reshape(-1)
        self._check(self._avec(P.h, int(degree), len(a), _i32(a), c, out))
        return out

    def integrand_eval(self, code, nreg, consts, xs, ys):
        code = _i32(code).reshape(-1)
        xs, ys = _f64(xs), _f64(ys)
        out = np.empty(xs.size)
        self._check(self._ie(code.size // 4, code, int(nreg), _f64(consts) if len(consts) else np.zeros(1), xs.size, xs, ys, out))
        return out

    def integrand_eval_n(self, code, nreg, consts, xs, ys, nxs, nys):
        code = _i32(code).reshape(-1)
        xs, ys, nxs, nys = _f64(xs), _f64(ys), _f64(nxs), _f64(nys)
        out = np.empty(xs.size)
        self._check(self._ien(code.size // 4, code, int(nreg), _f64(consts) if len(consts) else np.zeros(1), xs.size, xs, ys,
                              nxs, nys, out))
        return out

    def integrand_eval_in(self, code, nreg, consts, xs, ys, inputs):
        """The program with inputs: inputs is (ninput, npts)."""
        code = _i32(code).reshape(-1)
        xs, ys = _f64(xs), _f64(ys)
        inp = _f64(inputs).reshape(-1, xs.size) if len(inputs) else np.zeros((0, xs.size))
        out = np.empty(xs.size)
        self._check(self._iei(code.size // 4, code, int(nreg), _f64(consts) if len(consts) else np.zeros(1), xs.size, xs, ys,
                              inp.shape[0], np.ascontiguousarray(inp).reshape(-1) if inp.size else np.zeros(1), out))
        return out

    def _eval_field(self, P, degree, c, codes):
        npts = P.cache_points(degree)[0].size
        codes = _i32(codes)
        out = np.empty((codes.size, npts))
        self._check(self._evf(P.h, int(degree), _f64(c), codes.size, codes, out.reshape(-1)))
        return out

    def ilu0_factor(self, ip, ix, v, ordering="rcm"):
        ip, ix, v = _i32(ip), _i32(ix), _f64(v)
        n = ip.size - 1
        perm, ip2, ix2, lu = np.empty(n, dtype=np.int32), np.empty(n + 1, dtype=np.int32), np.empty_like(ix), np.empty_like(v)
        mp = _D(0.0)
        self._check(self._ilu(n, ip, ix, v, int(ordering == "rcm"), perm, ip2, ix2, lu, C.byref(mp)))
        return perm, ip2, ix2, lu, mp.value


class Ref(_Lib):
    prefix = "ref_"

    def __init__(self, path=REF_SO):
        super().__init__(path)
        self._mrect = self._fn("mesh_rectangle", _P, _D, _D, _D, _D, _I, _I, C.c_char_p)
        self._marr = self._fn("mesh_from_arrays", _P, _I, _dp, _I, _ip, _I, _ip, C.c_void_p)
        self._mfree = self._fn("mesh_free", None, _P)
        self._msizes = self._fn("mesh_sizes", _I, _P, _ip)
        self._mget = self._fn("mesh_arrays", _I, _P, _dp, _ip, _ip, _ip, _ip)
        self._screate = self._fn("space_create", _P, _P, _I, _I, _ip, _I, _I)
        self._sfree = self._fn("space_free", None, _P)
        self._ssizes = self._fn("space_sizes", _I, _P, _ip)
        self._snum = self._fn("space_numbering", _I, _P, _ip, _ip, _ip, _dp)
        self._pnnz = self._fn("pattern_nnz", _L, _P)
        self._ppat = self._fn("pattern", _I, _P, _ip, _ip)
        self._quad = self._fn("quadrature", _I, _I, C.POINTER(_I), _dp, _dp, _dp)
        self._basis = self._fn("basis_eval", _I, _I, _I, _D, _D, _dp)
        self._cpts = self._fn("cache_points", _I, _P, _I, _dp, _dp, _dp)
        self._amat = self._fn("assemble_matrix", _I, _P, _I, _I, _ip, _ip, _dp, _dp)
        self._avec = self._fn("assemble_vector", _I, _P, _I, _I, _ip, _dp, _dp)
        self._ie = self._fn("integrand_eval", _I, _I, _ip, _I, _I, _dp, _L, _dp, _dp, _dp)
        self._asym = self._fn("asymmetry", _D, _I, _ip, _ip, _dp)
        self._qcreate = self._fn("quad_space_create", _P, _I, _dp, _I, _ip, _I, _ip, C.c_void_p, _I, _I, _ip, _I, _I)
        self._qquads = self._fn("quad_mesh_quads", _I, _P, _ip)
        self._qquad = self._fn("quad_quadrature", _I, _I, C.POINTER(_I), _dp, _dp, _dp)
        self._qbasis = self._fn("quad_basis_eval", _I, _I, _I, _D, _D, _dp)
        self._nodes = self._fn("nodes", _I, _I, _I, _I, _dp, _dp)
        self._lgl = self._fn("lobatto_points", _I, _I, _dp)
        self._fsz = self._fn("facet_sizes", _I, _P, _I, _I, _ip, _I, _ip)
        self._fpts = self._fn("facet_points", _I, _P, _I, _I, _ip, _I, _dp, _dp, _dp, _dp, _dp)
        self._amg = self._fn("assemble_matrix_groups", _I, _P, _I, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp)
        self._avg = self._fn("assemble_vector_groups", _I, _P, _I, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp)
        self._ien = self._fn("integrand_eval_n", _I, _I, _ip, _I, _I, _dp, _L, _dp, _dp, _dp, _dp, _dp)
        self._iei = self._fn("integrand_eval_in", _I, _I, _ip, _I, _I, _dp, _L, _dp, _dp, _I, _dp, _dp)
        self._evf = self._fn("eval_field", _I, _P, _I, _dp, _I, _ip, _dp)

    def quad_from_arrays(self, points, quads, segments=None, markers=None, k=1, dirichlet=(), all_exterior=True,
                         equispaced=True):
        pts, q = _f64(points).reshape(-1), _i32(quads).reshape(-1)
        segs = _i32(segments).reshape(-1) if segments is not None else np.zeros(0, dtype=np.int32)
        mk = _i32(markers) if markers is not None else None
        d = _i32(list(dirichlet)) if len(dirichlet) else np.zeros(1, dtype=np.int32)
        h = self._qcreate(pts.size // 2, pts, q.size // 4, q, segs.size // 2, segs, mk.ctypes.data if mk is not None else None,
                          int(k), len(dirichlet), d, int(all_exterior), int(equispaced))
        if not h:
            raise RuntimeError(self._err())
        return OracleProblem(self, h, self._sfree, quad=True)

    def quad_rectangle(self, nx, ny, k, dirichlet=(), all_exterior=True, x0=0.0, y0=0.0, x1=1.0, y1=1.0, equispaced=True):
        """rectangle_quad_mesh (python/femd/quadmesh.py:146-170): the grid points and the sides of
        the reference's own rectangle_mesh, quad (i, j) with vertices (i,j), (i+1,j), (i+1,j+1), (i,j+1)."""
        m = self._mrect(x0, y0, x1, y1, nx, ny, b"right")
        if not m:
            raise RuntimeError(self._err())
        s = np.zeros(3, dtype=np.int32)
        self._check(self._msizes(m, s))
        pts, tris, nbr = np.empty(2 * s[0]), np.empty(3 * s[1], dtype=np.int32), np.empty(3 * s[1], dtype=np.int32)
        segs, mk = np.empty(2 * s[2], dtype=np.int32), np.empty(max(s[2], 1), dtype=np.int32)
        self._check(self._mget(m, pts, tris, nbr, segs, mk))
        self._mfree(m)
        j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        v = (j * (nx + 1) + i).ravel()
        quads = np.stack([v, v + 1, v + nx + 2, v + nx + 1], axis=1)
        return self.quad_from_arrays(pts[: 2 * (nx + 1) * (ny + 1)], quads, segs, mk[: s[2]], k, dirichlet, all_exterior,
                                     equispaced)

    def quad_quadrature(self, degree):
        xi, eta, w = np.empty(4096), np.empty(4096), np.empty(4096)
        nq = _I(0)
        self._check(self._qquad(int(degree), C.byref(nq), xi, eta, w))
        return xi[: nq.value].copy(), eta[: nq.value].copy(), w[: nq.value].copy()

    def quad_basis_eval(self, k, xi, eta, equispaced=True):
        n = (k + 1) ** 2
        out = np.empty(3 * n)
        self._check(self._qbasis(int(k), int(equispaced), float(xi), float(eta), out))
        return out.reshape(3, n)

    def quad_mesh_quads(self, P):
        out = np.empty(4 * P.ncells, dtype=np.int32)
        self._check(self._qquads(P.h, out))
        return out.reshape(-1, 4)

    def _space(self, mesh, k, dirichlet, all_exterior, equispaced=True):
        d = _i32(list(dirichlet)) if len(dirichlet) else np.zeros(1, dtype=np.int32)
        h = self._screate(mesh, int(k), len(dirichlet), d, int(all_exterior), int(equispaced))
        self._mfree(mesh)
        if not h:
            raise RuntimeError(self._err())
        return OracleProblem(self, h, self._sfree)

    def rectangle(self, nx, ny, k, diagonal="right", dirichlet=(), all_exterior=True, x0=0.0, y0=0.0, x1=1.0, y1=1.0,
                  equispaced=True):
        m = self._mrect(x0, y0, x1, y1, nx, ny, diagonal.encode())
        if not m:
            raise RuntimeError(self._err())
        return self._space(m, k, dirichlet, all_exterior, equispaced)

    def from_arrays(self, points, triangles, segments=None, markers=None, k=1, dirichlet=(), all_exterior=True,
                    equispaced=True):
        pts, tris = _f64(points).reshape(-1), _i32(triangles).reshape(-1)
        segs = _i32(segments).reshape(-1) if segments is not None else np.zeros(0, dtype=np.int32)
        mk = _i32(markers) if markers is not None else None
        m = self._marr(pts.size // 2, pts, tris.size // 3, tris, segs.size // 2, segs, mk.ctypes.data if mk is not None else None)
        if not m:
            raise RuntimeError(self._err())
        return self._space(m, k, dirichlet, all_exterior, equispaced)

    def _sizes(self, h):
        out = np.zeros(6, dtype=np.int32)
        self._check(self._ssizes(h, out))
        return out

    def _numbering(self, P):
        cd = np.empty(P.ncells * P.nloc, dtype=np.int32)
        cr = np.empty_like(cd)
        a2r = np.empty(max(P.dim, 1), dtype=np.int32)
        xy = np.empty(2 * P.raw_dim)
        self._check(self._snum(P.h, cd, cr, a2r, xy))
        return dict(cell_dof=cd.reshape(P.ncells, P.nloc), cell_raw=cr.reshape(P.ncells, P.nloc),
                    ad2raw=a2r[: P.dim], raw_xy=xy.reshape(-1, 2))

    _pattern = Oracle._pattern
    quadrature = Oracle.quadrature
    _npts = Oracle._npts
    _cache_points = Oracle._cache_points
    _assemble_matrix = Oracle._assemble_matrix
    _assemble_vector = Oracle._assemble_vector

    def basis_eval(self, k, xi, eta, equispaced=True):
        n = (k + 1) * (k + 2) // 2
        out = np.empty(3 * n)
        self._check(self._basis(int(k), int(equispaced), float(xi), float(eta), out))
        return out.reshape(3, n)

    def integrand_eval(self, code, nreg, consts, xs, ys):
        code = _i32(code).reshape(-1)
        xs, ys = _f64(xs), _f64(ys)
        out = np.empty(xs.size)
        self._check(self._ie(code.size // 4, code, int(nreg), len(consts), _f64(consts) if len(consts) else np.zeros(1),
                             xs.size, xs, ys, out))
        return out

    def integrand_eval_n(self, code, nreg, consts, xs, ys, nxs, nys):
        code = _i32(code).reshape(-1)
        xs, ys, nxs, nys = _f64(xs), _f64(ys), _f64(nxs), _f64(nys)
        out = np.empty(xs.size)
        self._check(self._ien(code.size // 4, code, int(nreg), len(consts), _f64(consts) if len(consts) else np.zeros(1),
                              xs.size, xs, ys, nxs, nys, out))
        return out

    def integrand_eval_in(self, code, nreg, consts, xs, ys, inputs):
        code = _i32(code).reshape(-1)
        xs, ys = _f64(xs), _f64(ys)
        inp = _f64(inputs).reshape(-1, xs.size) if len(inputs) else np.zeros((0, xs.size))
        out = np.empty(xs.size)
        self._check(self._iei(code.size // 4, code, int(nreg), len(consts), _f64(consts) if len(consts) else np.zeros(1),
                              xs.size, xs, ys, inp.shape[0], np.ascontiguousarray(inp).reshape(-1) if inp.size else np.zeros(1),
                              out))
        return out

    _eval_field = Oracle._eval_field

    def asymmetry(self, ip, ix, v):
        ip, ix, v = _i32(ip), _i32(ix), _f64(v)
        return float(self._asym(ip.size - 1, ip, ix, v))


def oracle():
    """The restatement; raises if it is not built (make -C oracle)."""
    if not ORACLE_SO.exists():
        raise FileNotFoundError(f"{ORACLE_SO} is not built: run make -C oracle")
    return Oracle()


def reference():
    """The reference compiled in place, or None when oracle/_ref was not built."""
    return Ref() if REF_SO.exists() else None
