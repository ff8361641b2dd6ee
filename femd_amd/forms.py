"""The form language on 2D meshes, the subset the assembly kernels take (python/femd/forms.py).

    u, v = TrialFunction(V), TestFunction(V)
    A = form((dot(grad(u), grad(v)) + u*v) * dx).assemble()      # SparseMatrix, on the GPU
    b = form(f * v * dx).assemble()                             # numpy vector

A form is canonicalized into terms  c(x) (D^a v)(D^b u)  exactly as the reference does
(forms.py:1364-1415: the same expansion and term order), its quadrature degree inferred by the
same rules (forms.py:2443-2500), and each coefficient compiled into the same register program
(_Program, forms.py:1822-1938), which the GPU evaluates at the quadrature points.  Known fields
(Function arguments inside coefficients), ds / dS measures and 1D meshes are outside this path
and raise NotImplementedError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


# ============================================================================ expressions
class Expr:
    def __add__(self, o):  return Sum([self, _lift(o)])
    def __radd__(self, o): return Sum([_lift(o), self])
    def __sub__(self, o):  return Sum([self, Prod([Const(-1.0), _lift(o)])])
    def __rsub__(self, o): return Sum([_lift(o), Prod([Const(-1.0), self])])
    def __neg__(self):     return Prod([Const(-1.0), self])
    def __pos__(self):     return self

    def __mul__(self, o):
        if isinstance(o, Measure):
            return Integral(self, o)
        if isinstance(o, ListTensor):
            return o.__rmul__(self)
        return Prod([self, _lift(o)])

    def __rmul__(self, o): return Prod([_lift(o), self])

    def __truediv__(self, o):
        o = _lift(o)
        if _has_arguments(o):
            raise TypeError("cannot divide by an expression containing a test or trial function")
        return Prod([self, Pow(o, -1)])

    def __rtruediv__(self, o): return Prod([_lift(o), Pow(self, -1)])

    def __pow__(self, n):
        if not isinstance(n, (int, float)):
            raise TypeError("exponent must be a number")
        return Pow(self, n)

    __array_priority__ = 1000


def _lift(o):
    if isinstance(o, Expr):
        return o
    if isinstance(o, ListTensor):
        raise TypeError(f"a vector {o!r} where a scalar is needed: use dot(), inner() or a component")
    if isinstance(o, (int, float, np.floating, np.integer)):
        return Const(float(o))
    raise TypeError(f"cannot use {o!r} in a form")


class Const(Expr):
    def __init__(self, value): self.value = float(value)
    def __repr__(self): return f"{self.value:g}"


class Constant(Expr):
    """A named number read when the form is assembled (forms.py:108-137)."""
    def __init__(self, value=0.0, name="c"):
        self.value, self.name = float(value), str(name)

    def assign(self, value):
        self.value = float(value)

    def __float__(self): return self.value
    def __repr__(self): return self.name


class Spatial(Expr):
    def __init__(self, index): self.index = int(index)
    def __repr__(self): return "xy"[self.index]


x = Spatial(0)
y = Spatial(1)


class Normal(Expr):
    """A component of the outward unit normal on the exterior facets, for ds terms
    (python/femd/forms.py:101-104, 634-639)."""
    def __init__(self, index: int): self.index = int(index)
    def __repr__(self): return f"n[{self.index}]"


def FacetNormal(mesh=None):
    """The outward unit normal (n[0], n[1]) on the exterior facets, for ds terms."""
    return ListTensor([Normal(0), Normal(1)])


class Sum(Expr):
    def __init__(self, terms): self.terms = list(terms)
    def __repr__(self): return "(" + " + ".join(map(repr, self.terms)) + ")"


class Prod(Expr):
    def __init__(self, factors): self.factors = list(factors)
    def __repr__(self): return "*".join(map(repr, self.factors))


class Pow(Expr):
    def __init__(self, base, n): self.base, self.n = base, n
    def __repr__(self): return f"({self.base!r})**{self.n}"


class Func(Expr):
    _NAMES = ("sin", "cos", "exp", "log", "tanh", "sqrt", "sinh", "cosh", "sech", "abs", "sign")

    def __init__(self, name, arg):
        if name not in self._NAMES:
            raise ValueError(f"unknown function {name}")
        self.name, self.arg = name, arg

    def __repr__(self): return f"{self.name}({self.arg!r})"


def sin(e):  return Func("sin", _lift(e))
def cos(e):  return Func("cos", _lift(e))
def exp(e):  return Func("exp", _lift(e))
def log(e):  return Func("log", _lift(e))
def tanh(e): return Func("tanh", _lift(e))
def sqrt(e): return Func("sqrt", _lift(e))
def sinh(e): return Func("sinh", _lift(e))
def cosh(e): return Func("cosh", _lift(e))
def sech(e): return Func("sech", _lift(e))
def sign(e): return Func("sign", _lift(e))
def abs_(e): return Func("abs", _lift(e))


class Argument(Expr):
    """Test or trial function of a 2D space, with a partial-derivative multi-index."""
    def __init__(self, kind, space, deriv=(0, 0)):
        self.kind, self.space, self.deriv = kind, space, tuple(deriv)

    def __repr__(self):
        base = {"test": "v", "trial": "u"}[self.kind]
        if any(self.deriv):
            base = "D" + "x" * self.deriv[0] + "y" * self.deriv[1] + f"({base})"
        return base

    def code(self):
        """The kernel's derivative code (forms.py:181-184): 0 value, 1 d/dx, 2 d/dy."""
        d = self.deriv
        return 0 if d[0] + d[1] == 0 else (1 if d[0] else 2)


def TestFunction(V):  return Argument("test", V)
def TrialFunction(V): return Argument("trial", V)


class ListTensor:
    """A vector of scalar expressions (grad(u))."""
    def __init__(self, components): self.components = [_lift(c) for c in components]
    def __len__(self): return len(self.components)
    def __getitem__(self, i): return self.components[i]
    def __iter__(self): return iter(self.components)
    def __repr__(self): return "[" + ", ".join(map(repr, self.components)) + "]"
    def __add__(self, o): return ListTensor([a + b for a, b in zip(self, o)])
    def __sub__(self, o): return ListTensor([a - b for a, b in zip(self, o)])
    def __neg__(self): return ListTensor([-a for a in self])
    def __mul__(self, o): return ListTensor([a * o for a in self])
    def __rmul__(self, o): return ListTensor([o * a for a in self])


def as_vector(c): return c if isinstance(c, ListTensor) else ListTensor(c)


def _dfunc(f):
    """d f(g) / dg for Dx of a function (forms.py:368-377)."""
    a = f.arg
    table = {
        "sin": lambda: cos(a), "cos": lambda: Prod([Const(-1.0), sin(a)]), "exp": lambda: exp(a),
        "log": lambda: Pow(a, -1), "sqrt": lambda: Prod([Const(0.5), Pow(sqrt(a), -1)]),
        "sinh": lambda: cosh(a), "cosh": lambda: sinh(a),
        "tanh": lambda: Sum([Const(1.0), Prod([Const(-1.0), Pow(tanh(a), 2)])]),
    }
    if f.name not in table:
        raise TypeError(f"cannot differentiate {f.name}")
    return table[f.name]()


def Dx(e, i: int):
    """d e / d x_i, linear with the product and chain rules (forms.py:380-416)."""
    if i not in (0, 1):
        raise ValueError(f"Dx(e, i): i is 0 (x) or 1 (y), got {i!r}")
    if isinstance(e, ListTensor):
        return ListTensor([Dx(c, i) for c in e.components])
    if isinstance(e, (int, float, np.floating, np.integer)):
        return Const(0.0)
    if isinstance(e, Argument):
        d = list(e.deriv)
        d[i] += 1
        return Argument(e.kind, e.space, tuple(d))
    if isinstance(e, Sum):
        return Sum([Dx(t, i) for t in e.terms])
    if isinstance(e, Prod):
        consts = [f for f in e.factors if isinstance(f, Const)]
        rest = [f for f in e.factors if not isinstance(f, Const)]
        if not rest:
            return Const(0.0)
        if len(rest) == 1:
            return Prod(consts + [Dx(rest[0], i)])
        return Sum([Prod(consts + rest[:j] + [Dx(f, i)] + rest[j + 1:]) for j, f in enumerate(rest)])
    if isinstance(e, (Const, Constant)):
        return Const(0.0)
    if isinstance(e, Spatial):
        return Const(1.0 if e.index == i else 0.0)
    if isinstance(e, Pow):
        return Prod([Const(float(e.n)), Pow(e.base, e.n - 1), Dx(e.base, i)])
    if isinstance(e, Func):
        return Prod([_dfunc(e), Dx(e.arg, i)])
    raise TypeError(f"cannot differentiate {e!r}")


def grad(e):
    """(Dx(e, 0), Dx(e, 1)) of a scalar (forms.py:521-532)."""
    if isinstance(e, ListTensor):
        raise NotImplementedError("grad() of a vector is outside this path")
    e = _lift(e)
    return ListTensor([Dx(e, 0), Dx(e, 1)])


def div(v):
    v = as_vector(v)
    return Sum([Dx(c, i) for i, c in enumerate(v.components)])


def dot(a, b):
    """a . b: a b for scalars, sum_i a_i b_i for vectors (forms.py:565-583)."""
    a = as_vector(a) if isinstance(a, (list, tuple)) else a
    b = as_vector(b) if isinstance(b, (list, tuple)) else b
    ra, rb = isinstance(a, ListTensor), isinstance(b, ListTensor)
    if not ra and not rb:
        return _lift(a) * _lift(b)
    if ra != rb:
        raise TypeError("dot(a, b): a scalar and a vector; use a * b for scaling")
    if len(a) != len(b):
        raise TypeError(f"dot(a, b): lengths {len(a)} and {len(b)} differ")
    return Sum([p * q for p, q in zip(a.components, b.components)])


inner = dot


# ============================================================================ measures
class Measure:
    """dx (cells), ds (exterior facets: every one, or those with the given side markers) and dS
    (interior facets, outside this path); forms.py:1269-1309."""
    def __init__(self, kind, where=None, quad_degree=None):
        self.kind, self.where, self.quad_degree = kind, where, quad_degree

    def __call__(self, *args, quad_degree=None, **kw):
        if self.kind == "dS":
            raise NotImplementedError("interior-facet (dS) integrals are outside this path")
        if kw:
            raise NotImplementedError(f"{self.kind} options other than the degree are outside this path")
        if self.kind == "dx":
            if args:
                quad_degree = args[0]
            if quad_degree is not None and int(quad_degree) < 1:
                raise ValueError(f"dx(n): n must be at least 1, got {quad_degree}")
            return Measure("dx", None, quad_degree)
        where = args[0] if args else None
        if len(args) > 1:
            where = tuple(args)
        if isinstance(where, (int, np.integer)) and not isinstance(where, bool):
            return Measure("ds", (int(where),), quad_degree)
        if isinstance(where, (tuple, list)) and where and all(isinstance(m, (int, np.integer)) for m in where):
            return Measure("ds", tuple(sorted({int(m) for m in where})), quad_degree)
        if where in ("left", "right"):
            raise TypeError(f"ds('{where}') is for 1D meshes; on a 2D mesh use ds(marker), ds((m1, m2)) or ds")
        if where not in (None, "both"):
            raise ValueError("ds takes a marker or markers on a 2D mesh")
        return Measure("ds", "both", quad_degree)

    def __rmul__(self, e):
        return Integral(_lift(e), self)


dx = Measure("dx")
ds = Measure("ds", "both")
dS = Measure("dS")


class Integral:
    def __init__(self, expr, measure):
        if measure.kind == "dS":
            raise NotImplementedError("interior-facet (dS) integrals are outside this path")
        self.expr, self.measure = expr, measure

    def __add__(self, o): return FormExpr([self]) + o
    def __radd__(self, o): return FormExpr([self]) + o
    def __sub__(self, o): return FormExpr([self]) - o
    def __neg__(self): return Integral(Prod([Const(-1.0), self.expr]), self.measure)
    def __mul__(self, c): return Integral(Prod([_lift(c), self.expr]), self.measure)
    __rmul__ = __mul__


class FormExpr:
    def __init__(self, integrals): self.integrals = list(integrals)

    def __add__(self, o):
        if isinstance(o, Integral): return FormExpr(self.integrals + [o])
        if isinstance(o, FormExpr): return FormExpr(self.integrals + o.integrals)
        if o == 0: return self
        raise TypeError("can only add integrals to a form")
    __radd__ = __add__

    def __sub__(self, o):
        o = o if isinstance(o, FormExpr) else FormExpr([o])
        return FormExpr(self.integrals + [-i for i in o.integrals])

    def __neg__(self): return FormExpr([-i for i in self.integrals])


# ============================================================================ analysis
def _has_arguments(e):
    if isinstance(e, Argument): return True
    if isinstance(e, Sum): return any(_has_arguments(t) for t in e.terms)
    if isinstance(e, Prod): return any(_has_arguments(f) for f in e.factors)
    if isinstance(e, Pow): return _has_arguments(e.base)
    if isinstance(e, Func): return _has_arguments(e.arg)
    return False


def _expand(e):
    """Sum of products over the test / trial structure (forms.py:1364-1381)."""
    if not _has_arguments(e) or isinstance(e, Argument):
        return [[e]]
    if isinstance(e, Sum):
        return [t for term in e.terms for t in _expand(term)]
    if isinstance(e, Prod):
        out = [[]]
        for f in e.factors:
            out = [a + b for a in out for b in _expand(f)]
        return out
    if isinstance(e, Pow):
        if e.n == 1:
            return _expand(e.base)
        raise TypeError(f"{e!r}: a form must be linear in each test and trial function")
    raise TypeError(f"cannot expand {e!r}")


class Term:
    def __init__(self, test, trial, coeff, measure):
        self.test, self.trial, self.coeff, self.measure = test, trial, coeff, measure

    @property
    def rank(self): return (self.test is not None) + (self.trial is not None)

    def __repr__(self):
        m = self.measure
        where = "" if m.kind == "dx" or m.where == "both" else f"({', '.join(map(str, m.where))})"
        return "*".join([repr(self.coeff)] + [repr(a) for a in (self.test, self.trial) if a is not None]) + f"*{m.kind}{where}"


def _trivially_zero(e):
    if isinstance(e, Const): return e.value == 0.0
    if isinstance(e, Prod): return any(_trivially_zero(f) for f in e.factors)
    if isinstance(e, Sum): return all(_trivially_zero(t) for t in e.terms)
    return False


def canonicalize(fe):
    """Terms coeff * test * trial in the reference's order (forms.py:1398-1415)."""
    terms = []
    for integral in fe.integrals:
        for factors in _expand(integral.expr):
            tests = [f for f in factors if isinstance(f, Argument) and f.kind == "test"]
            trials = [f for f in factors if isinstance(f, Argument) and f.kind == "trial"]
            rest = [f for f in factors if not isinstance(f, Argument)]
            if len(tests) > 1 or len(trials) > 1:
                raise TypeError(f"term {factors!r} is not linear in the test/trial functions")
            if trials and not tests:
                raise TypeError("a term with a trial function but no test function has no row index")
            coeff = Prod(rest) if len(rest) != 1 else rest[0]
            if not rest:
                coeff = Const(1.0)
            terms.append(Term(tests[0] if tests else None, trials[0] if trials else None, coeff, integral.measure))
    nonzero = [t for t in terms if not (t.test is None and t.trial is None and _trivially_zero(t.coeff))]
    return nonzero if nonzero else terms


def poly_degree(e):
    """Polynomial degree of a pointwise expression, None if not polynomial (forms.py:1429-1448)."""
    if isinstance(e, (Const, Constant)): return 0
    if isinstance(e, Spatial): return 1
    if isinstance(e, Normal): return 0
    if isinstance(e, Argument): return _arg_degree(e)
    if isinstance(e, Sum):
        ds_ = [poly_degree(t) for t in e.terms]
        return None if any(d is None for d in ds_) else max(ds_, default=0)
    if isinstance(e, Prod):
        ds_ = [poly_degree(f) for f in e.factors]
        return None if any(d is None for d in ds_) else sum(ds_)
    if isinstance(e, Pow):
        d = poly_degree(e.base)
        if d is None or not float(e.n).is_integer() or e.n < 0: return None
        return int(e.n) * d
    return None


def _arg_degree(a):
    """Total degree on triangles; on quadrilaterals the per-variable degree, which a derivative
    does not lower (forms.py:654-662)."""
    if getattr(a.space, "nverts", 3) == 4:
        return a.space.degree
    return max(a.space.degree - sum(a.deriv), 0)


def quad_degree_estimate(e, kmax):
    """forms.py:1481-1501: non-polynomial data counts as degree kmax."""
    d = poly_degree(e)
    if d is not None: return d
    if not _has_arguments(e): return kmax
    if isinstance(e, Argument): return _arg_degree(e)
    if isinstance(e, Sum): return max((quad_degree_estimate(t, kmax) for t in e.terms), default=0)
    if isinstance(e, Prod): return sum(quad_degree_estimate(f, kmax) for f in e.factors)
    if isinstance(e, Pow):
        if float(e.n).is_integer() and e.n >= 0:
            return int(e.n) * quad_degree_estimate(e.base, kmax)
        return kmax
    return kmax


_default_quad_degree = None


def set_quadrature_degree(degree=None):
    """One rule for every form made from now on (forms.py:1451-1466)."""
    global _default_quad_degree
    if degree is None:
        _default_quad_degree = None
        return
    n = int(degree)
    if n != degree or n < 1:
        raise ValueError(f"set_quadrature_degree: degree must be a positive integer or None, got {degree!r}")
    _default_quad_degree = n


def quadrature_degree():
    return _default_quad_degree


# ============================================================================ programs
class Program:
    """A coefficient compiled into the register program of forms/integrand.hpp (forms.py:1822-1909):
    post-order, one register per node, shared subexpressions emitted once (by identity)."""
    _FUNCS = {"sin": 9, "cos": 10, "exp": 11, "log": 12, "tanh": 13, "sqrt": 14, "sinh": 15, "cosh": 16,
              "sech": 17, "abs": 18, "sign": 19}
    CONST, X, Y, NX, NY, INPUT, ADD, MUL, POW = range(9)

    def __init__(self, expr):
        self.code, self.consts, self.nreg, self.memo = [], [], 0, {}
        self._emit(expr)

    def _reg(self):
        self.nreg += 1
        return self.nreg - 1

    def _const(self, c):
        self.consts.append(c)
        return len(self.consts) - 1

    def _emit(self, e):
        got = self.memo.get(id(e))
        if got is not None and got[0] is e:
            return got[1]
        r = self._node(e)
        self.memo[id(e)] = (e, r)
        return r

    def _node(self, e):
        if isinstance(e, Const):
            r = self._reg(); self.code.append((self.CONST, r, self._const(e.value), 0)); return r
        if isinstance(e, Constant):
            r = self._reg(); self.code.append((self.CONST, r, self._const(e), 0)); return r
        if isinstance(e, Spatial):
            r = self._reg(); self.code.append((self.X if e.index == 0 else self.Y, r, 0, 0)); return r
        if isinstance(e, Normal):
            r = self._reg(); self.code.append((self.NX if e.index == 0 else self.NY, r, 1, 0)); return r
        if isinstance(e, Sum):
            if not e.terms:
                r = self._reg(); self.code.append((self.CONST, r, self._const(0.0), 0)); return r
            acc = self._emit(e.terms[0])
            for t in e.terms[1:]:
                b = self._emit(t); r = self._reg(); self.code.append((self.ADD, r, acc, b)); acc = r
            return acc
        if isinstance(e, Prod):
            if not e.factors:
                r = self._reg(); self.code.append((self.CONST, r, self._const(1.0), 0)); return r
            acc = self._emit(e.factors[0])
            for f in e.factors[1:]:
                b = self._emit(f); r = self._reg(); self.code.append((self.MUL, r, acc, b)); acc = r
            return acc
        if isinstance(e, Pow):
            a = self._emit(e.base); r = self._reg(); self.code.append((self.POW, r, a, self._const(float(e.n)))); return r
        if isinstance(e, Func):
            a = self._emit(e.arg); r = self._reg(); self.code.append((self._FUNCS[e.name], r, a, 0)); return r
        if isinstance(e, Argument):
            raise NotImplementedError("known fields inside coefficients are outside this path")
        raise NotImplementedError(f"cannot compile {e!r}")

    def arrays(self):
        """(code int32 (n, 4), nreg, consts float64) with the Constants' current values."""
        consts = [c.value if isinstance(c, Constant) else c for c in self.consts]
        return np.asarray(self.code, dtype=np.int32).reshape(-1, 4), max(self.nreg, 1), np.asarray(consts, dtype=np.float64)


_NP_FUNCS = {"sin": np.sin, "cos": np.cos, "exp": np.exp, "log": np.log, "tanh": np.tanh, "sqrt": np.sqrt,
             "sinh": np.sinh, "cosh": np.cosh, "sech": lambda v: 1.0 / np.cosh(v), "abs": np.abs, "sign": np.sign}


def evaluate(e, x, y):
    """An expression in fd.x, fd.y, numbers and Constants at the points (x, y), with NumPy (the
    boundary data of V.lift; python/femd/forms.py evaluate)."""
    if isinstance(e, (Const, Constant)): return e.value
    if isinstance(e, Spatial): return x if e.index == 0 else y
    if isinstance(e, Sum):
        out = 0.0
        for t in e.terms:
            out = out + evaluate(t, x, y)
        return out
    if isinstance(e, Prod):
        out = 1.0
        for f in e.factors:
            out = out * evaluate(f, x, y)
        return out
    if isinstance(e, Pow): return np.power(evaluate(e.base, x, y), e.n)
    if isinstance(e, Func): return _NP_FUNCS[e.name](evaluate(e.arg, x, y))
    raise TypeError(f"boundary data: an expression may contain fd.x, fd.y, numbers and Constants, not {e!r}")


def _has_normal(e):
    if isinstance(e, Normal): return True
    if isinstance(e, Sum): return any(_has_normal(t) for t in e.terms)
    if isinstance(e, Prod): return any(_has_normal(f) for f in e.factors)
    if isinstance(e, Pow): return _has_normal(e.base)
    if isinstance(e, Func): return _has_normal(e.arg)
    return False


def _rebind(f, U):
    """The same form with its test and trial functions on space U (same mesh and degree)."""
    def arg(a):
        return None if a is None else Argument(a.kind, U, a.deriv)
    g = Form.__new__(Form)
    g.__dict__.update(f.__dict__)
    g.terms = [Term(arg(t.test), arg(t.trial), t.coeff, t.measure) for t in f.terms]
    g.test_space = U
    g.trial_space = U if f.trial_space is not None else None
    g._result = None
    return g


# ============================================================================ the Form
class Form:
    def __init__(self, fe):
        if isinstance(fe, Integral):
            fe = FormExpr([fe])
        if not isinstance(fe, FormExpr):
            raise TypeError("form() takes expr*dx, or a sum of those")
        self.terms = canonicalize(fe)
        if not self.terms:
            raise ValueError("empty form")
        ranks = {t.rank for t in self.terms}
        if len(ranks) != 1:
            raise TypeError(f"all terms must have the same rank, got {sorted(ranks)}: {self.terms}")
        self.rank = ranks.pop()
        if self.rank == 0:
            raise NotImplementedError("rank-0 forms (functionals) are outside this path")
        tests = [t.test for t in self.terms if t.test is not None]
        trials = [t.trial for t in self.terms if t.trial is not None]
        spaces = {id(a.space): a.space for a in tests + trials}
        if len({id(a.space) for a in tests}) > 1 or len({id(a.space) for a in trials}) > 1:
            raise TypeError("the test (trial) functions of a form must come from one space")
        self.test_space = tests[0].space
        self.trial_space = trials[0].space if trials else None
        if self.trial_space is not None and self.trial_space is not self.test_space:
            raise NotImplementedError("rectangular forms (test space != trial space) are outside this path")
        for t in self.terms:
            for a in (t.test, t.trial):
                if a is not None and sum(a.deriv) > 1:
                    raise ValueError(f"{a!r}: derivatives of order {sum(a.deriv)} are not available on a C^0 "
                                     "Lagrange space; integrate by parts (only first derivatives enter a form)")
        for t in self.terms:
            if t.measure.kind == "dx" and _has_normal(t.coeff):
                raise TypeError("the facet normal n appears in a dx term; it exists on ds terms only")
        present = set(np.asarray(self.test_space.mesh.segment_markers).tolist())
        for t in self.terms:
            if t.measure.kind == "ds" and t.measure.where != "both":
                missing = [m for m in t.measure.where if m not in present]
                if missing:
                    raise ValueError(f"ds({', '.join(str(m) for m in t.measure.where)}): no side of the mesh carries "
                                     f"marker(s) {missing}; the markers present are {sorted(present)}")
        # quadrature degrees per measure (forms.py:2487-2500)
        kmax = max(sp.degree for sp in spaces.values())
        deg = {"dx": 0, "ds": 0}
        for t in self.terms:
            k = t.measure.kind
            if t.measure.quad_degree is not None:
                deg[k] = max(deg[k], int(t.measure.quad_degree)); continue
            if _default_quad_degree is not None:
                deg[k] = max(deg[k], _default_quad_degree); continue
            d = poly_degree(t.coeff)
            arg_d = sum(_arg_degree(a) for a in (t.test, t.trial) if a is not None)
            if d is None:
                d = quad_degree_estimate(t.coeff, kmax + 1)
            deg[k] = max(deg[k], arg_d + d)
        self.quad_degree = max(deg["dx"], 1)
        self.facet_degree = max(deg["ds"], 1)
        # the point sets in order of first use: dx, or ds(where) (forms.py:2313-2336); the groups
        # add into one matrix (vector) in this order, as FormAssembler2D does
        self.groups = {}
        for i, t in enumerate(self.terms):
            key = ("dx",) if t.measure.kind == "dx" else ("ds", t.measure.where)
            self.groups.setdefault(key, []).append(i)
        self.programs = [Program(t.coeff) for t in self.terms]
        self.constant = not any(isinstance(c, Constant) for p in self.programs for c in p.consts)
        self._result = None

    def _coefficients(self, idx=None):
        items = []
        for i in (range(len(self.terms)) if idx is None else idx):
            code, nreg, consts = self.programs[i].arrays()
            items.append(((code, nreg, consts), None))
        return L.coefficient_array(items)

    def assemble(self, reuse: bool = True):
        """Rank 2: a SparseMatrix on the space's pattern (symmetric flag set as forms.py:2093-2094);
        rank 1: the vector (numpy, float64).  Assembled on the GPU, one group of terms per point
        set (dx, ds(markers)) in the form's order, each adding into the same matrix (vector) as
        FormAssembler2D does (forms/assembler_2d.hpp:166-188, 249-266).  A form without Constants
        is assembled once and the result reused, as the reference does (forms.py:2752-2760),
        unless reuse=False."""
        if reuse and self.constant and self._result is not None:
            return self._result
        V = self.test_space
        lib = L.load()
        if self.rank == 1:
            out = np.empty(V.dim)
        else:
            from .sparse import SparseMatrix
            out = SparseMatrix._on_space(V)
        for g, (key, idx) in enumerate(self.groups.items()):
            coeffs, keep = self._coefficients(idx)
            a = L.i32([self.terms[i].test.code() for i in idx])
            b = L.i32([self.terms[i].trial.code() for i in idx]) if self.rank == 2 else None
            if key[0] == "dx":
                if self.rank == 1:
                    fn = lib.femd_assemble_vector if g == 0 else lib.femd_assemble_vector_add
                    L.check(fn(V._h, self.quad_degree, len(idx), a, coeffs, out))
                else:
                    fn = lib.femd_assemble_matrix if g == 0 else lib.femd_assemble_matrix_add
                    L.check(fn(V._h, self.quad_degree, len(idx), a, b, coeffs, out._h))
            else:
                F = V.facet_set(self.facet_degree, None if key[1] == "both" else key[1])
                if self.rank == 1:
                    L.check(lib.femd_assemble_vector_facets(F, len(idx), a, coeffs, out, 0 if g == 0 else 1))
                else:
                    L.check(lib.femd_assemble_matrix_facets(F, len(idx), a, b, coeffs, out._h, 0 if g == 0 else 1))
            del keep
        if self.rank == 2:
            out.symmetric = bool(out.asymmetry() <= 1e-13)
        if self.constant:
            self._result = out
        return out

    def action(self, u):
        """The linear form v -> a(u, v) of a bilinear form for a known function u given by its
        values on the space's unconstrained nodes (raw numbering), assembled: a vector over the
        free nodes.  The reference assembles a.action(u) with u as a field; this forms the same
        sums through the matrix of the unconstrained space (equal to rounding)."""
        if self.rank != 2:
            raise TypeError("action() is for bilinear forms")
        V = self.test_space
        U = V.unconstrained
        u = np.asarray(u, dtype=np.float64).reshape(-1)
        if u.size != U.dim:
            raise ValueError(f"action(): u has {u.size} values, the unconstrained space has {U.dim}")
        KU = _rebind(self, U).assemble()
        return KU.matvec(u)[V.adapted_to_raw()]

    def __repr__(self):
        return f"Form(rank {self.rank}, {len(self.terms)} terms, quadrature degree {self.quad_degree}: {self.terms})"


def form(fe) -> Form:
    """Compile expr*dx (+ ...) into a Form (forms.py:3092-3095)."""
    return Form(fe)
