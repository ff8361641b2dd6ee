"""The form language on 2D meshes, the subset the assembly kernels take (python/femd/forms.py).

    u, v = TrialFunction(V), TestFunction(V)
    A = form((dot(grad(u), grad(v)) + u*v) * dx).assemble()      # SparseMatrix, on the GPU
    b = form(f * v * dx).assemble()                             # numpy vector

A form becomes a list of terms  c(x) (D^a v)(D^b u)  over dx or ds(markers), in the order the
reference's canonicalization gives them (python/femd/forms.py:1364-1415), with the quadrature
degree its rules infer (forms.py:2443-2500); each coefficient becomes a register program of the
Integrand instruction set (forms/integrand.hpp:33-46) in the reference's emission order
(forms.py:1822-1909), which the GPU evaluates at the quadrature points.  Because the expression
trees have the reference's shapes, the assembled values are the reference's bit for bit.  Known
fields (Function) enter coefficients of dx terms through INPUT instructions.  dS measures, known
fields on ds terms and 1D meshes are outside this path and raise NotImplementedError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


# ============================================================================ expression nodes
# An expression is a tree of nodes.  Its shape is part of the contract: the terms a form expands into,
# their order, and the instruction order of each coefficient's program follow from it, and the
# assembled values follow from those bit for bit.  So a + b is always a two-term sum, a - b the sum
# of a and (-1) * b, -a the product (-1) * a, a / b the product of a and b ** -1, and sums and
# products are never flattened (python/femd/forms.py builds the same trees).  The algorithms over
# trees (expansion, derivatives, degrees, program emission, NumPy evaluation) are tables keyed by
# node class, below.

class Expr:
    """Base of the scalar expression nodes; operators build nodes (see _OPERATORS)."""
    __array_priority__ = 1000

    def __pos__(self):
        return self

    def __neg__(self):
        return _negated(self)

    def __pow__(self, n):
        if not isinstance(n, (int, float)):
            raise TypeError("exponent must be a number")
        return Pow(self, n)


class Const(Expr):
    """A number fixed when the form is made."""
    def __init__(self, value):
        self.value = float(value)

    def __repr__(self):
        return repr(self.value)


class Constant(Expr):
    """A named number read each time the form is assembled (python/femd/forms.py:108-137)."""
    def __init__(self, value=0.0, name="c"):
        self.value, self.name = float(value), str(name)

    def assign(self, value):
        self.value = float(value)

    def __float__(self):
        return self.value

    def __repr__(self):
        return f"Constant({self.name}={self.value!r})"


class Spatial(Expr):
    """A coordinate: index 0 is x, 1 is y."""
    def __init__(self, index):
        self.index = int(index)

    def __repr__(self):
        return ("x", "y")[self.index]


x = Spatial(0)
y = Spatial(1)


class Normal(Expr):
    """A component of the outward unit normal on exterior facets (ds terms only;
    python/femd/forms.py:101-104)."""
    def __init__(self, index: int):
        self.index = int(index)

    def __repr__(self):
        return ("n_x", "n_y")[self.index]


class Sum(Expr):
    """terms[0] + terms[1] + ..., added left to right."""
    def __init__(self, terms):
        self.terms = list(terms)

    def __repr__(self):
        return "Sum(" + ", ".join(repr(t) for t in self.terms) + ")"


class Prod(Expr):
    """factors[0] * factors[1] * ..., multiplied left to right."""
    def __init__(self, factors):
        self.factors = list(factors)

    def __repr__(self):
        return "Prod(" + ", ".join(repr(f) for f in self.factors) + ")"


class Pow(Expr):
    """base ** n for a number n."""
    def __init__(self, base, n):
        self.base, self.n = base, n

    def __repr__(self):
        return f"Pow({self.base!r}, {self.n!r})"


# The pointwise functions of the instruction set (forms/integrand.hpp:33-46) with their opcodes.
_OPCODE = {"sin": 9, "cos": 10, "exp": 11, "log": 12, "tanh": 13, "sqrt": 14, "sinh": 15, "cosh": 16,
           "sech": 17, "abs": 18, "sign": 19}


class Func(Expr):
    """name(arg) for a function of the instruction set."""
    def __init__(self, name, arg):
        if name not in _OPCODE:
            raise ValueError(f"unknown function {name}")
        self.name, self.arg = name, arg

    def __repr__(self):
        return f"{self.name}({self.arg!r})"


class Argument(Expr):
    """The test or trial function of a 2D space, with a partial-derivative multi-index."""
    def __init__(self, kind, space, deriv=(0, 0)):
        self.kind, self.space, self.deriv = kind, space, tuple(deriv)

    def __repr__(self):
        name = "v" if self.kind == "test" else "u"
        dx_, dy_ = self.deriv
        return name if dx_ + dy_ == 0 else f"d{name}/{'dx' * dx_}{'dy' * dy_}"

    def code(self):
        """The kernels' derivative code of a first-order argument: 0 value, 1 d/dx, 2 d/dy."""
        dx_, dy_ = self.deriv
        if dx_ + dy_ == 0:
            return 0
        return 1 if dx_ else 2


def TestFunction(V):
    return Argument("test", V)


def TrialFunction(V):
    return Argument("trial", V)


def _deriv_code(deriv, what):
    dx_, dy_ = deriv
    if dx_ + dy_ > 1:
        raise ValueError(f"{what}: derivatives of order {dx_ + dy_} are not available on a C^0 Lagrange space")
    return 0 if dx_ + dy_ == 0 else (1 if dx_ else 2)


class Field(Expr):
    """A known field, or one of its first derivatives, inside a coefficient: the Function whose
    coefficient vector is read and a partial-derivative multi-index.  The program reads it through
    an INPUT instruction (forms/integrand.hpp:39); the GPU evaluates it at the quadrature points as
    Cache2D::eval_multi does (fe/cache_2d.hpp:245-288)."""
    def __init__(self, owner, deriv=(0, 0)):
        self._owner, self.deriv = owner, tuple(deriv)

    @property
    def owner(self):
        return self._owner

    @property
    def space(self):
        return self.owner.space

    def slot(self):
        """What an INPUT slot stands for: the vector read and the derivative."""
        return (id(self.owner), self.deriv)

    def code(self):
        return _deriv_code(self.deriv, repr(self))

    def __repr__(self):
        dx_, dy_ = self.deriv
        name = self.owner.name
        return name if dx_ + dy_ == 0 else f"d{name}/{'dx' * dx_}{'dy' * dy_}"


class Function(Field):
    """A known field in a Lagrange space that carries its coefficient vector (V.dim values;
    python/femd/forms.py:806-866).  Inside a form it is read when the form is assembled, so a form
    with a Function is assembled anew each time.

        u = fd.Function(V, "u")           # zero
        u.vector[:] = ...  /  u.interpolate(f)  /  u.assign(w)
        form(u * dot(grad(w), grad(v)) * dx)"""
    def __init__(self, V, name=None, vector=None):
        if isinstance(name, (np.ndarray, list, tuple)) and vector is None:
            name, vector = None, name
        self._space = V
        self.name = "w" if name is None else str(name)
        self._vector = np.zeros(V.dim) if vector is None else self._checked(vector)
        super().__init__(None, (0, 0))

    @property
    def owner(self):
        # a Function reads its own vector (held as a property: an attribute would be a reference cycle)
        return self

    @property
    def space(self):
        return self._space

    def _checked(self, a):
        v = np.array(a, dtype=np.float64).reshape(-1)
        if v.shape != (self._space.dim,):
            raise ValueError(f"Function: vector length {v.size} != V.dim = {self._space.dim}")
        return v

    @property
    def vector(self):
        return self._vector

    @vector.setter
    def vector(self, a):
        self._vector = self._checked(a)

    def __array__(self, dtype=None, copy=None):
        return self._vector if dtype is None else self._vector.astype(dtype)

    def assign(self, other):
        """Copy another Function's (or an array's) coefficients into this one."""
        self.vector = other.vector if isinstance(other, Function) else other
        return self

    def interpolate(self, f):
        """Nodal interpolation of f (a callable of x, y, or an expression in fd.x, fd.y) onto the
        space's free nodes."""
        self.vector = self._space.interpolate(f)
        return self


def _as_expr(o):
    """A node for an operand: nodes as they are, numbers as Const."""
    if isinstance(o, Expr):
        return o
    if isinstance(o, ListTensor):
        raise TypeError(f"a vector {o!r} where a scalar is needed: use dot(), inner() or a component")
    if isinstance(o, (int, float, np.floating, np.integer)):
        return Const(float(o))
    raise TypeError(f"cannot use {o!r} in a form")


def _negated(e):
    return Prod([Const(-1.0), e])


def _mul(a, b):
    if isinstance(b, Measure):
        return Integral(a, b)
    if isinstance(b, ListTensor):
        return ListTensor([a * c for c in b])
    return Prod([a, _as_expr(b)])


def _div(a, b):
    b = _as_expr(b)
    if _mentions(b, Argument):
        raise TypeError("cannot divide by an expression containing a test or trial function")
    return Prod([a, Pow(b, -1)])


# operator -> node builder (self, other); "r" variants have the operand order swapped
_OPERATORS = {
    "__add__": lambda a, b: Sum([a, _as_expr(b)]),
    "__radd__": lambda a, b: Sum([_as_expr(b), a]),
    "__sub__": lambda a, b: Sum([a, _negated(_as_expr(b))]),
    "__rsub__": lambda a, b: Sum([_as_expr(b), _negated(a)]),
    "__mul__": _mul,
    "__rmul__": lambda a, b: Prod([_as_expr(b), a]),
    "__truediv__": _div,
    "__rtruediv__": lambda a, b: Prod([_as_expr(b), Pow(a, -1)]),
}
for _name, _build in _OPERATORS.items():
    setattr(Expr, _name, _build)
del _name, _build


def _function(name):
    def make(e):
        return Func(name, _as_expr(e))
    make.__name__ = name
    make.__doc__ = f"{name}(e) in a coefficient (evaluated at the quadrature points on the GPU)."
    return make


sin, cos, exp, log, tanh, sqrt, sinh, cosh, sech, sign = (_function(f) for f in
                                                           ("sin", "cos", "exp", "log", "tanh", "sqrt", "sinh",
                                                            "cosh", "sech", "sign"))
abs_ = _function("abs")


class ListTensor:
    """A vector of scalar expressions (grad(u), the facet normal)."""
    def __init__(self, components):
        self.components = [_as_expr(c) for c in components]

    def __len__(self):
        return len(self.components)

    def __getitem__(self, i):
        return self.components[i]

    def __iter__(self):
        return iter(self.components)

    def __repr__(self):
        return "Vector(" + ", ".join(repr(c) for c in self.components) + ")"

    def _zip(self, o, fn):
        return ListTensor([fn(a, b) for a, b in zip(self.components, o)])

    def __add__(self, o): return self._zip(o, lambda a, b: a + b)
    def __sub__(self, o): return self._zip(o, lambda a, b: a - b)
    def __neg__(self): return ListTensor([-a for a in self.components])
    def __mul__(self, o): return ListTensor([a * o for a in self.components])
    def __rmul__(self, o): return ListTensor([o * a for a in self.components])


def as_vector(c):
    return c if isinstance(c, ListTensor) else ListTensor(c)


def FacetNormal(mesh=None):
    """The outward unit normal (n_x, n_y) on the exterior facets, for ds terms."""
    return ListTensor([Normal(0), Normal(1)])


# ============================================================================ tree walks
def _children(e):
    if isinstance(e, Sum):
        return e.terms
    if isinstance(e, Prod):
        return e.factors
    if isinstance(e, Pow):
        return (e.base,)
    if isinstance(e, Func):
        return (e.arg,)
    return ()


def _mentions(e, kind):
    """Whether the tree holds a node of class `kind` (iteratively: deep coefficient sums)."""
    todo = [e]
    while todo:
        node = todo.pop()
        if isinstance(node, kind):
            return True
        todo.extend(_children(node))
    return False


def _has_arguments(e):
    return _mentions(e, Argument)


def _has_normal(e):
    return _mentions(e, Normal)


# ---- derivatives (python/femd/forms.py:368-416 give the same rules and the same tree shapes)
# d f(g) / dg for the differentiable functions
_OUTER_DERIVATIVE = {
    "sin": lambda g: cos(g),
    "cos": lambda g: _negated(sin(g)),
    "exp": lambda g: exp(g),
    "log": lambda g: Pow(g, -1),
    "sqrt": lambda g: Prod([Const(0.5), Pow(sqrt(g), -1)]),
    "sinh": lambda g: cosh(g),
    "cosh": lambda g: sinh(g),
    "tanh": lambda g: Sum([Const(1.0), _negated(Pow(tanh(g), 2))]),
}


def _d_product(e, i):
    # constant factors stay in front; one product per non-constant factor, that factor differentiated
    fixed = [f for f in e.factors if isinstance(f, Const)]
    varying = [f for f in e.factors if not isinstance(f, Const)]
    if not varying:
        return Const(0.0)
    if len(varying) == 1:
        return Prod(fixed + [Dx(varying[0], i)])
    parts = []
    for j, f in enumerate(varying):
        parts.append(Prod(fixed + varying[:j] + [Dx(f, i)] + varying[j + 1:]))
    return Sum(parts)


def _d_function(e, i):
    outer = _OUTER_DERIVATIVE.get(e.name)
    if outer is None:
        raise TypeError(f"cannot differentiate {e.name}")
    return Prod([outer(e.arg), Dx(e.arg, i)])


def _d_argument(e, i):
    d = list(e.deriv)
    d[i] += 1
    return Argument(e.kind, e.space, tuple(d))


def _d_field(e, i):
    d = list(e.deriv)
    d[i] += 1
    return Field(e.owner, tuple(d))


_DERIVATIVE = {
    Argument: _d_argument,
    Field: _d_field,
    Function: _d_field,
    Sum: lambda e, i: Sum([Dx(t, i) for t in e.terms]),
    Prod: _d_product,
    Const: lambda e, i: Const(0.0),
    Constant: lambda e, i: Const(0.0),
    Spatial: lambda e, i: Const(1.0 if e.index == i else 0.0),
    Pow: lambda e, i: Prod([Const(float(e.n)), Pow(e.base, e.n - 1), Dx(e.base, i)]),
    Func: _d_function,
}


def Dx(e, i: int):
    """d e / d x_i by the sum, product and chain rules."""
    if i not in (0, 1):
        raise ValueError(f"Dx(e, i): i is 0 (x) or 1 (y), got {i!r}")
    if isinstance(e, ListTensor):
        return ListTensor([Dx(c, i) for c in e.components])
    if isinstance(e, (int, float, np.floating, np.integer)):
        return Const(0.0)
    rule = _DERIVATIVE.get(type(e))
    if rule is None:
        raise TypeError(f"cannot differentiate {e!r}")
    return rule(e, i)


def grad(e):
    """(Dx(e, 0), Dx(e, 1)) of a scalar."""
    if isinstance(e, ListTensor):
        raise NotImplementedError("grad() of a vector is outside this path")
    e = _as_expr(e)
    return ListTensor([Dx(e, 0), Dx(e, 1)])


def div(v):
    v = as_vector(v)
    return Sum([Dx(c, i) for i, c in enumerate(v.components)])


def dot(a, b):
    """a . b: the product of two scalars, or sum_i a_i b_i of two vectors of one length."""
    a = as_vector(a) if isinstance(a, (list, tuple)) else a
    b = as_vector(b) if isinstance(b, (list, tuple)) else b
    va, vb = isinstance(a, ListTensor), isinstance(b, ListTensor)
    if va != vb:
        raise TypeError("dot(a, b): a scalar and a vector; use a * b for scaling")
    if not va:
        return _as_expr(a) * _as_expr(b)
    if len(a) != len(b):
        raise TypeError(f"dot(a, b): lengths {len(a)} and {len(b)} differ")
    return Sum([p * q for p, q in zip(a.components, b.components)])


inner = dot


# ============================================================================ measures
class Measure:
    """dx (cells), ds (exterior facets: all of them, or those with given side markers) and dS
    (interior facets, outside this path).  where: None for dx, "both" for every exterior facet,
    else a sorted tuple of markers; quad_degree overrides the inferred degree."""
    def __init__(self, kind, where=None, quad_degree=None):
        self.kind, self.where, self.quad_degree = kind, where, quad_degree

    def __call__(self, *args, quad_degree=None, **kw):
        if self.kind == "dS":
            raise NotImplementedError("interior-facet (dS) integrals are outside this path")
        if kw:
            raise NotImplementedError(f"{self.kind} options other than the degree are outside this path")
        if self.kind == "dx":
            degree = args[0] if args else quad_degree
            if degree is not None and int(degree) < 1:
                raise ValueError(f"dx(n): n must be at least 1, got {degree}")
            return Measure("dx", None, degree)
        return Measure("ds", _facet_selection(args), quad_degree)

    def __rmul__(self, e):
        return Integral(_as_expr(e), self)


def _facet_selection(args):
    """ds(m), ds((m1, m2)), ds(m1, m2) -> the sorted markers; ds() or ds("both") -> every facet."""
    sel = args[0] if len(args) == 1 else (tuple(args) if args else None)
    if sel is None or sel == "both":
        return "both"
    if isinstance(sel, str):
        if sel in ("left", "right"):
            raise TypeError(f"ds('{sel}') is for 1D meshes; on a 2D mesh use ds(marker), ds((m1, m2)) or ds")
        raise ValueError("ds takes a marker or markers on a 2D mesh")
    markers = (sel,) if not isinstance(sel, (tuple, list)) else tuple(sel)
    if not markers or not all(isinstance(m, (int, np.integer)) and not isinstance(m, bool) for m in markers):
        raise ValueError("ds takes a marker or markers on a 2D mesh")
    return tuple(sorted({int(m) for m in markers}))


dx = Measure("dx")
ds = Measure("ds", "both")
dS = Measure("dS")


class Integral:
    """expr * measure."""
    def __init__(self, expr, measure):
        if measure.kind == "dS":
            raise NotImplementedError("interior-facet (dS) integrals are outside this path")
        self.expr, self.measure = expr, measure

    def __add__(self, o): return FormExpr([self]) + o
    def __radd__(self, o): return FormExpr([self]) + o
    def __sub__(self, o): return FormExpr([self]) - o
    def __neg__(self): return Integral(_negated(self.expr), self.measure)
    def __mul__(self, c): return Integral(Prod([_as_expr(c), self.expr]), self.measure)
    __rmul__ = __mul__


class FormExpr:
    """A sum of integrals, kept in the order written."""
    def __init__(self, integrals):
        self.integrals = list(integrals)

    def __add__(self, o):
        if isinstance(o, Integral):
            return FormExpr(self.integrals + [o])
        if isinstance(o, FormExpr):
            return FormExpr(self.integrals + o.integrals)
        if o == 0:
            return self
        raise TypeError("can only add integrals to a form")

    __radd__ = __add__

    def __sub__(self, o):
        o = o if isinstance(o, FormExpr) else FormExpr([o])
        return FormExpr(self.integrals + [-i for i in o.integrals])

    def __neg__(self):
        return FormExpr([-i for i in self.integrals])


# ============================================================================ analysis
def _expand(e):
    """The integrand as a list of products (lists of factors), distributing over the sums that hold
    test or trial functions: sums in term order, products as the ordered cross product of their
    factors' expansions (the reference's order, python/femd/forms.py:1364-1381)."""
    if isinstance(e, Argument) or not _has_arguments(e):
        return [[e]]
    if isinstance(e, Sum):
        out = []
        for t in e.terms:
            out.extend(_expand(t))
        return out
    if isinstance(e, Prod):
        rows = [[]]
        for f in e.factors:
            rows = [head + tail for head in rows for tail in _expand(f)]
        return rows
    if isinstance(e, Pow) and e.n == 1:
        return _expand(e.base)
    if isinstance(e, Pow):
        raise TypeError(f"{e!r}: a form must be linear in each test and trial function")
    raise TypeError(f"cannot expand {e!r}")


class Term:
    """coeff * (test) * (trial) * measure; test and trial are Arguments or None."""
    def __init__(self, test, trial, coeff, measure):
        self.test, self.trial, self.coeff, self.measure = test, trial, coeff, measure

    @property
    def rank(self):
        return int(self.test is not None) + int(self.trial is not None)

    def __repr__(self):
        m = self.measure
        where = "" if m.kind == "dx" or m.where == "both" else str(m.where)
        parts = [repr(self.coeff)] + [repr(a) for a in (self.test, self.trial) if a is not None]
        return " * ".join(parts) + f" {m.kind}{where}"


def _is_zero(e):
    """Zero by construction: a zero constant, a product with one, a sum of them."""
    if isinstance(e, Const):
        return e.value == 0.0
    if isinstance(e, Prod):
        return any(_is_zero(f) for f in e.factors)
    if isinstance(e, Sum):
        return all(_is_zero(t) for t in e.terms)
    return False


def canonicalize(fe):
    """The form's terms coeff * test * trial, in expansion order (python/femd/forms.py:1398-1415):
    a product's factors other than the arguments form its coefficient (one factor as itself,
    several as their product, none as 1); terms with no arguments whose coefficient is zero by
    construction are dropped unless nothing else is left."""
    terms = []
    for integral in fe.integrals:
        for factors in _expand(integral.expr):
            args = {"test": [], "trial": []}
            rest = []
            for f in factors:
                (args[f.kind] if isinstance(f, Argument) else rest).append(f)
            if len(args["test"]) > 1 or len(args["trial"]) > 1:
                raise TypeError(f"term {factors!r} is not linear in the test/trial functions")
            if args["trial"] and not args["test"]:
                raise TypeError("a term with a trial function but no test function has no row index")
            coeff = Const(1.0) if not rest else (rest[0] if len(rest) == 1 else Prod(rest))
            test = args["test"][0] if args["test"] else None
            trial = args["trial"][0] if args["trial"] else None
            terms.append(Term(test, trial, coeff, integral.measure))
    kept = [t for t in terms if t.rank or not _is_zero(t.coeff)]
    return kept or terms


# ---- degrees (python/femd/forms.py:1429-1501, 654-662)
def _arg_degree(a):
    """Total degree of an argument on triangles (a derivative lowers it); on quadrilaterals the
    per-variable degree, which a derivative does not lower."""
    if getattr(a.space, "nverts", 3) == 4:
        return a.space.degree
    return max(a.space.degree - sum(a.deriv), 0)


def _pow_degree(n, d):
    if d is None or not float(n).is_integer() or n < 0:
        return None
    return int(n) * d


_POLY_DEGREE = {
    Const: lambda e: 0,
    Constant: lambda e: 0,
    Normal: lambda e: 0,
    Spatial: lambda e: 1,
    Argument: _arg_degree,
    Field: _arg_degree,
    Function: _arg_degree,
    Sum: lambda e: _combine([poly_degree(t) for t in e.terms], max),
    Prod: lambda e: _combine([poly_degree(f) for f in e.factors], sum),
    Pow: lambda e: _pow_degree(e.n, poly_degree(e.base)),
}


def _combine(degrees, how):
    if any(d is None for d in degrees):
        return None
    return how(degrees) if degrees else 0


def poly_degree(e):
    """Polynomial degree of a pointwise expression, None when it is not a polynomial."""
    rule = _POLY_DEGREE.get(type(e))
    return rule(e) if rule else None


def quad_degree_estimate(e, kmax):
    """The degree a quadrature rule needs for e: its polynomial degree, with every non-polynomial
    part that holds no argument counted as degree kmax."""
    d = poly_degree(e)
    if d is not None:
        return d
    if not _mentions(e, (Argument, Field)):          # data alone
        return kmax
    if isinstance(e, (Argument, Field)):
        return _arg_degree(e)
    if isinstance(e, Sum):
        return max((quad_degree_estimate(t, kmax) for t in e.terms), default=0)
    if isinstance(e, Prod):
        return sum(quad_degree_estimate(f, kmax) for f in e.factors)
    if isinstance(e, Pow) and float(e.n).is_integer() and e.n >= 0:
        return int(e.n) * quad_degree_estimate(e.base, kmax)
    return kmax


_default_quad_degree = None


def set_quadrature_degree(degree=None):
    """One rule for every form made from now on (python/femd/forms.py:1451-1466); None restores
    the inferred degrees."""
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


_ASSEMBLY_MODES = ("exact", "mfma", "mfma_f16", "mfma_bf16")


def assembly_mode(mode=None) -> str:
    """How matrices form their dx element matrices: "exact" (the default) in the reference's
    operation order, bit-identical to it, or "mfma" on the FP64 matrix cores as one GEMM over the
    cells (the same sums in another order, within a few units of the last place); "mfma_f16" and
    "mfma_bf16" round the GEMM's operands to 16 bits and accumulate in FP32 (low-precision local
    matrices, for experiments).  With `mode` set it first switches; returns the previous mode."""
    if mode is not None and mode not in _ASSEMBLY_MODES:
        raise ValueError(f"assembly_mode: mode must be one of {_ASSEMBLY_MODES}, got {mode!r}")
    prev = L.load().femd_assembly_mode(-1 if mode is None else _ASSEMBLY_MODES.index(mode))
    return _ASSEMBLY_MODES[prev]


# ============================================================================ programs
class Program:
    """A coefficient as a register program of the instruction set of forms/integrand.hpp:33-46
    (4 ints per instruction: op, out, a, b), which the GPU runs at every quadrature point.

    Emission order (the reference's, python/femd/forms.py:1822-1909): a node's operands before
    the node; a sum (product) emits its first term, then each further term followed at once by an
    ADD (MUL) of the running value and that term; a power emits its base, then POW with the
    exponent in the constant table; a node met again (the same object) reuses its register; a
    known field emits INPUT with the slot of its (vector, derivative), slots numbered in order of
    first use.  Every instruction writes a new register, so register k holds instruction k's result."""
    CONST, X, Y, NX, NY, INPUT, ADD, MUL, POW = range(9)
    _FUNCS = dict(_OPCODE)

    def __init__(self, expr):
        self.code, self.consts, self.memo = [], [], {}
        self.inputs, self.input_slots = [], {}       # the fields read, one per (vector, derivative)
        self._compile(expr)

    @property
    def nreg(self):
        return len(self.code)

    def _put(self, op, a=0, b=0):
        self.code.append((op, len(self.code), a, b))
        return len(self.code) - 1

    def _constant(self, c):
        self.consts.append(c)
        return len(self.consts) - 1

    def _leaf(self, e):
        if isinstance(e, (Const, Constant)):
            # the register is taken before the constant's slot, as each instruction takes it
            reg = len(self.code)
            self.code.append((self.CONST, reg, self._constant(e.value if isinstance(e, Const) else e), 0))
            return reg
        if isinstance(e, Spatial):
            return self._put(self.X if e.index == 0 else self.Y)
        if isinstance(e, Normal):
            return self._put(self.NX if e.index == 0 else self.NY, 1)
        if isinstance(e, Field):
            key = e.slot()
            if key not in self.input_slots:
                self.input_slots[key] = len(self.inputs)
                self.inputs.append(e)
            return self._put(self.INPUT, self.input_slots[key])
        raise NotImplementedError(f"cannot compile {e!r}")

    def _compile(self, root):
        """Post-order emission with an explicit stack of frames [node, operands taken, running
        register]; a node finished (or met again) hands its register to the frame below."""
        stack = [[root, 0, None]]
        while stack:
            frame = stack[-1]
            node, taken, _ = frame
            reg = None
            if taken == 0:
                seen = self.memo.get(id(node))
                if seen is not None and seen[0] is node:
                    stack.pop()
                    self._handover(stack, seen[1])
                    continue
            ops = _children(node)
            if taken < len(ops):
                frame[1] = taken + 1
                stack.append([ops[taken], 0, None])
                continue
            if ops:
                reg = frame[2]                              # the folded value
            elif isinstance(node, (Sum, Prod)):             # an empty sum (product) is 0 (1)
                reg = len(self.code)
                self.code.append((self.CONST, reg, self._constant(0.0 if isinstance(node, Sum) else 1.0), 0))
            else:
                reg = self._leaf(node)
            self.memo[id(node)] = (node, reg)
            stack.pop()
            self._handover(stack, reg)

    def _handover(self, stack, reg):
        if stack:
            self._absorb(stack[-1], reg)
        else:
            self.root = reg

    def _absorb(self, frame, reg):
        """An operand of frame's node is done (in register reg): fold it into the node's value."""
        node, k, acc = frame
        if isinstance(node, (Sum, Prod)):
            if acc is None:
                frame[2] = reg
            else:
                frame[2] = self._put(self.ADD if isinstance(node, Sum) else self.MUL, acc, reg)
        elif isinstance(node, Pow):
            out = len(self.code)
            self.code.append((self.POW, out, reg, self._constant(float(node.n))))
            frame[2] = out
        else:
            frame[2] = self._put(self._FUNCS[node.name], reg)

    def fields(self):
        """The known fields the INPUT slots read: (space handle, coefficient vector, derivative code)."""
        return [(f.space._h, f.owner.vector, f.code()) for f in self.inputs]

    def arrays(self):
        """(code int32 (n, 4), registers, constants float64 with the Constants' current values)."""
        consts = [c.value if isinstance(c, Constant) else c for c in self.consts]
        code = np.asarray(self.code, dtype=np.int32).reshape(-1, 4)
        return code, max(self.nreg, 1), np.asarray(consts, dtype=np.float64)


# ============================================================================ NumPy evaluation
_NUMPY = {"sin": np.sin, "cos": np.cos, "exp": np.exp, "log": np.log, "tanh": np.tanh, "sqrt": np.sqrt,
          "sinh": np.sinh, "cosh": np.cosh, "sech": lambda v: 1.0 / np.cosh(v), "abs": np.abs, "sign": np.sign}


def _eval_fold(values, op, start):
    out = start
    for v in values:
        out = op(out, v)
    return out


_EVALUATE = {
    Const: lambda e, x_, y_: e.value,
    Constant: lambda e, x_, y_: e.value,
    Spatial: lambda e, x_, y_: x_ if e.index == 0 else y_,
    Sum: lambda e, x_, y_: _eval_fold((evaluate(t, x_, y_) for t in e.terms), lambda a, b: a + b, 0.0),
    Prod: lambda e, x_, y_: _eval_fold((evaluate(f, x_, y_) for f in e.factors), lambda a, b: a * b, 1.0),
    Pow: lambda e, x_, y_: np.power(evaluate(e.base, x_, y_), e.n),
    Func: lambda e, x_, y_: _NUMPY[e.name](evaluate(e.arg, x_, y_)),
}


def evaluate(e, x, y):
    """An expression in fd.x, fd.y, numbers and Constants at the points (x, y), with NumPy (the
    Dirichlet data of V.lift)."""
    rule = _EVALUATE.get(type(e))
    if rule is None:
        raise TypeError(f"boundary data: an expression may contain fd.x, fd.y, numbers and Constants, not {e!r}")
    return rule(e, x, y)


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
        for p in self.programs:
            for f in p.inputs:
                f.code()                                      # first derivatives at most
                if f.space.mesh is not self.test_space.mesh:
                    raise ValueError(f"{f!r}: a known field must live on the form's mesh")
        for t, p in zip(self.terms, self.programs):
            if p.inputs and t.measure.kind != "dx":
                raise NotImplementedError("known fields in ds terms are outside this path")
        # a form reading Constants or known fields is assembled anew each time
        self.constant = not any(isinstance(c, Constant) for p in self.programs for c in p.consts) and \
            not any(p.inputs for p in self.programs)
        self._result = None

    def _coefficients(self, idx=None):
        items = []
        for i in (range(len(self.terms)) if idx is None else idx):
            code, nreg, consts = self.programs[i].arrays()
            items.append(((code, nreg, consts), None, self.programs[i].fields()))
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
