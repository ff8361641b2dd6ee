"""femd_amd -- FEMd's 2D assemble -> SpMV -> Krylov path on an AMD Instinct MI355X (gfx950).

The public names follow the reference's flat namespace (python/femd/__init__.py):

    import femd_amd as fd
    m = fd.rectangle_mesh(0, 0, 1, 1, 32, 32)
    V = fd.LagrangeSpace(m, 2, bc="dirichlet")
    u, v = fd.TrialFunction(V), fd.TestFunction(V)
    A = fd.form((fd.dot(fd.grad(u), fd.grad(v)) + u*v) * fd.dx).assemble()
    b = fd.form(f * v * fd.dx).assemble()
    x, info = fd.cg(A, b, M=A.preconditioner("jacobi"))

Everything numeric runs in libfemd_hip.so (hand-written HIP kernels); without it the import fails.
"""
from . import _lib

_lib.load()        # no CPU fallback: the HIP library must be built

from .mesh import Mesh2D, QuadMesh, rectangle_mesh, mesh_from_arrays, rectangle_quad_mesh, quad_mesh_from_arrays
from .spaces import LagrangeSpace, LagrangeSpace2D
from .forms import (Const, Constant, Dx, FacetNormal, Function, TestFunction, TrialFunction, abs_, as_vector, assembly_mode, cos, cosh, div, dot,
                    ds, dS, dx, exp, form, grad, inner, log, quadrature_degree, sech, set_quadrature_degree, sign, sin,
                    sinh, sqrt, tanh, x, y)

n = FacetNormal()
from .sparse import Preconditioner, SparseMatrix, SparseSolver
from .krylov import KrylovInfo, cg, gmres, lgmres, solve_refined
from .linalg import solve
from .batch import Batch, device_setup, inner_order, value_dictionary
from .timestep import BackwardEuler
from .serve import Server

__version__ = "0.9.0"
