"""fealpy_amd -- MI355X-native Lagrange finite-element assemble-and-solve.

Drop-in for the path of weihuayi/fealpy that goes
LagrangeFESpace -> BilinearForm/LinearForm (diffusion, mass, source) ->
DirichletBC -> CSRTensor -> cg, on torch device tensors, with every numerical
step in hand-written HIP kernels for gfx950 (libfealpy_amd.so, C-ABI in
include/fealpy_amd.h).
"""
from .decorator import barycentric, cartesian
from .fem import BilinearForm, DirichletBC, GroupIntegrator, LinearForm, ScalarDiffusionIntegrator, \
    ScalarMassIntegrator, ScalarSourceIntegrator
from .functionspace import Function, LagrangeFESpace
from .mesh import TetrahedronMesh, TriangleMesh
from .solver import cg
from .sparse import COOTensor, CSRTensor

__version__ = "0.1.0"
