"""fealpy_amd -- MI355X-native Lagrange finite-element assemble-and-solve.

Drop-in for the path of weihuayi/fealpy that goes
LagrangeFESpace / TensorFunctionSpace -> BilinearForm/LinearForm (diffusion,
mass, source, linear elasticity, vector source) ->
DirichletBC -> CSRTensor -> cg, on torch device tensors, with every numerical
step in hand-written HIP kernels for gfx950 (libfealpy_amd.so, C-ABI in
include/fealpy_amd.h).
"""
from . import expr
from .decorator import barycentric, cartesian
from .fem import BilinearForm, DirichletBC, GroupIntegrator, LinearElasticityIntegrator, LinearForm, \
    ScalarDiffusionIntegrator, ScalarMassIntegrator, ScalarSourceIntegrator, VectorSourceIntegrator
from .functionspace import Function, LagrangeFESpace, TensorFunctionSpace
from .material import LinearElasticMaterial
from .mesh import TetrahedronMesh, TriangleMesh
from .solver import GeometricMG, cg
from .sparse import COOTensor, CSRTensor

__version__ = "0.1.0"
