from .dirichlet_bc import DirichletBC
from .elasticity import LinearElasticityIntegrator, VectorSourceIntegrator
from .forms import BilinearForm, LinearForm
from .integrators import GroupIntegrator, Integrator, ScalarDiffusionIntegrator, ScalarMassIntegrator, \
    ScalarSourceIntegrator

__all__ = ["BilinearForm", "LinearForm", "DirichletBC", "Integrator", "GroupIntegrator",
           "ScalarDiffusionIntegrator", "ScalarMassIntegrator", "ScalarSourceIntegrator",
           "LinearElasticityIntegrator", "VectorSourceIntegrator"]
