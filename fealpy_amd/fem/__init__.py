from .dirichlet_bc import DirichletBC
from .forms import BilinearForm, LinearForm
from .integrators import GroupIntegrator, Integrator, ScalarDiffusionIntegrator, ScalarMassIntegrator, \
    ScalarSourceIntegrator

__all__ = ["BilinearForm", "LinearForm", "DirichletBC", "Integrator", "GroupIntegrator",
           "ScalarDiffusionIntegrator", "ScalarMassIntegrator", "ScalarSourceIntegrator"]
