from .cg import cg

__all__ = ["cg"]
from .mg import GeometricMG, pcg_workgroups

__all__ += ["GeometricMG", "pcg_workgroups"]
from .amg import AlgebraicMG

__all__.append("AlgebraicMG")
from .refine import cg_refine

__all__.append("cg_refine")
