from .cg import cg

__all__ = ["cg"]
from .mg import GeometricMG

__all__.append("GeometricMG")
from .amg import AlgebraicMG

__all__.append("AlgebraicMG")
from .refine import cg_refine

__all__.append("cg_refine")
