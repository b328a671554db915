from .cg import cg

__all__ = ["cg"]
from .mg import GeometricMG

__all__.append("GeometricMG")
