from .lagrange import Function, LagrangeFESpace

__all__ = ["Function", "LagrangeFESpace"]
