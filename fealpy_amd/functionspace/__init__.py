from .lagrange import Function, LagrangeFESpace
from .tensor_space import TensorFunctionSpace

__all__ = ["Function", "LagrangeFESpace", "TensorFunctionSpace"]
