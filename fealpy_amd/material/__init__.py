"""Materials (fealpy/material/)."""
from .elastic_material import LinearElasticMaterial

__all__ = ["LinearElasticMaterial"]
