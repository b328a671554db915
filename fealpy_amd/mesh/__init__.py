from .simplex import SimplexMesh, TetrahedronMesh, TriangleMesh

__all__ = ["SimplexMesh", "TriangleMesh", "TetrahedronMesh"]
