from .tensors import COOTensor, CSRTensor, SparseTensor, csr_matrix, spdiags

__all__ = ["COOTensor", "CSRTensor", "SparseTensor", "csr_matrix", "spdiags"]
