"""COO / CSR sparse tensors on device memory (fealpy/sparse/coo_tensor.py, csr_tensor.py).

CSR index arrays are int32 (crow, col); values FP64 or FP32.  Matrix-vector
products run on the HIP SpMV kernel.  Conversions that are not on the
assemble-and-solve path (COO coalesce/tocsr, CSR + CSR of different patterns,
``to_scipy``) are plain tensor programs.
"""
import numpy as np
import torch

from .. import _native as nat


def _spmv(A, x):
    L = nat.lib()
    if x.dtype != A.values.dtype:
        raise TypeError(f"SpMV operand dtype {x.dtype} differs from matrix values {A.values.dtype}")
    if x.ndim != 1 or x.shape[0] != A.shape[1]:
        raise ValueError(f"SpMV operand of shape {tuple(x.shape)} for a matrix of shape {A.shape}")
    x = x.contiguous()
    y = torch.empty(A.shape[0], dtype=x.dtype, device=x.device)
    nat.check(L.fa_spmv(nat.dtype_code(x.dtype), A.shape[0], A.nnz, nat.ptr(A.crow), nat.ptr(A.col),
                        nat.ptr(A.values), nat.ptr(x), nat.ptr(y), nat.stream()), "fa_spmv")
    return y


def _spmm(A, X):
    """A X for a 2-D X (ncols, nvec): one launch over all columns (fa_spmm)."""
    L = nat.lib()
    if X.dtype != A.values.dtype:
        raise TypeError(f"SpMV operand dtype {X.dtype} differs from matrix values {A.values.dtype}")
    if X.shape[0] != A.shape[1]:
        raise ValueError(f"SpMV operand of shape {tuple(X.shape)} for a matrix of shape {A.shape}")
    X = X.contiguous()
    Y = torch.empty((A.shape[0], X.shape[1]), dtype=X.dtype, device=X.device)
    if X.shape[1] == 0:
        return Y
    nat.check(L.fa_spmm(nat.dtype_code(X.dtype), A.shape[0], A.nnz, X.shape[1], nat.ptr(A.crow), nat.ptr(A.col),
                        nat.ptr(A.values), nat.ptr(X), nat.ptr(Y), nat.stream()), "fa_spmm")
    return Y


class SparseTensor:
    pass


class CSRTensor(SparseTensor):
    def __init__(self, crow, col, values, spshape=None):
        self._crow = crow if crow.dtype == torch.int32 else crow.to(torch.int32)
        self._col = col if col.dtype == torch.int32 else col.to(torch.int32)
        self._values = values
        if spshape is None:
            ncol = int(self._col.max().item()) + 1 if self._col.numel() else 0
            spshape = (self._crow.shape[0] - 1, ncol)
        self._spshape = tuple(int(s) for s in spshape)
        if values is not None and values.shape[-1] != self._col.shape[0]:
            raise ValueError("values and col must have the same number of non-zeros")

    # -- properties (csr_tensor.py) --------------------------------------------
    @property
    def crow(self): return self._crow
    indptr = crow

    @property
    def col(self): return self._col
    indices = col

    @property
    def values(self): return self._values
    data = values

    @property
    def nnz(self): return int(self._col.shape[0])

    @property
    def shape(self): return self._spshape

    @property
    def sparse_shape(self): return self._spshape

    @property
    def dtype(self): return self._values.dtype

    @property
    def device(self): return self._col.device

    @property
    def itype(self): return self._col.dtype

    @property
    def ftype(self): return self._values.dtype

    def values_context(self):
        return {'dtype': self._values.dtype, 'device': self._values.device}

    def __repr__(self):
        return f"CSRTensor(shape={self.shape}, nnz={self.nnz}, dtype={self.dtype}, device={self.device})"

    # -- products ---------------------------------------------------------------
    def matmul(self, other):
        if isinstance(other, SparseTensor):
            raise NotImplementedError("sparse @ sparse is not provided")
        if other.ndim == 1:
            return _spmv(self, other)
        if other.ndim == 2:
            return _spmm(self, other)
        raise ValueError("`other` must be a 1-D or 2-D array.")

    __matmul__ = matmul

    # -- conversions -------------------------------------------------------------
    def astype(self, dtype=None, /, *, copy=True):
        if dtype is None or dtype == self.dtype:
            return CSRTensor(self._crow, self._col, self._values.clone() if copy else self._values, self._spshape)
        out = torch.empty(self.nnz, dtype=dtype, device=self.device)
        nat.check(nat.lib().fa_cast(nat.dtype_code(self.dtype), nat.dtype_code(dtype), self.nnz,
                                    nat.ptr(self._values), nat.ptr(out), nat.stream()), "fa_cast")
        return CSRTensor(self._crow, self._col, out, self._spshape)

    def device_put(self, device=None):
        return CSRTensor(self._crow.to(device), self._col.to(device), self._values.to(device), self._spshape)

    def row_index(self):
        counts = (self._crow[1:] - self._crow[:-1]).long()
        return torch.repeat_interleave(torch.arange(self.shape[0], device=self.device), counts)

    def tocoo(self):
        idx = torch.stack([self.row_index().to(torch.int32), self._col])
        return COOTensor(idx, self._values, self._spshape, is_coalesced=True)

    def to_dense(self):
        D = torch.zeros(self.shape, dtype=self.dtype, device=self.device)
        D.index_put_((self.row_index(), self._col.long()), self._values, accumulate=True)
        return D

    toarray = to_dense

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self._values.detach().cpu().numpy(), self._col.cpu().numpy(),
                              self._crow.cpu().numpy()), shape=self.shape)

    def diagonal(self):
        rows = self.row_index()
        d = torch.zeros(self.shape[0], dtype=self.dtype, device=self.device)
        on = rows == self._col.long()
        d[rows[on]] = self._values[on]
        return d

    def add(self, other, alpha=1):
        """CSRTensor.add (csr_tensor.py:272-329)."""
        if isinstance(other, CSRTensor):
            if other.shape != self.shape:
                raise ValueError("shape mismatch")
            if other.nnz == self.nnz and torch.equal(other.crow, self.crow) and torch.equal(other.col, self.col):
                return CSRTensor(self._crow, self._col, self._values + alpha * other.values, self._spshape)
            return self.tocoo().add(other.tocoo(), alpha).coalesce().tocsr()
        if torch.is_tensor(other):
            return self.to_dense() + alpha * other
        if isinstance(other, (int, float)):
            return CSRTensor(self._crow.clone(), self._col.clone(), self._values + alpha * other, self._spshape)
        raise TypeError(f"Unsupported type {type(other).__name__} in addition")

    __add__ = add

    def tril(self, k=0):
        keep = self._col.long() <= self.row_index() + k
        return self.tocoo()._select(keep).tocsr()

    def triu(self, k=0):
        keep = self._col.long() >= self.row_index() + k
        return self.tocoo()._select(keep).tocsr()


class COOTensor(SparseTensor):
    def __init__(self, indices, values, spshape=None, *, is_coalesced=None):
        if indices.ndim != 2:
            raise ValueError(f"indices must be a 2D tensor, but got {indices.ndim}D")
        if values is not None and values.shape[-1] != indices.shape[1]:
            raise ValueError("values must have the same size as indices in the last dimension")
        self._indices = indices
        self._values = values
        self.is_coalesced = is_coalesced
        if spshape is None:
            spshape = tuple(int(v) + 1 for v in indices.max(dim=1).values.tolist())
        self._spshape = tuple(int(s) for s in spshape)

    @property
    def indices(self): return self._indices

    @property
    def values(self): return self._values

    @property
    def nnz(self): return int(self._indices.shape[1])

    @property
    def shape(self): return self._spshape

    sparse_shape = shape

    @property
    def device(self): return self._indices.device

    def _select(self, keep):
        return COOTensor(self._indices[:, keep], self._values[keep], self._spshape)

    def coalesce(self, accumulate=True):
        """Stable sort by (row, col) and sum duplicates (coo_tensor.py:184-213)."""
        if self.is_coalesced or self.nnz == 0:
            return self
        n1 = self._spshape[1] if len(self._spshape) > 1 else 1
        key = self._indices[0].long() * n1 + (self._indices[1].long() if len(self._spshape) > 1 else 0)
        key_s, order = torch.sort(key, stable=True)
        uniq = torch.ones_like(key_s, dtype=torch.bool)
        uniq[1:] = key_s[1:] != key_s[:-1]
        grp = torch.cumsum(uniq.long(), 0) - 1
        vals = torch.zeros(int(uniq.sum()), dtype=self._values.dtype, device=self.device)
        vals.index_add_(0, grp, self._values[order])
        return COOTensor(self._indices[:, order][:, uniq], vals, self._spshape, is_coalesced=True)

    def tocsr(self):
        """Row counts -> crow; stable ordering by row (coo_tensor.py:137-157)."""
        n = self._spshape[0]
        rows = self._indices[0].long()
        crow = torch.zeros(n + 1, dtype=torch.int64, device=self.device)
        crow[1:] = torch.cumsum(torch.bincount(rows, minlength=n), 0)
        order = torch.sort(rows, stable=True).indices
        return CSRTensor(crow.to(torch.int32), self._indices[-1, order].to(torch.int32), self._values[order],
                         self._spshape)

    def to_dense(self):
        D = torch.zeros(self._spshape, dtype=self._values.dtype, device=self.device)
        D.index_put_(tuple(i.long() for i in self._indices), self._values, accumulate=True)
        return D

    def add(self, other, alpha=1):
        if isinstance(other, COOTensor):
            return COOTensor(torch.cat([self._indices, other.indices.to(self._indices.dtype)], dim=1),
                             torch.cat([self._values, other.values * alpha]), self._spshape)
        raise TypeError(f"Unsupported type {type(other).__name__} in addition")

    def matmul(self, x):
        return self.coalesce().tocsr().matmul(x)

    __matmul__ = matmul

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.coo_matrix((self._values.cpu().numpy(), tuple(self._indices.cpu().numpy())), shape=self._spshape)


def spdiags(data, diags, M, N, format='csr'):
    """Main-diagonal-only form of fealpy/sparse/ops.py:18-86 (non-zero entries only)."""
    if diags != 0:
        raise NotImplementedError("only the main diagonal is provided")
    n = min(M, N, data.shape[0])
    keep = torch.nonzero(data[:n] != 0)[:, 0]
    idx = torch.stack([keep, keep]).to(torch.int32)
    coo = COOTensor(idx, data[keep], (M, N), is_coalesced=True)
    return coo if format == 'coo' else coo.tocsr()


def csr_matrix(arg, shape=None):
    """Build a CSRTensor from (values, (row, col)) or (values, col, crow), scipy style."""
    values, idx = arg[0], arg[1:]
    if len(idx) == 1:
        row, col = idx[0]
        return COOTensor(torch.stack([torch.as_tensor(row), torch.as_tensor(col)]), torch.as_tensor(values),
                         shape).coalesce().tocsr()
    col, crow = idx
    return CSRTensor(torch.as_tensor(crow), torch.as_tensor(col), torch.as_tensor(values), shape)


def as_numpy_csr(A):
    """(indptr int64, indices int64, data) of a CSRTensor on the host."""
    return (A.crow.cpu().numpy().astype(np.int64), A.col.cpu().numpy().astype(np.int64),
            A.values.detach().cpu().numpy())
