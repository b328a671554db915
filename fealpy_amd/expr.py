"""Separable coefficient expressions, evaluated inside the element kernels.

The reference takes coefficients and sources as Python callables
(``@cartesian`` functions of the points, fealpy/decorator/coordinates.py:10-22,
dispatched by process_coef_func, fealpy/utils/utils.py:21-53).  A callable
run through torch needs the (NC, NQ, GD) quadrature points in HBM and a
chain of elementwise kernels.  A function that is a sum of products of
one-dimensional factors can instead be handed to the kernels as a short
table (``fa_sep_t`` in include/fealpy_amd.h) and evaluated in registers at
each quadrature point:

    f(x) = sum_t c_t * [k_t(cell)] * prod_k g_k(x[d_k]),
    g = sin(a s + b) | cos(a s + b) | exp(a s + b) | a s + b | s^n.

Build one from the coordinate symbols::

    from fealpy_amd.expr import x, y, z, sin, cell
    u = sin(pi * x) * sin(pi * y) * sin(pi * z)
    f = 3 * pi**2 * cell(k) * u          # k: per-cell FP64 tensor

An expression is also an ordinary ``@cartesian`` callable: ``f(p)`` evaluates
it with torch on points ``p[..., GD]``, so it can be passed anywhere the
reference takes a cartesian function (Dirichlet data, ``mesh.error``,
integrators whose kernel does not read the table).
"""
import ctypes
import math
import numbers

import torch

__all__ = ["x", "y", "z", "sin", "cos", "exp", "cell", "Sep", "Affine"]

SIN, COS, EXP, AFFINE, POW, SINPI, COSPI = range(7)
MAX_TERMS, MAX_FACTORS = 16, 48


class FaSep(ctypes.Structure):
    """ctypes image of fa_sep_t (include/fealpy_amd.h)."""
    _fields_ = [("nterms", ctypes.c_int32), ("nfactors", ctypes.c_int32),
                ("term_end", ctypes.c_int32 * MAX_TERMS),
                ("kind", ctypes.c_int32 * MAX_FACTORS), ("dim", ctypes.c_int32 * MAX_FACTORS),
                ("coef", ctypes.c_double * MAX_TERMS),
                ("a", ctypes.c_double * MAX_FACTORS), ("b", ctypes.c_double * MAX_FACTORS),
                ("field", ctypes.c_void_p * MAX_TERMS), ("bcs", ctypes.c_void_p)]


class Affine:
    """a * s + b for one coordinate s = x[dim]."""
    __slots__ = ("dim", "a", "b")

    def __init__(self, dim, a=1.0, b=0.0):
        self.dim, self.a, self.b = int(dim), float(a), float(b)

    def __mul__(self, o):
        if isinstance(o, numbers.Real):
            return Affine(self.dim, self.a * o, self.b * o)
        return Sep.of(self) * o

    __rmul__ = __mul__

    def __add__(self, o):
        if isinstance(o, numbers.Real):
            return Affine(self.dim, self.a, self.b + o)
        if isinstance(o, Affine) and o.dim == self.dim:
            return Affine(self.dim, self.a + o.a, self.b + o.b)
        return Sep.of(self) + o

    __radd__ = __add__

    def __neg__(self):
        return Affine(self.dim, -self.a, -self.b)

    def __sub__(self, o):
        return self + (-o)

    def __rsub__(self, o):
        return (-self) + o

    def __truediv__(self, o):
        if not isinstance(o, numbers.Real):
            return NotImplemented
        return self * (1.0 / o)

    def __pow__(self, n):
        return Sep.of(self) ** n

    def __call__(self, p):
        return Sep.of(self)(p)


class Sep:
    """Sum of terms (coef, field or None, ((kind, dim, a, b), ...))."""
    coordtype = 'cartesian'

    def __init__(self, terms):
        self.terms = list(terms)
        if len(self.terms) > MAX_TERMS:
            raise ValueError(f"a separable expression holds at most {MAX_TERMS} terms ({len(self.terms)})")
        if sum(len(t[2]) for t in self.terms) > MAX_FACTORS:
            raise ValueError(f"a separable expression holds at most {MAX_FACTORS} factors")
        self._dev = {}

    @staticmethod
    def of(o):
        if isinstance(o, Sep):
            return o
        if isinstance(o, numbers.Real):
            return Sep([(float(o), None, ())])
        if isinstance(o, Affine):
            if o.b == 0.0:
                return Sep([(o.a, None, ((POW, o.dim, 1.0, 0.0),))])
            return Sep([(1.0, None, ((AFFINE, o.dim, o.a, o.b),))])
        raise TypeError(f"cannot use {type(o).__name__} in a separable expression")

    # -- algebra ----------------------------------------------------------------
    def __add__(self, o):
        try:
            o = Sep.of(o)
        except TypeError:
            return NotImplemented
        return Sep(self.terms + o.terms)

    __radd__ = __add__

    def __neg__(self):
        return Sep([(-c, f, fs) for c, f, fs in self.terms])

    def __sub__(self, o):
        return self + (-Sep.of(o))

    def __rsub__(self, o):
        return Sep.of(o) + (-self)

    def __mul__(self, o):
        try:
            o = Sep.of(o)
        except TypeError:
            return NotImplemented
        out = []
        for c1, f1, g1 in self.terms:
            for c2, f2, g2 in o.terms:
                if f1 is not None and f2 is not None:
                    raise ValueError("a term of a separable expression holds at most one cell field")
                out.append((c1 * c2, f1 if f1 is not None else f2, g1 + g2))
        return Sep(out)

    __rmul__ = __mul__

    def __truediv__(self, o):
        if not isinstance(o, numbers.Real):
            return NotImplemented
        return self * (1.0 / o)

    def __pow__(self, n):
        if not (isinstance(n, numbers.Integral) and n >= 0):
            raise ValueError("separable expressions take whole powers >= 0")
        out = Sep.of(1.0)
        for _ in range(int(n)):
            out = out * self
        return out

    # -- evaluation with torch ------------------------------------------------------
    def __call__(self, p, index=None):
        """f at cartesian points p[..., GD].  A cell field is taken along the
        first axis of p, which must then be the cell axis ((NC, NQ, GD) or
        (NC, GD)), or the cells `index` selects."""
        p = torch.as_tensor(p)
        lead = p.shape[:-1]
        if (p.is_cuda and p.dtype == torch.float64 and 1 <= p.shape[-1] <= 3 and self.max_dim() < p.shape[-1]
                and not self.fields() and len(self.terms) <= 16 and sum(len(fs) for _, _, fs in self.terms) <= 48):
            # field-free f at device points: one kernel (fa_sep_points), each
            # factor's argument rounded as below, then sin / cos / exp of it
            from . import _native as nat
            L = nat.lib()
            q = p.reshape(-1, p.shape[-1]).contiguous()
            out = torch.empty(q.shape[0], dtype=torch.float64, device=p.device)
            nat.check(L.fa_sep_points(nat.ptr(self.point_table(p.device)), q.shape[0], q.shape[1], nat.ptr(q),
                                      nat.ptr(out), nat.stream()), "fa_sep_points")
            return out.reshape(lead)
        out = torch.zeros(lead, dtype=torch.float64, device=p.device)
        for c, fld, fs in self.terms:
            v = torch.full(lead, c, dtype=torch.float64, device=p.device)
            if fld is not None:
                kf = fld if index is None else fld[index]
                if p.ndim < 2 or kf.shape[0] != p.shape[0]:
                    raise ValueError("a cell field needs points with the cell axis first")
                v = v * kf.to(p.device).reshape((kf.shape[0],) + (1,) * (p.ndim - 2))
            for kind, d, a, b in fs:
                s = p[..., d]
                if kind == SIN:
                    g = torch.sin(a * s + b)
                elif kind == COS:
                    g = torch.cos(a * s + b)
                elif kind == EXP:
                    g = torch.exp(a * s + b)
                elif kind == AFFINE:
                    g = a * s + b
                else:
                    g = s ** int(a)
                v = v * g
            out = out + v
        return out

    # -- device table -----------------------------------------------------------------
    def fields(self):
        return [f for _, f, _ in self.terms if f is not None]

    def device_table(self, bcs, nc=None, sinpi=True):
        """fa_sep_t in device memory for quadrature points `bcs` (NQ, TD+1,
        FP64, on the device): a uint8 tensor that owns the copy (cached per
        bcs table)."""
        key = (bcs.data_ptr(), bcs.device)
        t = self._dev.get(key)
        if t is not None:
            return t
        st = FaSep()
        st.nterms = len(self.terms)
        k = 0
        keep = []
        for i, (c, fld, fs) in enumerate(self.terms):
            st.coef[i] = c
            if fld is not None:
                fld = fld.to(device=bcs.device, dtype=torch.float64).contiguous()
                if fld.ndim != 1 or (nc is not None and fld.shape[0] != nc):
                    raise ValueError(f"a cell field must have shape (NC,) = ({nc},)")
                keep.append(fld)
                st.field[i] = fld.data_ptr()
            for kind, d, a, b in fs:
                if sinpi and kind in (SIN, COS):  # sin(a s + b) = sinpi(a/pi s + b/pi): a cheap argument reduction
                    kind, a, b = (SINPI if kind == SIN else COSPI), a / math.pi, b / math.pi
                st.kind[k], st.dim[k], st.a[k], st.b[k] = kind, d, a, b
                k += 1
            st.term_end[i] = k
        st.nfactors = k
        st.bcs = bcs.data_ptr()
        raw = bytes(st)
        t = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(bcs.device)
        t._fa_keep = (keep, bcs)          # the table points into these tensors
        self._dev[key] = t
        return t

    def point_table(self, device):
        """fa_sep_t in device memory for evaluation at given points (no
        quadrature table, no fields), cached per device."""
        key = ('points', str(device))
        t = self._dev.get(key)
        if t is None:
            if self.fields():
                raise ValueError("point evaluation takes a field-free expression")
            z = torch.zeros((1, 1), dtype=torch.float64, device=device)
            t = self.device_table(z, sinpi=False)
            self._dev.pop((z.data_ptr(), z.device), None)
            self._dev[key] = t
        return t

    def max_dim(self):
        return max((d for _, _, fs in self.terms for _, d, _, _ in fs), default=-1)

    def __repr__(self):
        return f"Sep({len(self.terms)} terms)"


def _affine(s):
    if isinstance(s, Affine):
        return s
    raise TypeError("sin/cos/exp of a separable expression need an affine argument a*x + b of one coordinate")


def sin(s):
    if isinstance(s, numbers.Real):
        return math.sin(s)
    s = _affine(s)
    return Sep([(1.0, None, ((SIN, s.dim, s.a, s.b),))])


def cos(s):
    if isinstance(s, numbers.Real):
        return math.cos(s)
    s = _affine(s)
    return Sep([(1.0, None, ((COS, s.dim, s.a, s.b),))])


def exp(s):
    if isinstance(s, numbers.Real):
        return math.exp(s)
    s = _affine(s)
    return Sep([(1.0, None, ((EXP, s.dim, s.a, s.b),))])


def cell(k):
    """A per-cell factor k[c] (FP64 tensor of shape (NC,))."""
    if not torch.is_tensor(k) or k.ndim != 1:
        raise ValueError("cell() takes a 1-D tensor of per-cell values")
    return Sep([(1.0, k, ())])


x, y, z = Affine(0), Affine(1), Affine(2)
