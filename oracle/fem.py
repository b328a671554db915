"""Integrators, COO->CSR assembly, Dirichlet BC, CG and L2 error (NumPy FP64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Coefficients follow fealpy/utils/utils.py:21-53 (``process_coef_func``): a
number, an ``(NC,)`` or ``(NC, NQ)`` array, or a callable -- called on physical
points ``(NC, NQ, GD)`` when it carries ``coordtype == 'cartesian'``, otherwise
as ``f(bcs, index=slice(None))``.
"""
import numpy as np

from .basis import multi_index_matrix, simplex_grad_shape_function, simplex_shape_function


def cartesian(f):
    """fealpy/decorator/coordinates.py:10-15 (tag only)."""
    f.coordtype = 'cartesian'
    return f


def barycentric(f):
    """fealpy/decorator/coordinates.py:17-22 (tag only)."""
    f.coordtype = 'barycentric'
    return f


def process_coef(coef, bcs, mesh):
    """fealpy/utils/utils.py:21-53."""
    if callable(coef):
        if getattr(coef, 'coordtype', 'barycentric') == 'barycentric':
            return coef(bcs, index=slice(None))
        return coef(mesh.bc_to_point(bcs))
    return coef


def _fill_axis(x, ndim):
    """fealpy/utils/utils.py:96-106."""
    return x.reshape(x.shape + (1,) * (ndim - x.ndim))


def _is_scalar(x):
    return np.isscalar(x) or (isinstance(x, np.ndarray) and x.size == 1)


def bilinear_integral(b1, b2, ws, cm, coef=None):
    """fealpy/functional.py:68-106 (unbatched)."""
    b1 = b1.reshape(*b1.shape[:3], -1)
    b2 = b2.reshape(*b2.shape[:3], -1)
    if coef is None:
        return np.einsum('q, c, cqid, cqjd -> cij', ws, cm, b1, b2, optimize=True)
    if _is_scalar(coef):
        return np.einsum('q, c, cqid, cqjd -> cij', ws, cm, b1, b2, optimize=True) * coef
    coef = _fill_axis(np.asarray(coef), 3)
    return np.einsum('q, c, cqid, cqjd, ...cqd -> ...cij', ws, cm, b1, b2, coef, optimize=True)


def linear_integral(basis, ws, cm, source=None):
    """fealpy/functional.py:27-65 (unbatched, scalar dof)."""
    if source is None:
        return np.einsum('c, q, cq... -> c...', cm, ws, basis, optimize=True)
    if _is_scalar(source):
        return np.einsum('c, q, cq... -> c...', cm, ws, basis, optimize=True) * source
    dof_shape = basis.shape[3:]
    basis = basis.reshape(*basis.shape[:3], -1)
    source = np.asarray(source)
    if source.ndim <= 2:
        source = _fill_axis(source, 2)
        r = np.einsum('c, q, cqid, ...cq -> ...cid', cm, ws, basis, source, optimize=True)
        return r.reshape(r.shape[:-1] + dof_shape)
    source = _fill_axis(source, 3)
    return np.einsum('c, q, cqid, ...cqd -> ...ci', cm, ws, basis, source, optimize=True)


def diffusion_element(mesh, p, coef=None, q=None):
    """ScalarDiffusionIntegrator.assembly, default variant:
    fealpy/fem/scalar_diffusion_integrator.py:54-63 (q defaults to p+3, :32-37)."""
    q = p + 3 if q is None else q
    bcs, ws = mesh.quadrature_formula(q)
    cm = mesh.entity_measure('cell')
    c = process_coef(coef, bcs, mesh)
    gphi = mesh.grad_shape_function(bcs, p, 'x')
    return bilinear_integral(gphi, gphi, ws, cm, c)


def mass_element(mesh, p, coef=None, q=None):
    """ScalarMassIntegrator.assembly: fealpy/fem/scalar_mass_integrator.py:35-54."""
    q = p + 3 if q is None else q
    bcs, ws = mesh.quadrature_formula(q)
    cm = mesh.entity_measure('cell')
    phi = mesh.shape_function(bcs, p)[None, ...]
    c = process_coef(coef, bcs, mesh)
    return bilinear_integral(phi, phi, ws, cm, c)


def source_element(mesh, p, source, q=None):
    """ScalarSourceIntegrator.assembly: fealpy/fem/scalar_source_integrator.py:33-58."""
    q = p + 3 if q is None else q
    bcs, ws = mesh.quadrature_formula(q)
    cm = mesh.entity_measure('cell')
    phi = mesh.shape_function(bcs, p)[None, ...]
    f = process_coef(source, bcs, mesh)
    return linear_integral(phi, ws, cm, f)


def tensor_cell2dof(cell2dof, numel, gdof, dof_priority):
    """to_tensor_dof: fealpy/functionspace/utils.py:65-95 (array branch)."""
    idx = np.arange(gdof * numel)
    if dof_priority:
        idx = idx.reshape(numel, gdof)[:, cell2dof].swapaxes(0, 1)
    else:
        idx = idx.reshape(gdof, numel)[cell2dof, :]
    return idx.reshape(cell2dof.shape[0], -1)


def elasticity_element(mesh, p, lam, mu, q=None, dof_priority=False):
    """LinearElasticityIntegrator.assembly for simplices:
    fealpy/fem/linear_elasticity_integrator.py:60-156, isotropic D of
    fealpy/material/elastic_material.py:236-249 ('3D' / 'plane_strain').
    Blocks A_ab[c,i,j] = sum_q w_q |T| d_a phi_i d_b phi_j; the entry of
    (i, a), (j, b) is c_aa(A_aa, others) on diagonal blocks and
    lam A_ab + mu A_ba off the diagonal."""
    q = p + 3 if q is None else q
    bcs, ws = mesh.quadrature_formula(q)
    cm = mesh.entity_measure('cell')
    gphi = mesh.grad_shape_function(bcs, p, 'x')            # (NC, NQ, ldof, GD)
    NC, _, ldof, GD = gphi.shape
    A = [[np.einsum('q, cqi, cqj, c -> cij', ws, gphi[..., a], gphi[..., b], cm, optimize=True)
          for b in range(GD)] for a in range(GD)]
    KK = np.zeros((NC, GD * ldof, GD * ldof))

    def put(a, b, val):
        if dof_priority:
            KK[:, a * ldof:(a + 1) * ldof, b * ldof:(b + 1) * ldof] = val
        else:
            KK[:, a::GD, b::GD] = val
    if GD == 3:
        D00, D01, D55 = 2 * mu + lam, lam, mu
        dd = D00 if dof_priority else 2 * D55 + D01       # linear_elasticity_integrator.py:130-165
        put(0, 0, dd * A[0][0] + D55 * (A[1][1] + A[2][2]))
        put(1, 1, dd * A[1][1] + D55 * (A[0][0] + A[2][2]))
        put(2, 2, dd * A[2][2] + D55 * (A[0][0] + A[1][1]))
        for a in range(3):
            for b in range(3):
                if a != b:
                    put(a, b, D01 * A[a][b] + D55 * A[b][a])
    else:
        D00, D01, D22 = 2 * mu + lam, lam, mu
        put(0, 0, D00 * A[0][0] + D22 * A[1][1])
        put(1, 1, D00 * A[1][1] + D22 * A[0][0])
        put(0, 1, D01 * A[0][1] + D22 * A[1][0])
        put(1, 0, D01 * A[1][0] + D22 * A[0][1])
    return KK


def vector_source_element(mesh, p, source, q=None, dof_priority=False):
    """VectorSourceIntegrator.assembly: fealpy/fem/vector_source_integrator.py:45-52
    -> linear_integral with a tensor basis (functional.py:27-65,
    functionspace/functional.py:13-40); source values (NC, NQ, GD)."""
    q = p + 3 if q is None else q
    bcs, ws = mesh.quadrature_formula(q)
    cm = mesh.entity_measure('cell')
    phi = mesh.shape_function(bcs, p)                       # (NQ, ldof)
    f = process_coef(source, bcs, mesh)                     # (NC, NQ, GD)
    r = np.einsum('c, q, qi, cqd -> cid', cm, ws, phi, f, optimize=True)
    if dof_priority:
        r = r.swapaxes(1, 2)
    return r.reshape(r.shape[0], -1)


class CSR:
    """Minimal CSR triple (int64 indptr, int32/int64 indices, float64 values)."""

    def __init__(self, indptr, indices, data, shape):
        self.indptr, self.indices, self.data, self.shape = indptr, indices, data, shape

    def matvec(self, x):
        """scipy ``csr_matvec`` semantics (row-wise sum in storage order)."""
        import scipy.sparse as sp
        A = sp.csr_matrix((self.data, self.indices, self.indptr), shape=self.shape)
        return A @ x

    def canonical(self):
        """Sort columns inside every row, carrying values (the reference's
        ``tocsr`` uses an unstable argsort, coo_tensor.py:148)."""
        rows = np.repeat(np.arange(self.shape[0]), np.diff(self.indptr))
        order = np.lexsort((self.indices, rows))
        return CSR(self.indptr.copy(), self.indices[order], self.data[order], self.shape)

    def todense(self):
        A = np.zeros(self.shape)
        rows = np.repeat(np.arange(self.shape[0]), np.diff(self.indptr))
        np.add.at(A, (rows, self.indices), self.data)
        return A


def coo_coalesce(I, J, V):
    """COOTensor.coalesce: fealpy/sparse/coo_tensor.py:184-213 -- stable
    lexsort by (row, col), then sequential ``np.add.at`` of duplicates."""
    order = np.lexsort((J, I))
    Is, Js, Vs = I[order], J[order], V[order]
    uniq = np.concatenate([[True], (Is[1:] != Is[:-1]) | (Js[1:] != Js[:-1])])
    add_index = np.cumsum(uniq) - 1
    vals = np.zeros(int(uniq.sum()), dtype=V.dtype)
    np.add.at(vals, add_index, Vs)
    return Is[uniq], Js[uniq], vals


def coo_tocsr(I, J, V, n):
    """COOTensor.tocsr: coo_tensor.py:137-157 (bincount/cumsum crow).  The
    reference's ``argsort`` is unstable; a stable sort is used here, which
    keeps the coalesced (sorted) column order."""
    count = np.bincount(I, minlength=n)
    crow = np.concatenate([[0], np.cumsum(count)])
    order = np.argsort(I, kind='stable')
    return CSR(crow, J[order], V[order], (n, n))


def assemble_matrix(cell2dof, elems, gdof):
    """BilinearForm._scalar_assembly + assembly(format='csr'):
    fealpy/fem/bilinear_form.py:46-105.  ``elems`` is a list of (NC, ldof, ldof)
    group tensors (one per integrator group, concatenated in order)."""
    Is, Js, Vs = [], [], []
    for K in elems:
        NC, ldof = cell2dof.shape
        Is.append(np.broadcast_to(cell2dof[:, :, None], K.shape).ravel())
        Js.append(np.broadcast_to(cell2dof[:, None, :], K.shape).ravel())
        Vs.append(K.ravel())
    I = np.concatenate(Is).astype(np.int64)
    J = np.concatenate(Js).astype(np.int64)
    V = np.concatenate(Vs)
    I, J, V = coo_coalesce(I, J, V)
    return coo_tocsr(I, J, V, gdof)


def assemble_vector(cell2dof, elems, gdof):
    """LinearForm.assembly(format='dense'): fealpy/fem/linear_form.py:36-86 ->
    COOTensor.to_dense (coo_tensor.py:116-130, sequential ``np.add.at``)."""
    F = np.zeros(gdof)
    for b in elems:
        np.add.at(F, cell2dof.ravel(), b.ravel())
    return F


def dirichlet_apply(A, F, is_bd, ipoints, gd):
    """DirichletBC.apply: fealpy/fem/dirichlet_bc.py:101-234.

    apply_vector: uh = g on boundary DOFs (boundary_interpolate,
    lagrange_fe_space.py:109-142), f <- f - A uh, f[bd] <- uh[bd].
    apply_matrix: drop entries in boundary rows or columns (``_mul``,
    :213-229), then add the unit boundary diagonal (spdiags + CSRTensor.add ->
    COO concat -> tocsr, csr_tensor.py:272-306)."""
    uh = np.zeros_like(F)
    uh[is_bd] = gd(ipoints[is_bd]) if callable(gd) else np.asarray(gd)[is_bd]
    f = F - A.matvec(uh)
    bd_idx = np.nonzero(is_bd)[0]
    f[bd_idx] = uh[bd_idx]
    # _mul: CSR filter
    is_int = ~is_bd
    nnz_row = np.diff(A.indptr)
    keep = np.repeat(is_int, nnz_row) & is_int[A.indices]
    rows = np.repeat(np.arange(A.shape[0]), nnz_row)[keep]
    cols = A.indices[keep].astype(np.int64)
    vals = A.data[keep]
    # + D1 (boundary unit diagonal); concat then tocsr (no coalesce)
    I = np.concatenate([rows, bd_idx])
    J = np.concatenate([cols, bd_idx])
    V = np.concatenate([vals, np.ones(len(bd_idx))])
    return coo_tocsr(I, J, V, A.shape[0]).canonical(), f


def cg(A, b, x0=None, M=None, atol=1e-12, rtol=1e-8, maxit=10000):
    """fealpy/solver/cg.py:14-123.  b is one vector (n,) or a batch (n, B):
    the columns iterate in lockstep with their own alpha and beta (sums over
    axis 0, cg.py:85-110) and stop together on sqrt(sum_k r_k . z_k) against
    atol and rtol * ||b|| (Frobenius norm for a batch, cg.py:98-113).
    Returns (x, info)."""
    info = {'residual': 0.0, 'niter': 0}
    if np.linalg.norm(b) < 1e-15:
        return np.zeros_like(b), info
    matvec = A.matvec if hasattr(A, 'matvec') else (lambda v: A @ v)
    prec = (lambda v: v) if M is None else M

    def apply(f, v):
        return f(v) if v.ndim == 1 else np.stack([f(v[:, k]) for k in range(v.shape[1])], axis=1)
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - apply(matvec, x)
    z = apply(prec, r)
    p = z
    b_norm = np.linalg.norm(b)
    rTr = np.sum(r * z, axis=0)
    n_iter = 0
    while True:
        Ap = apply(matvec, p)
        alpha = rTr / np.sum(p * Ap, axis=0)
        x = x + alpha * p
        r_new = r - alpha * Ap
        z_new = apply(prec, r_new)
        rTr_new = np.sum(r_new * z_new, axis=0)
        r_norm = np.sqrt(np.sum(rTr_new))
        n_iter += 1
        info['residual'], info['niter'] = r_norm, n_iter
        if r_norm < atol or r_norm < rtol * b_norm or (maxit is not None and n_iter >= maxit):
            break
        beta = rTr_new / rTr
        p = z_new + beta * p
        r, z, rTr = r_new, z_new, rTr_new
    return x, info


def l2_error(mesh, p, u, uh, cell2dof, q=3):
    """HomogeneousMesh.error(u, uh) with power 2: fealpy/mesh/mesh_base.py:545-608;
    ``uh`` is evaluated as a barycentric Function (lagrange_fe_space.py:206-215)."""
    bcs, ws = mesh.quadrature_formula(q)
    ps = mesh.bc_to_point(bcs)
    uv = u(ps)
    phi = mesh.shape_function(bcs, p)
    vv = np.einsum('ql, cl -> cq', phi, uh[cell2dof])
    cm = mesh.entity_measure('cell')
    e = np.einsum('q, cq, c -> c', ws, np.abs(uv - vv) ** 2, cm)
    return np.sqrt(np.sum(e))
