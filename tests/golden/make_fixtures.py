"""Write golden fixtures by running the reference (weihuayi/fealpy, numpy backend).

Only for what the reference's own test data and closed forms cannot pin:
assembled FP64 values with variable coefficients through the reference's
quadrature and coefficient-evaluation order, the linear-elasticity integrator
(which has no reference test), post-``DirichletBC`` patterns, CG iteration
counts / residuals and L2 errors.

Runs ONLY in a container that has the reference at ``/root/reference``; the
outputs are small ``.npz`` files committed next to this script.  The GPU tests
read only those files; nothing of the reference is needed where they run.

``fealpy/fem/__init__.py`` and ``fealpy/solver/__init__.py`` import unrelated
model modules that need ``gmsh`` (absent in this image); the hot-path modules
are imported directly with those two aggregators registered as bare packages,
so no reference code is replaced or stubbed.

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_fixtures.py
"""
import os
import sys
import types

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
PI = np.pi


def _import_reference():
    sys.path.insert(0, REF)
    import fealpy  # noqa: F401
    for name in ("fealpy.fem", "fealpy.solver"):
        mod = types.ModuleType(name)
        mod.__path__ = [os.path.join(REF, *name.split("."))]
        sys.modules[name] = mod
    from fealpy.backend import backend_manager as bm
    bm.set_backend("numpy")
    from fealpy.mesh import TriangleMesh, TetrahedronMesh
    from fealpy.functionspace import LagrangeFESpace, TensorFunctionSpace
    from fealpy.fem.bilinear_form import BilinearForm
    from fealpy.fem.linear_form import LinearForm
    from fealpy.fem.scalar_diffusion_integrator import ScalarDiffusionIntegrator
    from fealpy.fem.scalar_mass_integrator import ScalarMassIntegrator
    from fealpy.fem.scalar_source_integrator import ScalarSourceIntegrator
    from fealpy.fem.vector_source_integrator import VectorSourceIntegrator
    # fealpy/fem/utils.py:9 imports these two names from the package itself
    sys.modules["fealpy.fem"].ScalarMassIntegrator = ScalarMassIntegrator
    sys.modules["fealpy.fem"].ScalarDiffusionIntegrator = ScalarDiffusionIntegrator
    from fealpy.fem.linear_elasticity_integrator import LinearElasticityIntegrator
    from fealpy.fem.dirichlet_bc import DirichletBC
    from fealpy.material.elastic_material import LinearElasticMaterial
    from fealpy.solver.cg import cg
    from fealpy.decorator import cartesian, barycentric
    return types.SimpleNamespace(**locals())


def canon(A):
    """CSR with columns sorted inside each row (the reference's tocsr uses an
    unstable argsort, fealpy/sparse/coo_tensor.py:148)."""
    indptr = np.asarray(A.crow).astype(np.int64)
    indices = np.asarray(A.col).astype(np.int64)
    data = np.asarray(A.values)
    rows = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    order = np.lexsort((indices, rows))
    return indptr, indices[order].astype(np.int32), data[order]


def put(out, prefix, A):
    ip, ix, d = canon(A)
    out[prefix + "_indptr"], out[prefix + "_indices"], out[prefix + "_data"] = ip, ix, d


def scalar_case(R, mesh, p, integrators_A, integrators_F, u, q_err=3, cg_kw=None):
    """Assemble, apply Dirichlet BC with gd=u, solve tightly, measure L2 error."""
    space = R.LagrangeFESpace(mesh, p)
    out = {"cell2dof": np.asarray(space.cell_to_dof()), "gdof": space.number_of_global_dofs()}
    bform = R.BilinearForm(space)
    for grp in integrators_A:
        bform.add_integrator(*grp)
    A = bform.assembly()
    put(out, "A", A)
    lform = R.LinearForm(space)
    for I in integrators_F:
        lform.add_integrator(I)
    F = lform.assembly()
    out["F"] = np.asarray(F)
    bc = R.DirichletBC(space, gd=u)
    out["is_bd"] = np.asarray(bc.is_boundary_dof)
    A2, F2 = bc.apply(A, F)
    put(out, "Abc", A2)
    out["Fbc"] = np.asarray(F2)
    x, info = R.cg(A2, F2, atol=1e-14, rtol=1e-14, maxit=20000, returninfo=True)
    out["x"], out["x_niter"], out["x_res"] = np.asarray(x), info["niter"], info["residual"]
    if cg_kw is not None:
        x2, info2 = R.cg(A2, F2, returninfo=True, **cg_kw)
        out["x_cg"], out["cg_niter"], out["cg_res"] = np.asarray(x2), info2["niter"], info2["residual"]
    uh = space.function()
    uh[:] = x
    out["l2_error"] = mesh.error(u, uh, q=q_err)
    return out


def poisson2d_p1(R):
    """Tri 64x64, P1, coefficient 1, f = 2 pi^2 sin sin, q=2, cg rtol 1e-10."""
    @R.cartesian
    def u(p):
        return np.sin(PI * p[..., 0]) * np.sin(PI * p[..., 1])

    @R.cartesian
    def f(p):
        return 2 * PI**2 * np.sin(PI * p[..., 0]) * np.sin(PI * p[..., 1])

    mesh = R.TriangleMesh.from_box([0, 1, 0, 1], 64, 64)
    return scalar_case(R, mesh, 1, [(R.ScalarDiffusionIntegrator(q=2),)],
                       [R.ScalarSourceIntegrator(f, q=2)], u, cg_kw=dict(rtol=1e-10))


def diffusion3d_p1_lognormal(R, n=8, seed=20261202):
    """Tet n^3, P1, per-cell k = exp(N(0,1)), barycentric f_c = 3 pi^2 k_c u, q=2."""
    mesh = R.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], n, n, n)
    NC = mesh.number_of_cells()
    k = np.exp(np.random.default_rng(seed).standard_normal(NC))

    @R.cartesian
    def u(p):
        return np.sin(PI * p[..., 0]) * np.sin(PI * p[..., 1]) * np.sin(PI * p[..., 2])

    @R.barycentric
    def f(bcs, index=None):
        ps = mesh.bc_to_point(bcs)
        return 3 * PI**2 * k[:, None] * u(ps)

    out = scalar_case(R, mesh, 1, [(R.ScalarDiffusionIntegrator(k, q=2),)],
                      [R.ScalarSourceIntegrator(f, q=2)], u)
    out["k"], out["n"], out["seed"] = k, n, seed
    return out


def _sha(*arrays):
    import hashlib
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def diffusion3d_p1_summary(R, n=128, seed=20261202, nsample=65536):
    """The lognormal P1 problem at full size (from_box 128^3, q=2): too large
    to store, so the fixture holds the SHA-256 of the canonical CSR patterns
    (int64 indptr bytes, int32 indices bytes) before and after the Dirichlet
    condition, value statistics (sum, sum |.|, sum of squares, min, max) of
    A, F and their post-BC forms, values at seeded random positions, a tight
    solution's norm and samples, and the L2 error."""
    mesh = R.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], n, n, n)
    NC = mesh.number_of_cells()
    k = np.exp(np.random.default_rng(seed).standard_normal(NC))

    @R.cartesian
    def u(p):
        return np.sin(PI * p[..., 0]) * np.sin(PI * p[..., 1]) * np.sin(PI * p[..., 2])

    @R.barycentric
    def f(bcs, index=None):
        return 3 * PI**2 * k[:, None] * u(mesh.bc_to_point(bcs))

    space = R.LagrangeFESpace(mesh, 1)
    bform = R.BilinearForm(space)
    bform.add_integrator(R.ScalarDiffusionIntegrator(k, q=2))
    A = bform.assembly()
    lform = R.LinearForm(space)
    lform.add_integrator(R.ScalarSourceIntegrator(f, q=2))
    F = np.asarray(lform.assembly())
    bc = R.DirichletBC(space, gd=u)
    A2, F2 = bc.apply(A, F)
    out = {"n": n, "seed": seed, "gdof": space.number_of_global_dofs()}
    g = np.random.default_rng(seed + 1)
    for name, M in (("A", A), ("Abc", A2)):
        ip, ix, d = canon(M)
        out[name + "_nnz"] = len(ix)
        out[name + "_pattern_sha256"] = _sha(ip.astype(np.int64), ix.astype(np.int32))
        out[name + "_stats"] = np.array([d.sum(), np.abs(d).sum(), (d * d).sum(), d.min(), d.max()])
        pos = np.sort(g.choice(len(ix), nsample, replace=False))
        out[name + "_sample_pos"], out[name + "_sample_val"] = pos, d[pos]
        del ip, ix, d
    for name, v in (("F", F), ("Fbc", np.asarray(F2))):
        out[name + "_stats"] = np.array([v.sum(), np.abs(v).sum(), (v * v).sum(), v.min(), v.max()])
        pos = np.sort(g.choice(len(v), nsample, replace=False))
        out[name + "_sample_pos"], out[name + "_sample_val"] = pos, v[pos]
    x, info = R.cg(A2, F2, atol=1e-14, rtol=1e-13, maxit=50000, returninfo=True)
    x = np.asarray(x)
    out["x_norm"], out["x_niter"], out["x_res"] = np.linalg.norm(x), info["niter"], info["residual"]
    pos = np.sort(g.choice(len(x), nsample, replace=False))
    out["x_sample_pos"], out["x_sample_val"] = pos, x[pos]
    uh = space.function()
    uh[:] = x
    out["l2_error"] = mesh.error(u, uh, q=3)
    return out


def _k_smooth(p):
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    return 1 + 0.5 * np.sin(2 * PI * x) * np.cos(2 * PI * y) * z


def highorder3d(R, p, q, n):
    """Tet n^3, P3/P4, diffusion k(x) + mass 1 (one group), f = -div(k grad u) + u."""
    mesh = R.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], n, n, n)

    @R.cartesian
    def u(pt):
        return np.sin(PI * pt[..., 0]) * np.sin(PI * pt[..., 1]) * np.sin(PI * pt[..., 2])

    @R.cartesian
    def kfun(pt):
        return _k_smooth(pt)

    @R.cartesian
    def f(pt):
        x, y, z = pt[..., 0], pt[..., 1], pt[..., 2]
        sx, sy, sz = np.sin(PI * x), np.sin(PI * y), np.sin(PI * z)
        cx, cy, cz = np.cos(PI * x), np.cos(PI * y), np.cos(PI * z)
        uu = sx * sy * sz
        gk = (PI * np.cos(2 * PI * x) * np.cos(2 * PI * y) * z,
              -PI * np.sin(2 * PI * x) * np.sin(2 * PI * y) * z,
              0.5 * np.sin(2 * PI * x) * np.cos(2 * PI * y))
        gu = (PI * cx * sy * sz, PI * sx * cy * sz, PI * sx * sy * cz)
        return 3 * PI**2 * _k_smooth(pt) * uu - sum(a * b for a, b in zip(gk, gu)) + uu

    out = scalar_case(R, mesh, p, [(R.ScalarDiffusionIntegrator(kfun, q=q), R.ScalarMassIntegrator(1.0, q=q))],
                      [R.ScalarSourceIntegrator(f, q=q)], u, q_err=q)
    out["n"], out["q"] = n, q
    return out


def poisson2d_p2_fourier(R, n=8, seed=7):
    """Tri n x n, P2, per-cell k = exp(N(0,1)), f = sum_m a_m sin(k_m pi x) sin(l_m pi y), g = 0."""
    rng = np.random.default_rng(seed)
    mesh = R.TriangleMesh.from_box([0, 1, 0, 1], n, n)
    NC = mesh.number_of_cells()
    k = np.exp(rng.standard_normal(NC))
    a = rng.standard_normal(8)
    km = rng.integers(1, 9, size=8)
    lm = rng.integers(1, 9, size=8)

    @R.cartesian
    def f(p):
        x, y = p[..., 0], p[..., 1]
        return sum(a[m] * np.sin(km[m] * PI * x) * np.sin(lm[m] * PI * y) for m in range(8))

    @R.cartesian
    def zero(p):
        return np.zeros(p.shape[:-1])

    out = scalar_case(R, mesh, 2, [(R.ScalarDiffusionIntegrator(k),)], [R.ScalarSourceIntegrator(f)], zero)
    out.update(k=k, a=a, km=km, lm=lm, n=n, seed=seed)
    return out


def elasticity3d_p2(R, n=2, lam=1.0, mu=0.5):
    """Tet n^3, vector P2 (shape=(-1, 3)), LinearElasticityIntegrator, q=5."""
    mesh = R.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], n, n, n)
    space = R.LagrangeFESpace(mesh, 2)
    tspace = R.TensorFunctionSpace(space, shape=(-1, 3))
    mat = R.LinearElasticMaterial(name="E", lame_lambda=lam, shear_modulus=mu)
    integ = R.LinearElasticityIntegrator(material=mat, q=5)
    out = {"ke": np.asarray(integ.assembly(tspace)), "cell2dof": np.asarray(tspace.cell_to_dof()),
           "gdof": tspace.number_of_global_dofs()}
    bform = R.BilinearForm(tspace)
    bform.add_integrator(integ)
    put(out, "A", bform.assembly())
    c = np.array([1.0, 2.0, 3.0])

    @R.cartesian
    def body(p):
        x, y, z = p[..., 0], p[..., 1], p[..., 2]
        s = np.sin(PI * x) * np.sin(PI * y) * np.sin(PI * z)
        H = np.zeros(p.shape[:-1] + (3, 3))
        H[..., 0, 0] = H[..., 1, 1] = H[..., 2, 2] = -PI**2 * s
        H[..., 0, 1] = H[..., 1, 0] = PI**2 * np.cos(PI * x) * np.cos(PI * y) * np.sin(PI * z)
        H[..., 0, 2] = H[..., 2, 0] = PI**2 * np.cos(PI * x) * np.sin(PI * y) * np.cos(PI * z)
        H[..., 1, 2] = H[..., 2, 1] = PI**2 * np.sin(PI * x) * np.cos(PI * y) * np.cos(PI * z)
        return 3 * PI**2 * mu * s[..., None] * c - (lam + mu) * np.einsum('...ij, j -> ...i', H, c)

    lform = R.LinearForm(tspace)
    lform.add_integrator(R.VectorSourceIntegrator(body, q=5))
    F = lform.assembly()
    out["F"] = np.asarray(F)

    @R.cartesian
    def zero(p):
        return np.zeros(p.shape[:-1] + (3,))

    bc = R.DirichletBC(tspace, gd=zero)              # clamped on every face
    out["is_bd"] = np.asarray(bc.is_boundary_dof)
    A2, F2 = bc.apply(bform.assembly(), F)
    put(out, "Abc", A2)
    out["Fbc"] = np.asarray(F2)
    x, info = R.cg(A2, F2, atol=1e-14, rtol=1e-14, maxit=20000, returninfo=True)
    out["x"], out["x_niter"], out["x_res"] = np.asarray(x), info["niter"], info["residual"]
    out.update(n=n, lam=lam, mu=mu)
    return out


def small_patterns(R):
    """P1..P4 tet 4^3 and tri P1/P2 patterns (sizes pinned in closed form too)."""
    out = {}
    for p in (1, 2, 3, 4):
        mesh = R.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], 4, 4, 4)
        space = R.LagrangeFESpace(mesh, p)
        b = R.BilinearForm(space)
        b.add_integrator(R.ScalarDiffusionIntegrator(q=p + 1))
        A = b.assembly()
        out[f"tet4_p{p}_gdof"] = space.number_of_global_dofs()
        out[f"tet4_p{p}_nnz"] = int(np.asarray(A.crow)[-1])
        out[f"tet4_p{p}_cell2dof"] = np.asarray(space.cell_to_dof())
        ip, ix, d = canon(A)
        out[f"tet4_p{p}_indptr"], out[f"tet4_p{p}_indices"], out[f"tet4_p{p}_data"] = ip, ix, d
    return out


def main():
    R = _import_reference()
    jobs = {
        "fixtures_poisson2d_p1.npz": lambda: poisson2d_p1(R),
        "fixtures_diffusion3d_p1_lognormal.npz": lambda: diffusion3d_p1_lognormal(R),
        "fixtures_diffusion3d_p1_lognormal_24.npz": lambda: diffusion3d_p1_lognormal(R, n=24),
        "fixtures_diffusion3d_p1_128_summary.npz": lambda: diffusion3d_p1_summary(R),
        "fixtures_highorder3d_p3.npz": lambda: highorder3d(R, 3, 5, 2),
        "fixtures_highorder3d_p4.npz": lambda: highorder3d(R, 4, 6, 2),
        "fixtures_poisson2d_p2_fourier.npz": lambda: poisson2d_p2_fourier(R),
        "fixtures_elasticity3d_p2.npz": lambda: elasticity3d_p2(R),
        "fixtures_patterns.npz": lambda: small_patterns(R),
    }
    only = set(sys.argv[1:])
    for name, job in jobs.items():
        if only and name not in only:
            continue
        out = job()
        np.savez_compressed(os.path.join(HERE, name), **out)
        print(name, {k: np.shape(v) for k, v in out.items()}, file=sys.stderr)


if __name__ == "__main__":
    main()
