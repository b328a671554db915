"""Batched solves for serving.

Workgroup solves (pcg_workgroups, fa_pcg_wg_batched): many small systems in
one launch, one workgroup per system running its whole MG-preconditioned CG:
- every result solves its own system to the requested tolerance (against a
  tight FP64 CG of the same assembled system) and agrees with the ordinary
  multi-kernel MG-PCG on the same hierarchy;
- a system's result is bitwise the same alone and in a batch of others;
- the server with batches equals the one-at-a-time results within the
  solve tolerance; a small resident budget evicts and rebuilds sizes with
  the same result.

Packed batches (PackedBatch, fa_pcg_batched) are switched off with their
test skipped (PACKED_BATCHES_OFF in fealpy_amd/serving.py): the packing
kernels reproduce each request's operator (SpMV of a packed operator equals
the requests' own SpMVs, segment by segment)."""
import numpy as np
import pytest
import torch

import fealpy_amd as fa
from fealpy_amd.serving import (MAX_BATCH, PACKED_BATCHES_ENABLED, PACKED_BATCHES_OFF, PackedBatch, SolveServer,
                                fourier_source, make_requests)

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
# the packed path is switched off until the cause of its stall on the
# saturated 4096-request trace is found; its tests stay, skipped
packed = pytest.mark.skipif(not PACKED_BATCHES_ENABLED, reason=PACKED_BATCHES_OFF)


def _tight(srv, r):
    space = srv.context(r.nx, r.p).space
    k = r.k.to("cuda")
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=3)).assembly()
    F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(fourier_source(r.modes, "cuda"), q=3)).assembly()
    A, F = fa.DirichletBC(space, gd=0.0).apply(A, F)
    return fa.cg(A, F, atol=1e-30, rtol=1e-13, maxit=100000)


def _rel(a, b):
    return float(torch.linalg.norm(a - b)) / max(float(torch.linalg.norm(b)), 1e-300)


def test_workgroup_batch_accuracy_and_alone_equals_mix():
    from fealpy_amd.serving import WG_MAX_GDOF
    srv = SolveServer(slots=1, rtol=1e-10)
    reqs = [r for r in make_requests(40, seed=21, nx_range=(8, 96)) if (r.p * r.nx + 1) ** 2 <= WG_MAX_GDOF][:12]
    assert len(reqs) >= 6
    srv.solve_batch(reqs)
    for r in reqs:
        assert r.info["status"] == 1 and r.info["batch"] == len(reqs), r.info
        assert _rel(r.x, _tight(srv, r)) < 1e-8, (r.nx, r.p, r.info)
    by_id = {r.rid: r for r in reqs}
    for r in reqs[:4]:
        alone = make_requests(40, seed=21, nx_range=(8, 96))[r.rid]
        srv.solve_batch([alone])
        assert alone.info["batch"] == 1
        assert torch.equal(alone.x, by_id[r.rid].x), (r.nx, r.p)          # bitwise: its own workgroup


def test_workgroup_solve_matches_multi_kernel_pcg():
    """pcg_workgroups against GeometricMG.pcg on the same hierarchy and
    system: the same solution to the tolerance, iteration counts within one;
    2D P1/P2 and a small 3D P1 system in one batch."""
    from fealpy_amd.solver import GeometricMG, pcg_workgroups
    systems, refs = [], []
    g = np.random.default_rng(5)
    for mesh, p in ((fa.TriangleMesh.from_box([0, 1, 0, 1], nx=40, ny=40, device=DEV), 1),
                    (fa.TriangleMesh.from_box([0, 1, 0, 1], nx=23, ny=23, device=DEV), 2),
                    (fa.TetrahedronMesh.from_box(nx=10, ny=9, nz=11, device=DEV), 1)):
        space = fa.LagrangeFESpace(mesh, p)
        k = torch.as_tensor(np.exp(g.standard_normal(mesh.number_of_cells())), device=DEV)
        A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=p + 2)).assembly()
        F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=p + 2)).assembly()
        bc = fa.DirichletBC(space, gd=fa.cartesian(lambda x: x[..., 0] + 2 * x[..., 1]))
        A, F = bc.apply(A, F)
        mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32, compress=False)
        mg.setup(A)
        x_ref, info_ref = mg.pcg(A, F, atol=0.0, rtol=1e-10, maxit=500)
        systems.append((mg, A, F))
        refs.append((x_ref, info_ref))
    out = pcg_workgroups(systems, atol=0.0, rtol=1e-10, maxit=500)
    for (x, info), (x_ref, info_ref) in zip(out, refs):
        assert info["status"] == 1, info
        assert abs(info["niter"] - info_ref["niter"]) <= 1, (info, info_ref)
        assert _rel(x, x_ref) < 1e-8


@packed
def test_packed_operator_equals_request_operators():
    srv = SolveServer(slots=1, rtol=1e-8)
    reqs = [r for r in make_requests(10, seed=3, nx_range=(20, 60)) if r.p == 1][:4]
    mgs, Fs = [], []
    for r in reqs:
        ctx = srv.context(r.nx, r.p)
        mg = ctx.checkout(packed=True)
        A, F = srv._system(r, ctx, mg)
        mgs.append(mg)
        Fs.append(F)
    depth = len(mgs[0].levels)
    same = [i for i, m in enumerate(mgs) if len(m.levels) == depth]
    mgs = [mgs[i] for i in same]
    pb = PackedBatch(mgs, DEV)
    xs = [torch.rand(m.levels[0].n, dtype=torch.float64, device=DEV) for m in mgs]
    xp = torch.empty(pb.N[0], dtype=torch.float64, device=DEV)
    pb.copy_in(xs, 0, xp)
    from fealpy_amd.sparse.sell import SellMatrix
    S = SellMatrix.__new__(SellMatrix)
    sp, cols, vals = pb.A64[1]
    S.shape, S.slice_ptr, S.cols, S.entries, S.vals = (pb.N[0], pb.N[0]), sp, cols, int(cols.numel()), \
        {torch.float64: vals}
    yp = S.matvec(xp)
    outs = [torch.empty_like(x) for x in xs]
    pb.copy_out(yp, outs, 0)
    for m, x, y in zip(mgs, xs, outs):
        ref = m.levels[0].As.matvec(x)
        assert torch.equal(y, ref)
    for r, m in zip(reqs, mgs):
        srv.context(r.nx, r.p).checkin(m)


def test_server_batches_match_one_at_a_time():
    reqs = make_requests(40, seed=9, nx_range=(16, 128))
    srv = SolveServer(slots=2, rtol=1e-10, batch=MAX_BATCH)
    alone = [srv.solve(r).x.clone() for r in make_requests(40, seed=9, nx_range=(16, 128))]
    srv.start()
    for r in reqs:
        srv.submit(r)
    for r in reqs:
        assert r.event.wait(120), "request did not finish"
        assert r.error is None, r.error
    srv.close()
    for a, r in zip(alone, reqs):
        assert _rel(r.x, a) < 1e-8, (r.nx, r.p)
    assert max(r.info.get("batch", 1) for r in reqs) >= 2


def test_resident_budget_evicts_least_recently_used():
    """A small resident budget drops the least recently used sizes (never one
    in use); a dropped size is rebuilt on its next request with the same result."""
    from fealpy_amd.serving import request_bytes
    reqs = make_requests(8, seed=4, nx_range=(24, 80))
    budget = 3 * max(request_bytes(r.nx, r.p) for r in reqs)
    srv = SolveServer(slots=1, rtol=1e-10, resident_budget=budget)
    first = [srv.solve(r).x.clone() for r in reqs]
    assert srv.evictions > 0
    assert srv.resident_bytes() <= budget + 2 * max(request_bytes(r.nx, r.p) for r in reqs)
    again = [srv.solve(r).x for r in make_requests(8, seed=4, nx_range=(24, 80))]
    for a, b in zip(first, again):
        assert torch.equal(a, b)
