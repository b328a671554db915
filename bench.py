"""Benchmark: DOF/s assembled + solved to 1e-8 relative error on one MI355X.

Workload (one "step"): 3D scalar diffusion -div(k grad u) = f on the unit
cube, TetrahedronMesh.from_box 128^3 (12,582,912 tets, 2,146,689 P1 DOFs),
per-cell coefficient k = exp(g), g ~ N(0,1) i.i.d. from numpy
default_rng(20261202), manufactured u = sin(pi x) sin(pi y) sin(pi z),
source f_c = 3 pi^2 k_c u, q = 2.  A step assembles A and F, applies the
Dirichlet condition, sets up the multigrid hierarchy (Galerkin products,
smoother bounds, coarse factor) and solves with multigrid-preconditioned CG.
Inputs (mesh, coefficient, the sparsity pattern) are resident in HBM before
the timed region.  The source and the Dirichlet data are separable
expressions (fealpy_amd.expr), which the load-vector kernel evaluates at
the quadrature points itself (``--torch-source``: the same source as a
barycentric torch callable, evaluated on (NC, NQ, 3) points first).

Reported ``value``: DOF/s of the mixed-precision form: A assembled in FP32,
FP32 Dirichlet filter and multigrid setup, FP32 multigrid-preconditioned CG
corrections (FP16 D^-1 A smoothing on the finest levels), and FP64 residuals
b - A x formed matrix-free from the FP64 geometry and coefficient (iterative
refinement; no FP64 matrix is assembled).  The plain FP64 GPU path (FP64
assembly, CG and V-cycle) and the FP64-CG form with an FP32 cycle are timed
beside it; every solution is checked against a tight FP64 solve (relative
2-norm error <= 1e-8).  The same line carries the serving figures (solves/s,
p50/p99) of the 4096-request mixed-size trace.

    python bench.py [--steps K] [--warmup W] [--n 128] [--no-serving] [--no-cpu]
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of k_sell_spmv<float> on the 128^3 operator: FETCH_SIZE
# and WRITE_SIZE counter passes over `bench.py --spmv-only`, each scaled by the
# same counter's reading for a 256 MiB device copy (profiles/sell_spmv_f32_128cube_counters.md)
SPMV_F32_TRAFFIC_128 = 267.3e6
# HBM bytes per launch on the 128^3 problem, from the same two counter passes
# over `bench.py --spmv-only` scaled the same way
# (profiles/step_128cube_counters.md): k_pcg_spmv<double> on the FP64
# operator; the finest level's FP16 smoother steps and residual (mean over
# the family's launches); the structured P1 matrix kernel on the tensor grid;
# the structured load vector (both passes, with its cell-order buffer)
PCG_SPMV_TRAFFIC_128 = 403.0e6
SMOOTH0_TRAFFIC_128 = 164.0e6
BOX_A_TRAFFIC_128 = 403.8e6
BOX_F_TRAFFIC_128 = 1261.2e6
# serving: slots (streams) of the continuous-batching server and the hardware
# queues the process asks for (the slots' streams, the default stream and
# torch's side streams each get one); read by HIP when it initialises
SERVING_SLOTS = 4
# relative tolerance of each FP32 inner solve of the mixed (refinement) form
INNER_RTOL = 1e-4
SERVING_HW_QUEUES = 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--rtol", type=float, default=1e-8,
                    help="PCG relative residual; every run checks the solution error against 1e-8")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-n", type=int, default=48, help="CPU baseline: from_box size of the bounded sample")
    ap.add_argument("--cpu-runs", type=int, default=3, help="CPU baseline: runs (the median is reported)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving figures of the headline line")
    ap.add_argument("--no-solve", action="store_true", help="highorder: contraction figures only")
    ap.add_argument("--torch-source", action="store_true",
                    help="source as a barycentric torch callable instead of a separable expression")
    ap.add_argument("--requests", type=int, default=4096, help="serving: number of requests")
    ap.add_argument("--serving-batch", type=int, default=1,
                    help="serving: most waiting requests a slot takes at once, the small ones solved in one launch "
                         "with one workgroup each (1: one request per slot)")
    ap.add_argument("--slots", type=int, default=SERVING_SLOTS, help="serving: concurrent request slots (streams)")
    ap.add_argument("--resident-gb", type=float, default=200.0,
                    help="serving: device memory (estimate) kept for resident per-size data")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="serving: Poisson arrival rate per second (0: 0.8 x the measured server throughput)")
    ap.add_argument("--workload", choices=("diffusion", "elasticity", "highorder", "serving", "workgroups"),
                    default="diffusion",
                    help="diffusion: 3D P1 lognormal diffusion (the headline line); elasticity: vector P2 "
                         "linear elasticity on from_box n^3 (n defaults to 40); highorder: P3/P4 diffusion + "
                         "mass element matrices on the matrix cores, from_box n^3 (n defaults to 48)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="assemble once, then run only the FP32 SELL SpMV, a few mixed PCG iterations and a "
                         "plain copy kernel of known size (for hardware-counter runs)")
    return ap.parse_args()


def make_problem(n, dev, torch_source=False):
    import torch
    import fealpy_amd as fa
    from fealpy_amd.expr import cell, sin, x, y, z
    mesh = fa.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], nx=n, ny=n, nz=n, device=dev)
    g = np.random.default_rng(20261202).standard_normal(mesh.number_of_cells())
    k = torch.as_tensor(np.exp(g), device=dev)
    pi = np.pi
    c3 = 3 * pi ** 2
    u = sin(pi * x) * sin(pi * y) * sin(pi * z)
    if torch_source:
        @fa.barycentric
        def f(bcs, index=None):
            return (c3 * k)[:, None] * torch.sin(pi * mesh.bc_to_point(bcs)).prod(dim=-1)
    else:
        f = c3 * cell(k) * u
    space = fa.LagrangeFESpace(mesh, 1)
    return mesh, space, k, u, f


def step(space, k, u, f, mg, rtol, maxit=400):
    """The FP64 form (and the FP64-outer form with an FP32 cycle): FP64
    assembly, Dirichlet condition, multigrid setup, MG-preconditioned CG."""
    import fealpy_amd as fa
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=2)).assembly()
    F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(f, q=2)).assembly()
    A2, F2 = fa.DirichletBC(space, gd=u).apply(A, F)
    mg.setup(A2)
    x, info = fa.cg(A2, F2, M=mg, atol=0.0, rtol=rtol, maxit=maxit, returninfo=True)
    return x, info, A2, F2


def step_refine(space, k, u, f, mg, rtol, inner_rtol=INNER_RTOL):
    """The mixed form: A assembled in FP64 (the operator of the FP64
    residuals), its Dirichlet-filtered FP32 copy made in the filter pass
    (each value rounded once), FP32 multigrid setup, FP32 MG-PCG corrections
    and FP64 residuals b - A x by one SpMV pass (iterative refinement)."""
    import torch
    import fealpy_amd as fa
    from fealpy_amd.solver.refine import cg_refine
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=2)).assembly()
    F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(f, q=2)).assembly()
    bc = fa.DirichletBC(space, gd=u)
    A2 = bc.apply_matrix(A, dtype=torch.float32)
    mg.setup(A2)
    x, info = cg_refine(A, F, A2, mg, bc, rtol=rtol, inner_rtol=inner_rtol, returninfo=True)
    return x, info, A2, F


def timed(fn, K, W):
    import torch
    for _ in range(W):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(K):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / K, out


def phase_breakdown(space, k, u, f, mg, rtol):
    import torch
    import fealpy_amd as fa
    ph = {}

    def t(name, fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        ph[name] = round((time.perf_counter() - t0) * 1e3, 3)
        return r
    A = t("assemble_A_ms", lambda: fa.BilinearForm(space).add_integrator(
        fa.ScalarDiffusionIntegrator(k, q=2)).assembly())
    F = t("assemble_F_ms", lambda: fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(f, q=2)).assembly())
    A2, F2 = t("dirichlet_ms", lambda: fa.DirichletBC(space, gd=u).apply(A, F))
    t("mg_setup_ms", lambda: mg.setup(A2))
    x, info = t("pcg_ms", lambda: fa.cg(A2, F2, M=mg, atol=0.0, rtol=rtol, maxit=400, returninfo=True))
    ph["pcg_iterations"] = info["niter"]
    return ph


def phase_breakdown_refine(space, k, u, f, mg, rtol):
    import torch
    import fealpy_amd as fa
    from fealpy_amd.solver.refine import cg_refine
    ph = {}

    def t(name, fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        ph[name] = round((time.perf_counter() - t0) * 1e3, 3)
        return r
    A = t("assemble_A_ms", lambda: fa.BilinearForm(space).add_integrator(
        fa.ScalarDiffusionIntegrator(k, q=2)).assembly())
    F = t("assemble_F_ms", lambda: fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(f, q=2)).assembly())
    bc = fa.DirichletBC(space, gd=u)
    A2 = t("dirichlet_ms", lambda: bc.apply_matrix(A, dtype=torch.float32))
    t("mg_setup_ms", lambda: mg.setup(A2))
    x, info = t("pcg_ms", lambda: cg_refine(A, F, A2, mg, bc, rtol=rtol, inner_rtol=INNER_RTOL, returninfo=True))
    ph["pcg_iterations"] = info["niter"]
    ph["refinements"] = info["refinements"]
    return ph


def in_step_kernels(stepfn, space, k, u, f, mg, rtol):
    """One more step of the form with the in-step kernel timer armed: HIP
    events on the launching stream around every launch of the step's
    dominant kernels (fa_timer_arm / fa_timer_read).  -> {kind: (launches, ms)}"""
    import torch
    from fealpy_amd import _native as nat
    kinds = {"pcg_spmv": nat.TIMER_PCG_SPMV, "a_pass1": nat.TIMER_P1_ROWS, "a_pass2": nat.TIMER_P1_SUM_ROWS,
             "f_pass1": nat.TIMER_P1_VECTORS, "f_pass2": nat.TIMER_P1_SUM_VALUES, "dirichlet": nat.TIMER_BC_APPLY,
             "vcycle": nat.TIMER_VCYCLE, "p1_apply": nat.TIMER_P1_APPLY, "smooth0": nat.TIMER_SMOOTH0,
             "a_box": nat.TIMER_P1_BOX}
    torch.cuda.synchronize()
    nat.timer_arm(*kinds.values())
    try:
        stepfn(space, k, u, f, mg, rtol)
        torch.cuda.synchronize()
        return {name: nat.timer_read(kd) for name, kd in kinds.items()}
    finally:
        nat.timer_arm()


def _roof(kernel, algo_bytes, launches, ms, traffic=None, **extra):
    avg = ms / max(launches, 1)
    gbs = algo_bytes / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
    out = {"bound": "hbm", "kernel": kernel, "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "algorithmic_bytes": int(algo_bytes),
           "avg_us": round(avg * 1e3, 2), "launches": int(launches)}
    out.update(extra)
    return out


def rooflines(kt, mesh, gdof, nnz_a, nnz_bc, n, vb, vb_a=8):
    """Roofline objects from the in-step kernel times of one step.
    Algorithmic bytes are the compulsory traffic of each kernel or phase
    (inputs read once, outputs written once); vb = bytes per value of the
    step's CG operator, vb_a of the assembled matrix (8 FP64, 4 FP32):
    - CG operator A p on int16 column offsets: (vb + 2) nnz + 4 (n+1) + 2 vb n;
    - A values: 4 NC ldof (cells) + 8 GD NN (coordinates) + 8 NC (k) + vb_a nnz;
    - F (FP64): 4 NC ldof + 8 GD NN + 8 NC (k) + 8 gdof;
    - matrix-free residual b - A x (FP64 in, FP32 out): 4 NC ldof + 8 GD NN +
      8 NC + 8 gdof (x) + 8 gdof (b) + gdof (boundary flags) + 4 gdof (r).
    The dominant kernel (largest in-step time among the CG operator and the
    matrix-free residual) goes first."""
    nc, nn = mesh.number_of_cells(), mesh.number_of_nodes()
    tname = "double" if vb == 8 else "float"
    # the CG operator reads its values through the int16 column offsets of the
    # finest smoother slices: vb + 2 bytes an entry (the int32-index formula
    # has vb + 4), plus the row pointer and the vectors p (in) and Ap (out)
    b_spmv = (vb + 2) * nnz_bc + 4 * (n + 1) + 2 * vb * n
    pl, pms = kt["pcg_spmv"]
    tr = PCG_SPMV_TRAFFIC_128 if (n == 2146689 and vb == 8) else None
    spmv = _roof(f"k_pcg_spmv<{tname}>", b_spmv, pl, pms, tr,
                 note=f"CG operator: values and int16 column offsets ({vb + 2} B an entry); with int32 indices "
                      f"the count would be {(vb + 4) * nnz_bc + 4 * (n + 1) + 2 * vb * n} B")
    b_r = 16 * nc + 24 * nn + 8 * nc + 8 * gdof + 8 * gdof + gdof + 4 * gdof
    al, ams = kt["p1_apply"]
    resid = _roof("k_p1_apply<3> (residual form)", b_r, al, ams) if al else None
    # finest-level smoothing step on FP16 D^-1 A with int16 column offsets:
    # (2 + 2) nnz + 5 FP32 vectors (dinv, b, x in, x out, d)
    b_s = 4 * nnz_bc + 20 * n
    sl, sms = kt["smooth0"]
    big = n == 2146689
    smooth = _roof("k_cheb16* / k_residual16 (finest level)", b_s, sl, sms, SMOOTH0_TRAFFIC_128 if big else None,
                   note="FP16 D^-1 A values and int16 column offsets: (2 + 2) B an entry + 5 FP32 vectors") \
        if sl else None
    b_a = 16 * nc + 24 * nn + 8 * nc + vb_a * nnz_a
    b_f = 16 * nc + 24 * nn + 8 * nc + 8 * gdof
    (l1, m1), (l2, m2) = kt["a_pass1"], kt["a_pass2"]
    (l3, m3), (l4, m4) = kt["f_pass1"], kt["f_pass2"]
    lb, mb = kt["a_box"]
    if lb:      # the structured from_box kernel: no cell array (cells come from the grid position)
        axes = mesh.box_axes()
        if axes is not None:    # tensor grid: the grid lines instead of the node rows
            b_box = 8 * axes.numel() + 8 * nc + 4 * (gdof + 1) + vb_a * nnz_a
            asm_a = _roof("k_p1_box_rows_tensor", b_box, lb, mb, BOX_A_TRAFFIC_128 if big else None,
                          note="one pass on a tensor grid: the grid lines, k and the row pointer in, CSR values out "
                               "(no node rows, cell array or incidence buffer); with the node coordinates the count "
                               f"would be {24 * nn + 8 * nc + 4 * (gdof + 1) + vb_a * nnz_a} B, the generic kernels' "
                               f"with cells {b_a} B",
                          generic_algorithmic_bytes=int(b_a))
        else:
            b_box = 24 * nn + 8 * nc + 4 * (gdof + 1) + vb_a * nnz_a
            asm_a = _roof("k_p1_box_rows", b_box, lb, mb, None,
                          note="one pass: coordinates, k and the row pointer in, CSR values out (no cell array or "
                               f"incidence buffer); the generic kernels' compulsory count with cells would be {b_a} B",
                          generic_algorithmic_bytes=int(b_a))
    else:
        asm_a = _roof("k_p1_incidence_rows<3> + k_sum_incidence_rows", b_a, 1, m1 + m2,
                      pass1_us=round(m1 * 1e3 / max(l1, 1), 2), pass2_us=round(m2 * 1e3 / max(l2, 1), 2))
    fname = ("k_p1_box_cell_vectors + k_p1_box_sum_vectors" if lb else
             "k_p1_incidence_vectors<3> + k_sum_incidence_values")
    # the structured kernels read no cell array: coordinates, k, the vector out
    asm_f = _roof(fname, 24 * nn + 8 * nc + 8 * gdof if lb else b_f, 1, m3 + m4, BOX_F_TRAFFIC_128 if (big and lb) else None,
                  pass1_us=round(m3 * 1e3 / max(l3, 1), 2), pass2_us=round(m4 * 1e3 / max(l4, 1), 2))
    vl, vms = kt["vcycle"]
    dl, dms = kt["dirichlet"]
    cands = [(pms, spmv), (ams, resid), (sms, smooth)]
    cands = [c for c in cands if c[1] is not None]
    cands.sort(key=lambda c: -c[0])
    main = cands[0][1]
    extra = {"A": asm_a, "F": asm_f, "dirichlet_us": round(dms * 1e3 / max(dl, 1), 2),
             "vcycle_us": round(vms * 1e3 / max(vl, 1), 2), "vcycles": vl,
             "other": [c[1] for c in cands[1:]]}
    return main, extra


def spmv_roofline(A, dtype):
    """Average duration of the SELL SpMV on torch's current stream, timed with
    HIP events over 50 stand-alone launches (an extra figure: the solvers'
    own operator is k_pcg_spmv)."""
    import torch
    from fealpy_amd.sparse.sell import sell_of
    S = sell_of(A)
    S.set_values(A.values, dtype)
    x = torch.rand(A.shape[0], dtype=dtype, device=A.device)
    y = torch.empty_like(x)
    for _ in range(3):
        S.matvec(x, y)
    reps = 50
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        S.matvec(x, y)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    n, nnz = A.shape[0], A.nnz
    eb = 4 if dtype == torch.float32 else 8
    algo = (eb + 4) * nnz + (4 + 2 * eb) * n   # values+indices, indptr + x + y
    gbs = algo / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": f"k_sell_spmv<{'float' if eb == 4 else 'double'}>", "achieved": round(gbs, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": None, "algorithmic_bytes": int(algo), "avg_us": round(ms * 1e3, 2),
            "sell_entries": int(S.entries)}


def spmv_only(args):
    """For hardware-counter runs: assemble the 128^3 system once, run 50
    FP32 SELL SpMVs, one mixed-form MG-PCG solve (its k_pcg_spmv<double>
    launches) and a 256 MiB device-to-device copy (torch elementwise copy)
    as the counter calibration: FETCH_SIZE / WRITE_SIZE of the copy divided
    by its 256 MiB gives the counter's bytes-per-byte factor."""
    import torch
    import fealpy_amd as fa
    from fealpy_amd.solver import GeometricMG
    dev = torch.device("cuda:0")
    mesh, space, k, u, f = make_problem(args.n, dev)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=2)).assembly()
    F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(f, q=2)).assembly()
    bc = fa.DirichletBC(space, gd=u)
    A2, F2 = bc.apply(A, F)
    roof = spmv_roofline(A2, torch.float32)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
    mg.setup(A2)
    x, info = fa.cg(A2, F2, M=mg, atol=0.0, rtol=args.rtol, maxit=400, returninfo=True)
    nbytes = 256 << 20
    xb = torch.rand(nbytes // 4, dtype=torch.float32, device=dev)
    yb = torch.empty_like(xb)
    for _ in range(10):
        yb.copy_(xb)
    torch.cuda.synchronize()
    print(json.dumps({"spmv": roof, "pcg_iterations": info["niter"], "copy_bytes_each_way": nbytes}))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(n, runs):
    """The oracle (NumPy/SciPy FP64 restatement of the reference path) on a
    bounded sample of the same workload, from_box n^3: assemble A and F
    (einsum element tensors, lexsort-coalesce, add.at), Dirichlet filter,
    and the reference's CG loop (unpreconditioned, rtol 1e-10, which reaches
    the 1e-8 solution error).  BLAS and OpenMP may use every host core; the
    sorts, add.at and SciPy's csr_matvec run on one.  Median of `runs`."""
    from oracle import fem as ofem
    from oracle.mesh import tet_from_box
    try:
        from threadpoolctl import threadpool_info
        threads = max((int(p.get("num_threads", 1)) for p in threadpool_info()), default=1)
    except ImportError:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    mesh = tet_from_box(nx=n, ny=n, nz=n)
    k = np.exp(np.random.default_rng(20261202).standard_normal(mesh.number_of_cells()))
    u = ofem.cartesian(lambda p: np.sin(np.pi * p[..., 0]) * np.sin(np.pi * p[..., 1]) * np.sin(np.pi * p[..., 2]))
    f = ofem.barycentric(lambda bcs, index=None: 3 * np.pi ** 2 * k[:, None] * u(mesh.bc_to_point(bcs)))
    c2d = mesh.cell_to_ipoint(1)
    gdof = mesh.number_of_global_ipoints(1)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        A = ofem.assemble_matrix(c2d, [ofem.diffusion_element(mesh, 1, k, 2)], gdof)
        F = ofem.assemble_vector(c2d, [ofem.source_element(mesh, 1, f, 2)], gdof)
        A2, F2 = ofem.dirichlet_apply(A, F, mesh.is_boundary_dof(1), mesh.interpolation_points(1), u)
        t_asm = time.perf_counter() - t0
        t0 = time.perf_counter()
        x, info = ofem.cg(A2, F2, atol=0.0, rtol=1e-10, maxit=100000)
        t_solve = time.perf_counter() - t0
        times.append((t_asm + t_solve, t_asm, t_solve, info["niter"]))
    tot, t_asm, t_solve, it = sorted(times)[len(times) // 2]
    env = {v: os.environ.get(v) for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    return {"value": round(gdof / tot, 1), "unit": "DOF/s", "cores": threads, "kind": "port",
            "sample": f"tet from_box {n}^3 ({gdof} DOFs) of the same workload (lognormal k, q=2, Dirichlet, "
                      f"CG to rtol 1e-10: {it} iterations); median of {runs} runs: assemble+BC {t_asm:.2f} s + "
                      f"CG {t_solve:.2f} s; the oracle (NumPy/SciPy FP64 restatement of the reference path)",
            "runs_s": [round(t[0], 3) for t in times], "host_cpus": os.cpu_count(), "cpu_model": _cpu_model(),
            "thread_env": env, "blas_threads": threads,
            "single_threaded_steps": "lexsort, np.add.at and SciPy csr_matvec run on one core"}


def make_elasticity(n, dev, lam=1.0, mu=0.5):
    """Vector P2 on TetrahedronMesh.from_box n^3, lam = 1, mu = 0.5; body force
    of the manufactured displacement u_i = (i+1) sin(pi x) sin(pi y) sin(pi z):
    f = 3 pi^2 mu s c - (lam + mu) H(s) c with c = (1, 2, 3), H the Hessian of s.
    u vanishes on the boundary (clamped)."""
    import torch
    import fealpy_amd as fa
    mesh = fa.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], nx=n, ny=n, nz=n, device=dev)
    ts = fa.TensorFunctionSpace(fa.LagrangeFESpace(mesh, 2), (-1, 3))
    mat = fa.LinearElasticMaterial("E", lame_lambda=lam, shear_modulus=mu)
    c = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64, device=dev)
    pi = np.pi

    @fa.cartesian
    def f(p):
        sn, cs = torch.sin(pi * p), torch.cos(pi * p)
        s = sn.prod(dim=-1)
        hxy = pi ** 2 * cs[..., 0] * cs[..., 1] * sn[..., 2]
        hxz = pi ** 2 * cs[..., 0] * sn[..., 1] * cs[..., 2]
        hyz = pi ** 2 * sn[..., 0] * cs[..., 1] * cs[..., 2]
        hc = torch.stack([-pi ** 2 * s * c[0] + hxy * c[1] + hxz * c[2],
                          hxy * c[0] - pi ** 2 * s * c[1] + hyz * c[2],
                          hxz * c[0] + hyz * c[1] - pi ** 2 * s * c[2]], dim=-1)
        return 3 * pi ** 2 * mu * s[..., None] * c - (lam + mu) * hc

    @fa.cartesian
    def u(p):
        return torch.sin(pi * p).prod(dim=-1)[..., None] * c

    return mesh, ts, mat, f, u


def elasticity_step(ts, mat, f, mg, rtol, maxit=1000):
    import fealpy_amd as fa
    A = fa.BilinearForm(ts).add_integrator(fa.LinearElasticityIntegrator(mat, q=5)).assembly()
    F = fa.LinearForm(ts).add_integrator(fa.VectorSourceIntegrator(f, q=5)).assembly()
    A2, F2 = fa.DirichletBC(ts, gd=0.0).apply(A, F)
    mg.setup(A2)
    x, info = fa.cg(A2, F2, M=mg, atol=0.0, rtol=rtol, maxit=maxit, returninfo=True)
    return x, info, A2, F2


def elasticity_main(args):
    import torch
    import fealpy_amd as fa
    from fealpy_amd.solver import GeometricMG
    n = args.n if args.n != 128 else 40
    dev = torch.device("cuda:0")
    mesh, ts, mat, f, u = make_elasticity(n, dev)
    gdof = ts.number_of_global_dofs()
    from fealpy_amd.fem.pattern import pattern_of
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pat = pattern_of(ts)
    torch.cuda.synchronize()
    t_pattern = time.perf_counter() - t0
    bd = fa.DirichletBC(ts, gd=0.0).is_boundary_dof
    mg32 = GeometricMG(ts, bd, dtype=torch.float32)
    mg64 = GeometricMG(ts, bd, dtype=torch.float64)
    x_ref, info_ref, A2, _ = elasticity_step(ts, mat, f, mg64, 1e-14)
    nref = float(torch.linalg.norm(x_ref))
    t_m, (x_m, info_m, _, _) = timed(lambda: elasticity_step(ts, mat, f, mg32, args.rtol), args.steps, args.warmup)
    t_d, (x_d, info_d, _, _) = timed(lambda: elasticity_step(ts, mat, f, mg64, args.rtol), args.steps, args.warmup)
    err_m = float(torch.linalg.norm(x_m - x_ref)) / nref
    err_d = float(torch.linalg.norm(x_d - x_ref)) / nref
    err_l2 = float(mesh.error(lambda p: u(p), ts.function(x_m)))
    print(json.dumps({
        "metric": "DOF/s assembled+solved to 1e-8 rel err, mixed vs FP64 GPU; serving p99",
        "value": round(gdof / t_m, 1), "unit": "DOF/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_m * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": f"3D linear elasticity, vector P2, TetrahedronMesh.from_box {n}^3, lam=1, mu=0.5, "
                               f"q=5, clamped, multigrid-preconditioned CG", "gdof": gdof,
                   "ncell": mesh.number_of_cells(), "nnz": int(A2.nnz), "rtol": args.rtol},
        "rel_err_ok": bool(max(err_m, err_d) <= 1e-8),
        "fp64": {"value": round(gdof / t_d, 1), "ms_per_step": round(t_d * 1e3, 3), "rel_err": err_d,
                 "pcg_iterations": info_d["niter"]},
        "mixed": {"rel_err": err_m, "pcg_iterations": info_m["niter"], "speedup_vs_fp64": round(t_d / t_m, 3)},
        "l2_error_vs_manufactured": err_l2,
        "cold": {"pattern_ms": round(t_pattern * 1e3, 3)},
        "reference_solve": {"pcg_iterations": info_ref["niter"]},
        "levels": [lv.n for lv in mg32.levels],
    }))


MFMA_PEAK_TFLOPS = {"bf16": 2500.0, "fp32": 157.3}
FP64_PEAK_TFLOPS = 78.6


def highorder_main(args):
    """Element matrices of diffusion k(x) = 1 + 0.5 sin(2 pi x) cos(2 pi y) z
    plus mass 1 on TetrahedronMesh.from_box n^3, P3 (q=5, 35 points) and P4
    (q=6, 56 points): FP64 element kernel vs the MFMA contraction in FP32 and
    BF16 (FP32 accumulation).  Kernel time by HIP events on the launching
    stream; FLOPs = 2 NC NQ 7 ldof^2."""
    import torch
    import fealpy_amd as fa
    from fealpy_amd.fem.highorder import element_matrices_mfma, mfma_flops
    from fealpy_amd.fem.integrators import GroupIntegrator, coef_spec
    n = args.n if args.n != 128 else 48
    dev = torch.device("cuda:0")
    mesh = fa.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], nx=n, ny=n, nz=n, device=dev)
    pi = np.pi

    @fa.cartesian
    def kx(p):
        return 1 + 0.5 * torch.sin(2 * pi * p[..., 0]) * torch.cos(2 * pi * p[..., 1]) * p[..., 2]

    def ev_time(fn, reps=5):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            out = fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, out

    results = []
    for p, q in ((3, 5), (4, 6)):
        space = fa.LagrangeFESpace(mesh, p)
        nq = space.tables(q)[1].shape[0]
        ldof = space.number_of_local_dofs()
        nc = mesh.number_of_cells()
        diff, mass = coef_spec(kx, space, q), coef_spec(1.0, space, q)      # coefficient sampled once
        kq = diff[2]
        g64 = GroupIntegrator(fa.ScalarDiffusionIntegrator(kq, q=q), fa.ScalarMassIntegrator(1.0, q=q))
        ms64, ref = ev_time(lambda: g64.assembly(space), reps=2)
        flops = mfma_flops(nc, nq, ldof)
        print(f"# highorder contraction p={p}: FP64 element kernel {ms64:.2f} ms", file=sys.stderr, flush=True)
        row = {"p": p, "q": q, "nq": nq, "ldof": ldof, "ncell": nc, "gflop": round(flops / 1e9, 2),
               "fp64_element_ms": round(ms64, 3),
               "fp64_tflops": round(flops / (ms64 * 1e-3) / 1e12, 2)}
        scale = float(ref.abs().max())
        for prec in ("fp32", "bf16"):
            ms, got = ev_time(lambda: element_matrices_mfma(space, diff, mass, q, prec))
            tf = flops / (ms * 1e-3) / 1e12
            out_gbs = nc * ldof * ldof * 4 / (ms * 1e-3) / 1e9
            row[prec] = {"ms": round(ms, 3), "tflops": round(tf, 1), "peak_tflops": MFMA_PEAK_TFLOPS[prec],
                         "frac": round(tf / MFMA_PEAK_TFLOPS[prec], 4),
                         "output_write_gbs": round(out_gbs, 1),
                         "max_rel_err_vs_fp64": float((got.double() - ref).abs().max()) / scale}
            del got
        del ref
        results.append(row)
        torch.cuda.empty_cache()
    solves = []
    if not args.no_solve:
        for p, q in ((3, 5), (4, 6)):
            try:
                solves.append(highorder_solve(n, p, q, args.rtol, max(1, min(args.steps, 3))))
            except Exception as e:               # report and keep the contraction figures
                solves.append({"p": p, "error": f"{type(e).__name__}: {e}"})
            print(f"# highorder solve p={p}: {solves[-1]}", file=sys.stderr, flush=True)
    best = results[-1]["bf16"]
    print(json.dumps({
        "metric": "MFMA element contraction TFLOP/s (high-order diffusion + mass)", "value": best["tflops"],
        "unit": "TFLOP/s", "n_gpus": 1, "higher_is_better": True, "dtype": "bf16 in, f32 acc", "data": "synthetic",
        "config": {"workload": f"TetrahedronMesh.from_box {n}^3, P3 (q=5) and P4 (q=6), k(x)=1+0.5 sin(2 pi x) "
                               f"cos(2 pi y) z + mass 1"},
        "roofline": {"bound": "mfma", "achieved": best["tflops"], "peak": MFMA_PEAK_TFLOPS["bf16"], "unit": "TFLOP/s",
                     "frac": best["frac"], "traffic": None},
        "results": results, "solve": solves}))


def _highorder_exprs():
    """k(x) = 1 + 0.5 sin(2 pi x) cos(2 pi y) z, u = sin sin sin and
    f = -div(k grad u) + u, as separable expressions (in-kernel)."""
    from fealpy_amd.expr import cos, sin, x, y, z
    pi = np.pi
    k = 1 + 0.5 * sin(2 * pi * x) * cos(2 * pi * y) * z
    u = sin(pi * x) * sin(pi * y) * sin(pi * z)
    dk = (pi * cos(2 * pi * x) * cos(2 * pi * y) * z, -pi * sin(2 * pi * x) * sin(2 * pi * y) * z,
          0.5 * sin(2 * pi * x) * cos(2 * pi * y))
    du = (pi * cos(pi * x) * sin(pi * y) * sin(pi * z), pi * sin(pi * x) * cos(pi * y) * sin(pi * z),
          pi * sin(pi * x) * sin(pi * y) * cos(pi * z))
    f = 3 * pi ** 2 * k * u - (dk[0] * du[0] + dk[1] * du[1] + dk[2] * du[2]) + u
    return k, u, f


def highorder_solve(n, p, q, rtol, steps):
    """cfg3 assembled and solved: diffusion k(x) + mass 1 on from_box n^3 with
    P p, degree-2p rule q; assemble A (FP64 element kernel + fixed-order
    gather), F, the Dirichlet condition with g = u, the multigrid hierarchy
    (Pp -> P1 on the same mesh, then geometric), and MG-preconditioned CG
    (FP64 outer, FP32 cycle) to rtol.  The error is against a tight FP64
    solve of the same system."""
    import torch
    import fealpy_amd as fa
    from fealpy_amd.fem.pattern import pattern_of
    from fealpy_amd.solver import GeometricMG
    dev = torch.device("cuda:0")
    k, u, f = _highorder_exprs()
    mesh = fa.TetrahedronMesh.from_box([0, 1, 0, 1, 0, 1], nx=n, ny=n, nz=n, device=dev)
    space = fa.LagrangeFESpace(mesh, p)
    gdof = space.number_of_global_dofs()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pat = pattern_of(space)
    torch.cuda.synchronize()
    t_pat = time.perf_counter() - t0
    bc = fa.DirichletBC(space, gd=u)
    mg = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
    ph = {}

    def one(rt, record):
        def t(name, fn):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            if record:
                ph[name] = ph.get(name, 0.0) + (time.perf_counter() - t1) * 1e3
            return r
        A = t("assemble_A_ms", lambda: fa.BilinearForm(space).add_integrator(
            fa.ScalarDiffusionIntegrator(k, q=q), fa.ScalarMassIntegrator(1.0, q=q)).assembly())
        F = t("assemble_F_ms", lambda: fa.LinearForm(space).add_integrator(
            fa.ScalarSourceIntegrator(f, q=q)).assembly())
        A2, F2 = t("dirichlet_ms", lambda: bc.apply(A, F))
        del A
        t("mg_setup_ms", lambda: mg.setup(A2))
        xs, info = t("pcg_ms", lambda: fa.cg(A2, F2, M=mg, atol=0.0, rtol=rt, maxit=2000, returninfo=True))
        return xs, info, A2
    print(f"# highorder solve p={p}: {gdof} DOFs, pattern {t_pat * 1e3:.1f} ms; reference solve", file=sys.stderr,
          flush=True)
    x_ref, info_ref, A2 = one(1e-13, False)
    print(f"# highorder solve p={p}: reference solve {info_ref['niter']} iterations; timed steps", file=sys.stderr,
          flush=True)
    nnz = int(A2.nnz)
    del A2
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        xs, info, _ = one(rtol, True)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / steps
    err = float(torch.linalg.norm(xs - x_ref) / torch.linalg.norm(x_ref))
    l2 = float(mesh.error(u, space.function(xs), q=q))
    out = {"p": p, "q": q, "gdof": gdof, "nnz": nnz, "ms_per_step": round(t_step * 1e3, 2),
           "dof_per_s": round(gdof / t_step, 1), "rel_err": err, "rel_err_ok": bool(err <= 1e-8),
           "pcg_iterations": info["niter"], "reference_iterations": info_ref["niter"], "l2_error": l2,
           "phases_ms": {kk: round(v / steps, 3) for kk, v in ph.items()},
           "cold_pattern_ms": round(t_pat * 1e3, 2), "levels": [lv.n for lv in mg.levels]}
    del mg, bc, pat, space, mesh, xs, x_ref
    torch.cuda.empty_cache()
    return out


def _pcts(lat):
    lat = np.sort(np.asarray(lat))
    return {"p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
            "mean_ms": round(float(lat.mean()) * 1e3, 2)}


def _wait_all(reqs, label):
    """Wait for every request in order, with a progress line per 256 done."""
    for i, r in enumerate(reqs):
        r.event.wait()
        if (i + 1) % 256 == 0:
            print(f"# serving {label}: {i + 1}/{len(reqs)} done", file=sys.stderr, flush=True)


def run_serving(args):
    """Mixed-size serving: `requests` 2D diffusion requests (TriangleMesh.from_box
    nx^2, nx log-uniform in [16, 512], P1/P2 50:50, k = exp(N(0,1)) per cell,
    8 Fourier modes, u = 0 on the boundary), solved to rtol 1e-8 with
    multigrid-preconditioned CG.  Resident per-size data is built first
    (reported as cold time).  Then: (1) the plain one-at-a-time loop on one
    stream, (2) the continuous-batching server with `slots` streams, both
    saturated (every request queued at t=0) for solves/s, (3) the server under
    Poisson arrivals at `rate` for p50/p99 latency (arrival to result); the
    plain loop's latency at that rate follows from its measured service times
    (a FIFO single server: start = max(arrival, previous done)).  Needs
    GPU_MAX_HW_QUEUES >= 2 x slots, set before HIP initialises (main does)."""
    import torch
    from fealpy_amd.serving import MAX_BATCH, WG_MAX_GDOF, SolveServer, make_requests
    n = args.requests
    seed = 20261202
    srv = SolveServer(slots=args.slots, rtol=args.rtol, resident_budget=int(args.resident_gb * (1 << 30)))
    plain = make_requests(n, seed=seed)
    sizes = sorted({(r.nx, r.p) for r in plain})
    t_warm = 0.0
    for i in range(0, len(sizes), 64):
        t_warm += srv.warm(sizes[i:i + 64])
        print(f"# serving warm {min(i + 64, len(sizes))}/{len(sizes)} sizes, {t_warm:.1f}s", file=sys.stderr,
              flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, r in enumerate(plain):
        srv.solve(r)
        if (i + 1) % 256 == 0:
            print(f"# serving plain loop {i + 1}/{n} solved", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    print(f"# serving plain loop {n / t_plain:.1f} solves/s", file=sys.stderr, flush=True)
    mix = make_requests(n, seed=seed)
    srv.start()
    t0 = time.perf_counter()
    for r in mix:
        srv.submit(r)
    _wait_all(mix, "server")
    t_srv = time.perf_counter() - t0
    thr = n / t_srv
    print(f"# serving server {thr:.1f} solves/s", file=sys.stderr, flush=True)
    same = sum(int(torch.equal(a.x, b.x)) for a, b in zip(plain, mix) if b.error is None)
    errors = sum(1 for r in mix if r.error is not None)
    rate = args.rate if args.rate > 0 else 0.8 * thr
    timed_reqs = make_requests(n, seed=seed, rate=rate)
    t0 = time.perf_counter()
    for r in timed_reqs:
        while True:
            now = time.perf_counter() - t0
            if now >= r.arrival:
                break
            time.sleep(min(r.arrival - now, 0.002))
        r.arrival += t0
        srv.submit(r)
    _wait_all(timed_reqs, "server, Poisson arrivals")
    srv.close()
    lat_srv = [r.done - r.arrival for r in timed_reqs]
    # the same server with batches: a slot takes the requests waiting when
    # it frees up and solves the small ones in one launch, one workgroup each
    batched = None
    if args.serving_batch > 1:
        # small sizes also keep an uncompressed hierarchy for workgroup solves
        small = [(nx, p) for nx, p in sizes if (p * nx + 1) ** 2 <= WG_MAX_GDOF]
        t_wb = 0.0
        for i in range(0, len(small), 64):
            t_wb += srv.warm(small[i:i + 64], packed=True)
        print(f"# serving warm (workgroup hierarchies, {len(small)} sizes) {t_wb:.1f}s", file=sys.stderr, flush=True)
        srv.batch = min(args.serving_batch, MAX_BATCH)
        mixb = make_requests(n, seed=seed)
        srv.start()
        t0 = time.perf_counter()
        for r in mixb:
            srv.submit(r)
        _wait_all(mixb, "server, batches")
        t_b = time.perf_counter() - t0
        thr_b = n / t_b
        print(f"# serving server, batches {thr_b:.1f} solves/s", file=sys.stderr, flush=True)
        close = sum(int(float(torch.linalg.norm(b.x - a.x)) <= 1e-8 * float(torch.linalg.norm(a.x)))
                    for a, b in zip(plain, mixb) if b.error is None)
        errors_b = sum(1 for r in mixb if r.error is not None)
        rate_b = args.rate if args.rate > 0 else 0.8 * thr_b
        timed_b = make_requests(n, seed=seed, rate=rate_b)
        t0 = time.perf_counter()
        for r in timed_b:
            while True:
                now = time.perf_counter() - t0
                if now >= r.arrival:
                    break
                time.sleep(min(r.arrival - now, 0.002))
            r.arrival += t0
            srv.submit(r)
        _wait_all(timed_b, "server, batches, Poisson arrivals")
        srv.close()
        lat_b = [r.done - r.arrival for r in timed_b]
        sizes_b = [r.info.get("batch", 1) for r in mixb if r.info]
        batched = {"solves_per_s": round(thr_b, 2), "speedup_vs_plain": round(thr_b * t_plain / n, 3),
                   "arrival_rate": round(rate_b, 2), **_pcts(lat_b), "max_batch": srv.batch,
                   "mean_requests_per_workgroup_launch": round(float(np.mean(sizes_b)), 2) if sizes_b else None,
                   "results_within_1e-8_of_one_at_a_time": f"{close}/{n}", "request_errors": errors_b,
                   "cold_workgroup_warm_s": round(t_wb, 2)}
    service = [r.done - r.start for r in plain]
    done, lat_plain = 0.0, []
    for r, sv in zip(timed_reqs, service):
        arr = r.arrival - t0
        done = max(arr, done) + sv
        lat_plain.append(done - arr)
    gdofs = [(r.p * r.nx + 1) ** 2 for r in plain]
    iters = [r.info["niter"] for r in plain]
    head = {"mode": "streams", "solves_per_s": round(thr, 2), "speedup_vs_plain": round(thr * t_plain / n, 3),
            "p50_ms": _pcts(lat_srv)["p50_ms"], "p99_ms": _pcts(lat_srv)["p99_ms"]}
    if batched is not None and batched["solves_per_s"] > thr:
        head = {"mode": "batches (workgroup solves)", "solves_per_s": batched["solves_per_s"],
                "speedup_vs_plain": batched["speedup_vs_plain"], "p50_ms": batched["p50_ms"],
                "p99_ms": batched["p99_ms"]}
    return {
        **head,
        "config": {"workload": f"{n} requests, TriangleMesh.from_box nx^2 with nx log-uniform in [16, 512], P1/P2, "
                               f"k=exp(N(0,1)) per cell, 8 Fourier modes, u=0 on the boundary",
                   "slots": args.slots, "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                   "rtol": args.rtol, "distinct_sizes": len(sizes), "resident_gb": args.resident_gb,
                   "batch": args.serving_batch, "workgroup_max_gdof": WG_MAX_GDOF,
                   "mean_gdof": round(float(np.mean(gdofs)), 1), "max_gdof": int(max(gdofs)),
                   "pcg_iterations_mean": round(float(np.mean(iters)), 2), "pcg_iterations_max": int(max(iters))},
        "plain_loop": {"solves_per_s": round(n / t_plain, 2), **_pcts(lat_plain)},
        "server": {"solves_per_s": round(thr, 2), "arrival_rate": round(rate, 2), **_pcts(lat_srv)},
        "server_batches": batched,
        "results_identical_alone_and_in_mix": f"{same}/{n}", "request_errors": errors,
        "cold": {"warm_s": round(t_warm, 2)}, "resident_evictions": srv.evictions,
        "dof_per_s_server": round(sum(gdofs) / t_srv, 1)}


def workgroups_main(args):
    """Per-request GPU time of small 2D requests (P1 nx 16-160, P2 nx 8-80,
    lognormal k, rtol 1e-8): one request on the multi-kernel MG-PCG path
    (compressed hierarchy) against fa_pcg_wg_batched with one request and
    with 32 requests of the same size in one launch (uncompressed hierarchy,
    one workgroup each).  One JSON line with a row per size."""
    import torch
    import fealpy_amd as fa
    from fealpy_amd.solver import GeometricMG, pcg_workgroups
    dev = torch.device("cuda:0")

    def system(nx, p, seed):
        mesh = fa.TriangleMesh.from_box([0, 1, 0, 1], nx=nx, ny=nx, device=dev)
        space = fa.LagrangeFESpace(mesh, p)
        k = torch.as_tensor(np.exp(np.random.default_rng(seed).standard_normal(mesh.number_of_cells())), device=dev)
        A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=3)).assembly()
        F = fa.LinearForm(space).add_integrator(fa.ScalarSourceIntegrator(1.0, q=3)).assembly()
        bc = fa.DirichletBC(space, gd=0.0)
        A, F = bc.apply(A, F)
        mgc = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
        mgu = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32, compress=False)
        mgc.setup(A)
        mgu.setup(A)
        return A, F, mgc, mgu

    def timed_ms(fn, reps):
        out = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3, out

    rows = []
    for p, nxs in ((1, (16, 32, 64, 96, 128, 160)), (2, (8, 16, 32, 48, 64, 80))):
        for nx in nxs:
            A, F, mgc, mgu = system(nx, p, nx)
            ms_k, (x1, i1) = timed_ms(lambda: mgc.pcg(A, F, atol=0.0, rtol=1e-8, maxit=500), 5)
            ms_1, out = timed_ms(lambda: pcg_workgroups([(mgu, A, F)], atol=0.0, rtol=1e-8, maxit=500), 5)
            batch = [system(nx, p, 100 + s) for s in range(32)]
            ms_32, _ = timed_ms(lambda: pcg_workgroups([(b[3], b[0], b[1]) for b in batch], atol=0.0, rtol=1e-8,
                                                       maxit=500), 3)
            x2, i2 = out[0]
            rows.append({"p": p, "nx": nx, "gdof": A.shape[0], "multi_kernel_ms": round(ms_k, 3),
                         "multi_kernel_iterations": i1["niter"], "workgroup_ms": round(ms_1, 3),
                         "workgroup_iterations": i2["niter"], "batch32_ms": round(ms_32, 3),
                         "batch32_ms_per_request": round(ms_32 / 32, 4),
                         "rel_diff": float(torch.linalg.norm(x2 - x1) / torch.linalg.norm(x1))})
            print(f"# workgroups {rows[-1]}", file=sys.stderr, flush=True)
            del batch
    print(json.dumps({"metric": "GPU time a small request, multi-kernel vs workgroup solves", "unit": "ms",
                      "n_gpus": 1, "data": "synthetic", "rows": rows}))


def serving_main(args):
    s = run_serving(args)
    print(json.dumps({
        "metric": "serving solves/s and p99 latency, mixed-size 2D diffusion to rtol 1e-8",
        "value": s["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f64+f32+f16", "data": "synthetic", "config": s.pop("config"), **s}))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: the multi-GPU path is a design only, not code (asked --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    # hardware queues for the serving streams: HIP reads this when it initialises
    want = min(32, max(SERVING_HW_QUEUES, 2 * args.slots))
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < want:
        os.environ["GPU_MAX_HW_QUEUES"] = str(want)
    if args.spmv_only:
        spmv_only(args)
        return
    if args.workload == "elasticity":
        elasticity_main(args)
        return
    if args.workload == "highorder":
        highorder_main(args)
        return
    if args.workload == "serving":
        serving_main(args)
        return
    if args.workload == "workgroups":
        workgroups_main(args)
        return
    import torch
    import fealpy_amd as fa
    from fealpy_amd.solver import GeometricMG
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    mesh, space, k, u, f = make_problem(args.n, dev, args.torch_source)
    torch.cuda.synchronize()
    t_mesh = time.perf_counter() - t0
    gdof = space.number_of_global_dofs()
    # cold pattern build (kept resident afterwards)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    from fealpy_amd.fem.pattern import pattern_of
    pat = pattern_of(space)
    torch.cuda.synchronize()
    t_pattern = time.perf_counter() - t0
    # tile shapes of the step's tunable kernels, measured on this problem (untimed)
    from fealpy_amd.tune import Autotuner
    tuner = Autotuner(path=None)
    tuned_lanes, tuned_ms = tuner.tune_assembly(space, k, q=2)
    bc = fa.DirichletBC(space, gd=u)
    mg32 = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)
    mg64 = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float64)

    mgr = GeometricMG(space, bc.is_boundary_dof, dtype=torch.float32)     # hierarchy of the FP32 operator

    # reference solution: FP64 multigrid CG to a tight tolerance (untimed)
    x_ref, info_ref, A2, F2 = step(space, k, u, f, mg64, 1e-14, maxit=1000)
    nref = float(torch.linalg.norm(x_ref))

    t_mixed, (x_m, info_m, _, _) = timed(lambda: step_refine(space, k, u, f, mgr, args.rtol), args.steps,
                                         args.warmup)
    err_m = float(torch.linalg.norm(x_m - x_ref)) / nref
    t_cyc, (x_c, info_c, _, _) = timed(lambda: step(space, k, u, f, mg32, args.rtol), args.steps, args.warmup)
    err_c = float(torch.linalg.norm(x_c - x_ref)) / nref
    t_fp64, (x_d, info_d, _, _) = timed(lambda: step(space, k, u, f, mg64, args.rtol), args.steps, args.warmup)
    err_d = float(torch.linalg.norm(x_d - x_ref)) / nref
    if max(err_m, err_d, err_c) > 1e-8:
        print(f"bench: solution error {max(err_m, err_d, err_c):.3e} exceeds 1e-8; the DOF/s figure is not valid",
              file=sys.stderr)
    phases_m = phase_breakdown_refine(space, k, u, f, mgr, args.rtol)
    phases_c = phase_breakdown(space, k, u, f, mg32, args.rtol)
    phases_d = phase_breakdown(space, k, u, f, mg64, args.rtol)
    kt = in_step_kernels(step_refine, space, k, u, f, mgr, args.rtol)
    kt64 = in_step_kernels(step, space, k, u, f, mg64, args.rtol)
    err_l2 = float(mesh.error(u, space.function(x_m)))
    nnz_a = pat.nnz
    roof, roof_asm = rooflines(kt, mesh, gdof, nnz_a, int(A2.nnz), gdof, 4)     # A itself is FP64 in both forms
    roof64, roof_asm64 = rooflines(kt64, mesh, gdof, nnz_a, int(A2.nnz), gdof, 8)
    roof32 = spmv_roofline(A2, torch.float32)
    if args.n == 128:
        roof32["traffic"] = SPMV_F32_TRAFFIC_128
    out = {
        "metric": "DOF/s assembled+solved to 1e-8 rel err, mixed vs FP64 GPU; serving p99",
        "value": round(gdof / t_mixed, 1),
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32+f16",
        "data": "synthetic",
        "config": {"workload": f"3D P1 diffusion, TetrahedronMesh.from_box {args.n}^3, lognormal per-cell k, "
                               f"q=2, multigrid-preconditioned CG", "gdof": gdof,
                   "ncell": mesh.number_of_cells(), "nnz": int(A2.nnz), "rtol": args.rtol,
                   "source": "torch callable" if args.torch_source else "separable expression, in-kernel"},
        "fp64": {"value": round(gdof / t_fp64, 1), "ms_per_step": round(t_fp64 * 1e3, 3), "rel_err": err_d,
                 "pcg_iterations": info_d["niter"], "phases": phases_d, "roofline": roof64,
                 "roofline_assembly": roof_asm64},
        "rel_err_ok": bool(max(err_m, err_d, err_c) <= 1e-8),
        "mixed": {"form": "A assembled in FP64; FP32 Dirichlet-filtered copy (rounded in the filter pass), FP32 "
                          f"MG setup and MG-PCG corrections to rtol {INNER_RTOL:g}; FP64 residuals b - A x by one "
                          "SpMV pass (iterative refinement)",
                  "rel_err": err_m, "pcg_iterations": info_m["niter"], "refinements": info_m["refinements"],
                  "phases": phases_m, "speedup_vs_fp64": round(t_fp64 / t_mixed, 3)},
        "fp64_cg_fp32_cycle": {"form": "FP64 assembly and CG, FP32/FP16 V-cycle", "value": round(gdof / t_cyc, 1),
                               "ms_per_step": round(t_cyc * 1e3, 3), "rel_err": err_c,
                               "pcg_iterations": info_c["niter"], "phases": phases_c,
                               "speedup_vs_fp64": round(t_fp64 / t_cyc, 3)},
        "l2_error_vs_manufactured": err_l2,
        "cold": {"pattern_ms": round(t_pattern * 1e3, 3), "mesh_s": round(t_mesh, 3),
                 "dof_per_s_with_pattern": round(gdof / (t_mixed + t_pattern), 1)},
        "reference_solve": {"pcg_iterations": info_ref["niter"]},
        "autotune": {"assembly_pass2_lanes_per_row": tuned_lanes,
                     "candidates_ms": {str(c): round(v, 4) for c, v in tuned_ms.items()}},
        "roofline": roof,
        "roofline_assembly": roof_asm,
        "roofline_sell_spmv_f32": roof32,
    }
    del x_ref, x_m, x_d, x_c, A2, F2, mg32, mg64, mgr, pat, bc, space, mesh, k
    torch.cuda.empty_cache()
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(args.cpu_n, args.cpu_runs)
    if not args.no_serving:
        try:
            out["serving"] = run_serving(args)
        except Exception as e:          # the headline figures stand on their own
            out["serving"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
