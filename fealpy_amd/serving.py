"""Continuous-batching solve server for many Lagrange problems of mixed size.

A request is a 2D diffusion problem -div(k grad u) = f on the unit square
(TriangleMesh.from_box nx x nx, P1 or P2, per-cell coefficient k, source a
sum of Fourier modes, u = 0 on the boundary).  The server keeps, per
(nx, p), the mesh, the space, its sparsity pattern and boundary data
resident on the device, and a pool of multigrid hierarchies (one per request
of that size in flight), so repeated sizes skip every symbolic step.

Continuous batching: ``slots`` worker threads, each with its own HIP stream
(default priority), take admitted requests as soon as they finish the last
one; requests that arrive while others are on the GPU are admitted as slots
and the HBM budget free up, and their kernels overlap on the device.  The
solve of a request is the same sequence of deterministic kernels whether it
runs alone or in the mix, so its result is bitwise the same.

``SolveServer.solve`` runs one request on the calling thread's current
stream; a plain one-at-a-time loop is ``for r in reqs: server.solve(r)``.

HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 by
default, read when HIP initialises); streams beyond that share queues and
their kernels serialise, so a process serving with ``slots`` streams should
set GPU_MAX_HW_QUEUES to at least slots + 2 before its first GPU call.
"""
import math
import queue
import threading
import time

import numpy as np
import torch

from .fem import BilinearForm, DirichletBC, LinearForm, ScalarDiffusionIntegrator, ScalarSourceIntegrator
from . import expr as ex
from .expr import sin
from .functionspace import LagrangeFESpace
from .mesh import TriangleMesh
from .solver import GeometricMG, cg

PI = math.pi


class SolveRequest:
    """nx, p, per-cell coefficient k (NC,) and Fourier modes (M, 3) rows
    (amplitude, kx, ky): f = sum_m a_m sin(kx_m pi x) sin(ky_m pi y)."""

    __slots__ = ("rid", "nx", "p", "k", "modes", "arrival", "start", "done", "x", "info", "event", "error",
                 "nbytes")

    def __init__(self, rid, nx, p, k, modes, arrival=0.0):
        self.rid, self.nx, self.p, self.k, self.modes, self.arrival = rid, int(nx), int(p), k, modes, arrival
        self.start = self.done = None
        self.x = self.info = self.error = None
        self.event = threading.Event()
        self.nbytes = 0

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def fourier_source(modes, device=None):
    """f = sum_m a_m sin(kx_m pi x) sin(ky_m pi y) as a separable expression:
    the load-vector kernel evaluates it at the quadrature points itself."""
    f = None
    for a, kx, ky in np.asarray(modes, dtype=np.float64):
        t = float(a) * sin(float(kx) * PI * ex.x) * sin(float(ky) * PI * ex.y)
        f = t if f is None else f + t
    return f


def request_bytes(nx, p):
    """Device memory one request of this size holds while in flight (operator,
    filtered operator, SELL copies, hierarchy values, vectors); an estimate
    from the DOF count with 7 (P1) / 15 (P2) nonzeros per row."""
    gdof = (p * nx + 1) ** 2
    per_row = 7 if p == 1 else 15
    return int(gdof * (per_row * 40 + 200))


class _SizeContext:
    """Resident data of one (nx, p): mesh, space, pattern, boundary flags,
    and the pool of multigrid hierarchies."""

    def __init__(self, nx, p, device, mg_kwargs):
        from .fem.pattern import pattern_of
        self.mesh = TriangleMesh.from_box([0, 1, 0, 1], nx=nx, ny=nx, device=device)
        self.space = LagrangeFESpace(self.mesh, p)
        pattern_of(self.space)
        self.space._boundary()
        self.space.tables(3)
        self.is_bd = DirichletBC(self.space, gd=0.0).is_boundary_dof
        torch.cuda.current_stream(self.mesh.node.device).synchronize()   # resident data complete before sharing
        self.mg_kwargs = mg_kwargs
        self._pool = []
        self._lock = threading.Lock()

    def checkout(self):
        with self._lock:
            if self._pool:
                return self._pool.pop()
        return GeometricMG(self.space, self.is_bd, **self.mg_kwargs)

    def checkin(self, mg):
        with self._lock:
            self._pool.append(mg)


class SolveServer:
    def __init__(self, *, slots=4, hbm_budget=64 << 30, rtol=1e-8, maxit=2000, q=3, mg_dtype=torch.float32,
                 graphs=False, device=None):
        self.device = torch.device("cuda") if device is None else torch.device(device)
        self.slots, self.budget, self.rtol, self.maxit, self.q = slots, int(hbm_budget), rtol, maxit, q
        self.mg_kwargs = {"dtype": mg_dtype, "graphs": graphs}
        self._ctx = {}
        self._ctx_lock = threading.Lock()
        self._key_locks = {}
        self._queue = queue.Queue()
        self._in_flight = 0
        self._budget_cv = threading.Condition()
        self._threads = []
        self._closed = False

    # -- resident per-size data --------------------------------------------------
    def context(self, nx, p):
        key = (nx, p)
        with self._ctx_lock:
            ctx = self._ctx.get(key)
            if ctx is not None:
                return ctx
            klock = self._key_locks.setdefault(key, threading.Lock())
        with klock:
            with self._ctx_lock:
                ctx = self._ctx.get(key)
            if ctx is None:
                ctx = _SizeContext(nx, p, self.device, self.mg_kwargs)
                with self._ctx_lock:
                    self._ctx[key] = ctx
        return ctx

    def warm(self, sizes):
        """Build the resident data (and one hierarchy with its symbolic
        products) of every (nx, p) in `sizes`; returns seconds spent."""
        t0 = time.perf_counter()
        for nx, p in sizes:
            ctx = self.context(nx, p)
            mg = ctx.checkout()
            nc = ctx.mesh.number_of_cells()
            r = SolveRequest(-1, nx, p, torch.ones(nc, dtype=torch.float64, device=self.device),
                             [[1.0, 1.0, 1.0]])
            self._run(r, ctx, mg)
            ctx.checkin(mg)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # -- one request ---------------------------------------------------------------
    def _run(self, req, ctx, mg):
        space = ctx.space
        k = req.k.to(self.device, torch.float64, non_blocking=True)
        f = fourier_source(req.modes, self.device)
        A = BilinearForm(space).add_integrator(ScalarDiffusionIntegrator(k, q=self.q)).assembly()
        F = LinearForm(space).add_integrator(ScalarSourceIntegrator(f, q=self.q)).assembly()
        A, F = DirichletBC(space, gd=0.0).apply(A, F)
        mg.setup(A)
        x, info = cg(A, F, M=mg, atol=0.0, rtol=self.rtol, maxit=self.maxit, returninfo=True)
        return x, info

    def solve(self, req):
        """Solve on the calling thread's current stream; fills req.x / req.info."""
        req.start = time.perf_counter()
        ctx = self.context(req.nx, req.p)
        mg = ctx.checkout()
        try:
            req.x, req.info = self._run(req, ctx, mg)
        finally:
            ctx.checkin(mg)
        req.done = time.perf_counter()
        return req

    # -- continuous batching ----------------------------------------------------
    def start(self):
        for i in range(self.slots):
            t = threading.Thread(target=self._worker, name=f"fa-slot-{i}", daemon=True)
            t.start()
            self._threads.append(t)
        return self

    def submit(self, req):
        """Queue a request (admission happens when a slot takes it and the
        HBM budget has room); returns the request, whose .event is set when done."""
        if self._closed:
            raise RuntimeError("server is closed")
        req.nbytes = request_bytes(req.nx, req.p)
        self._queue.put(req)
        return req

    def _admit(self, req):
        with self._budget_cv:
            while self._in_flight > 0 and self._in_flight + req.nbytes > self.budget:
                self._budget_cv.wait()
            self._in_flight += req.nbytes

    def _release(self, req):
        with self._budget_cv:
            self._in_flight -= req.nbytes
            self._budget_cv.notify_all()

    def _worker(self):
        stream = torch.cuda.Stream(device=self.device)
        with torch.cuda.stream(stream):
            while True:
                req = self._queue.get()
                if req is None:
                    break
                self._admit(req)
                try:
                    self.solve(req)
                except Exception as e:          # reported on the request, the slot keeps serving
                    req.error = e
                    req.done = time.perf_counter()
                finally:
                    stream.synchronize()        # the result is complete before the request is released
                    self._release(req)
                    req.event.set()

    def close(self):
        self._closed = True
        for _ in self._threads:
            self._queue.put(None)
        for t in self._threads:
            t.join()
        self._threads = []


def make_requests(n, *, seed=20261202, nx_range=(16, 512), p_choices=(1, 2), modes=8, rate=None, device="cpu"):
    """Synthetic trace: nx log-uniform in nx_range, p uniform in p_choices,
    k = exp(N(0,1)) per cell, `modes` Fourier modes with amplitude N(0,1) and
    wave numbers uniform in 1..8, Poisson arrivals at `rate` per second
    (all at 0 when rate is None).  Per-request generators derive from `seed`."""
    out = []
    t = 0.0
    arr_rng = np.random.default_rng([seed, 0])
    lo, hi = math.log(nx_range[0]), math.log(nx_range[1])
    for i in range(n):
        g = np.random.default_rng([seed, i + 1])
        nx = int(round(math.exp(g.uniform(lo, hi))))
        p = int(p_choices[g.integers(0, len(p_choices))])
        nc = 2 * nx * nx
        k = torch.as_tensor(np.exp(g.standard_normal(nc)), dtype=torch.float64, device=device)
        m = np.stack([g.standard_normal(modes), g.integers(1, 9, modes), g.integers(1, 9, modes)], axis=1)
        if rate:
            t += arr_rng.exponential(1.0 / rate)
        out.append(SolveRequest(i, nx, p, k, m.astype(np.float64), arrival=t))
    return out
