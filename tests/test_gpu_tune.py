"""The project autotuner (fealpy_amd.tune): it picks one of the candidate
tiles by measurement, stores it in the library's table and its cache file,
and the choice never changes assembled values (pass 2 sums in incidence
order for every lane count); the CSR SpMV agrees across choices to rounding."""
import json

import numpy as np
import pytest
import torch

import fealpy_amd as fa
from fealpy_amd import _native as nat
from fealpy_amd.tune import CANDIDATES, SPMV_LANES, SUM_ROWS_LANES, Autotuner

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def test_assembly_lanes_bitwise_and_cached(tmp_path, monkeypatch):
    from fealpy_amd.fem import integrators
    monkeypatch.setattr(integrators, "BOX_ASSEMBLY", False)     # the two-pass kernels, not the from_box one
    L = nat.lib()
    mesh = fa.TetrahedronMesh.from_box(nx=24, ny=24, nz=20, device=DEV)     # 69,120 cells: two-pass path
    space = fa.LagrangeFESpace(mesh, 1)
    k = torch.as_tensor(np.exp(np.random.default_rng(9).standard_normal(mesh.number_of_cells())), device=DEV)
    from fealpy_amd.fem.pattern import pattern_of
    key = L.fa_tune_key(SUM_ROWS_LANES, 3, pattern_of(space).max_row)
    vals = {}
    try:
        for g in CANDIDATES[SUM_ROWS_LANES]:
            nat.check(L.fa_tune_set(SUM_ROWS_LANES, key, g), "fa_tune_set")
            vals[g] = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(k, q=2)).assembly().values
        ref = vals[8]
        for g, v in vals.items():
            assert torch.equal(v, ref), g
        path = tmp_path / "tune.json"
        t = Autotuner(path=str(path), reps=3)
        best, ms = t.tune_assembly(space, k, q=2)
        assert best in CANDIDATES[SUM_ROWS_LANES] and set(ms) == set(CANDIDATES[SUM_ROWS_LANES])
        assert L.fa_tune_get(SUM_ROWS_LANES, key) == best
        data = json.loads(path.read_text())
        dev = data[torch.cuda.get_device_name(DEV)]
        assert dev["sum_rows_lanes"][str(key)] == best
        nat.check(L.fa_tune_set(SUM_ROWS_LANES, key, 0), "fa_tune_set")
        assert Autotuner(path=str(path)).apply_cached(DEV) >= 1
        assert L.fa_tune_get(SUM_ROWS_LANES, key) == best
    finally:
        nat.check(L.fa_tune_set(SUM_ROWS_LANES, key, 0), "fa_tune_set")


def test_spmv_lanes(tmp_path):
    L = nat.lib()
    mesh = fa.TetrahedronMesh.from_box(nx=12, ny=12, nz=12, device=DEV)
    space = fa.LagrangeFESpace(mesh, 2)
    A = fa.BilinearForm(space).add_integrator(fa.ScalarDiffusionIntegrator(q=3)).assembly()
    key = L.fa_tune_key(SPMV_LANES, A.shape[0], A.nnz)
    x = torch.rand(A.shape[0], dtype=torch.float64, device=DEV)
    try:
        ys = {}
        for g in CANDIDATES[SPMV_LANES]:
            nat.check(L.fa_tune_set(SPMV_LANES, key, g), "fa_tune_set")
            ys[g] = A @ x
        ref = ys[CANDIDATES[SPMV_LANES][0]]
        for g, y in ys.items():
            assert float((y - ref).abs().max()) <= 1e-13 * float(ref.abs().max()), g
        best, _ = Autotuner(path=str(tmp_path / "t.json"), reps=3).tune_spmv(A)
        assert best in CANDIDATES[SPMV_LANES] and L.fa_tune_get(SPMV_LANES, key) == best
    finally:
        nat.check(L.fa_tune_set(SPMV_LANES, key, 0), "fa_tune_set")


def test_tune_set_rejects_bad_family():
    L = nat.lib()
    assert L.fa_tune_set(7, 0, 8) == nat.FA_EINVAL
    assert L.fa_tune_key(7, 1, 1) == -1
