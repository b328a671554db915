"""Project autotuner: tile shapes measured on the caller's own data.

Each tunable kernel family reads its tile (lanes per CSR row, ...) from a
process-wide table in the library (fa_tune_set / fa_tune_get), keyed by a
shape class (fa_tune_key).  ``Autotuner.tune`` times every candidate tile on
the real operator with HIP events on the current stream (median of a few
runs after a warm-up launch), keeps the fastest in that table, and records
it in a JSON cache per device name, so a later process starts with the
measured choices (``Autotuner.apply_cached``).  A choice changes only the
work decomposition.  Assembly pass 2 sums every entry in incidence order
for any lane count, so its values are bitwise the same for every choice;
the CSR SpMV adds its lane partials by a tree whose shape follows the lane
count, so its results move at the rounding level with the choice (a cached
choice keeps them identical from run to run).
"""
import json
import os
import statistics
import threading

import torch

from . import _native as nat

SPMV_LANES = 0           # FA_TUNE_SPMV_LANES
SUM_ROWS_LANES = 1       # FA_TUNE_SUM_ROWS_LANES

CANDIDATES = {SPMV_LANES: (4, 8, 16, 32, 64), SUM_ROWS_LANES: (4, 8, 16, 32)}
FAMILY_NAMES = {SPMV_LANES: "spmv_lanes", SUM_ROWS_LANES: "sum_rows_lanes"}

_DEFAULT_PATH = os.path.join(os.path.expanduser("~"), ".cache", "fealpy_amd", "tune.json")


def _device_name(device=None):
    return torch.cuda.get_device_name(device)


class Autotuner:
    def __init__(self, path=_DEFAULT_PATH, reps=5):
        self.path, self.reps = path, reps
        self._lock = threading.Lock()
        self.results = {}        # {(family, key): {"best": tile, "ms": {tile: median ms}}}

    # -- cache ------------------------------------------------------------------
    def _load(self):
        if not self.path:
            return {}
        try:
            with open(self.path) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}

    def _save(self, device):
        if not self.path:
            return
        data = self._load()
        dev = data.setdefault(_device_name(device), {})
        for (fam, key), r in self.results.items():
            dev.setdefault(FAMILY_NAMES[fam], {})[str(key)] = r["best"]
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            tmp = self.path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass                 # a read-only home keeps the in-process choices only

    def apply_cached(self, device=None):
        """Load this device's cached choices into the library's table; returns
        how many were applied."""
        L = nat.lib()
        dev = self._load().get(_device_name(device), {})
        n = 0
        for fam, name in FAMILY_NAMES.items():
            for key, val in dev.get(name, {}).items():
                if int(val) in CANDIDATES[fam]:
                    nat.check(L.fa_tune_set(fam, int(key), int(val)), "fa_tune_set")
                    n += 1
        return n

    # -- measurement ----------------------------------------------------------------
    def tune(self, family, key, run, candidates=None, device=None):
        """Time `run()` (one launch of the family's kernel on the caller's
        data) under every candidate tile; keep and cache the fastest."""
        L = nat.lib()
        cands = tuple(candidates or CANDIDATES[family])
        ms = {}
        with self._lock:
            old = L.fa_tune_get(family, key)
            try:
                for c in cands:
                    nat.check(L.fa_tune_set(family, key, int(c)), "fa_tune_set")
                    run()                                   # warm-up
                    times = []
                    for _ in range(self.reps):
                        e0 = torch.cuda.Event(enable_timing=True)
                        e1 = torch.cuda.Event(enable_timing=True)
                        e0.record()
                        run()
                        e1.record()
                        e1.synchronize()
                        times.append(e0.elapsed_time(e1))
                    ms[c] = statistics.median(times)
            except BaseException:
                L.fa_tune_set(family, key, old)
                raise
            best = min(ms, key=ms.get)
            nat.check(L.fa_tune_set(family, key, int(best)), "fa_tune_set")
            self.results[(family, key)] = {"best": int(best), "ms": {int(c): round(t, 5) for c, t in ms.items()}}
        self._save(device)
        return int(best), ms

    # -- the tunable families on real data ---------------------------------------
    def tune_spmv(self, A):
        """Lanes per row of fa_spmv for the row-length class of the CSR matrix A."""
        L = nat.lib()
        key = L.fa_tune_key(SPMV_LANES, A.shape[0], A.nnz)
        x = torch.rand(A.shape[1], dtype=A.dtype, device=A.device)
        y = torch.empty(A.shape[0], dtype=A.dtype, device=A.device)

        def run():
            nat.check(L.fa_spmv(nat.dtype_code(A.dtype), A.shape[0], A.nnz, nat.ptr(A.crow), nat.ptr(A.col),
                                nat.ptr(A.values), nat.ptr(x), nat.ptr(y), nat.stream()), "fa_spmv")
        return self.tune(SPMV_LANES, key, run, device=A.device)

    def tune_assembly(self, space, coef=None, q=None, dtype=torch.float64):
        """Lanes per row of P1 assembly pass 2 (fa_sum_incidence_rows) on the
        space's pattern, timing a whole two-pass assembly of a diffusion form
        (also on a from_box mesh, whose own matrices take the structured
        kernel: the choice serves the meshes with this row length that do
        not).  Not thread-safe: the two-pass route is forced for the duration."""
        from .fem import BilinearForm, ScalarDiffusionIntegrator, integrators
        from .fem.pattern import pattern_of
        L = nat.lib()
        pat = pattern_of(space)
        key = L.fa_tune_key(SUM_ROWS_LANES, space.TD, pat.max_row)
        form = BilinearForm(space).add_integrator(ScalarDiffusionIntegrator(coef, q=q))

        def run():
            form.assembly(dtype=dtype)
        box, two = integrators.BOX_ASSEMBLY, integrators.TWO_PASS_MIN_CELLS
        integrators.BOX_ASSEMBLY, integrators.TWO_PASS_MIN_CELLS = False, 0
        try:
            return self.tune(SUM_ROWS_LANES, key, run, device=space.device)
        finally:
            integrators.BOX_ASSEMBLY, integrators.TWO_PASS_MIN_CELLS = box, two


_default = None


def default_tuner():
    global _default
    if _default is None:
        _default = Autotuner()
    return _default
